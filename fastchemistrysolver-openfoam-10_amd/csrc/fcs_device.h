// fcs_device.h -- gfx950 device code of the per-cell stiff chemistry
// integrator: one cell per 64-lane wavefront for systems up to 64 rows
// (matrix in LDS), one cell per workgroup of G = 256 threads (four
// wavefronts) for up to 256 rows (matrix in HBM, LU in fcs_lu_large.h).
//
// Layout of one cell (NP = padded system size, a multiple of 4; a multiple
// of 16 for the workgroup form; thread = lane of the wavefront or thread of
// the workgroup):
//   lane s < ns        species s: state y[s] = Y_s, matrix row s in a[0..NP-1]
//   ns <= lane < NP-1  padding: zero state, identity-like rows
//   lane NP-1          temperature: y = T, matrix row of dT/dt
// Matrix columns follow the same order, so every register index is a
// compile-time constant; the temperature column is a[NP-1].  Placing T at
// NP-1 instead of ns leaves the pivot decisions of the reference's block-4
// LU unchanged: padding rows hold zeros in every real column, so they never
// win a pivot search (src/LUsolver/LUsolver.C:106-246,503-650).
//
// Reaction-level work (rate constants, equilibrium constants) is spread over
// lanes (lane l handles reactions l, l+64, ...); species-level gathers use a
// per-species reaction list, so no floating-point atomics are used and every
// result is bit-reproducible for a fixed launch shape and for any other.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fcs {

constexpr double RU = 8314.470066505449722171761095523834228515625;  // OptReaction.H:404
constexpr double PSTD = 1.0e5;                                         // OptReaction.H:409
constexpr double KC_FLOOR = 1.4901171103413047e-8;                     // ReactionFunctionI.H:17
constexpr double TC_KC_FLOOR = 1.49011611938476e-08;                   // Tc.C:599
constexpr double TROE_SMALL = 2.2e-16;                                 // ReactionRate.C:475
constexpr double PIVOT_EPS = 2.2204460492503131e-16;                   // LUsolver.C:155
constexpr double OF_SMALL = 1.0e-15;    // OpenFOAM-10 doubleScalarSmall
constexpr double OF_GREAT = 1.0e15;     // OpenFOAM-10 doubleScalarGreat
constexpr double OF_VGREAT = 1.0e300;   // OpenFOAM-10 doubleScalarVGreat
constexpr int WAVE = 64;
// Column stride (doubles) of the per-cell matrix in LDS.  65 rather than 64:
// a row-wise access (lane = row, one column) is conflict-free either way,
// and a column-wise one (lane = column, one row) spreads over the banks
// instead of hitting one bank with every lane.
constexpr int JS = 65;

// Workgroup-form sizes up to this many rows keep the iteration matrix in LDS
// (NP*NP + NP doubles: 98.9 KiB at NP = 112) instead of an HBM slice.
constexpr int kJLdsMaxNP = 112;
__host__ __device__ constexpr bool jlds_of(int NP) { return NP > 64 && NP <= kJLdsMaxNP; }

// Geometry of one cell for a padded system size NP.
template <int NP>
struct Cell {
    static constexpr bool kLarge = NP > 64;                       // workgroup per cell
    static constexpr bool kJLds = jlds_of(NP);                    // workgroup form with the matrix in LDS
    // threads per cell: one wavefront, or a workgroup of one thread per row
    // (two wavefronts up to 128 rows, so two cells share a CU; four beyond)
    static constexpr int G = !kLarge ? 64 : (NP <= 128 ? 128 : 256);
    static constexpr int NW = G / 64;                             // wavefronts per cell
    static constexpr int SW = G;                                  // species stride of per-species tables
    static constexpr int ONE = kLarge ? 255 : 64;                 // concentration slot holding 1.0
    static constexpr int CS = kLarge ? 256 : 65;                  // concentration slots
    static constexpr int JSt = kLarge ? NP : JS;                  // matrix column stride
    // wavefronts per SIMD the step kernels (Rosenbrock and Seulex) are
    // compiled for: two for the HBM-matrix workgroup form (two cells share a
    // CU, hiding the matrix round trips of one behind the other; registers
    // capped at 256)
    static constexpr int kMinWaves = (kLarge && !kJLds) ? 2 : 1;
};

enum { RXN_ARR = 0, RXN_TB = 1, RXN_LIND = 2, RXN_TROE = 3, RXN_PLOG = 4, RXN_SRI = 5 };
enum { SOLVER_RODAS34 = 0, SOLVER_ROS34 = 1, SOLVER_SEULEX = 2, SOLVER_RODAS23 = 3, SOLVER_ROS23 = 4 };
enum { CELL_SKIPPED = 1, CELL_MAXSTEPS = 2, CELL_NONFINITE = 4 };

// Pointer to a read-only table in global memory.  Kernels read the tables
// through a copy of DevMech in LDS, and a plain pointer loaded from there is
// generic: its reads compile to flat loads, which count against the LDS
// counter too, so every wait for an LDS read would also wait for them.
// Reads through address space 1 are global loads, tracked by the vector
// memory counter alone.
template <class T>
struct gptr {
    T* p;
    __host__ __device__ gptr& operator=(T* q)
    {
        p = q;
        return *this;
    }
    __device__ __forceinline__ T operator[](long long i) const
    {
        return ((const __attribute__((address_space(1))) T*)p)[i];
    }
    __host__ __device__ gptr operator+(long long i) const { return gptr{p + i}; }
};

// Device-resident mechanism tables (built on the host by fcs_mech_create).
// Reactions are sorted by rate form so that lanes of one reaction round
// mostly take the same branch.
struct DevMech {
    int ns, nr, NPm;          // species, reactions, padded size used for tables
    int nTB;                  // third-body + falloff rows
    int pmax;                 // PLOG points per row
    int maxDeg;               // per-species list length, padded to a multiple of 16
    int nTBpad;               // row stride of effT: nTB padded to a multiple of 128
    double TlowMin, ThighMax, TcMin, TcMax;
    // per-species tables are padded to SW = Cell<NP>::SW species (64, or the
    // workgroup size): "[SW]" below
    gptr<const double> W;          // [SW]  (0 beyond ns)
    gptr<const double> invW;       // [SW]  (0 beyond ns)
    gptr<const double> Tc;         // [SW]
    gptr<const double> clo;        // [7*SW] coefficient-major: c[k*SW + s]
    gptr<const double> chi;        // [7*SW]
    // [nr] one word per reaction, one load in a rate round: third-body row
    // + 1 (0: none) in bits 0-23, rate form (RXN_*) in bits 24-27, reverse
    // kind in bits 28-31: 0 irreversible, 1 kf/Kc, 2 own reverse Arrhenius
    // in rA0/rbeta0/rTa0, 3 kf/Kc over a product side of more than 3 slots
    // (gen row rgen[r])
    gptr<const int> rmeta;
    gptr<const int> rnl;           // [nr]
    gptr<const int> rnr;           // [nr]
    gptr<const int> rsp;           // [6*nr] slot-major: rsp[q*nr + r], q<3 lhs, q>=3 rhs
    gptr<const int> raux;          // [nr]  troe / plog row
    gptr<const double> rA;         // [nr]  forward (k_inf)
    gptr<const double> rbeta;
    gptr<const double> rTa;
    gptr<const double> rA0;        // [nr]  k_0 (falloff) or the reverse Arrhenius (nonEquilibrium)
    gptr<const double> rbeta0;
    gptr<const double> rTa0;
    gptr<const double> troe;       // [4*nTroe] (alpha, Tsss, Ts, Tss) per row
    gptr<const double> rtroe;      // [5*nr] coefficient-major rtroe[q*nr + r]: Troe (alpha, Tsss, Ts, Tss, 0)
                              // or SRI (a, b, c, d, e); zero for other rows
    gptr<const double> eff;        // [nTB*SW] row-major, zero beyond ns
    gptr<const double> effT;       // [SW*nTBpad] species-major for M sums (zero padding)
    gptr<const int> plog_n;        // [nPlog]
    gptr<const double> plog_lnp;   // [nPlog*pmax]
    gptr<const double> plog_lnA;
    gptr<const double> plog_beta;
    gptr<const double> plog_Ta;
    gptr<const int> sp_rn;         // [maxDeg*SW] thread-interleaved: reaction | (nu + 8) << 16, nu = net
                              // stoichiometric coefficient (padding: reaction 0, nu 0)
    int nTBz;                 // third-body/falloff (species, row) entries, padded to a multiple of 16
    gptr<const int> tbz_row;       // [nTBz] species, ascending (padding repeats the last)
    gptr<const int> tbz_tb;        // [nTBz] third-body row (padding: 0)
    gptr<const double> tbz_nu;     // [nTBz] net stoichiometric coefficient (padding: 0)
    gptr<const int> tbz_blk;       // [NPm/16 + 1] first entry of each 16-row block (HBM-matrix form)
    gptr<const unsigned long long> rpack;    // [nr] species of slots 0-5 in bytes 0-5, nl, nr in bytes 6, 7
    gptr<const unsigned long long> sp_pack;  // [maxDeg*SW] rpack of each per-species list entry
    gptr<const int> sp_deg;   // [SW] length of each species' list (0 beyond ns): the Jacobian's
                              // production walk skips the padding blocks of shorter lists
    int maxTerms;             // Jacobian terms per thread (jacobian's sparse-partial walk), multiple of 16
    gptr<const unsigned long long> sp_term;  // [maxTerms*SW] thread-interleaved terms, see jterm_*: every
                                        // element's terms in one thread, the threads' loads balanced
    gptr<const double> Hf;         // [SW] mass-specific enthalpy of formation at 298.15 K (0 beyond ns)
    // Kc-reversible reactions with more than 3 product slots (rrev 3): their
    // Jacobian takes one reverse partial kr prod_{b != q} C_b per product slot
    // q, kept in the reaction arrays of LDS past the nr reactions, so nrq =
    // nr + those partials sizes the reaction arrays
    int nrq;
    // PLOG rows are the reactions plogBase .. plogBase + nPlog - 1 (reactions
    // are sorted by rate form); plog row = r - plogBase
    int plogBase, nPlog;
    // first reaction that reads the k_0 / reverse Arrhenius triple (falloff
    // forms, nonEquilibrium reverse rates), and the first that reads the Troe
    // / SRI coefficients (nr when none does)
    int extLo, troeLo;
    gptr<const int> rgen;          // [nr] row of gen (rrev 3 only; -1 elsewhere)
    gptr<const int> gen;           // [rows * GEN_STRIDE]: product slots n, LDS index of the first partial
                              // (>= nr), the n product species
#ifdef FCS_PHASE_TIMING
    unsigned long long* phase;  // [PH_SLOTS * PH_SPREAD] shader-clock cycles per phase
#endif
};

// Jacobian term: nu_ir * d q_r / d C_j for one reactant (forward) or product
// (reverse) slot of reaction r, added to matrix element (row i, column j).
// Bits 0-15 r (index into the reaction arrays), 16-23 column j, 24-31 row i
// (16-31 together: the element key), 32-39 and 40-47 the other species of the
// slot's side (slot order; ONE when absent), 48-55 the signed coefficient + 8
// (negated for reverse terms; 0 = padding), 56 reverse (kr instead of kf),
// 57 dump (a padding term of a thread with no element of its own: its store
// goes to a slot no phase reads).
__device__ __forceinline__ int jterm_r(unsigned long long w) { return (int)(w & 0xffffu); }
__device__ __forceinline__ int jterm_key(unsigned long long w) { return (int)((w >> 16) & 0xffffu); }
__device__ __forceinline__ int jterm_o1(unsigned long long w) { return (int)((w >> 32) & 255u); }
__device__ __forceinline__ int jterm_o2(unsigned long long w) { return (int)((w >> 40) & 255u); }
__device__ __forceinline__ int jterm_nu(unsigned long long w) { return (int)((w >> 48) & 255u) - 8; }
__device__ __forceinline__ bool jterm_rev(unsigned long long w) { return (w >> 56) & 1u; }
__device__ __forceinline__ bool jterm_dump(unsigned long long w) { return (w >> 57) & 1u; }
__device__ __forceinline__ int rn_r(int w) { return w & 0xffff; }
__device__ __forceinline__ int rn_nu(int w) { return (w >> 16) - 8; }
__device__ __forceinline__ int pk_species(unsigned long long w, int q) { return (int)((w >> (8 * q)) & 255u); }
__device__ __forceinline__ int pk_nl(unsigned long long w) { return (int)((w >> 48) & 255u); }
__device__ __forceinline__ int pk_nr(unsigned long long w) { return (int)(w >> 56); }

// Ints per row of DevMech::gen: slot count, partial base, up to 16 species.
constexpr int GEN_STRIDE = 18;

// LDS doubles at the end of a cell's LDS that hold a copy of its DevMech
// (see stage_mech in fcs_kernels.h).
constexpr int kMechLdsDoubles = (int)((sizeof(DevMech) + 7) / 8);

struct DevOpts {
    int solver, jac, maxSteps;
    double absTol, relTol, Treact, dtChemMax;
};

// ---------------------------------------------------------------- lane utils
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Wave reductions that leave the same bits in every lane: within each
// 16-lane row, DPP exchanges (quad xor 1, quad xor 2, half-row mirror, row
// mirror) add commutative pairs, so all lanes of a row hold the row total;
// the four row totals are then combined as (r0 + r1) + (r2 + r3) from
// uniform (scalar) copies.  No LDS traffic.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<DPP_XOR1>(v);
    v += dpp_d<DPP_XOR2>(v);
    v += dpp_d<DPP_HALF_MIRROR>(v);
    v += dpp_d<DPP_MIRROR>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_d<DPP_XOR1>(v));
    v = fmax(v, dpp_d<DPP_XOR2>(v));
    v = fmax(v, dpp_d<DPP_HALF_MIRROR>(v));
    v = fmax(v, dpp_d<DPP_MIRROR>(v));
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}
__device__ __forceinline__ int first_lane(bool pred)
{
    const unsigned long long b = __ballot(pred);
    return b ? (__ffsll((long long)b) - 1) : -1;
}
// Orders LDS traffic between the lanes of the (single-wave) workgroup.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// ---------------------------------------------------------------- phase clock
// Profiling builds (-DFCS_PHASE_TIMING) add the shader-clock cycles of each
// phase of an attempt into M.phase[slot * PH_SPREAD + block % PH_SPREAD]
// (once per cell, by its first thread; spread over PH_SPREAD counters per
// slot so the atomics do not queue on one address); in other builds the
// clock compiles to nothing.
enum {
    PH_JAC_RATES, PH_JAC_TERMS, PH_JAC_PROD, PH_JAC_TB, PH_JAC_CONV,   // Jacobian evaluation
    PH_FORM, PH_LU, PH_SOLVE,                                            // iteration matrix, LU, stage solves
    PH_RHS_SETUP, PH_RHS_TB, PH_RHS_RATES, PH_RHS_PROD, PH_RHS_REST,    // derivatives()
    PH_STAGE, PH_ERR,                                                    // stage sums, error norm
    PH_LU_PANEL, PH_LU_UROWS, PH_LU_TRAIL,                               // parts of PH_LU
    PH_SLOTS = 24, PH_SPREAD = 256
};
#ifdef FCS_PHASE_TIMING
struct PhaseClock {
    unsigned long long t;
    __device__ PhaseClock() : t(__builtin_amdgcn_s_memtime()) {}
    __device__ void mark(const DevMech& M, bool first, int slot)
    {
        const unsigned long long n = __builtin_amdgcn_s_memtime();
        if (first) atomicAdd(M.phase + slot * PH_SPREAD + (blockIdx.x & (PH_SPREAD - 1)), n - t);
        t = n;
    }
};
#else
struct PhaseClock {
    __device__ void mark(const DevMech&, bool, int) {}
};
#endif

// ---------------------------------------------------------------- LDS layout
// The per-cell matrix lives in LDS, lane-interleaved column-major: element
// (row i, column j) at J[j*JS + i], row i owned by lane i.  Column accesses
// of one row index across the lanes are conflict-free; the pivot rows of the
// LU are read by all lanes at one address (broadcast).
struct Lds {
    double* C;     // [CS] concentrations, C[ONE] = 1
    double* expG;  // [G] exp(-G/RT)
    double* dBdT;  // [G] d(-G/RT)/dT
    double* rq;    // [nrP] q (rates) or kf (Jacobian)
    double* rk;    // [nrP] kr (Jacobian)
    double* rdT;   // [nrP] dq/dT (Jacobian)
    double* H;     // [G] species enthalpy (T-row column sums), over rk once the
    double* Q;     // [G] dY/dt * cp (T-row column sums)    sparse partials are in
    double* rw;    // [nTBP] third-body partial (Jacobian)
    double* Mtb;   // [nTBP] third-body concentration M of each third-body/falloff row
    double* J;     // [NP*JSt] Jacobian, then iteration matrix and its LU factors
                   // (LDS; for the workgroup form a per-workgroup slice in HBM
                   // of NP*NP + NP doubles, the last NP holding 1/pivot)
    double* tbt;   // workgroup form, matrix in HBM: [16][NP+1] tile of third-body partials
    double* plog;  // [3*nPlog] per-cell PLOG rows as Arrhenius (ln A, beta, Ta) at the
                   // cell's pressure (stage_plog), kept for the whole kernel
    int* ip;       // [G] thread holding each logical row after the LU
    double* lu;    // workgroup form: LU / solve scratch over rq..Mtb (dead by then)
    double* red;   // workgroup form: [16] cross-wavefront reduction / broadcast slots
};

__host__ __device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Panel width of the workgroup-form LU (columns per MFMA trailing update).
constexpr int LU_PW = 16;
// LU / solve scratch of the workgroup form (doubles): the 4 pivot rows of a
// 4-block (double-buffered, 2 x 4 x LU_PW), the panel's unit-lower block
// (LU_PW^2), -L21 and U12 tiles (2 x LU_PW x NP), 3 solve vectors of
// LU_PW, the panel's logical-to-physical map (LU_PW ints).
__host__ __device__ inline size_t lu_scratch_doubles(int NP)
{
    return 8 * LU_PW + LU_PW * LU_PW + 2 * (size_t)LU_PW * NP + 3 * LU_PW + LU_PW / 2 + 8;
}

__host__ __device__ inline int group_of(int NP) { return NP > 64 ? (NP <= 128 ? 128 : 256) : 64; }

// LDS doubles per cell, the staged DevMech last.  Wavefront form: four
// wavefronts of the 53-species mechanism fit the 160 KiB of a CU (40,504 B
// each).  Workgroup form: the
// LU scratch reuses the reaction arrays; the matrix (and its reciprocal
// pivots) follows them in LDS up to kJLdsMaxNP rows, else it is in HBM.
// Tile of the third-body partials in the HBM-matrix workgroup form (16 rows
// by NP columns, row stride NP + 1 so that a column's 16 entries fall in
// different banks).
__host__ __device__ inline size_t tbt_doubles(int NP) { return NP > 64 && !jlds_of(NP) ? 16 * ((size_t)NP + 1) : 0; }

// nrq: slots of the per-reaction arrays (DevMech::nrq); nPlog: PLOG rows.
__host__ __device__ inline size_t lds_doubles(int NP, int nrq, int nTB, int nPlog)
{
    const size_t nrP = round_up(nrq, 2), plP = round_up(3 * nPlog, 2);
    if (NP <= 64) {
        const size_t kq = 2 * nrP > 128 ? 2 * nrP : 128;   // rk + rdT, later H + Q
        return 3 * 66 + nrP + kq + 2 * (size_t)round_up(nTB > 0 ? nTB : 1, 2) + plP + (size_t)NP * JS + 32 +
               kMechLdsDoubles;
    }
    const size_t G = group_of(NP);
    const size_t kq = 2 * nrP > 2 * G ? 2 * nrP : 2 * G;
    const size_t react = nrP + kq + 2 * (size_t)round_up(nTB > 0 ? nTB : 1, 2);
    const size_t lu = lu_scratch_doubles(NP);
    const size_t rest = 256 + 2 * G + (react > lu ? react : lu) + G / 2 + 16 + plP;
    return (jlds_of(NP) ? ((rest + 1) & ~(size_t)1) + (size_t)NP * NP + NP : rest + tbt_doubles(NP)) +
           kMechLdsDoubles;
}

__device__ inline Lds make_lds(double* base, int NP, int nrq, int nTB, int nPlog, double* Jglobal = nullptr)
{
    Lds L;
    const int plP = round_up(3 * nPlog, 2);
    if (NP > 64) {
        const int G = group_of(NP);
        const int nrP = round_up(nrq, 2), nTBP = round_up(nTB > 0 ? nTB : 1, 2);
        const int kq = 2 * nrP > 2 * G ? 2 * nrP : 2 * G;
        const size_t react = nrP + kq + 2 * (size_t)nTBP;
        const size_t lu = lu_scratch_doubles(NP);
        L.C = base;
        L.expG = base + 256;
        L.dBdT = L.expG + G;
        L.rq = L.dBdT + G;
        L.rk = L.rq + nrP;
        L.rdT = L.rk + nrP;
        L.H = L.rk;
        L.Q = L.rk + G;
        L.rw = L.rk + kq;
        L.Mtb = L.rw + nTBP;
        L.lu = L.rq;
        L.ip = (int*)(L.rq + (react > lu ? react : lu));
        L.red = L.rq + (react > lu ? react : lu) + G / 2;
        L.plog = L.red + 16;   // past the LU scratch, which reuses rq .. Mtb
        const size_t rest = 256 + 2 * (size_t)G + (react > lu ? react : lu) + G / 2 + 16 + plP;
        L.J = jlds_of(NP) ? base + ((rest + 1) & ~(size_t)1) : Jglobal;
        L.tbt = jlds_of(NP) ? nullptr : base + rest;
        return L;
    }
    const int nrP = round_up(nrq, 2), nTBP = round_up(nTB > 0 ? nTB : 1, 2);
    const int kq = 2 * nrP > 128 ? 2 * nrP : 128;
    L.C = base;              // C[64] = 1.0: the "no other species" slot of the Jacobian terms
    L.expG = base + 66;      // expG[64] = 1.0 and dBdT[64] = 0: the identity slot unused
    L.dBdT = base + 132;     // reaction slots name (wavefront form)
    L.rq = base + 198;
    L.rk = L.rq + nrP;
    L.rdT = L.rk + nrP;
    L.H = L.rk;
    L.Q = L.rk + 64;
    L.rw = L.rk + kq;
    L.Mtb = L.rw + nTBP;
    L.plog = L.Mtb + nTBP;
    L.J = L.plog + plP;
    L.ip = (int*)(L.J + (size_t)NP * JS);
    L.lu = nullptr;
    L.red = nullptr;
    L.tbt = nullptr;
    return L;
}

// ---------------------------------------------------------------- group ops
// Thread index within the cell's group, and reductions / broadcasts over
// the group: the wavefront primitives for the wavefront form, LDS slots and
// barriers across the wavefronts of the workgroup form.  Every thread of the
// group must call them (uniform control flow).
template <int NP>
__device__ __forceinline__ int tid()
{
    if constexpr (Cell<NP>::kLarge) return (int)threadIdx.x;
    else return lane_id();
}

template <int NP>
__device__ __forceinline__ double gsum(double v, const Lds& L)
{
    if constexpr (!Cell<NP>::kLarge) {
        return wave_sum(v);
    } else {
        const double w = wave_sum(v);
        __syncthreads();
        if (lane_id() == 0) L.red[threadIdx.x >> 6] = w;
        __syncthreads();
        double r = L.red[0];
#pragma unroll
        for (int k = 1; k < Cell<NP>::NW; ++k) r += L.red[k];
        return r;
    }
}

template <int NP>
__device__ __forceinline__ double gmax(double v, const Lds& L)
{
    if constexpr (!Cell<NP>::kLarge) {
        return wave_max(v);
    } else {
        const double w = wave_max(v);
        __syncthreads();
        if (lane_id() == 0) L.red[threadIdx.x >> 6] = w;
        __syncthreads();
        double r = L.red[0];
#pragma unroll
        for (int k = 1; k < Cell<NP>::NW; ++k) r = fmax(r, L.red[k]);
        return r;
    }
}

template <int NP>
__device__ __forceinline__ bool gany(bool p, const Lds& L)
{
    if constexpr (!Cell<NP>::kLarge) return __ballot(p) != 0ull;
    else return gmax<NP>(p ? 1.0 : 0.0, L) > 0.0;
}

// Value of v held by thread src (uniform).
template <int NP>
__device__ __forceinline__ double gbcast(double v, int src, const Lds& L)
{
    if constexpr (!Cell<NP>::kLarge) {
        return readlane_d(v, src);
    } else {
        __syncthreads();
        if ((int)threadIdx.x == src) L.red[8] = v;
        __syncthreads();
        return L.red[8];
    }
}

// The wavefront form's identity slot 64 (C 1, exp(-G/RT) 1, d/dT 0), which
// the unused reaction slots name.
__device__ __forceinline__ void set_identity_slot(const Lds& L, double dslot = 0.0)
{
    L.C[64] = 1.0;
    L.expG[64] = 1.0;
    L.dBdT[64] = dslot;
}

// ---------------------------------------------------------------- thermo
// Mass-based NASA-7 thermo (src/Reaction/ThermoFunctionI.H:33-123); the
// coefficient choice follows setPtrCoeffs (ThermoFunctionI.H:1-31).
struct Thermo { double cp, ha, negG, dBdT, dcpdT; };

// Per-species table entries of lane s (both NASA-7 ranges), loaded before
// the temperature is known so that the loads overlap the density sum.
struct SpeciesData { double lo[7], hi[7], Tc, invW, W; };

template <int NP>
__device__ __forceinline__ SpeciesData load_species(const DevMech& M, int s)
{
    constexpr int SW = Cell<NP>::SW;
    SpeciesData d;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        d.lo[k] = M.clo[k * SW + s];
        d.hi[k] = M.chi[k * SW + s];
    }
    d.Tc = M.Tc[s];
    d.invW = M.invW[s];
    d.W = M.W[s];
    return d;
}

__device__ inline Thermo species_thermo(const DevMech& M, const SpeciesData& S, double T, double lnT, double iT)
{
    const bool low = (T < M.TcMin) ? true : (T > M.TcMax ? false : (T < S.Tc));
    const double a0 = low ? S.lo[0] : S.hi[0], a1 = low ? S.lo[1] : S.hi[1], a2 = low ? S.lo[2] : S.hi[2],
                 a3 = low ? S.lo[3] : S.hi[3], a4 = low ? S.lo[4] : S.hi[4], a5 = low ? S.lo[5] : S.hi[5],
                 a6 = low ? S.lo[6] : S.hi[6];
    const double r = RU * S.invW;
    Thermo t;
    t.cp = ((((a4 * T + a3) * T + a2) * T + a1) * T + a0) * r;
    t.ha = (((((a4 * 0.2 * T + a3 * 0.25) * T + a2 * (1.0 / 3.0)) * T + a1 * 0.5) * T + a0) * T + a5) * r;
    t.negG = (((a4 * 0.05 * T + a3 * (1.0 / 12.0)) * T + a2 * (1.0 / 6.0)) * T + a1 * 0.5) * T
             - a5 * iT + a6 + (lnT - 1.0) * a0;
    t.dBdT = ((a4 * 0.2 * T + a3 * 0.25) * T + a2 * (1.0 / 3.0)) * T + a1 * 0.5 + (a5 * iT + a0) * iT;
    t.dcpdT = (((4.0 * a4 * T + 3.0 * a3) * T + 2.0 * a2) * T + a1) * r;
    return t;
}

__device__ __forceinline__ double clampT(const DevMech& M, double T)
{
    T = T < M.TlowMin ? M.TlowMin : T;   // src/Reaction/ReactionRate.C:14-15
    return T > M.ThighMax ? M.ThighMax : T;
}

// ---------------------------------------------------------------- rates
struct RxnK { double kf, dkfdT, dkfdM; };

// Table entries of one reaction, loaded together and unconditionally (the
// tables cover every reaction), so the loads of a round are all in flight
// before the first use.
struct RxnData {
    int t, rev, tb;
    unsigned long long pk;
    double A, b, Ta, A0, b0, Ta0, alpha, Tsss, Ts, Tss, Mc;
};

// Table entries of reaction r (on: the lane has a reaction; an idle lane
// fetches nothing and gets a reaction that is never evaluated).  The k_0 /
// reverse-Arrhenius triple is fetched only from M.extLo on and the Troe /
// SRI coefficients only from M.troeLo on (every reaction that reads them
// lies there): a round whose lanes all lie below skips those loads, which
// otherwise occupy the CU's vector-memory pipeline that the four cells of
// a CU share.
__device__ __forceinline__ RxnData load_rxn(const DevMech& M, const Lds& L, int r, bool on = true)
{
    RxnData d{RXN_ARR, 0, -1, 0ull, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (on) {
        const unsigned m = (unsigned)M.rmeta[r];
        d.t = (int)((m >> 24) & 15u);
        d.rev = (int)(m >> 28);
        d.tb = (int)(m & 0xffffffu) - 1;
        d.pk = M.rpack[r];
        d.A = M.rA[r];
        d.b = M.rbeta[r];
        d.Ta = M.rTa[r];
        if (r >= M.extLo) {
            d.A0 = M.rA0[r];
            d.b0 = M.rbeta0[r];
            d.Ta0 = M.rTa0[r];
        }
        if (r >= M.troeLo) {
            d.alpha = M.rtroe[r];
            d.Tsss = M.rtroe[M.nr + r];
            d.Ts = M.rtroe[2 * M.nr + r];
            d.Tss = M.rtroe[3 * M.nr + r];
        }
    }
    d.Mc = L.Mtb[d.tb >= 0 ? d.tb : 0];
    return d;
}

// PLOG rows at the cell's pressure as Arrhenius triples (ln A, beta, Ta) in
// L.plog, one per row, for rate_constant.  The reference interpolates ln k
// linearly in ln p between the tabulated pressures around p, clamped to the
// end rows (ReactionRate.C:577-650, Jacobian.C:996 for dk/dT):
//   ln k = lnA_j + beta_j lnT - Ta_j/T + f (dlnA + dbeta lnT - dTa/T)
//        = (lnA_j + f dlnA) + (beta_j + f dbeta) lnT - (Ta_j + f dTa)/T,
// with j and f fixed by p, so one triple per row serves every temperature
// of the cell and the rate rounds read it from LDS instead of walking the
// tables.  Every thread of the cell's group calls this (it synchronises).
template <int NP>
__device__ inline void stage_plog(const DevMech& M, const Lds& L, double lnp)
{
    __syncthreads();   // earlier readers of L.plog (the previous cell) are done
    for (int row = tid<NP>(); row < M.nPlog; row += Cell<NP>::G) {
        const int np = M.plog_n[row];
        const auto lp = M.plog_lnp + (size_t)row * M.pmax;
        const auto la = M.plog_lnA + (size_t)row * M.pmax;
        const auto b = M.plog_beta + (size_t)row * M.pmax;
        const auto ta = M.plog_Ta + (size_t)row * M.pmax;
        int j = 0;
        double f = 0.0;
        if (lnp <= lp[0]) {
            j = 0;
        } else if (lnp >= lp[np - 1]) {
            j = np - 1;
        } else {
            for (int q = 0; q < np - 1; ++q)
                if (lp[q] <= lnp && lnp <= lp[q + 1]) { j = q; break; }
            f = (lnp - lp[j]) / (lp[j + 1] - lp[j]);
        }
        const int j1 = f != 0.0 ? j + 1 : j;
        L.plog[3 * row + 0] = la[j] + (la[j1] - la[j]) * f;
        L.plog[3 * row + 1] = b[j] + (b[j1] - b[j]) * f;
        L.plog[3 * row + 2] = ta[j] + (ta[j1] - ta[j]) * f;
    }
    __syncthreads();
}

// Rate constant of reaction r (kf includes M for third-body and falloff
// forms).  JAC selects the Jacobian-path variants, which also return
// dkf/dT and dkf/dM (src/Reaction/ReactionRate.C:84-672,
// src/Reaction/Jacobian.C:108-910).
template <bool JAC>
__device__ inline RxnK rate_constant(const DevMech& M, const Lds& L, const RxnData& D, int r, double T, double lnT,
                                     double iT)
{
    const int t = D.t;
    RxnK k{0.0, 0.0, 0.0};
    if (t == RXN_PLOG) {
        const double* e = L.plog + 3 * (r - M.plogBase);   // stage_plog
        const double lk = e[0] + e[1] * lnT - e[2] * iT;
        k.kf = exp(lk);
        if (JAC) k.dkfdT = k.kf * ((e[1] + e[2] * iT) * iT);
        return k;
    }
    const double kk = D.A * exp(D.b * lnT - D.Ta * iT);
    const double dk = JAC ? kk * (D.b + D.Ta * iT) * iT : 0.0;
    if (t == RXN_ARR) {
        k.kf = kk;
        k.dkfdT = dk;
        return k;
    }
    // third-body concentration M = sum_j eff_rj C_j (third_body_sums)
    const double Mc = D.Mc;
    if (t == RXN_TB) {
        k.kf = kk * Mc;
        k.dkfdT = dk * Mc;
        k.dkfdM = kk;
        return k;
    }
    const double kinf = kk, dkinf = dk;
    const double k0 = D.A0 * exp(D.b0 * lnT - D.Ta0 * iT);
    const double dk0 = JAC ? k0 * (D.b0 + D.Ta0 * iT) * iT : 0.0;
    if (t == RXN_LIND) {
        const double Pr = k0 * Mc / kinf;
        const double tmp = 1.0 / (1.0 + Pr);
        if (JAC) {
            const double dPrdT = (Mc * dk0 - Pr * dkinf) / kinf;
            k.dkfdT = tmp * Pr * dkinf + tmp * tmp * dPrdT * kinf;
            k.dkfdM = k0 * tmp * tmp;
        }
        k.kf = Mc * tmp * k0;
        return k;
    }
    if (t == RXN_SRI) {
        // SRI_F / SRI_F_dFdT_dFdPr (src/Reaction/pressureReactionFunctionI.H:
        // 2-70), blended as ReactionRate.C:506-518 / Jacobian.C:681-711: log10
        // of max(Pr, small) (OpenFOAM small), and a dF/dT without the
        // dF/dPr dPr/dT chain term, as the reference has it.
        const double a = D.alpha, b = D.Tsss, c = D.Ts, d = D.Tss;
        const double e = M.rtroe[4 * M.nr + r];   // SRI rows only: loaded here
        const double Pr = k0 * Mc / kinf;
        const double logPr = log10(Pr > OF_SMALL ? Pr : OF_SMALL);
        const double X = 1.0 / (1.0 + logPr * logPr);
        const double eb = exp(-b * iT), ec = exp(-T / c);
        const double psi = a * eb + ec;
        const double F = d * pow(psi, X) * pow(T, e);
        const double i1 = 1.0 / (1.0 + Pr);
        if (JAC) {
            const double dpsidT = a * b / (T * T) * eb - 1.0 / c * ec;
            const double dFdT = F * (X / psi * dpsidT + e / T);
            const double dlogPrdPr = Pr >= OF_SMALL ? 1.0 / (2.302585092994045684 * Pr) : 0.0;
            const double dXdPr = -(X * X) * 2.0 * logPr * dlogPrdPr;
            const double dFdPr = F * log(psi) * dXdPr;
            const double dPrdT = (Mc * dk0 - Pr * dkinf) / kinf;
            k.dkfdT = F * i1 * (Pr * dkinf) + F * i1 * i1 * dPrdT * kinf + k0 * i1 * dFdT * Mc;
            k.dkfdM = k0 * i1 * (F * i1 + Pr * dFdPr);
        }
        k.kf = Mc * (i1 * F * k0);
        return k;
    }
    // Troe
    const double alpha = D.alpha, Tsss = D.Tsss, Ts = D.Ts, Tss = D.Tss;
    const double e1 = exp(-T / Tsss), e2 = exp(-Tss * iT), e3 = exp(-T / Ts);
    const double Fcent = (1.0 - alpha) * e1 + alpha * e3 + e2;
    const double logFc = log10(Fcent > TROE_SMALL ? Fcent : TROE_SMALL);
    const double c = -0.4 - 0.67 * logFc;
    const double nn = 0.75 - 1.27 * logFc;
    if (!JAC) {
        double Pr = k0 * Mc / kinf;
        Pr = Pr > TROE_SMALL ? Pr : TROE_SMALL;
        const double logPr = log10(Pr);
        const double x1 = nn - 0.14 * (logPr + c);
        const double x2 = (logPr + c) / x1;
        const double F = exp10(logFc / (1.0 + x2 * x2));
        k.kf = Mc * (k0 * F / (1.0 + Pr));
        return k;
    }
    const double Pr0 = Mc * k0 / kinf;
    const bool prOk = Pr0 >= TROE_SMALL;
    const double Pr = Pr0 + 1e-100;
    const double logPr = log10(Pr > TROE_SMALL ? Pr : TROE_SMALL);
    const double x1 = nn - 0.14 * (logPr + c);
    const double x2 = (logPr + c) / x1;
    const double x3 = 1.0 + x2 * x2;
    const double x4 = logFc / x3;
    const double F = exp10(x4);
    const double ln10 = 2.302585092994045684;
    const double dFcdT = (alpha - 1.0) / Tsss * e1 - alpha / Ts * e3 + e2 * Tss * iT * iT;
    const double dlogFcdT = (Fcent >= TROE_SMALL) ? dFcdT / Fcent / ln10 : 0.0;
    const double dcdT = -0.67 * dlogFcdT, dndT = -1.27 * dlogFcdT;
    const double dx1dT = dndT - 0.14 * dcdT;
    const double dx2dT = (dcdT - x2 * dx1dT) / x1;
    const double dx3dT = 2.0 * x2 * dx2dT;
    const double dx4dT = (dlogFcdT - x4 * dx3dT) / x3;
    double dFdT = ln10 * F * dx4dT;
    const double dlogPrdPr = prOk ? 1.0 / (Pr * ln10) : 0.0;
    const double dx1dPr = -0.14 * dlogPrdPr;
    const double dx2dPr = (dlogPrdPr - x2 * dx1dPr) / x1;
    const double dx3dPr = 2.0 * x2 * dx2dPr;
    const double dx4dPr = -x4 * dx3dPr / x3;
    const double dFdPr = ln10 * F * dx4dPr;
    const double dPrdT = (Mc * dk0 - Pr * dkinf) / kinf;
    dFdT += dFdPr * dPrdT;
    const double i1 = 1.0 / (1.0 + Pr);
    k.dkfdT = F * i1 * (Pr * dkinf) + kinf * F * i1 * i1 * dPrdT + k0 * i1 * dFdT * Mc;
    k.dkfdM = k0 * i1 * (F * i1 + Pr * dFdPr);
    k.kf = (F * k0 * i1) * Mc;
    return k;
}

// Third-body concentrations M_t = sum_j eff_tj C_j of every third-body and
// falloff row (src/Reaction/ReactionRate.C:178-399 forms the same sums per
// row); L.C must be complete.  Passes of G rows, one row per thread; a last
// pass of at most G/2 (G/4) rows splits the species over 2 (4) lane groups
// of each wavefront, each summing H of them, and adds the groups' sums by
// xor shuffles, so it issues 1/2 (1/4) of the efficiency loads.  H is a
// multiple of 8 and S H <= SW, so every index stays inside the zero-padded
// table (nTBpad rows, SW species) and C[0..SW-1] (zero beyond ns, 1.0 in
// the workgroup form's ONE slot, which meets a zero efficiency).
template <int NP>
__device__ inline void third_body_sums(const DevMech& M, const Lds& L)
{
    constexpr int G = Cell<NP>::G;
    // batches of TS species: 16 in the wavefront form, 8 in the workgroup
    // form, whose two-workgroups-per-CU build has half the registers
    constexpr int TS = Cell<NP>::kLarge ? 8 : 16;
    const int lane = lane_id(), w = tid<NP>() >> 6;
    for (int t0 = 0; t0 < M.nTB; t0 += G) {
        const int R = M.nTB - t0;   // uniform
        const int S = R > G / 2 ? 1 : (R > G / 4 ? 2 : 4);
        const int rows = 64 / S;    // rows per wavefront
        const int t = t0 + w * rows + lane % rows;
        const int H = round_up((M.ns + S - 1) / S, 8);
        const int jb = (lane / rows) * H;
        double m = 0.0;
        int k = 0;
        for (; k + TS <= H; k += TS) {
            // unconditional loads, all 2 TS of a batch in flight at once
            double e[TS], c[TS];
#pragma unroll
            for (int u = 0; u < TS; ++u) {
                c[u] = L.C[jb + k + u];
                e[u] = M.effT[(size_t)(jb + k + u) * M.nTBpad + t];
            }
#pragma unroll
            for (int u = 0; u < TS; ++u) m += e[u] * c[u];   // no species test: padding adds 0 * c
        }
        if constexpr (TS > 8) {
            if (k < H) {   // H - k == 8
                double e[8], c[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    c[u] = L.C[jb + k + u];
                    e[u] = M.effT[(size_t)(jb + k + u) * M.nTBpad + t];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) m += e[u] * c[u];
            }
        }
        if (S >= 4) m += __shfl(m, lane ^ 16, 64);
        if (S >= 2) m += __shfl(m, lane ^ 32, 64);
        if (lane < rows && t < M.nTB) L.Mtb[t] = m;
    }
}

// Reaction round: lane l handles reactions l, l+64, ...
//  !JAC: rq[r] = q_r
//   JAC: rq[r] = kf, rk[r] = kr, rdT[r] = dq/dT, rw[tb] = dkf/dM CF - dkr/dM CR
// Kc in concentration units follows src/Reaction/ReactionFunctionI.H:1-52
// and ChemicalEquilibriumFunctionI.H:14-25, d ln Kc / dT
// src/Reaction/JacobianFunctionI.H:1150-1158.  The species of all six slots
// are gathered unconditionally (unused slots name the identity slot 64 in
// the wavefront form, species 0 in the workgroup form) and the products take
// the used slots in slot order.
// One reaction of a round: rates from the preloaded record D and the
// gathered slot values c (concentrations), g (exp(-G/RT)), db (d(-G/RT)/dT).
template <bool JAC, bool ONES>
__device__ __forceinline__ void reaction_eval(const DevMech& M, const Lds& L, const RxnData& D, int r,
                                              const double* c, const double* g, const double* db, double T,
                                              double lnT, double iT, double lnp, double PbyRT)
{
    const unsigned long long pk = D.pk;
    const int nl = pk_nl(pk), nrh = pk_nr(pk);
    const RxnK k = rate_constant<JAC>(M, L, D, r, T, lnT, iT);
    double kr = 0.0, dkrdT = 0.0, dkrdM = 0.0;
    double CRg = 1.0;   // product of every product slot's C (rrev 3)
    if (D.rev == 3) {
        // Kc-reversible with more than 3 product slots: the reference's
        // general path over every slot (RFG, ReactionFunctionI.H:648-720;
        // JFG, JacobianFunctionI.H:2965-3060).  RFG floors Kc at
        // 1.49011611938476e-8 (the value of Tc.C), JFG at the usual
        // 1.4901171103413047e-8.  The Jacobian's reverse partials
        // kr prod_{b != q} C_b go to the reaction arrays past nr, where the
        // term walk reads them.
        const auto gp = M.gen + M.rgen[r] * GEN_STRIDE;
        const int n = gp[0], base = gp[1];
        double num = 1.0, den = 1.0, sdb = 0.0;
        for (int a = 0; a < n; ++a) {
            const int s = gp[2 + a];
            num *= L.expG[s];
            CRg *= L.C[s];
            if (JAC) sdb += L.dBdT[s];
        }
        for (int a = 0; a < nl; ++a) {
            const int s = pk_species(pk, a);
            den *= L.expG[s];
            if (JAC) sdb -= L.dBdT[s];
        }
        const int dn = n - nl;   // >= 1: at least 4 product and at most 3 reactant slots
        double pw = 1.0;
        for (int q = 0; q < dn; ++q) pw *= PbyRT;
        const double kcMin = JAC ? KC_FLOOR : TC_KC_FLOOR;
        double Kc = (num / den) * pw;
        const bool floored = !(Kc > kcMin);
        Kc = floored ? kcMin : Kc;
        const double invKc = 1.0 / Kc;
        kr = k.kf * invKc;
        if (JAC) dkrdT = k.dkfdT * invKc - (floored ? 0.0 : kr * (sdb - dn * iT));
        dkrdM = k.dkfdM * invKc;
        if (JAC) {
            for (int q = 0; q < n; ++q) {
                double P = kr;
                for (int b = 0; b < n; ++b)
                    if (b != q) P *= L.C[gp[2 + b]];
                L.rq[base + q] = P;
            }
        }
    } else if (D.rev == 2) {
        // nonEquilibrium: the reaction's own reverse Arrhenius (in the k_0
        // slot), times M for third-body rows (ReactionFunctionI.H:20-23,
        // ReactionRate.C:362-367; reverse efficiencies = forward ones,
        // OptReaction.C:678)
        const double kra = D.A0 * exp(D.b0 * lnT - D.Ta0 * iT);
        const double dkra = JAC ? kra * (D.b0 + D.Ta0 * iT) * iT : 0.0;
        const bool tb = D.t == RXN_TB;
        kr = tb ? kra * D.Mc : kra;
        dkrdT = tb ? dkra * D.Mc : dkra;
        dkrdM = tb ? kra : 0.0;
    } else if (D.rev) {
        double num = 1.0, den = 1.0, sdb = 0.0;
        const int dn = nrh - nl;
        double pw = 1.0;
        if constexpr (ONES && !JAC) {
            // 1/Kc as a product: db holds exp(+G/RT) here; unused slots read
            // the identity slot; the floor test Kc > floor reads as
            // 1/Kc < 1/floor
            const double P2 = PbyRT * PbyRT, M1 = 1.0 / PbyRT, M2 = M1 / PbyRT;
            const double ipw = dn == 1 ? M1 : dn == 2 ? M2 : dn == -1 ? PbyRT : dn == -2 ? P2 : 1.0;
            const double inv = ((((db[3] * db[4]) * db[5]) * ((g[0] * g[1]) * g[2]))) * ipw;
            const double invKc = inv < 1.0 / KC_FLOOR ? inv : 1.0 / KC_FLOOR;
            kr = k.kf * invKc;
            dkrdM = k.dkfdM * invKc;
        } else if constexpr (ONES) {
            // unused slots read the identity slot (1, 1, 0): the products and
            // sums below equal the used-slots-only ones exactly, and the
            // Kc-reversible form has at most 3 slots a side (|dn| <= 2)
            num = (g[3] * g[4]) * g[5];
            den = (g[0] * g[1]) * g[2];
            if (JAC) sdb = ((((db[3] + db[4]) + db[5]) - db[0]) - db[1]) - db[2];
            const double P2 = PbyRT * PbyRT, M1 = 1.0 / PbyRT, M2 = M1 / PbyRT;
            pw = dn == 1 ? PbyRT : dn == 2 ? P2 : dn == -1 ? M1 : dn == -2 ? M2 : 1.0;
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (a < nrh) { num *= g[3 + a]; if (JAC) sdb += db[3 + a]; }
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (a < nl) { den *= g[a]; if (JAC) sdb -= db[a]; }
            if (dn > 0) for (int q = 0; q < dn; ++q) pw *= PbyRT;
            if (dn < 0) for (int q = 0; q < -dn; ++q) pw /= PbyRT;
        }
        if constexpr (!(ONES && !JAC)) {
            double Kc = (num / den) * pw;
            const bool floored = !(Kc > KC_FLOOR);
            Kc = floored ? KC_FLOOR : Kc;
            const double invKc = 1.0 / Kc;
            kr = k.kf * invKc;
            if (JAC) dkrdT = k.dkfdT * invKc - (floored ? 0.0 : kr * (sdb - dn * iT));
            dkrdM = k.dkfdM * invKc;
        }
    }
    double CF = 1.0, CR = 1.0;
    if constexpr (ONES) {
        CF = (c[0] * c[1]) * c[2];
        CR = (c[3] * c[4]) * c[5];
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a < nl) CF *= c[a];
            if (a < nrh) CR *= c[3 + a];
        }
    }
    if (D.rev == 3) CR = CRg;
    if (!JAC) {
        L.rq[r] = k.kf * CF - kr * CR;
    } else {
        L.rq[r] = k.kf;
        // (rrev 3: kr CR, which the production walk takes with CR = 1 -- the
        // per-species list entries of these reactions name no product slot)
        L.rk[r] = D.rev == 3 ? kr * CR : kr;
        L.rdT[r] = k.dkfdT * CF - dkrdT * CR;
        if (D.t == RXN_TB || D.t == RXN_LIND || D.t == RXN_TROE || D.t == RXN_SRI)
            L.rw[D.tb] = k.dkfdM * CF - dkrdM * CR;
    }
}

// Reaction round: lane l handles reactions l, l+64, ..., two per pass (their
// table loads and gathers are all issued before either is evaluated).
//  !JAC: rq[r] = q_r
//   JAC: rq[r] = kf, rk[r] = kr, rdT[r] = dq/dT, rw[tb] = dkf/dM CF - dkr/dM CR
// Kc in concentration units follows src/Reaction/ReactionFunctionI.H:1-52
// and ChemicalEquilibriumFunctionI.H:14-25, d ln Kc / dT
// src/Reaction/JacobianFunctionI.H:1150-1158.  The species of all six slots
// are gathered unconditionally (unused slots name the identity slot 64 in
// the wavefront form, species 0 in the workgroup form) and the products take
// the used slots in slot order.
template <bool JAC, int NP>
__device__ inline void reaction_round(const DevMech& M, const Lds& L, double T, double lnT, double iT,
                                      double lnp)
{
    constexpr int G = Cell<NP>::G;
    const double PbyRT = PSTD / (RU * T);
    // reactions from the last one down: the tables order them by rate form,
    // falloff, PLOG and SRI forms last, so the costliest forms share the first slot and
    // the partly filled slot of the last round holds the cheapest ones (a
    // form's code runs once per slot in which any lane holds it)
    for (int k = tid<NP>(); k < M.nr; k += 2 * G) {
        const int r0 = M.nr - 1 - k;
        const int r1 = r0 - G;
        const bool has1 = r1 >= 0;
        const RxnData D0 = load_rxn(M, L, r0);
        const RxnData D1 = load_rxn(M, L, has1 ? r1 : r0, has1);
        double c0[6], g0[6], db0[6], c1[6], g1[6], db1[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int s0 = pk_species(D0.pk, q), s1 = pk_species(D1.pk, q);
            c0[q] = L.C[s0];
            g0[q] = L.expG[s0];
            db0[q] = (JAC || !Cell<NP>::kLarge) ? L.dBdT[s0] : 0.0;
            c1[q] = L.C[s1];
            g1[q] = L.expG[s1];
            db1[q] = (JAC || !Cell<NP>::kLarge) ? L.dBdT[s1] : 0.0;
        }
        reaction_eval<JAC, !Cell<NP>::kLarge>(M, L, D0, r0, c0, g0, db0, T, lnT, iT, lnp, PbyRT);
        if (has1) reaction_eval<JAC, !Cell<NP>::kLarge>(M, L, D1, r1, c1, g1, db1, T, lnT, iT, lnp, PbyRT);
    }
}

// Molar production rate of this lane's species from the rates q_r in
// L.rq: sum over its reaction list of nu_ir q_r, in list order.
template <int NP>
__device__ __forceinline__ double production_rate(const DevMech& M, const Lds& L)
{
    constexpr int SW = Cell<NP>::SW;
    const int lane = tid<NP>();
    double om = 0.0;
    for (int e0 = 0; e0 < M.maxDeg; e0 += 16) {   // unconditional batches of 16 (maxDeg % 16 == 0)
        int w[16];
        double q[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) w[u] = M.sp_rn[(e0 + u) * SW + lane];
#pragma unroll
        for (int u = 0; u < 16; ++u) q[u] = L.rq[rn_r(w[u])];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int nu = rn_nu(w[u]);
            om += (double)nu * q[u];   // padding entries: nu = 0
        }
    }
    return om;
}

// Chemical time scale sums of OptReaction::Tc (src/Reaction/Tc.C:566-636,
// 752-832): per reaction the forward and reverse molar rates, each times the
// number of species slots it produces, enter sumW and (squared)
// sumWRateByCTot.  The reverse rate is kf/Kc with the floor Tc.C uses
// (1.49011611938476e-8) for Kc-reversible reactions and zero otherwise
// (Tc.C:583 tests isIrreversible == 0, so nonEquilibrium reactions count
// as irreversible there).  L.C, L.expG and L.Mtb must be complete.
// Returns (sumW, sumWRateByCTot) in every lane.
template <int NP>
__device__ inline double2 time_scale_sums(const DevMech& M, const Lds& L, double T, double lnT, double iT,
                                          double lnp)
{
    const double PbyRT = PSTD / (RU * T);
    double sw = 0.0, sw2 = 0.0;
    for (int r = tid<NP>(); r < M.nr; r += Cell<NP>::G) {
        const RxnData D = load_rxn(M, L, r);
        const RxnK k = rate_constant<false>(M, L, D, r, T, lnT, iT);
        const int nl = pk_nl(D.pk), nrh = pk_nr(D.pk);
        double CF = 1.0, CR = 1.0, num = 1.0, den = 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a < nl) { CF *= L.C[pk_species(D.pk, a)]; den *= L.expG[pk_species(D.pk, a)]; }
            if (a < nrh) { CR *= L.C[pk_species(D.pk, 3 + a)]; num *= L.expG[pk_species(D.pk, 3 + a)]; }
        }
        double kr = 0.0;
        if (D.rev == 3) {   // every product slot (Tc.C:578-600 walks the whole side)
            const auto gp = M.gen + M.rgen[r] * GEN_STRIDE;
            CR = 1.0;
            num = 1.0;
            for (int a = 0; a < gp[0]; ++a) {
                CR *= L.C[gp[2 + a]];
                num *= L.expG[gp[2 + a]];
            }
        }
        if (D.rev == 1 || D.rev == 3) {
            const int dn = nrh - nl;
            double pw = 1.0;
            if (dn > 0) for (int q = 0; q < dn; ++q) pw *= PbyRT;
            if (dn < 0) for (int q = 0; q < -dn; ++q) pw /= PbyRT;
            double Kc = (num / den) * pw;
            Kc = Kc > TC_KC_FLOOR ? Kc : TC_KC_FLOOR;
            kr = k.kf / Kc;
        }
        const double wf = (double)nrh * (k.kf * CF), wr = (double)nl * (kr * CR);
        sw += wf + wr;
        sw2 += wf * wf + wr * wr;
    }
    const double a = gsum<NP>(sw, L);
    return make_double2(a, gsum<NP>(sw2, L));
}

// ---------------------------------------------------------------- RHS
// FastChemistryModel::derivatives (src/ChemistryModel/FastChemistryModel.C:
// 194-273, 340-417): y (this lane's element) is clipped in place for species
// lanes; returns this lane's element of dy/dt.
template <int NP>
__device__ inline double derivatives(const DevMech& M, const Lds& L, double p, double lnp, double& y)
{
    const int lane = tid<NP>();
    const bool sp = lane < M.ns;
    const SpeciesData S = load_species<NP>(M, lane);  // tables are padded to the group size
    if (sp) y = y > 0.0 ? y : 0.0;
    PhaseClock clk;
    const double T = gbcast<NP>(y, NP - 1, L);
    const double RuTByP = RU * T / p;                 // unclamped T for the density
    const double invW = sp ? S.invW : 0.0;
    const double invrho = gsum<NP>(sp ? y * invW * RuTByP : 0.0, L);
    const double rho = 1.0 / invrho;
    const double Tc = clampT(M, T), lnT = log(Tc), iT = 1.0 / Tc;
    Thermo th{0, 0, 0, 0, 0};
    if (sp) th = species_thermo(M, S, Tc, lnT, iT);
    L.C[lane] = sp ? rho * invW * y : 0.0;
    const double eg = sp ? exp(th.negG) : 0.0;
    L.expG[lane] = eg;
    if constexpr (!Cell<NP>::kLarge) {
        // exp(+G/RT) in the dBdT slots (no derivative is formed here), for
        // the rate round's division-free 1/Kc
        L.dBdT[lane] = sp ? 1.0 / eg : 0.0;
        if (lane == 0) set_identity_slot(L, 1.0);
    }
    wave_sync();
    clk.mark(M, lane == 0, PH_RHS_SETUP);
    third_body_sums<NP>(M, L);
    wave_sync();
    clk.mark(M, lane == 0, PH_RHS_TB);
    reaction_round<false, NP>(M, L, Tc, lnT, iT, lnp);
    wave_sync();
    clk.mark(M, lane == 0, PH_RHS_RATES);
    const double om = sp ? production_rate<NP>(M, L) : 0.0;
    clk.mark(M, lane == 0, PH_RHS_PROD);
    const double Yd = sp ? om * (S.W * invrho) : 0.0;
    const double cpM = gsum<NP>(sp ? y * th.cp : 0.0, L);
    const double hY = gsum<NP>(sp ? th.ha * Yd : 0.0, L);
    wave_sync();
    clk.mark(M, lane == 0, PH_RHS_REST);
    return sp ? Yd : (lane == NP - 1 ? -hY / cpM : 0.0);
}

// ---------------------------------------------------------------- Jacobian
// FastChemistryModel::jacobian (src/ChemistryModel/FastChemistryModel.C:
// 489-572) -> OptReaction::ddNdtByVdcTp (src/Reaction/Jacobian.C:6-1001) and
// the dC/dY conversion.  The exact conversion uses the rank-1 identity
//   sum_k Jc_ik (rho/W_k)(delta_kj - (MW/W_j) Y_k)
//     = Jc_ij rho/W_j - (MW/W_j) sum_k Jc_ik C_k
// (O(n^2) instead of the reference's dense O(n^3) product,
// Jacobian.C:2814-2857); fast keeps the diagonal dC/dY only
// (Jacobian.C:4570-4841).  y is clipped in place; Jy is left in L.J (row i
// at lane i, columns 0..NP-1, temperature column NP-1); returns f(y) for
// this lane.
template <int NP>
__device__ inline double jacobian(const DevMech& M, const Lds& L, double p, double lnp, int jacType, double& y)
{
    constexpr int SW = Cell<NP>::SW, JSt = Cell<NP>::JSt, ONE = Cell<NP>::ONE, G = Cell<NP>::G;
    const int lane = tid<NP>();
    const int ns = M.ns;
    const bool sp = lane < ns;
    PhaseClock clk;
    const SpeciesData S = load_species<NP>(M, lane);  // tables are padded to the group size
    if (sp) y = y > 0.0 ? y : 0.0;
    const double T = clampT(M, gbcast<NP>(y, NP - 1, L));
    const double lnT = log(T), iT = 1.0 / T;
    const double invW = sp ? S.invW : 0.0;
    Thermo th{0, 0, 0, 0, 0};
    if (sp) th = species_thermo(M, S, T, lnT, iT);
    const double RuTByP = RU * T / p;
    const double invrho = gsum<NP>(sp ? RuTByP * (y * invW) : 0.0, L);
    const double MW = 1.0 / gsum<NP>(sp ? y * invW : 0.0, L);
    const double rho = 1.0 / invrho;
    const double cpM = gsum<NP>(sp ? y * th.cp : 0.0, L);
    const double dcpMdT = gsum<NP>(sp ? y * th.dcpdT : 0.0, L);
    const double Cl = sp ? rho * y * invW : 0.0;
    L.C[lane] = lane == ONE ? 1.0 : Cl;
    if (ONE >= G && lane == 0) L.C[ONE] = 1.0;
    if (!Cell<NP>::kLarge && lane == 0) set_identity_slot(L);
    L.expG[lane] = sp ? exp(th.negG) : 0.0;
    L.dBdT[lane] = th.dBdT;
    double* Jrow = L.J + lane;   // element j of this thread's row at Jrow[j*JSt]
    if (!Cell<NP>::kLarge || lane < NP)   // (wavefront form: every lane owns a slot of each column)
        for (int j = 0; j < NP; ++j) Jrow[j * JSt] = 0.0;
    wave_sync();
    third_body_sums<NP>(M, L);
    wave_sync();
    reaction_round<true, NP>(M, L, T, lnT, iT, lnp);
    wave_sync();
    clk.mark(M, lane == 0, PH_JAC_RATES);
    // species rows, sparse partials: the terms nu_ir dq_r/dC_j of every
    // matrix element, grouped by element (rows, then columns ascending;
    // within an element in the order of row i's reaction list) and split
    // over the threads at element boundaries in runs of equal length, so no
    // thread walks more terms than the longest run (a row-per-thread walk
    // takes the longest row's list on every lane).  Each thread keeps a
    // running sum of the current element in a register and stores it after
    // every term (the element's last store holds its sum): each element is
    // written by one thread, no read-modify-write, no branch.
    {
        const int dumpOff = Cell<NP>::kLarge ? NP * NP + lane % NP : (lane % NP) * JSt + 64;
        double acc = 0.0;
        int cur = jterm_key(M.sp_term[lane]);
        unsigned long long wn[16];   // next batch, loaded one batch ahead
#pragma unroll
        for (int u = 0; u < 16; ++u) wn[u] = M.sp_term[u * SW + lane];
        for (int t0 = 0; t0 < M.maxTerms; t0 += 16) {
            unsigned long long w[16];
            double k[16], c1[16], c2[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) w[u] = wn[u];
            const int tn = t0 + 16 < M.maxTerms ? t0 + 16 : t0;
#pragma unroll
            for (int u = 0; u < 16; ++u) wn[u] = M.sp_term[(tn + u) * SW + lane];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int r = jterm_r(w[u]);
                k[u] = jterm_rev(w[u]) ? L.rk[r] : L.rq[r];
                c1[u] = L.C[jterm_o1(w[u])];   // ONE (= 1.0) when there is no other species
                c2[u] = L.C[jterm_o2(w[u])];
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int nu = jterm_nu(w[u]);
                const int key = jterm_key(w[u]);
                const double t = (double)nu * ((k[u] * c1[u]) * c2[u]);
                acc = key == cur ? acc + t : 0.0 + t;
                cur = key;
                L.J[jterm_dump(w[u]) ? dumpOff : (key & 255) * JSt + (key >> 8)] = acc;
            }
        }
    }
    // (other threads' rows are read after the wave_sync below)
    clk.mark(M, lane == 0, PH_JAC_TERMS);
    // production rate and d(omega)/dT of this lane's species (its reaction
    // list, loads and gathers issued per block of 8 entries)
    double om = 0.0, jT = 0.0;
    if (sp) {
        const int deg = M.sp_deg[lane];
        for (int e0 = 0; e0 < M.maxDeg; e0 += 8) {
            double kf[8], kr[8], rd[8], c[8][6];
            unsigned long long pk[8];
            int wv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {   // padding: reaction 0, nu 0, species slots 0
                wv[u] = 8 << 16;
                pk[u] = 0ull;
            }
            if (e0 < deg) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    wv[u] = M.sp_rn[(e0 + u) * SW + lane];
                    pk[u] = M.sp_pack[(e0 + u) * SW + lane];
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int r = rn_r(wv[u]);
                kf[u] = L.rq[r];
                kr[u] = L.rk[r];
                rd[u] = L.rdT[r];
#pragma unroll
                for (int q = 0; q < 6; ++q) c[u][q] = L.C[pk_species(pk[u], q)];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int nui = rn_nu(wv[u]);
                const double nu = (double)nui;
                const int nl = pk_nl(pk[u]), nrh = pk_nr(pk[u]);
                double CF = 1.0, CR = 1.0;
                if constexpr (!Cell<NP>::kLarge) {   // unused slots: the identity slot
                    CF = (c[u][0] * c[u][1]) * c[u][2];
                    CR = (c[u][3] * c[u][4]) * c[u][5];
                } else {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        if (q < nl) CF *= c[u][q];
                        if (q < nrh) CR *= c[u][3 + q];
                    }
                }
                om += nu * (kf[u] * CF - kr[u] * CR);
                jT += nu * rd[u];
            }
        }
    }
    wave_sync();
    clk.mark(M, lane == 0, PH_JAC_PROD);
    // third-body / falloff partials: element (i, j) gains sum_r nu_ir w_r
    // eff_rj over the third-body/falloff entries of species i.  Column-wise
    // over a flat entry list sorted by row (uniform loop, 16 entries per
    // batch, coalesced efficiency rows): lane j forms row i's sum and writes
    // old + sum once when the row ends.  A batch's entries are fetched with
    // one per-lane load (entry k0 + (lane mod 16)) and broadcast with
    // readlane, so its efficiency rows, w_t and the current value of each
    // entry's row are all requested before any of them is used; a row first
    // seen in this batch still holds its sparse-partial value (rows are
    // visited once, in ascending order), and the value read for a row
    // continuing from the last batch is not used.
    if constexpr (Cell<NP>::kLarge && !Cell<NP>::kJLds) {
        // matrix in HBM, where a column-wise read-modify-write per row reads
        // one cache line per thread: the rows go in blocks of 16 -- each
        // thread sums its column's partials of the block's rows into an LDS
        // tile, then the group adds the tile to the slice with 16 threads on
        // the 16 rows of one column (one 128-byte line), all lines in flight
        constexpr int TS = NP + 1;
        for (int b = 0; b < NP / 16; ++b) {
            const int e0 = __builtin_amdgcn_readfirstlane(M.tbz_blk[b]);
            const int e1 = __builtin_amdgcn_readfirstlane(M.tbz_blk[b + 1]);
            if (e0 == e1) continue;   // no entry in this block (uniform)
            if (lane < NP)
#pragma unroll
                for (int q = 0; q < 16; ++q) L.tbt[q * TS + lane] = 0.0;
            double acc = 0.0;
            int cur = -1;
            for (int k0 = e0; k0 < e1; k0 += 16) {
                const int e = k0 + (lane & 15) < M.nTBz ? k0 + (lane & 15) : M.nTBz - 1;
                const int vt = M.tbz_tb[e], vrow = M.tbz_row[e];
                const double vnu = M.tbz_nu[e];
                int rw_[16];
                double cf[16], ef[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const bool in = k0 + u < e1;   // uniform
                    const int t = readlane_i(vt, u);
                    rw_[u] = in ? readlane_i(vrow, u) : -2;   // -2: past the block, adds nothing
                    cf[u] = in ? readlane_d(vnu, u) * L.rw[t] : 0.0;
                    ef[u] = M.eff[t * SW + lane];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    if (rw_[u] != cur && rw_[u] != -2) {
                        if (cur >= 0 && lane < NP) L.tbt[(cur - 16 * b) * TS + lane] = acc;
                        acc = 0.0;
                        cur = rw_[u];
                    }
                    acc += cf[u] * ef[u];
                }
            }
            if (cur >= 0 && lane < NP) L.tbt[(cur - 16 * b) * TS + lane] = acc;
            __syncthreads();
            const int i = 16 * b + (lane & 15);
            for (int j = lane >> 4; j < NP; j += G / 16) L.J[(size_t)j * NP + i] += L.tbt[(lane & 15) * TS + j];
            __syncthreads();   // the tile is cleared by the next block
        }
    } else if (M.nTBz > 0) {
        double acc = 0.0, oldv = 0.0;
        int cur = -1;
        int nt = M.tbz_tb[lane & 15], nrow = M.tbz_row[lane & 15];
        double nnu = M.tbz_nu[lane & 15];
        for (int k0 = 0; k0 < M.nTBz; k0 += 16) {   // nTBz % 16 == 0
            const int vt = nt, vrow = nrow;
            const double vnu = nnu;
            if (k0 + 16 < M.nTBz) {   // the next batch's entries, one load round trip ahead
                const int e = k0 + 16 + (lane & 15);
                nt = M.tbz_tb[e];
                nrow = M.tbz_row[e];
                nnu = M.tbz_nu[e];
            }
            int rw_[16];
            double cf[16], ef[16], ov[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int t = readlane_i(vt, u);
                rw_[u] = readlane_i(vrow, u);
                cf[u] = readlane_d(vnu, u) * L.rw[t];
                ef[u] = M.eff[t * SW + lane];
                ov[u] = sp ? L.J[lane * JSt + rw_[u]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                if (rw_[u] != cur) {
                    if (sp && cur >= 0) L.J[lane * JSt + cur] = oldv + acc;
                    acc = 0.0;
                    cur = rw_[u];
                    oldv = ov[u];
                }
                acc += cf[u] * ef[u];
            }
        }
        if (sp) L.J[lane * JSt + cur] = oldv + acc;
    }
    wave_sync();
    clk.mark(M, lane == 0, PH_JAC_TB);
    // dC/dY conversion of the species rows, T column
    const double WiByRho = sp ? S.W * invrho : 0.0;
    const double Yd = sp ? om * WiByRho : 0.0;
    // (every column up to NP without a species test: the padding columns
    // and C are zero there, and the T column is rewritten after the loop)
    // (batches of CB elements, all loads of a batch in flight together: in
    // the HBM-matrix form every batch costs a round trip to the slice)
    constexpr int CB = Cell<NP>::kLarge ? 16 : (NP % 8 == 0 ? 8 : 4);
    // In the wavefront form every lane runs the conversion, so that the
    // readlanes of 1/W_j (lane j holds it) sit in wave-uniform control flow:
    // WiByRho and Yd are zero outside the species lanes, so the padding rows
    // and the temperature row get zeros (the latter is rewritten below), and
    // lanes past NP write only column slots no phase reads.  The workgroup
    // form reads 1/W_j from the table in the species threads.
    const bool conv = Cell<NP>::kLarge ? sp : true;
    if (conv) {
        double sC = 0.0;
        for (int j0 = 0; j0 < NP; j0 += CB) {
            double v[CB], c[CB];
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                v[u] = Jrow[(j0 + u) * JSt];
                c[u] = L.C[j0 + u];
            }
#pragma unroll
            for (int u = 0; u < CB; ++u) sC += v[u] * c[u];
        }
        const double sCx = jacType == 0 ? sC : 0.0;   // fast: diagonal dC/dY only
        for (int j0 = 0; j0 < NP; j0 += CB) {
            double v[CB], iw[CB];
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                v[u] = Jrow[(j0 + u) * JSt];
                iw[u] = Cell<NP>::kLarge ? M.invW[j0 + u] : readlane_d(S.invW, j0 + u);
            }
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                const double iWj = iw[u];
                Jrow[(j0 + u) * JSt] = jacType == 0
                    ? WiByRho * (v[u] * (rho * iWj) - (MW * iWj) * sCx) + (MW * iWj) * Yd
                    : WiByRho * (v[u] * (rho * iWj)) + (MW * iWj) * Yd;
            }
        }
        Jrow[(NP - 1) * JSt] = WiByRho * (jT - sC * iT) + iT * Yd;
    }
    // temperature row (src/Reaction/Jacobian.C:3743-3791): lane j < NP forms
    // column j's sum over the species rows and writes row NP-1, column j
    const double dTdt = -gsum<NP>(sp ? Yd * th.ha : 0.0, L) / cpM;
    L.H[lane] = sp ? th.ha : 0.0;
    L.Q[lane] = sp ? Yd * th.cp : 0.0;
    wave_sync();
    // (the sums read species rows only, so row NP-1 can be written at once)
    if (lane < NP) {
        const bool tcol = lane == NP - 1;
        double colsum = 0.0;
        for (int q0 = 0; q0 < NP; q0 += CB) {   // every row: H and Q are zero beyond ns
            double v[CB], h[CB], qq[CB];
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                const int i = q0 + u;
                v[u] = L.J[lane * JSt + i];
                h[u] = L.H[i];
                qq[u] = tcol ? L.Q[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CB; ++u) colsum += v[u] * h[u] + qq[u];
        }
        const double rowT = !tcol ? (sp ? (-colsum - th.cp * dTdt) / cpM : 0.0)
                                  : (-colsum - dTdt * dcpMdT) / cpM;
        L.J[lane * JSt + (NP - 1)] = rowT;
    }
    wave_sync();
    clk.mark(M, lane == 0, PH_JAC_CONV);
    return sp ? Yd : (lane == NP - 1 ? dTdt : 0.0);
}

// ---------------------------------------------------------------- LU
// Block-4 LU with pivoting restricted to each 4-row diagonal block
// (src/LUsolver/LUsolver.C:106-246 full blocks, :503-551 / :553-650 tails).
// Wavefront form: each lane holds its matrix row in registers for the whole
// factorisation (loaded from LDS, factors stored back for the solves).
// Rows never move: pos is the logical row a lane holds, swapped as the
// reference swaps rows, and ip[r] is the lane holding logical row r when the
// factorisation ends.  Right-looking elimination; a pivot row's entry is
// broadcast with readlane.
// invD: reciprocal pivot of logical row `lane`; ipl: lane that held logical row `lane` during the factorisation
struct LuState { double invD; int pos; int ipl; };

template <int NP>
__device__ inline LuState lu_factor(const DevMech& M, const Lds& L)
{
    static_assert(NP % 4 == 0 && NP <= 64, "wavefront-form LU: padded size a multiple of 4, at most 64");
    const int lane = lane_id();
    PhaseClock clk;
    const bool row = lane < NP;
    double* Jrow = L.J + lane;
    double R[NP];            // this lane's row
#pragma unroll
    for (int j = 0; j < NP; ++j) R[j] = Jrow[j * JS];
    int pos = lane;          // lanes >= NP hold no row
    double invD = 0.0;
#pragma unroll
    for (int k0 = 0; k0 < NP; k0 += 4) {
        // ---- panel: columns k0..k0+3.  Each column takes the pivot choice,
        // swap, eps replacement and elimination of the reference, restricted
        // to the panel columns.
        // Swaps only ever exchange rows within a block, so at the block's
        // start lane k0+q holds logical row k0+q; lr tracks the lane of each
        // of the block's logical rows (uniform) through the block's swaps.
        int lr[4] = {k0, k0 + 1, k0 + 2, k0 + 3};
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int k = k0 + d;
            const double ak = R[k];
            double v[4];     // |column-k entry| of logical rows k0+q, q >= d
#pragma unroll
            for (int q = d; q < 4; ++q) v[q] = fabs(readlane_d(ak, lr[q]));
            int im = d;
            if (d == 0) {
                double temp = 0.0;
                im = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (temp < v[q]) { im = q; temp = v[q]; }
            } else if (d == 1) {
                double t = v[1];
                if (t < v[2]) {
                    im = 2;
                    t = readlane_d(ak, lr[2]);   // signed, LUsolver.C:177
                }
                if (t < v[3]) im = 3;
            } else if (d == 2) {
                if (v[2] < v[3]) im = 3;
            }
            const int pick = k0 + im;
            if (pick != k) {
                if (pos == k) pos = pick;
                else if (pos == pick) pos = k;
#pragma unroll
                for (int q = d + 1; q < 4; ++q)
                    if (q == im) { const int t = lr[d]; lr[d] = lr[q]; lr[q] = t; }
            }
            const int p = lr[d];
            double piv = readlane_d(ak, p);
            if (piv == 0.0) {                        // LUsolver.C:155
                piv = PIVOT_EPS;
                if (lane == p) R[k] = PIVOT_EPS;
            }
            const double inv = 1.0 / piv;
            if (lane == p) invD = inv;
            const bool elim = row && pos > k;
            const double l = ak * inv;
            if (elim) R[k] = l;
#pragma unroll
            for (int e = d + 1; e < 4; ++e) {
                const double pe = readlane_d(R[k0 + e], p);
                if (elim) R[k0 + e] = R[k0 + e] - l * pe;
            }
        }
        const int pl[4] = {lr[0], lr[1], lr[2], lr[3]};   // lane holding logical row k0+d
        clk.mark(M, lane == 0, PH_LU_PANEL);
        // ---- trailing columns: step q subtracts l_iq u_qj from every row
        // below logical row k0+q, u_qj read from that row after steps 0..q-1
        // -- the column-by-column elimination's operations in its order (the
        // pivot rows' own entries become the U rows on the way).  Rows not
        // below take a zero multiplier, which leaves them unchanged exactly.
        double m[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) m[q] = (row && pos > k0 + q) ? R[k0 + q] : 0.0;
#pragma unroll
        for (int j = k0 + 4; j < NP; ++j) {
            double a = R[j];
#pragma unroll
            for (int q = 0; q < 4; ++q) a = a - m[q] * readlane_d(a, pl[q]);
            R[j] = a;
        }
        clk.mark(M, lane == 0, PH_LU_TRAIL);
    }
    // factors stored in logical row order (row pos at slot pos of every
    // column), the reciprocal pivot in the diagonal slot, so the solves read
    // logical row r from lane r
    if (row) {
        double* Jlog = L.J + pos;
#pragma unroll
        for (int j = 0; j < NP; ++j) Jlog[j * JS] = R[j];
        Jlog[pos * JS] = invD;   // the diagonal slot holds 1/pivot
        L.ip[pos] = lane;
    }
    wave_sync();
    const int ipl = row ? L.ip[lane] : lane;
    const double invDl = row ? Jrow[lane * JS] : 0.0;   // reciprocal pivot of logical row `lane`
    wave_sync();
    return LuState{invDl, pos, ipl};
}

// Solves (L U) x = P b; b is this lane's element in the original row order
// (lane i = row i); returns x_i for lane i (src/LUsolver/LUsolver.C:
// 3922-4111: swaps, unit-L forward, U backward scaled by 1/pivot).  The
// factors are in logical row order (lu_factor), so after one permuting
// exchange of b every step reads a fixed lane: y_r = b_r, then every lane
// below subtracts its L entry times y_r; backward x_r = b_r / u_rr and every
// lane above subtracts its U entry times x_r.  A lane's entries on the other
// side of the diagonal enter as zero multipliers.
template <int NP>
__device__ inline double lu_solve(const Lds& L, const LuState& st, double b)
{
    const int lane = lane_id();
    const double* Jrow = L.J + lane;
    b = __shfl(b, st.ipl, WAVE);                  // lane r: b of logical row r
#pragma unroll
    for (int r = 0; r < NP - 1; ++r) {
        const double yr = readlane_d(b, r);
        const double lr = Jrow[r * JS];
        b = b - (lane > r ? lr : 0.0) * yr;
    }
    double x = 0.0;
#pragma unroll
    for (int r = NP - 1; r >= 0; --r) {
        const double xr = readlane_d(b * st.invD, r);
        const double ur = Jrow[r * JS];
        b = b - (lane < r ? ur : 0.0) * xr;
        if (lane == r) x = xr;
    }
    return x;
}

}  // namespace fcs
