// fcs_lu_large.h -- block-4-pivoted LU factor / solve of the workgroup form
// (systems of more than 64 rows; the matrix lives in HBM, one slice per
// workgroup, column-major with column stride NP).
//
// Pivoting follows the reference's Block4LUDecompose
// (src/LUsolver/LUsolver.C:106-246, :3058-3141): the pivot of column k is
// searched among the four rows of k's 4-row diagonal block only, with the
// reference's comparison rules (including the signed running maximum in
// the block's second column, LUsolver.C:177), zero pivots are replaced by
// 2.2204460492503131e-16, and rows never move: each thread owns one
// physical row and carries its logical position `pos`.  Because swaps stay
// inside aligned 4-row blocks, the physical rows of every aligned row range
// hold exactly the logical rows of that range.
//
// The factorisation runs in panels of LU_PW columns:
//  1. panel: each thread keeps its row's LU_PW panel values in registers.
//     For every 4-block the four pivot rows are exchanged through LDS and
//     every thread factors that 4 x LU_PW strip redundantly (pivot choice,
//     swaps, 1/pivot, multipliers and the pivot rows' updated values), then
//     forms its own row's multipliers and updates its remaining panel
//     columns in the elimination's order;
//  2. U12: thread j forms the panel's pivot-row entries of trailing column
//     j by forward substitution with the panel's unit-lower block (the same
//     subtractions, in the same order, as the column-by-column elimination);
//  3. trailing update A22 -= L21 U12 on FP64 MFMA (v_mfma_f64_16x16x4f64),
//     16 x 16 tiles spread over the workgroup's wavefronts, L21 and U12
//     staged in LDS; only this step sums in another order than the
//     reference.
// The triangular solves are blocked the same way: a 16-row diagonal block is
// solved by 16 lanes of wavefront 0, the other rows are updated in parallel.
#pragma once
#include "fcs_device.h"

namespace fcs {

typedef double dbl4 __attribute__((ext_vector_type(4)));

struct LuLarge { int pos; };

// LDS layout of the LU scratch (Lds::lu), LU_PW = 16.
struct LuScratch {
    double* blk;     // [2][4][LU_PW]  pivot rows of a 4-block (double-buffered)
    int* blkPos;     // [2][4]         their logical positions
    double* Lp;      // [LU_PW][LU_PW] unit-lower block of the panel, logical order
    double* negL;    // [LU_PW][NP]    -L21 (k-major), rows below the panel
    double* U12;     // [LU_PW][NP]    U12 (k-major), columns right of the panel
    double* v0;      // [LU_PW]        solve exchange vectors
    double* v1;      // [LU_PW]
    int* p2;         // [LU_PW]        physical row of each logical row of the panel
};

template <int NP>
__device__ __forceinline__ LuScratch lu_scratch(const Lds& L)
{
    LuScratch S;
    double* b = L.lu;
    S.blk = b;
    S.Lp = b + 8 * LU_PW;
    S.negL = S.Lp + LU_PW * LU_PW;
    S.U12 = S.negL + LU_PW * NP;
    S.v0 = S.U12 + LU_PW * NP;
    S.v1 = S.v0 + LU_PW;
    S.p2 = (int*)(S.v1 + 2 * LU_PW);
    S.blkPos = S.p2 + LU_PW;
    return S;
}

// One 4-block (columns k0 + 4B .. k0 + 4B + 3) of the panel factorisation;
// B is a template argument so that every register index is static.  The
// 4 x LU_PW strip of the block's pivot rows is factored by wavefront 0, lane
// (q, e) = (lane / 16, lane % 16) holding element e of the block's physical
// row kb + q: pivot choice, swaps, 1/pivot, multipliers and the pivot rows'
// updated values, with the reference's rules.  It leaves the factored strip,
// the reciprocal pivots, the picks and the pivot rows' places in LDS; every
// thread then applies the picks to its position and, when its row lies
// below the block, forms its multipliers and updates its panel columns, in
// the elimination's order.
template <int NP, int B>
__device__ __forceinline__ void lu_panel_block(const LuScratch& S, double* invDg, int t, bool act, int k0, int& pos,
                                               int& buf, double (&pa)[LU_PW])
{
    constexpr int PW = LU_PW;
    constexpr int b = B;
    const int kb = k0 + 4 * b;
    double* blk = S.blk + buf * 4 * PW;
    int* bpos = S.blkPos + buf * 4;
    double* invS = S.v0;                 // [4] reciprocal pivots (the solves' exchange vectors, free here)
    int* pickS = (int*)S.v1;             // [4] physical row picked for logical kb + d
    int* plS = pickS + 4;                // [4] place (0..3) of the pivot row of logical kb + d
    if (t >= kb && t < kb + 4) {
#pragma unroll
        for (int e = 4 * b; e < PW; ++e) blk[(t - kb) * PW + e] = pa[e];
        bpos[t - kb] = pos;
    }
    __syncthreads();
    if (t < 64) {
        const int q = t >> 4, e = t & 15;
        double v = e >= 4 * b ? blk[q * PW + e] : 0.0;
        int vp[4];   // logical position of the block's physical row kb + z (uniform)
#pragma unroll
        for (int z = 0; z < 4; ++z) vp[z] = __builtin_amdgcn_readfirstlane(bpos[z]);
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int k = kb + d;
            const int c = 4 * b + d;   // panel column
            double colv[4];            // column c of the block's rows (uniform)
#pragma unroll
            for (int z = 0; z < 4; ++z) colv[z] = readlane_d(v, z * 16 + c);
            // value in column k of the row holding logical kb+qq
            auto V = [&](int qq) -> double {
                double r = 0.0;
#pragma unroll
                for (int z = 0; z < 4; ++z)
                    if (vp[z] == kb + qq) r = colv[z];
                return r;
            };
            int pick = k;
            if (d == 0) {
                double temp = 0.0;
                int im = 0;
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) {
                    const double a = fabs(V(qq));
                    if (temp < a) { im = qq; temp = a; }
                }
                pick = kb + im;
            } else if (d == 1) {
                double tt = fabs(V(1));
                int im = 1;
                if (tt < fabs(V(2))) {
                    im = 2;
                    tt = V(2);   // signed, LUsolver.C:177
                }
                if (tt < fabs(V(3))) im = 3;
                pick = kb + im;
            } else if (d == 2) {
                if (fabs(V(2)) < fabs(V(3))) pick = kb + 3;
            }
            if (pick != k) {
#pragma unroll
                for (int z = 0; z < 4; ++z) {
                    if (vp[z] == k) vp[z] = pick;
                    else if (vp[z] == pick) vp[z] = k;
                }
            }
            int p = 0;
#pragma unroll
            for (int z = 0; z < 4; ++z)
                if (vp[z] == k) p = z;
            double piv = 0.0;
#pragma unroll
            for (int z = 0; z < 4; ++z)
                if (z == p) piv = colv[z];
            if (piv == 0.0) {                        // LUsolver.C:155
                piv = PIVOT_EPS;
                if (q == p && e == c) v = PIVOT_EPS;
            }
            const double inv = 1.0 / piv;
            if (t == 0) {
                invS[d] = inv;
                pickS[d] = pick;
                plS[d] = p;
                invDg[kb + p] = inv;
            }
            // block rows below the pivot in logical order: multiplier in
            // column c, the later columns less multiplier x pivot row
            const double pe = __shfl(v, p * 16 + e, 64);
            const double l = __shfl(v, q * 16 + c, 64) * inv;
            int vq = vp[0];
#pragma unroll
            for (int z = 1; z < 4; ++z)
                if (q == z) vq = vp[z];
            if (vq > k) {
                if (e == c) v = l;
                else if (e > c) v = v - l * pe;
            }
        }
        if (e >= 4 * b) blk[q * PW + e] = v;
    }
    __syncthreads();
    // positions after the block's swaps, then this thread's row
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const int k = kb + d, pick = pickS[d];
        if (pick != k) {
            if (pos == k) pos = pick;
            else if (pos == pick) pos = k;
        }
    }
    if (act && t >= kb + 4) {
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int c = 4 * b + d;
            const double* pr = blk + plS[d] * PW;   // the pivot row: final beyond column c
            const double l = pa[c] * invS[d];
            pa[c] = l;
#pragma unroll
            for (int e = c + 1; e < PW; ++e) pa[e] = pa[e] - l * pr[e];
        }
    }
    // pivot-row threads take their final values
    if (t >= kb && t < kb + 4) {
#pragma unroll
        for (int e = 4 * b; e < PW; ++e) pa[e] = blk[(t - kb) * PW + e];
    }
    buf ^= 1;
}

template <int NP>
__device__ inline LuLarge lu_factor_large(const DevMech& M, const Lds& L)
{
    static_assert(NP % LU_PW == 0 && NP % 16 == 0 && LU_PW == 16, "workgroup-form LU: NP a multiple of 16");
    constexpr int PW = LU_PW;
    const int t = (int)threadIdx.x;
    const bool row = t < NP;
    double* J = L.J;
    double* invDg = L.J + (size_t)NP * NP;
    const LuScratch S = lu_scratch<NP>(L);
    int pos = t;
    int buf = 0;
    PhaseClock clk;
    for (int k0 = 0; k0 < NP; k0 += PW) {
        // ---- 1. panel factorisation (rows >= k0)
        double pa[PW];
        const bool act = row && t >= k0;
#pragma unroll
        for (int d = 0; d < PW; ++d) pa[d] = act ? J[(size_t)(k0 + d) * NP + t] : 0.0;
        lu_panel_block<NP, 0>(S, invDg, t, act, k0, pos, buf, pa);
        lu_panel_block<NP, 1>(S, invDg, t, act, k0, pos, buf, pa);
        lu_panel_block<NP, 2>(S, invDg, t, act, k0, pos, buf, pa);
        lu_panel_block<NP, 3>(S, invDg, t, act, k0, pos, buf, pa);
        // ---- panel to HBM; -L21 and the panel's unit-lower block to LDS
        if (act) {
#pragma unroll
            for (int d = 0; d < PW; ++d) J[(size_t)(k0 + d) * NP + t] = pa[d];
        }
        if (act && t >= k0 + PW) {
#pragma unroll
            for (int d = 0; d < PW; ++d) S.negL[d * NP + t] = -pa[d];
        }
        if (t >= k0 && t < k0 + PW) {
            const int q = pos - k0;
            S.p2[q] = t;
#pragma unroll
            for (int d = 0; d < PW; ++d) S.Lp[q * PW + d] = d < q ? pa[d] : 0.0;
        }
        __syncthreads();
        clk.mark(M, t == 0, PH_LU_PANEL);
        if (k0 + PW >= NP) break;
        // ---- 2. U12: thread j forms the pivot rows' entries of column j
        const int j = t;
        if (j >= k0 + PW && j < NP) {
            double u[PW];
#pragma unroll
            for (int q = 0; q < PW; ++q) u[q] = J[(size_t)j * NP + S.p2[q]];
#pragma unroll
            for (int q = 1; q < PW; ++q) {
                double a = u[q];
#pragma unroll
                for (int r = 0; r < q; ++r) a = a - S.Lp[q * PW + r] * u[r];
                u[q] = a;
            }
#pragma unroll
            for (int q = 0; q < PW; ++q) {
                J[(size_t)j * NP + S.p2[q]] = u[q];
                S.U12[q * NP + j] = u[q];
            }
        }
        __syncthreads();
        clk.mark(M, t == 0, PH_LU_UROWS);
        // ---- 3. trailing update on FP64 MFMA: tiles of rows r0.., columns c0..
        const int m0 = k0 + PW, nT = (NP - m0) / 16;
        const int lane = lane_id(), w = t >> 6;
        const int li = lane & 15, lk = lane >> 4;
        // tiles of this wavefront in groups of 4: the C loads of a group are
        // all in flight before its MFMAs
        const int nTiles = nT * nT;
        for (int base = w; base < nTiles; base += 4 * Cell<NP>::NW) {
            dbl4 acc[4];
            int r0[4], c0[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int tile = base + u * Cell<NP>::NW;   // uniform over the wavefront
                r0[u] = m0 + 16 * (tile % nT);
                c0[u] = m0 + 16 * (tile / nT);
                if (tile < nTiles) {
                    // C / D layout of v_mfma_f64_16x16x4f64: lane l, element i ->
                    // row l / 16 + 4 i, column l % 16 (checked on the device)
                    const double* cp = J + (size_t)(c0[u] + li) * NP + r0[u] + lk;
                    acc[u] = dbl4{cp[0], cp[4], cp[8], cp[12]};
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (base + u * Cell<NP>::NW < nTiles) {
#pragma unroll
                    for (int s4 = 0; s4 < PW / 4; ++s4) {
                        // A (16 x 4): lane l -> row l % 16, k l / 16;  B (4 x 16): lane l -> k l / 16, column l % 16
                        const double a = S.negL[(4 * s4 + lk) * NP + r0[u] + li];
                        const double bb = S.U12[(4 * s4 + lk) * NP + c0[u] + li];
                        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc[u], 0, 0, 0);
                    }
                    double* cp = J + (size_t)(c0[u] + li) * NP + r0[u] + lk;
                    cp[0] = acc[u][0];
                    cp[4] = acc[u][1];
                    cp[8] = acc[u][2];
                    cp[12] = acc[u][3];
                }
            }
        }
        __syncthreads();
        clk.mark(M, t == 0, PH_LU_TRAIL);
    }
    if (row) L.ip[pos] = t;
    __syncthreads();
    return LuLarge{pos};
}

// Solves (L U) x = P b (src/LUsolver/LUsolver.C:3922-4111: unit-L forward,
// U backward scaled by 1/pivot); b is this thread's element in the original
// row order; returns x_t for thread t.  Each panel's factors are loaded at
// the start of its step: a prefetch of the next panel into a second
// register set cost more than it hid, through the spills it added under the
// 256-register cap of the two-workgroups-per-CU build.
template <int NP>
__device__ inline double lu_solve_large(const Lds& L, const LuLarge& st, double b)
{
    constexpr int PW = LU_PW;
    const int t = (int)threadIdx.x;
    const bool row = t < NP;
    const double* J = L.J;
    const double* invDg = L.J + (size_t)NP * NP;
    const LuScratch S = lu_scratch<NP>(L);
    const int lane = lane_id();
    const bool diag = t < PW;   // wavefront 0, lanes 0..15: the diagonal block
    // row updates (lr: this thread's factors in the panel's columns) and the
    // diagonal block's factors (dq: lane q = logical row k0 + q)
    double lr[PW], dq[PW];
    auto load_fwd = [&](int k0, double (&l)[PW], double (&d)[PW]) {
#pragma unroll
        for (int r = 0; r < PW; ++r) l[r] = (row && t >= k0 + PW) ? J[(size_t)(k0 + r) * NP + t] : 0.0;
        const int ph = diag ? L.ip[k0 + lane] : 0;
#pragma unroll
        for (int r = 0; r < PW; ++r) d[r] = (diag && r < lane) ? J[(size_t)(k0 + r) * NP + ph] : 0.0;
    };
    auto load_bwd = [&](int k0, double (&l)[PW], double (&d)[PW]) {
#pragma unroll
        for (int r = 0; r < PW; ++r) l[r] = (row && t < k0) ? J[(size_t)(k0 + r) * NP + t] : 0.0;
        const int ph = diag ? L.ip[k0 + lane] : 0;
#pragma unroll
        for (int r = 0; r < PW; ++r) d[r] = (diag && r > lane) ? J[(size_t)(k0 + r) * NP + ph] : 0.0;
    };
    // forward: y = L^-1 P b, panel by panel in logical order
    for (int k0 = 0; k0 < NP; k0 += PW) {
        load_fwd(k0, lr, dq);
        if (t >= k0 && t < k0 + PW) S.v0[st.pos - k0] = b;
        __syncthreads();
        if (diag) {
            double y = S.v0[lane];
#pragma unroll
            for (int r = 0; r < PW; ++r) {
                const double yr = __shfl(y, r, 16);
                if (lane > r) y = y - dq[r] * yr;
            }
            S.v1[lane] = y;
        }
        __syncthreads();
        if (row && t >= k0 + PW) {
            double acc = b;
#pragma unroll
            for (int r = 0; r < PW; ++r) acc = acc - lr[r] * S.v1[r];
            b = acc;
        } else if (t >= k0 && t < k0 + PW) {
            b = S.v1[st.pos - k0];
        }
    }
    // backward: x = U^-1 y, panels in descending logical order
    double x = 0.0;
    for (int k0 = NP - PW; k0 >= 0; k0 -= PW) {
        load_bwd(k0, lr, dq);
        const double id = diag ? invDg[L.ip[k0 + lane]] : 0.0;
        __syncthreads();
        if (t >= k0 && t < k0 + PW) S.v0[st.pos - k0] = b;
        __syncthreads();
        if (diag) {
            double v = S.v0[lane];
#pragma unroll
            for (int r = PW - 1; r >= 0; --r) {
                if (lane == r) v = v * id;
                const double xr = __shfl(v, r, 16);
                if (lane < r) v = v - dq[r] * xr;
            }
            S.v1[lane] = v;
        }
        __syncthreads();
        if (row && t < k0) {
            double acc = b;
#pragma unroll
            for (int r = PW - 1; r >= 0; --r) acc = acc - lr[r] * S.v1[r];
            b = acc;
        }
        if (t >= k0 && t < k0 + PW) x = S.v1[t - k0];
    }
    __syncthreads();
    return x;
}

}  // namespace fcs
