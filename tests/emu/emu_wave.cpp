// CPU emulation of the wavefront-form step kernel for the tests: one
// 64-lane wavefront, one thread per lane, running the library's own
// k_step<NP, Rodas34> source (fcs_kernels.h) over a list of cells until
// every cell has finished, with the mechanism tables built by the library's
// own fcs_mech_tables.h.  Every cross-lane operation is a rendezvous of the
// 64 lane threads (include/hip/hip_runtime.h); at each one lane 0 checks
// that all lanes came from the same call site with the same lane argument,
// and a rendezvous that some lane never reaches ends the run.  The counts
// of those checks and of their failures are returned with the results.
#include <condition_variable>
#include <chrono>
#include <cstdio>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fcs.h"
#include "fcs_kernels.h"
#include "fcs_mech_tables.h"

double lds_raw[1 << 16];   // the workgroup's LDS (one workgroup runs at a time)

namespace emu {
thread_local int t_lane = 0;
thread_local unsigned t_block = 0;
unsigned g_grid = 1;

struct Abort : std::exception {};

struct Wave {
    std::mutex m;
    std::condition_variable cv;
    int arrived = 0;
    unsigned long long gen = 0;
    bool abort = false;
    long long val[64];
    long long arg[64];
    const char* file[64];
    int line[64];
    long long checks = 0, violations = 0;
    std::string first;
};
static Wave* g_wave = nullptr;
static double g_timeout_s = 60.0;

// Generation barrier of the 64 lane threads: waiting lanes yield the CPU
// until the last one arrives (a futex round per barrier would dominate the
// run time); a barrier not completed within g_timeout_s ends the run.
static void barrier(Wave& w, const char* file, int line)
{
    if (__atomic_load_n(&w.abort, __ATOMIC_ACQUIRE)) throw Abort();
    const unsigned long long g = __atomic_load_n(&w.gen, __ATOMIC_ACQUIRE);
    if (__atomic_add_fetch(&w.arrived, 1, __ATOMIC_ACQ_REL) == 64) {
        __atomic_store_n(&w.arrived, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&w.gen, g + 1, __ATOMIC_RELEASE);
        return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; ++spin) {
        if (__atomic_load_n(&w.gen, __ATOMIC_ACQUIRE) != g) return;
        if (__atomic_load_n(&w.abort, __ATOMIC_ACQUIRE)) throw Abort();
        std::this_thread::yield();
        if ((spin & 1023) == 1023 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > g_timeout_s) {
            std::lock_guard<std::mutex> lk(w.m);
            if (!w.abort) {
                ++w.violations;
                if (w.first.empty())
                    w.first = std::string("rendezvous not reached by every lane (") + file + ":" + std::to_string(line) + ")";
                __atomic_store_n(&w.abort, true, __ATOMIC_RELEASE);
            }
            throw Abort();
        }
    }
}

// All 64 lanes deposit (value, lane argument, call site); lane 0 checks that
// the sites and arguments agree; every lane then reads what it needs.
template <class F>
static long long collective(long long v, long long a, bool argUniform, const char* file, int line, F&& pick)
{
    Wave& w = *g_wave;
    const int l = t_lane;
    w.val[l] = v;
    w.arg[l] = a;
    w.file[l] = file;
    w.line[l] = line;
    barrier(w, file, line);
    bool same = true;
    for (int k = 1; k < 64; ++k)
        if (w.line[k] != w.line[0] || std::strcmp(w.file[k], w.file[0]) != 0 || (argUniform && w.arg[k] != w.arg[0]))
            same = false;
    if (l == 0) {
        std::lock_guard<std::mutex> lk(w.m);
        ++w.checks;
        if (!same) {
            ++w.violations;
            int k = 1;
            while (k < 64 && w.line[k] == w.line[0] && std::strcmp(w.file[k], w.file[0]) == 0) ++k;
            if (w.first.empty())
                w.first = std::string("lanes at different call sites or with different lane arguments: lane 0 at ") +
                          w.file[0] + ":" + std::to_string(w.line[0]) + (k < 64 ? std::string(", lane ") +
                          std::to_string(k) + " at " + w.file[k] + ":" + std::to_string(w.line[k]) : std::string());
        }
    }
    if (!same) {   // the lanes' values are not one operation's: stop the run
        barrier(w, file, line);
        __atomic_store_n(&w.abort, true, __ATOMIC_RELEASE);
        throw Abort();
    }
    const long long r = pick(w.val, l);
    barrier(w, file, line);
    return r;
}

int readlane(int v, int lane, const char* file, int line)
{
    return (int)collective(v, lane, true, file, line, [&](const long long* vals, int) { return vals[lane & 63]; });
}
int readfirstlane(int v, const char* file, int line)
{
    return (int)collective(v, 0, true, file, line, [](const long long* vals, int) { return vals[0]; });
}
int update_dpp(int src, int ctrl, const char* file, int line)
{
    return (int)collective(src, ctrl, true, file, line, [&](const long long* vals, int l) {
        int from = l;
        if (ctrl < 0x100) {   // quad_perm: lane 4q + i reads lane 4q + sel_i
            from = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
        } else if (ctrl == 0x140) {   // row_mirror: within 16 lanes
            from = (l & ~15) | (15 - (l & 15));
        } else if (ctrl == 0x141) {   // row_half_mirror: within 8 lanes
            from = (l & ~7) | (7 - (l & 7));
        }
        return vals[from];
    });
}
unsigned long long ballot(bool p, const char* file, int line)
{
    return (unsigned long long)collective(p ? 1 : 0, 0, true, file, line, [](const long long* vals, int) {
        unsigned long long b = 0;
        for (int k = 0; k < 64; ++k)
            if (vals[k]) b |= 1ull << k;
        return (long long)b;
    });
}
void sync(const char* file, int line)
{
    (void)collective(0, 0, true, file, line, [](const long long*, int) { return 0ll; });
}
long long shfl64(long long v, int src, int width, const char* file, int line)
{
    return collective(v, width, true, file, line, [&](const long long* vals, int l) {
        return vals[(l & ~(width - 1)) + (src & (width - 1))];
    });
}
int shfl32(int v, int src, int width, const char* file, int line)
{
    return (int)shfl64(v, src, width, file, line);
}
}  // namespace emu

using namespace fcs;

// Runs body() on the 64 lane threads of one workgroup (block b of g).
template <class F>
static bool run_wave(emu::Wave& w, unsigned b, unsigned g, F&& body)
{
    std::vector<std::thread> th;
    std::exception_ptr err = nullptr;
    std::mutex em;
    emu::g_grid = g;
    for (int l = 0; l < 64; ++l)
        th.emplace_back([&, l] {
            emu::t_lane = l;
            emu::t_block = b;
            try {
                body();
                emu::sync(__FILE__, __LINE__);   // every lane leaves the kernel together
            } catch (const emu::Abort&) {
            } catch (...) {
                std::lock_guard<std::mutex> lk(em);
                err = std::current_exception();
            }
        });
    for (auto& t : th) t.join();
    return !w.abort && !err;
}

template <int NP>
static bool run_launch(int solver, const DevMech* M, DevOpts o, CellArrays a, int* lin, unsigned* cin, int* lout,
                       unsigned* cout, unsigned* czero, emu::Wave& w)
{
    return run_wave(w, 0, 1, [&] {
        switch (solver) {
        case SOLVER_RODAS34: k_step<NP, SOLVER_RODAS34>(M, o, a, lin, cin, lout, cout, czero); break;
        case SOLVER_ROS34: k_step<NP, SOLVER_ROS34>(M, o, a, lin, cin, lout, cout, czero); break;
        case SOLVER_RODAS23: k_step<NP, SOLVER_RODAS23>(M, o, a, lin, cin, lout, cout, czero); break;
        case SOLVER_ROS23: k_step<NP, SOLVER_ROS23>(M, o, a, lin, cin, lout, cout, czero); break;
        default: k_step_seulex<NP>(M, o, a, lin, cin, lout, cout, czero); break;
        }
    });
}

static bool run_verify(int NP, const DevMech* M, DevOpts o, CellArrays a, const int* list, const int* listG,
                       const unsigned* countG, const float* P, int seed, int* listOut, unsigned* countOut,
                       unsigned long long* nAccepted, emu::Wave& w)
{
    switch (NP) {
    case 12: return run_wave(w, 0, 1, [&] { k_verify<12>(M, o, a, list, listG, countG, P, seed, listOut, countOut, nAccepted); });
    case 24: return run_wave(w, 0, 1, [&] { k_verify<24>(M, o, a, list, listG, countG, P, seed, listOut, countOut, nAccepted); });
    case 40: return run_wave(w, 0, 1, [&] { k_verify<40>(M, o, a, list, listG, countG, P, seed, listOut, countOut, nAccepted); });
    case 56: return run_wave(w, 0, 1, [&] { k_verify<56>(M, o, a, list, listG, countG, P, seed, listOut, countOut, nAccepted); });
    default: return run_wave(w, 0, 1, [&] { k_verify<64>(M, o, a, list, listG, countG, P, seed, listOut, countOut, nAccepted); });
    }
}

static int pick_np(int ns)
{
    for (int n : {12, 24, 40, 56, 64})
        if (n >= ns + 1) return n;
    return 0;
}

// The library's table build alone, for a padded size NP of either form
// (host images only): its return code.
extern "C" int emu_tables(const fcs_mech_desc* d, int NP)
{
    std::vector<unsigned char> image;
    DevMech dm{};
    std::vector<int> perm;
    const int SW = NP > 64 ? (NP <= 128 ? 128 : 256) : 64;
    return build_mech_tables(d, NP, SW, [&](const std::vector<unsigned char>& host) -> void* {
        image = host;
        return image.data();
    }, dm, perm);
}

// RHS of nS states (YT [ns+1][nS], clipped in place) through k_derivatives.
extern "C" int emu_derivatives(const fcs_mech_desc* d, int nS, const double* p, double* YT, double* out,
                               long long* checks, long long* violations, char* msg, int msgLen)
{
    const int NP = pick_np(d->n_species);
    if (!NP) return FCS_E_UNSUPPORTED;
    std::vector<unsigned char> image;
    DevMech dm{};
    std::vector<int> perm;
    const int rc = build_mech_tables(d, NP, 64, [&](const std::vector<unsigned char>& host) -> void* {
        image = host;
        return image.data();
    }, dm, perm);
    if (rc != FCS_OK) return rc;
    emu::Wave w;
    emu::g_wave = &w;
    bool ok = true;
    for (int sidx = 0; sidx < nS && ok; ++sidx) {
        switch (NP) {
        case 12: ok = run_wave(w, sidx, nS, [&] { k_derivatives<12>(&dm, nS, p, YT, out); }); break;
        case 24: ok = run_wave(w, sidx, nS, [&] { k_derivatives<24>(&dm, nS, p, YT, out); }); break;
        case 40: ok = run_wave(w, sidx, nS, [&] { k_derivatives<40>(&dm, nS, p, YT, out); }); break;
        case 56: ok = run_wave(w, sidx, nS, [&] { k_derivatives<56>(&dm, nS, p, YT, out); }); break;
        default: ok = run_wave(w, sidx, nS, [&] { k_derivatives<64>(&dm, nS, p, YT, out); }); break;
        }
    }
    *checks = w.checks;
    *violations = w.violations;
    if (msg && msgLen > 0) std::snprintf(msg, (size_t)msgLen, "%s", w.first.c_str());
    return ok ? FCS_OK : FCS_E_CELLS;
}

// Exact Jacobian of nS states through k_jacobian (f [n][nS], J [nS][n][n]).
extern "C" int emu_jacobian(const fcs_mech_desc* d, int jacType, int nS, const double* p, double* YT, double* f,
                            double* J, long long* checks, long long* violations, char* msg, int msgLen)
{
    const int NP = pick_np(d->n_species);
    if (!NP) return FCS_E_UNSUPPORTED;
    std::vector<unsigned char> image;
    DevMech dm{};
    std::vector<int> perm;
    const int rc = build_mech_tables(d, NP, 64, [&](const std::vector<unsigned char>& host) -> void* {
        image = host;
        return image.data();
    }, dm, perm);
    if (rc != FCS_OK) return rc;
    emu::Wave w;
    emu::g_wave = &w;
    bool ok = true;
    for (int sidx = 0; sidx < nS && ok; ++sidx) {
        switch (NP) {
        case 12: ok = run_wave(w, sidx, nS, [&] { k_jacobian<12>(&dm, jacType, nS, p, YT, f, J, nullptr); }); break;
        case 24: ok = run_wave(w, sidx, nS, [&] { k_jacobian<24>(&dm, jacType, nS, p, YT, f, J, nullptr); }); break;
        case 40: ok = run_wave(w, sidx, nS, [&] { k_jacobian<40>(&dm, jacType, nS, p, YT, f, J, nullptr); }); break;
        case 56: ok = run_wave(w, sidx, nS, [&] { k_jacobian<56>(&dm, jacType, nS, p, YT, f, J, nullptr); }); break;
        default: ok = run_wave(w, sidx, nS, [&] { k_jacobian<64>(&dm, jacType, nS, p, YT, f, J, nullptr); }); break;
        }
    }
    *checks = w.checks;
    *violations = w.violations;
    if (msg && msgLen > 0) std::snprintf(msg, (size_t)msgLen, "%s", w.first.c_str());
    return ok ? FCS_OK : FCS_E_CELLS;
}

// One Rodas34 attempt of step dx[s] from each state (k_attempt): YTnew, err.
extern "C" int emu_attempt(const fcs_mech_desc* d, double absTol, double relTol, int nS, const double* p,
                           const double* dx, double* YT, double* YTnew, double* err, long long* checks,
                           long long* violations, char* msg, int msgLen)
{
    const int NP = pick_np(d->n_species);
    if (!NP) return FCS_E_UNSUPPORTED;
    std::vector<unsigned char> image;
    DevMech dm{};
    std::vector<int> perm;
    const int rc = build_mech_tables(d, NP, 64, [&](const std::vector<unsigned char>& host) -> void* {
        image = host;
        return image.data();
    }, dm, perm);
    if (rc != FCS_OK) return rc;
    DevOpts o{SOLVER_RODAS34, 0, 10000, absTol, relTol, 0.0, 1.0e15};
    emu::Wave w;
    emu::g_wave = &w;
    bool ok = true;
    for (int sidx = 0; sidx < nS && ok; ++sidx) {
        switch (NP) {
        case 12: ok = run_wave(w, sidx, nS, [&] { k_attempt<12, SOLVER_RODAS34>(&dm, o, nS, p, dx, YT, YTnew, err, nullptr); }); break;
        case 24: ok = run_wave(w, sidx, nS, [&] { k_attempt<24, SOLVER_RODAS34>(&dm, o, nS, p, dx, YT, YTnew, err, nullptr); }); break;
        case 40: ok = run_wave(w, sidx, nS, [&] { k_attempt<40, SOLVER_RODAS34>(&dm, o, nS, p, dx, YT, YTnew, err, nullptr); }); break;
        case 56: ok = run_wave(w, sidx, nS, [&] { k_attempt<56, SOLVER_RODAS34>(&dm, o, nS, p, dx, YT, YTnew, err, nullptr); }); break;
        default: ok = run_wave(w, sidx, nS, [&] { k_attempt<64, SOLVER_RODAS34>(&dm, o, nS, p, dx, YT, YTnew, err, nullptr); }); break;
        }
    }
    *checks = w.checks;
    *violations = w.violations;
    if (msg && msgLen > 0) std::snprintf(msg, (size_t)msgLen, "%s", w.first.c_str());
    return ok ? FCS_OK : FCS_E_CELLS;
}

// Checks of the checker: kind 0 every lane reads lane 3 at one call site (no
// report); kind 1 half the lanes read at another call site; kind 2 the lane
// argument differs between lanes; kind 3 the last lane skips the read.
extern "C" int emu_selftest(int kind, double timeoutSeconds, long long* checks, long long* violations, char* msg,
                            int msgLen)
{
    emu::Wave w;
    emu::g_wave = &w;
    emu::g_timeout_s = timeoutSeconds;
    std::vector<int> got(64, -1);
    const bool ok = run_wave(w, 0, 1, [&] {
        const int l = emu::t_lane;
        int v = -1;
        if (kind == 0) {
            v = __builtin_amdgcn_readlane(100 + l, 3);
        } else if (kind == 1) {
            if (l < 32) v = __builtin_amdgcn_readlane(100 + l, 3);
            else v = __builtin_amdgcn_readlane(100 + l, 3);
        } else if (kind == 2) {
            v = __builtin_amdgcn_readlane(100 + l, l & 1);
        } else if (l < 63) {
            v = __builtin_amdgcn_readlane(100 + l, 3);
        }
        got[l] = v;
    });
    *checks = w.checks;
    *violations = w.violations;
    if (msg && msgLen > 0) std::snprintf(msg, (size_t)msgLen, "%s", w.first.c_str());
    if (ok && kind == 0)
        for (int l = 0; l < 64; ++l)
            if (got[l] != 103) return FCS_E_ARG;   // the emulated readlane itself is wrong
    return ok ? FCS_OK : FCS_E_CELLS;
}

extern "C" int emu_solve(const fcs_mech_desc* d, int solver, double absTol, double relTol, int maxSteps, double Treact,
                         double dtChemMax, int nc,
                         double dt, const double* T0, const double* p0, const double* rho0, const double* Y0,
                         const double* dtcIn, double* RR, double* Tend, double* dtcOut, int* status, int* nAtt,
                         long long* checks, long long* violations, char* msg, int msgLen, double timeoutSeconds,
                         const float* P, int seed, unsigned long long* nAccepted)
{
    const int ns = d->n_species;
    const int NP = pick_np(ns);
    if (!NP) return FCS_E_UNSUPPORTED;
    std::vector<unsigned char> image;
    DevMech dm{};
    std::vector<int> perm;
    auto place = [&](const std::vector<unsigned char>& host) -> void* {
        image = host;
        return image.data();
    };
    const int rc = build_mech_tables(d, NP, 64, place, dm, perm);
    if (rc != FCS_OK) return rc;
    if (solver < SOLVER_RODAS34 || solver > SOLVER_ROS23) return FCS_E_ARG;
    DevOpts o{solver, 0, maxSteps, absTol, relTol, Treact, dtChemMax};
    const size_t n = (size_t)nc;
    std::vector<double> Yw((ns + 1) * n), x(n), dxTry(n), dx(n), dxTry0(n), dxFirst(n);
    std::vector<int> flags(n), nStep(n), list0(n), list1(n), kTarg(n);
    // Seulex: step-start state and scale vector [ns+1][nc], theta, kTarg, the
    // per-cell Jacobian (NP columns of 64)
    std::vector<double> Y0s(solver == SOLVER_SEULEX ? (ns + 1) * n : 1), Scl(solver == SOLVER_SEULEX ? (ns + 1) * n : 1),
        theta(n), Jy(solver == SOLVER_SEULEX ? n * (size_t)NP * 64 : 1);
    unsigned long long dtMin = 0x7ff0000000000000ull;
    CellArrays a{};
    a.nc = nc;
    a.dt = dt;
    a.T0 = T0;
    a.p0 = p0;
    a.rho0 = rho0;
    a.Y0 = Y0;
    a.dtcIn = dtcIn;
    a.dtcOut = dtcOut;
    a.RR = RR;
    a.Tend = Tend;
    a.status = status;
    a.nAtt = nAtt;
    a.dtMinBits = &dtMin;
    a.Yw = Yw.data();
    a.x = x.data();
    a.dxTry = dxTry.data();
    a.dx = dx.data();
    a.dxTry0 = dxTry0.data();
    a.flags = flags.data();
    a.nStep = nStep.data();
    a.dxFirst = dxFirst.data();
    a.Y0s = Y0s.data();
    a.Scl = Scl.data();
    a.theta = theta.data();
    a.Jy = Jy.data();
    a.kTarg = kTarg.data();
    // initial state as the library's set-up kernel makes it
    unsigned counts[3] = {0, 0, 0};
    for (int c = 0; c < nc; ++c) {
        status[c] = 0;
        if (!(T0[c] > Treact)) {
            for (int i = 0; i < ns; ++i) RR[i * n + c] = 0.0;
            status[c] = CELL_SKIPPED;
            dtcOut[c] = dtcIn[c];
            Tend[c] = T0[c];
            nAtt[c] = 0;
            continue;
        }
        for (int i = 0; i < ns; ++i) Yw[i * n + c] = Y0[i * n + c] > 0.0 ? Y0[i * n + c] : 0.0;
        Yw[ns * n + c] = T0[c];
        x[c] = 0.0;
        dxTry[c] = dtcIn[c];
        dx[c] = 0.0;
        dxTry0[c] = 0.0;
        flags[c] = solver == SOLVER_SEULEX ? SX_FIRST : 0;
        dxFirst[c] = 0.0;
        theta[c] = 0.0;
        kTarg[c] = 0;
        nStep[c] = 0;
        nAtt[c] = 0;
        list0[counts[0]++] = c;
    }
    emu::Wave w;
    emu::g_wave = &w;
    emu::g_timeout_s = timeoutSeconds;
    emu::g_grid = 1;
    const long long maxLaunches = 101LL * maxSteps + 2;
    bool ok = true;
    long long it0 = 0;
    unsigned long long acc[2] = {0, 0};
    if (P && counts[0] > 0) {
        // learned seed as the library runs it: the gate (proposals with h >=
        // deltaT/2 to the gated list, the others, seeded when asked, straight
        // to list 1), then k_verify's one Rodas34 attempt per gated cell
        std::vector<int> listG(n);
        unsigned countG = 0;
        const unsigned n0 = counts[0];
        for (unsigned i = 0; i < n0; ++i) {
            const int c = list0[i];
            const double h = std::exp((double)P[(size_t)i * (ns + 2)]);
            if (h >= 0.5 * dt) {
                listG[countG++] = (int)i;
                ++acc[1];
                continue;
            }
            if (seed && h == h && h > 0.0) dxTry[c] = h < dt ? h : dt;
            list1[counts[1]++] = c;
        }
        ok = run_verify(NP, &dm, o, a, list0.data(), listG.data(), &countG, P, seed, list1.data(), counts + 1, acc, w);
        it0 = 1;
    }
    for (long long it = it0; it < maxLaunches && ok; ++it) {
        int* lin = (it & 1) ? list1.data() : list0.data();
        int* lout = (it & 1) ? list0.data() : list1.data();
        unsigned* cin = counts + it % 3;
        unsigned* cout = counts + (it + 1) % 3;
        unsigned* czero = counts + (it + 2) % 3;
        if (*cin == 0) break;
        switch (NP) {
        case 12: ok = run_launch<12>(solver, &dm, o, a, lin, cin, lout, cout, czero, w); break;
        case 24: ok = run_launch<24>(solver, &dm, o, a, lin, cin, lout, cout, czero, w); break;
        case 40: ok = run_launch<40>(solver, &dm, o, a, lin, cin, lout, cout, czero, w); break;
        case 56: ok = run_launch<56>(solver, &dm, o, a, lin, cin, lout, cout, czero, w); break;
        default: ok = run_launch<64>(solver, &dm, o, a, lin, cin, lout, cout, czero, w); break;
        }
    }
    *checks = w.checks;
    *violations = w.violations;
    if (nAccepted) {
        nAccepted[0] = acc[0];
        nAccepted[1] = acc[1];
    }
    if (msg && msgLen > 0) std::snprintf(msg, (size_t)msgLen, "%s", w.first.c_str());
    return ok ? FCS_OK : FCS_E_CELLS;
}
