// Host build of the library's device code for the tests (tests/emu): the
// HIP names the device sources use, mapped onto a CPU emulation of one
// 64-lane wavefront in which each lane is a thread (emu_wave.cpp).  Every
// cross-lane operation (readlane, readfirstlane, DPP, shuffle, ballot) and
// every barrier is a rendezvous of all 64 lanes that also records where
// each lane called it from; lanes arriving from different call sites, with
// different lane arguments, or not arriving at all are reported as
// non-uniform control flow.  Only the wavefront form (one 64-lane group per
// cell) is built this way.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__
#define __constant__

struct double2 {
    double x, y;
};
inline double2 make_double2(double x, double y) { return double2{x, y}; }
using std::isinf;
using std::isnan;
using std::isfinite;

struct emu_dim3 {
    unsigned x, y, z;
};

namespace emu {
extern thread_local int t_lane;   // lane of the calling thread
extern unsigned g_grid;           // workgroups of the emulated launch
extern thread_local unsigned t_block;
int readlane(int v, int lane, const char* file, int line);
int readfirstlane(int v, const char* file, int line);
int update_dpp(int src, int ctrl, const char* file, int line);
unsigned long long ballot(bool p, const char* file, int line);
void sync(const char* file, int line);
long long shfl64(long long v, int src, int width, const char* file, int line);
int shfl32(int v, int src, int width, const char* file, int line);
template <class T>
inline T shfl(T v, int src, int width, const char* file, int line)
{
    static_assert(sizeof(T) == 8 || sizeof(T) == 4, "shuffle of a 4- or 8-byte value");
    if constexpr (sizeof(T) == 8) {
        long long b;
        std::memcpy(&b, &v, 8);
        b = shfl64(b, src, width, file, line);
        T r;
        std::memcpy(&r, &b, 8);
        return r;
    } else {
        int b;
        std::memcpy(&b, &v, 4);
        b = shfl32(b, src, width, file, line);
        T r;
        std::memcpy(&r, &b, 4);
        return r;
    }
}
}  // namespace emu

#define threadIdx (emu_dim3{(unsigned)emu::t_lane, 0u, 0u})
#define blockIdx (emu_dim3{emu::t_block, 0u, 0u})
#define gridDim (emu_dim3{emu::g_grid, 1u, 1u})
#define blockDim (emu_dim3{64u, 1u, 1u})
#define __lane_id() (emu::t_lane)
#define __builtin_amdgcn_readlane(v, l) emu::readlane((v), (l), __FILE__, __LINE__)
#define __builtin_amdgcn_readfirstlane(v) emu::readfirstlane((v), __FILE__, __LINE__)
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) emu::update_dpp((src), (ctrl), __FILE__, __LINE__)
#define __builtin_amdgcn_s_memtime() 0ull
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) (c)   // workgroup form only: not built here
#define __syncthreads() emu::sync(__FILE__, __LINE__)
#define __ballot(p) emu::ballot((p), __FILE__, __LINE__)
#define __shfl(v, s, w) emu::shfl((v), (s), (w), __FILE__, __LINE__)
#define __ffsll(x) __builtin_ffsll(x)

inline long long __double_as_longlong(double d)
{
    long long r;
    std::memcpy(&r, &d, 8);
    return r;
}
inline double __longlong_as_double(long long b)
{
    double r;
    std::memcpy(&r, &b, 8);
    return r;
}
template <class T>
inline T atomicAdd(T* p, T v)
{
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v)
{
    unsigned long long cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
