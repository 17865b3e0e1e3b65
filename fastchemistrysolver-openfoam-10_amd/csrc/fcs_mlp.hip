// fcs_mlp.hip -- the learned step/state seed: a small fully connected
// network (SiLU hidden layers) evaluated and trained on gfx950 with BF16
// MFMA tiles and FP32 accumulation.  FP32 master weights, Adam.
//
// Every matrix product of the forward and backward passes is one call of
// k_gemm: C = A * B (+ beta C) over strided FP32 operands that are rounded
// to BF16 as they are loaded into the MFMA operand registers.  One 64-lane
// wavefront computes one 16 x 16 tile of C with
// v_mfma_f32_16x16x16_bf16 (operand layout: lane l holds A[l%16][4(l/16)+i]
// and B[4(l/16)+i][l%16], i < 4; D[4(l/16)+i][l%16]).
//
// This part has no counterpart in the reference; it is the learned seed of
// the integrator (see fcs_learned_* in include/fcs.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "fcs.h"
#include "fcs_mlp_internal.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round-to-nearest-even (NaN kept quiet).
__device__ __forceinline__ short to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[m][n] = sum_k A(m,k) B(k,n) (+ beta * C[m][n]); A(m,k) = A[m*sAm + k*sAk],
// B(k,n) = B[k*sBk + n*sBn], C row-major with leading dimension ldc.
// M, N, K are multiples of 16 (operands are zero-padded).  Grid: one
// wavefront per 16 x 16 tile, blockIdx.x over N tiles, blockIdx.y over M.
__global__ __launch_bounds__(64) void k_gemm(int K, const float* __restrict__ A, long sAm, long sAk,
                                             const float* __restrict__ B, long sBk, long sBn, float* __restrict__ C,
                                             long ldc, float beta)
{
    const int lane = threadIdx.x;
    const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    const int r = lane & 15, q = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + 4 * q + i;
            a[i] = to_bf16(A[(long)(m0 + r) * sAm + (long)k * sAk]);
            b[i] = to_bf16(B[(long)k * sBk + (long)(n0 + r) * sBn]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float* c = C + (long)(m0 + 4 * q + i) * ldc + n0 + r;
        *c = beta == 0.f ? acc[i] : acc[i] + beta * *c;
    }
}

__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + expf(-z)); }

// Z[m][n] += bias[n]; H = SiLU(Z) (hidden layers) or Z (output layer).
// Padding columns (n >= nValid) are held at zero.
__global__ void k_bias_act(int M, int N, int nValid, float* Z, float* H, const float* bias, int silu)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * N) return;
    const int n = (int)(i % N);
    const float z = n < nValid ? Z[i] + bias[n] : 0.f;
    Z[i] = z;
    H[i] = silu ? z * sigm(z) : z;
}

// dZ = dH * SiLU'(Z) for hidden layers.
__global__ void k_silu_back(long MN, const float* Z, float* dH)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MN) return;
    const float z = Z[i], s = sigm(z);
    dH[i] *= s * (1.f + z * (1.f - s));
}

// Mean-squared error over the valid block: dY = 2 (Y - T) / (rows * nValid),
// per-row squared errors into err[m] (summed on the host).
__global__ void k_mse(int M, int N, int rows, int nValid, const float* Y, const float* T, float* dY, float* err)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float e = 0.f;
    for (int n = 0; n < N; ++n) {
        const long i = (long)m * N + n;
        const float d = (m < rows && n < nValid) ? Y[i] - T[i] : 0.f;
        dY[i] = 2.f * d / ((float)rows * (float)nValid);
        e += d * d;
    }
    err[m] = e;
}

// db[n] = sum_m dZ[m][n]
__global__ void k_colsum(int M, int N, const float* dZ, float* db)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += dZ[(long)m * N + n];
    db[n] = s;
}

__global__ void k_adam(long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows of a device data set gathered into a padded batch (zero rows beyond n)
__global__ void k_gather(int rows, int Mpad, int cols, int ld, const float* src, const int* idx, float* dst)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Mpad * ld) return;
    const int m = (int)(i / ld), c = (int)(i % ld);
    dst[i] = (m < rows && c < cols) ? src[(long)idx[m] * cols + c] : 0.f;
}

// (x - mean) / std into a padded matrix, or its inverse (y * std + mean)
__global__ void k_normalize(int rows, int cols, int ld, const float* src, long sld, const float* mean, const float* sd,
                            float* dst, int inverse)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * ld) return;
    const int m = (int)(i / ld), c = (int)(i % ld);
    float v = 0.f;
    if (c < cols) {
        const float x = src[(long)m * sld + c];
        v = inverse ? x * sd[c] + mean[c] : (x - mean[c]) / sd[c];
    }
    dst[i] = v;
}

inline int pad16(int x) { return (x + 15) / 16 * 16; }
inline unsigned blocks(long n, int t) { return (unsigned)((n + t - 1) / t); }

#define MLP_HIP(x)                                 \
    do {                                           \
        if ((x) != hipSuccess) return FCS_E_HIP;   \
    } while (0)

}  // namespace

// Layer l maps width[l] -> width[l+1]; padded widths Wp[l] = pad16(width[l]).
// Parameters (FP32, device): for each layer W [Wp[l]][Wp[l+1]] row-major
// (input-major) then b [Wp[l+1]], padding entries zero and never updated.
struct fcs_mlp {
    int nL = 0;
    std::vector<int> w, wp;
    std::vector<size_t> offW, offB;
    size_t nParam = 0;
    float *param = nullptr, *grad = nullptr, *adm = nullptr, *adv = nullptr;
    float *inMean = nullptr, *inStd = nullptr, *outMean = nullptr, *outStd = nullptr;   // device, padded
    long adamStep = 0;
    hipStream_t stream = nullptr;
    // work buffers for a batch of up to cap rows: Z and H per layer, dH
    int cap = 0;
    std::vector<float*> Z, H;
    float *dA = nullptr, *dB = nullptr, *err = nullptr, *tgt = nullptr;

    void free_work()
    {
        for (auto p : Z) (void)hipFree(p);
        for (auto p : H) (void)hipFree(p);
        Z.clear();
        H.clear();
        for (float** p : {&dA, &dB, &err, &tgt})
            if (*p) (void)hipFree(*p), *p = nullptr;
        cap = 0;
    }
    int ensure(int rows)
    {
        const int M = pad16(std::max(rows, 1));
        if (M <= cap) return FCS_OK;
        free_work();
        int maxw = 0;
        for (int v : wp) maxw = std::max(maxw, v);
        for (int l = 0; l <= nL; ++l) {
            float *z = nullptr, *h = nullptr;
            if (hipMalloc(&z, (size_t)M * wp[l] * 4) != hipSuccess || hipMalloc(&h, (size_t)M * wp[l] * 4) != hipSuccess)
                return FCS_E_NOMEM;
            Z.push_back(z);
            H.push_back(h);
        }
        if (hipMalloc(&dA, (size_t)M * maxw * 4) != hipSuccess || hipMalloc(&dB, (size_t)M * maxw * 4) != hipSuccess ||
            hipMalloc(&err, (size_t)M * 4) != hipSuccess || hipMalloc(&tgt, (size_t)M * wp[nL] * 4) != hipSuccess)
            return FCS_E_NOMEM;
        cap = M;
        return FCS_OK;
    }
    // forward of M padded rows already normalised in H[0]
    int forward(int M) { return forward(M, stream); }
    int forward(int M, hipStream_t st)
    {
        for (int l = 0; l < nL; ++l) {
            const int K = wp[l], N = wp[l + 1];
            hipLaunchKernelGGL(k_gemm, dim3(N / 16, M / 16), dim3(64), 0, st, K, H[l], (long)K, 1L,
                               param + offW[l], (long)N, 1L, Z[l + 1], (long)N, 0.f);
            hipLaunchKernelGGL(k_bias_act, dim3(blocks((long)M * N, 256)), dim3(256), 0, st, M, N, w[l + 1],
                               Z[l + 1], H[l + 1], param + offB[l], l + 1 < nL ? 1 : 0);
        }
        return hipGetLastError() == hipSuccess ? FCS_OK : FCS_E_HIP;
    }
    // backward from dY (in dA) of M padded rows; gradients into grad
    int backward(int M)
    {
        float* dCur = dA;
        float* dNext = dB;
        for (int l = nL - 1; l >= 0; --l) {
            const int K = wp[l], N = wp[l + 1];
            if (l + 1 < nL)
                hipLaunchKernelGGL(k_silu_back, dim3(blocks((long)M * N, 256)), dim3(256), 0, stream, (long)M * N,
                                   Z[l + 1], dCur);
            // dW[K][N] = H[l]^T [K x M] * dZ [M x N]
            hipLaunchKernelGGL(k_gemm, dim3(N / 16, K / 16), dim3(64), 0, stream, M, H[l], 1L, (long)K, dCur,
                               (long)N, 1L, grad + offW[l], (long)N, 0.f);
            hipLaunchKernelGGL(k_colsum, dim3(blocks(N, 64)), dim3(64), 0, stream, M, N, dCur, grad + offB[l]);
            if (l > 0)   // dH[l] [M x K] = dZ [M x N] * W^T [N x K]
                hipLaunchKernelGGL(k_gemm, dim3(K / 16, M / 16), dim3(64), 0, stream, N, dCur, (long)N, 1L,
                                   param + offW[l], 1L, (long)N, dNext, (long)K, 0.f);
            std::swap(dCur, dNext);
        }
        return hipGetLastError() == hipSuccess ? FCS_OK : FCS_E_HIP;
    }
    // padding of the weight gradients back to zero (never trained)
    void release()
    {
        free_work();
        for (float** p : {&param, &grad, &adm, &adv, &inMean, &inStd, &outMean, &outStd})
            if (*p) (void)hipFree(*p), *p = nullptr;
        if (stream) (void)hipStreamDestroy(stream), stream = nullptr;
    }
};

namespace {

// zero the padded rows/columns of every weight block and bias of a host image
void zero_padding(const fcs_mlp& m, std::vector<float>& p)
{
    for (int l = 0; l < m.nL; ++l) {
        for (int i = 0; i < m.wp[l]; ++i)
            for (int j = 0; j < m.wp[l + 1]; ++j)
                if (i >= m.w[l] || j >= m.w[l + 1]) p[m.offW[l] + (size_t)i * m.wp[l + 1] + j] = 0.f;
        for (int j = m.w[l + 1]; j < m.wp[l + 1]; ++j) p[m.offB[l] + j] = 0.f;
    }
}

// normalised, padded inputs of n host rows into H[0]
int load_inputs(fcs_mlp* m, int64_t n, const float* X, int M)
{
    float* raw = nullptr;
    if (hipMalloc(&raw, std::max<size_t>((size_t)n * m->w[0] * 4, 4)) != hipSuccess) return FCS_E_NOMEM;
    int rc = FCS_OK;
    if (hipMemcpyAsync(raw, X, (size_t)n * m->w[0] * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess) rc = FCS_E_HIP;
    if (rc == FCS_OK) {
        MLP_HIP(hipMemsetAsync(m->H[0], 0, (size_t)M * m->wp[0] * 4, m->stream));
        hipLaunchKernelGGL(k_normalize, dim3(blocks((long)n * m->wp[0], 256)), dim3(256), 0, m->stream, (int)n,
                           m->w[0], m->wp[0], raw, (long)m->w[0], m->inMean, m->inStd, m->H[0], 0);
        if (hipStreamSynchronize(m->stream) != hipSuccess) rc = FCS_E_HIP;
    }
    (void)hipFree(raw);
    return rc;
}

}  // namespace

int fcs_mlp_forward_device(fcs_mlp* m, int rows, int w0, int wL, const float* X, float* Y, hipStream_t st)
{
    if (!m || rows < 1 || !X || !Y) return FCS_E_ARG;
    if (m->w[0] != w0 || m->w[m->nL] != wL) return FCS_E_ARG;
    const int M = pad16(rows), L = m->nL;
    int rc = m->ensure(rows);
    if (rc != FCS_OK) return rc;
    // normalised, zero-padded inputs (a kernel fill, not a memset node)
    hipLaunchKernelGGL(k_normalize, dim3(blocks((long)M * m->wp[0], 256)), dim3(256), 0, st, M, 0, m->wp[0], X,
                       (long)w0, m->inMean, m->inStd, m->H[0], 0);
    hipLaunchKernelGGL(k_normalize, dim3(blocks((long)rows * m->wp[0], 256)), dim3(256), 0, st, rows, w0, m->wp[0], X,
                       (long)w0, m->inMean, m->inStd, m->H[0], 0);
    rc = m->forward(M, st);
    if (rc != FCS_OK) return rc;
    hipLaunchKernelGGL(k_normalize, dim3(blocks((long)rows * wL, 256)), dim3(256), 0, st, rows, wL, wL, m->H[L],
                       (long)m->wp[L], m->outMean, m->outStd, Y, 1);
    return hipGetLastError() == hipSuccess ? FCS_OK : FCS_E_HIP;
}

extern "C" {

int fcs_mlp_create(int32_t nLayers, const int32_t* widths, uint64_t seed, fcs_mlp** out)
{
    if (nLayers < 1 || nLayers > 8 || !widths || !out) return FCS_E_ARG;
    for (int l = 0; l <= nLayers; ++l)
        if (widths[l] < 1 || widths[l] > 4096) return FCS_E_ARG;
    fcs_mlp* m = new (std::nothrow) fcs_mlp();
    if (!m) return FCS_E_NOMEM;
    m->nL = nLayers;
    for (int l = 0; l <= nLayers; ++l) {
        m->w.push_back(widths[l]);
        m->wp.push_back(pad16(widths[l]));
    }
    size_t off = 0;
    for (int l = 0; l < nLayers; ++l) {
        m->offW.push_back(off);
        off += (size_t)m->wp[l] * m->wp[l + 1];
        m->offB.push_back(off);
        off += m->wp[l + 1];
    }
    m->nParam = off;
    // He/Glorot-style uniform initialisation, seeded
    std::vector<float> p(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < nLayers; ++l) {
        const float a = std::sqrt(6.f / (float)(m->w[l] + m->w[l + 1]));
        std::uniform_real_distribution<float> U(-a, a);
        for (int i = 0; i < m->w[l]; ++i)
            for (int j = 0; j < m->w[l + 1]; ++j) p[m->offW[l] + (size_t)i * m->wp[l + 1] + j] = U(rng);
    }
    std::vector<float> zeros0(m->wp[0], 0.f), ones0(m->wp[0], 1.f), zerosL(m->wp[nLayers], 0.f),
        onesL(m->wp[nLayers], 1.f);
    bool ok = hipStreamCreate(&m->stream) == hipSuccess;
    for (float** q : {&m->param, &m->grad, &m->adm, &m->adv})
        ok = ok && hipMalloc(q, off * 4) == hipSuccess;
    ok = ok && hipMalloc(&m->inMean, m->wp[0] * 4) == hipSuccess && hipMalloc(&m->inStd, m->wp[0] * 4) == hipSuccess &&
         hipMalloc(&m->outMean, m->wp[nLayers] * 4) == hipSuccess &&
         hipMalloc(&m->outStd, m->wp[nLayers] * 4) == hipSuccess;
    ok = ok && hipMemcpy(m->param, p.data(), off * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(m->grad, 0, off * 4) == hipSuccess && hipMemset(m->adm, 0, off * 4) == hipSuccess &&
         hipMemset(m->adv, 0, off * 4) == hipSuccess &&
         hipMemcpy(m->inMean, zeros0.data(), m->wp[0] * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m->inStd, ones0.data(), m->wp[0] * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m->outMean, zerosL.data(), m->wp[nLayers] * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m->outStd, onesL.data(), m->wp[nLayers] * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        m->release();
        delete m;
        return FCS_E_HIP;
    }
    *out = m;
    return FCS_OK;
}

int fcs_mlp_destroy(fcs_mlp* m)
{
    if (!m) return FCS_E_ARG;
    m->release();
    delete m;
    return FCS_OK;
}

int64_t fcs_mlp_param_count(const fcs_mlp* m) { return m ? (int64_t)m->nParam : -1; }

int fcs_mlp_get_params(fcs_mlp* m, float* params)
{
    if (!m || !params) return FCS_E_ARG;
    MLP_HIP(hipMemcpy(params, m->param, m->nParam * 4, hipMemcpyDeviceToHost));
    return FCS_OK;
}

int fcs_mlp_set_params(fcs_mlp* m, const float* params)
{
    if (!m || !params) return FCS_E_ARG;
    std::vector<float> p(params, params + m->nParam);
    zero_padding(*m, p);
    MLP_HIP(hipMemcpy(m->param, p.data(), m->nParam * 4, hipMemcpyHostToDevice));
    return FCS_OK;
}

int fcs_mlp_set_norm(fcs_mlp* m, const float* inMean, const float* inStd, const float* outMean, const float* outStd)
{
    if (!m || !inMean || !inStd || !outMean || !outStd) return FCS_E_ARG;
    const int a = m->w[0], b = m->w[m->nL];
    std::vector<float> im(m->wp[0], 0.f), is(m->wp[0], 1.f), om(m->wp[m->nL], 0.f), os(m->wp[m->nL], 1.f);
    for (int i = 0; i < a; ++i) {
        if (!(inStd[i] > 0.f)) return FCS_E_ARG;
        im[i] = inMean[i];
        is[i] = inStd[i];
    }
    for (int i = 0; i < b; ++i) {
        if (!(outStd[i] > 0.f)) return FCS_E_ARG;
        om[i] = outMean[i];
        os[i] = outStd[i];
    }
    MLP_HIP(hipMemcpy(m->inMean, im.data(), im.size() * 4, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(m->inStd, is.data(), is.size() * 4, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(m->outMean, om.data(), om.size() * 4, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(m->outStd, os.data(), os.size() * 4, hipMemcpyHostToDevice));
    return FCS_OK;
}

// Inference: raw inputs X [n][w0] -> de-normalised outputs Y [n][wL].
int fcs_mlp_forward(fcs_mlp* m, int64_t n, const float* X, float* Y)
{
    if (!m || n < 1 || n > (1 << 26) || !X || !Y) return FCS_E_ARG;
    const int M = pad16((int)n);
    int rc = m->ensure((int)n);
    if (rc == FCS_OK) rc = load_inputs(m, n, X, M);
    if (rc == FCS_OK) rc = m->forward(M);
    if (rc != FCS_OK) return rc;
    const int L = m->nL, wl = m->w[L], wpl = m->wp[L];
    hipLaunchKernelGGL(k_normalize, dim3(blocks((long)n * wl, 256)), dim3(256), 0, m->stream, (int)n, wl, wl,
                       m->H[L], (long)wpl, m->outMean, m->outStd, m->dA, 1);
    MLP_HIP(hipMemcpyAsync(Y, m->dA, (size_t)n * wl * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return FCS_OK;
}

// One forward + backward on n rows: mean-squared error of the normalised
// outputs against normalised targets T [n][wL] (raw units), and the
// gradient of that loss with respect to every parameter (flat, padded
// layout of fcs_mlp_get_params).  Used to check the kernels.
int fcs_mlp_loss_grad(fcs_mlp* m, int64_t n, const float* X, const float* T, float* loss, float* grad)
{
    if (!m || n < 1 || n > (1 << 20) || !X || !T || !loss) return FCS_E_ARG;
    const int M = pad16((int)n), L = m->nL, wl = m->w[L], wpl = m->wp[L];
    int rc = m->ensure((int)n);
    if (rc == FCS_OK) rc = load_inputs(m, n, X, M);
    if (rc == FCS_OK) rc = m->forward(M);
    if (rc != FCS_OK) return rc;
    float* raw = nullptr;
    if (hipMalloc(&raw, (size_t)n * wl * 4) != hipSuccess) return FCS_E_NOMEM;
    rc = hipMemcpy(raw, T, (size_t)n * wl * 4, hipMemcpyHostToDevice) == hipSuccess ? FCS_OK : FCS_E_HIP;
    if (rc == FCS_OK) {
        (void)hipMemsetAsync(m->tgt, 0, (size_t)M * wpl * 4, m->stream);
        hipLaunchKernelGGL(k_normalize, dim3(blocks((long)n * wpl, 256)), dim3(256), 0, m->stream, (int)n, wl, wpl,
                           raw, (long)wl, m->outMean, m->outStd, m->tgt, 0);
        hipLaunchKernelGGL(k_mse, dim3(blocks(M, 64)), dim3(64), 0, m->stream, M, wpl, (int)n, wl, m->H[L], m->tgt,
                           m->dA, m->err);
        rc = m->backward(M);
    }
    (void)hipFree(raw);
    if (rc != FCS_OK) return rc;
    std::vector<float> e(M);
    MLP_HIP(hipMemcpyAsync(e.data(), m->err, (size_t)M * 4, hipMemcpyDeviceToHost, m->stream));
    if (grad) MLP_HIP(hipMemcpyAsync(grad, m->grad, m->nParam * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (int i = 0; i < (int)n; ++i) s += e[i];
    *loss = (float)(s / ((double)n * wl));
    return FCS_OK;
}

// Adam training on n host samples (inputs X [n][w0], targets T [n][wL], raw
// units; normalised with the stored statistics) for `epochs` passes in
// shuffled mini-batches.  lossOut[e] receives each epoch's mean loss.
int fcs_mlp_train(fcs_mlp* m, int64_t n, const float* X, const float* T, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, float* lossOut)
{
    if (!m || n < 1 || n > (1 << 26) || !X || !T || epochs < 1 || batch < 16 || batch > 65536 || !(lr > 0))
        return FCS_E_ARG;
    const int L = m->nL, w0 = m->w[0], wl = m->w[L], wp0 = m->wp[0], wpl = m->wp[L];
    const int M = pad16(batch);
    int rc = m->ensure(batch);
    if (rc != FCS_OK) return rc;
    float *dX = nullptr, *dT = nullptr, *nX = nullptr, *nT = nullptr;
    int* dIdx = nullptr;
    auto cleanup = [&]() {
        for (float* p : {dX, dT, nX, nT})
            if (p) (void)hipFree(p);
        if (dIdx) (void)hipFree(dIdx);
    };
    if (hipMalloc(&dX, (size_t)n * w0 * 4) != hipSuccess || hipMalloc(&dT, (size_t)n * wl * 4) != hipSuccess ||
        hipMalloc(&nX, (size_t)n * w0 * 4) != hipSuccess || hipMalloc(&nT, (size_t)n * wl * 4) != hipSuccess ||
        hipMalloc(&dIdx, (size_t)M * 4) != hipSuccess) {
        cleanup();
        return FCS_E_NOMEM;
    }
    hipStream_t st = m->stream;
    bool ok = hipMemcpyAsync(dX, X, (size_t)n * w0 * 4, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(dT, T, (size_t)n * wl * 4, hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok) {   // normalise the whole data set once (unpadded row layout)
        hipLaunchKernelGGL(k_normalize, dim3(blocks((long)n * w0, 256)), dim3(256), 0, st, (int)n, w0, w0, dX,
                           (long)w0, m->inMean, m->inStd, nX, 0);
        hipLaunchKernelGGL(k_normalize, dim3(blocks((long)n * wl, 256)), dim3(256), 0, st, (int)n, wl, wl, dT,
                           (long)wl, m->outMean, m->outStd, nT, 0);
    }
    std::vector<int> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = (int)i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<float> e(M);
    for (int ep = 0; ok && ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        double esum = 0.0;
        for (int64_t s0 = 0; ok && s0 < n; s0 += batch) {
            const int rows = (int)std::min<int64_t>(batch, n - s0);
            ok = hipMemcpyAsync(dIdx, order.data() + s0, (size_t)rows * 4, hipMemcpyHostToDevice, st) == hipSuccess;
            hipLaunchKernelGGL(k_gather, dim3(blocks((long)M * wp0, 256)), dim3(256), 0, st, rows, M, w0, wp0, nX, dIdx,
                               m->H[0]);
            hipLaunchKernelGGL(k_gather, dim3(blocks((long)M * wpl, 256)), dim3(256), 0, st, rows, M, wl, wpl, nT, dIdx,
                               m->tgt);
            ok = ok && m->forward(M) == FCS_OK;
            hipLaunchKernelGGL(k_mse, dim3(blocks(M, 64)), dim3(64), 0, st, M, wpl, rows, wl, m->H[L], m->tgt, m->dA,
                               m->err);
            ok = ok && m->backward(M) == FCS_OK;
            ++m->adamStep;
            const float c1 = 1.f - std::pow(b1, (float)m->adamStep), c2 = 1.f - std::pow(b2, (float)m->adamStep);
            hipLaunchKernelGGL(k_adam, dim3(blocks((long)m->nParam, 256)), dim3(256), 0, st, (long)m->nParam, m->param,
                               m->grad, m->adm, m->adv, lr, b1, b2, eps, c1, c2);
            ok = ok && hipMemcpyAsync(e.data(), m->err, (size_t)M * 4, hipMemcpyDeviceToHost, st) == hipSuccess &&
                 hipStreamSynchronize(st) == hipSuccess;
            for (int i = 0; i < rows; ++i) esum += e[i];
        }
        if (lossOut) lossOut[ep] = (float)(esum / ((double)n * wl));
    }
    cleanup();
    if (!ok) return FCS_E_HIP;
    // Adam moves padding entries only through zero gradients: keep them zero
    std::vector<float> p(m->nParam);
    MLP_HIP(hipMemcpy(p.data(), m->param, m->nParam * 4, hipMemcpyDeviceToHost));
    zero_padding(*m, p);
    MLP_HIP(hipMemcpy(m->param, p.data(), m->nParam * 4, hipMemcpyHostToDevice));
    return FCS_OK;
}

}  // extern "C"
