// surrogate.hip -- multilayer perceptron trained and evaluated with BF16 matrix-core (MFMA)
// products and FP32 accumulation on gfx950, behind include/spp_surrogate.h.
//
// Every matrix product of the forward and backward passes is one launch of k_gemm: a wavefront
// owns a 16x16 output tile and walks the reduction dimension 32 at a time with
// v_mfma_f32_16x16x32_bf16.  Operands are read with explicit strides, so the same kernel serves
// H W (forward), dZ W^T (backward data) and H^T dZ (weight gradients) without transposed copies;
// the epilogue adds the bias and applies SiLU (forward) or multiplies by SiLU' (backward) and
// writes the FP32 value and/or its BF16 rounding.  The network is small (4 -> 128 -> 128 -> 128
// -> 3), so the kernels are latency-bound; the matrix-core path keeps each product to a few MFMA
// instructions per wavefront.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/spp_surrogate.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;  // four wavefronts per workgroup, one 16x16 tile each

// FP32 -> BF16, round to nearest even (bit arithmetic, identical on host and device).
__host__ __device__ inline uint16_t to_bf16(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ inline float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

struct Gemm {
    int M, N, K;
    const uint16_t* A;   // A(m, k) = A[m * sam + k * sak]
    int64_t sam, sak;
    const uint16_t* B;   // B(k, n) = B[k * sbk + n * sbn]
    int64_t sbk, sbn;
    const float* bias;   // [N] or null
    int act;             // 0: none; 1: SiLU (o32 = pre-activation, o16 = activation); 2: times SiLU'(zin)
    const float* zin;    // act 2: pre-activation, zin[m * szm + n]
    int64_t szm;
    float* o32;          // o32[m * s32m + n] or null
    int64_t s32m;
    uint16_t* o16;       // o16[m * s16m + n] or null
    int64_t s16m;
};

// C = A B (+ bias, activation).  MFMA operand layout: lane l supplies row (A) / column (B) l % 16
// and eight reduction indices from group l / 16; the accumulator holds C[4 (l / 16) + i][l % 16].
__global__ __launch_bounds__(kThreads) void k_gemm(Gemm g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tn = (g.N + 15) >> 4, tm = (g.M + 15) >> 4;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= tm * tn) return;  // whole wavefront
    const int m0 = (tile / tn) << 4, n0 = (tile % tn) << 4;
    const int r = lane & 15, kg = lane >> 4;
    const int am = m0 + r, bn = n0 + r;
    const bool arow = am < g.M, bcol = bn < g.N;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        u16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + kg * 8 + e;
            const bool kin = k < g.K;
            a[e] = (arow && kin) ? g.A[am * g.sam + k * g.sak] : (unsigned short)0;
            b[e] = (bcol && kin) ? g.B[k * g.sbk + bn * g.sbn] : (unsigned short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                      __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + 4 * kg + i, n = n0 + r;
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i] + (g.bias ? g.bias[n] : 0.0f);
        float h = v;
        if (g.act == 1) {
            h = v * sigmoidf(v);
        } else if (g.act == 2) {
            const float z = g.zin[m * g.szm + n], s = sigmoidf(z);
            v = v * (s * (1.0f + z * (1.0f - s)));
            h = v;
        }
        if (g.o32) g.o32[m * g.s32m + n] = v;
        if (g.o16) g.o16[m * g.s16m + n] = to_bf16(h);
    }
}

__global__ __launch_bounds__(kThreads) void k_to_bf16(int n, const float* __restrict__ x, uint16_t* __restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// Output gradient of the weighted mean-square loss, and the loss itself (one workgroup, fixed
// summation order).
__global__ __launch_bounds__(kThreads) void k_out_grad(int batch, int n_out, const float* __restrict__ y,
                                                       const float* __restrict__ t, const float* __restrict__ w,
                                                       float* __restrict__ d32, uint16_t* __restrict__ d16,
                                                       float* __restrict__ loss)
{
    __shared__ float red[kThreads];
    const int n = batch * n_out;
    const float scale = 1.0f / (float)n;
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        const float wo = w ? w[i % n_out] : 1.0f, d = y[i] - t[i];
        s += wo * d * d;
        const float g = 2.0f * wo * d * scale;
        d32[i] = g;
        d16[i] = to_bf16(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] * scale;
}

// Bias gradient: column sums of dZ (FP32), one thread per column.
__global__ __launch_bounds__(kThreads) void k_colsum(int rows, int cols, const float* __restrict__ d,
                                                     float* __restrict__ out)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int b = 0; b < rows; ++b) s += d[(int64_t)b * cols + c];
    out[c] = s;
}

// Adam (Kingma & Ba), bias corrections c1 = 1 - beta1^t, c2 = 1 - beta2^t from the host; the BF16
// copy of the parameters is refreshed in the same pass.
__global__ __launch_bounds__(kThreads) void k_adam(int n, float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, float lr,
                                                   float b1, float b2, float eps, float c1, float c2,
                                                   uint16_t* __restrict__ p16)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    p16[i] = to_bf16(pi);
}

// Minibatch assembly: rows idx[0 .. batch) of x and t.
__global__ __launch_bounds__(kThreads) void k_gather_rows(int batch, int nx, int nt, const int* __restrict__ idx,
                                                          const float* __restrict__ X, const float* __restrict__ T,
                                                          float* __restrict__ xb, float* __restrict__ tb)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = nx + nt;
    if (i >= batch * w) return;
    const int b = i / w, c = i - b * w, r = idx[b];
    if (c < nx)
        xb[b * nx + c] = X[(int64_t)r * nx + c];
    else
        tb[b * nt + (c - nx)] = T[(int64_t)r * nt + (c - nx)];
}

thread_local std::string g_mlp_error;

}  // namespace

struct spp_mlp {
    spp_mlp_config c{};
    int L = 0;                    // layers (hidden + output)
    std::vector<int> dim;         // dim[0] = n_in, ..., dim[L] = n_out
    std::vector<int64_t> offW, offb;
    int64_t P = 0;
    int64_t adam_t = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grad = nullptr, *mom1 = nullptr, *mom2 = nullptr;
    uint16_t* p16 = nullptr;
    std::vector<float*> Z, dZ32;  // per layer: pre-activation (hidden layers), its gradient
    std::vector<uint16_t*> H, dZ16;
    float *x = nullptr, *t = nullptr, *y = nullptr, *w = nullptr, *loss = nullptr;
    // spp_mlp_train data
    float *X = nullptr, *T = nullptr, *losses = nullptr;
    int* idx = nullptr;
    int64_t data_n = 0, losses_n = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    std::string err;
};

namespace {

int mfail(spp_mlp* m, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (m)
        m->err = buf;
    else
        g_mlp_error = buf;
    return code;
}

#define MHIP(m, call)                                                                                \
    do {                                                                                             \
        hipError_t _st = (call);                                                                     \
        if (_st != hipSuccess) return mfail((m), -2, "%s: %s", #call, hipGetErrorString(_st));       \
    } while (0)

#define MTRY(x)                    \
    do {                           \
        int _rc = (x);             \
        if (_rc != 0) return _rc;  \
    } while (0)

template <typename T>
int malloc_dev(spp_mlp* m, T** p, int64_t n)
{
    MHIP(m, hipMalloc((void**)p, (size_t)std::max<int64_t>(n, 1) * sizeof(T)));
    return 0;
}

inline int blocks(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

template <typename F, typename... Args>
int mlaunch(spp_mlp* m, F kernel, int grid, Args... args)
{
    if (grid <= 0) return 0;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, m->stream, args...);
    MHIP(m, hipGetLastError());
    return 0;
}

int gemm(spp_mlp* m, const Gemm& g)
{
    const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    return mlaunch(m, k_gemm, (tiles + 3) / 4, g);
}

// Forward pass of `batch` rows already in m->x.
int forward(spp_mlp* m, int batch)
{
    const int L = m->L;
    MTRY(mlaunch(m, k_to_bf16, blocks((int64_t)batch * m->dim[0]), batch * m->dim[0], (const float*)m->x, m->H[0]));
    for (int l = 0; l < L; ++l) {
        const int din = m->dim[l], dout = m->dim[l + 1];
        Gemm g{};
        g.M = batch;
        g.N = dout;
        g.K = din;
        g.A = m->H[l];
        g.sam = din;
        g.sak = 1;
        g.B = m->p16 + m->offW[l];
        g.sbk = dout;
        g.sbn = 1;
        g.bias = m->params + m->offb[l];
        if (l < L - 1) {
            g.act = 1;
            g.o32 = m->Z[l];
            g.s32m = dout;
            g.o16 = m->H[l + 1];
            g.s16m = dout;
        } else {
            g.act = 0;
            g.o32 = m->y;
            g.s32m = dout;
        }
        MTRY(gemm(m, g));
    }
    return 0;
}

// Backward pass for the rows of the last forward pass against targets m->t; gradient into m->grad,
// loss into *loss_dev.
int backward(spp_mlp* m, int batch, const float* w, float* loss_dev)
{
    const int L = m->L;
    MTRY(mlaunch(m, k_out_grad, 1, batch, m->dim[L], (const float*)m->y, (const float*)m->t, w, m->dZ32[L - 1],
                 m->dZ16[L - 1], loss_dev));
    for (int l = L - 1; l >= 0; --l) {
        const int din = m->dim[l], dout = m->dim[l + 1];
        // weight gradient: H_l^T dZ_l, reduction over the batch
        Gemm gw{};
        gw.M = din;
        gw.N = dout;
        gw.K = batch;
        gw.A = m->H[l];
        gw.sam = 1;
        gw.sak = din;
        gw.B = m->dZ16[l];
        gw.sbk = dout;
        gw.sbn = 1;
        gw.o32 = m->grad + m->offW[l];
        gw.s32m = dout;
        MTRY(gemm(m, gw));
        MTRY(mlaunch(m, k_colsum, blocks(dout), batch, dout, (const float*)m->dZ32[l], m->grad + m->offb[l]));
        if (l == 0) break;
        // data gradient of the layer below: dZ_l W_l^T, times SiLU'(Z_{l-1})
        Gemm gd{};
        gd.M = batch;
        gd.N = din;
        gd.K = dout;
        gd.A = m->dZ16[l];
        gd.sam = dout;
        gd.sak = 1;
        gd.B = m->p16 + m->offW[l];
        gd.sbk = 1;
        gd.sbn = dout;
        gd.act = 2;
        gd.zin = m->Z[l - 1];
        gd.szm = din;
        gd.o32 = m->dZ32[l - 1];
        gd.s32m = din;
        gd.o16 = m->dZ16[l - 1];
        gd.s16m = din;
        MTRY(gemm(m, gd));
    }
    return 0;
}

int adam(spp_mlp* m)
{
    m->adam_t++;
    const float c1 = 1.0f - std::pow(m->c.beta1, (float)m->adam_t);
    const float c2 = 1.0f - std::pow(m->c.beta2, (float)m->adam_t);
    return mlaunch(m, k_adam, blocks(m->P), (int)m->P, m->params, (const float*)m->grad, m->mom1, m->mom2, m->c.lr,
                   m->c.beta1, m->c.beta2, m->c.eps, c1, c2, m->p16);
}

int check_batch(spp_mlp* m, int batch)
{
    if (batch <= 0 || batch > m->c.max_batch)
        return mfail(m, -1, "batch %d outside [1, %d]", batch, m->c.max_batch);
    return 0;
}

int upload_w(spp_mlp* m, const float* w)
{
    if (!w) return 0;
    MHIP(m, hipMemcpyAsync(m->w, w, m->dim[m->L] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    return 0;
}

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

void spp_mlp_config_default(spp_mlp_config* c)
{
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->n_in = 4;
    c->hidden = 128;
    c->n_hidden = 3;
    c->n_out = 3;
    c->max_batch = 4096;
    c->seed = 1;
    c->lr = 1e-3f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1e-8f;
}

const char* spp_mlp_last_error(const spp_mlp* m) { return m ? m->err.c_str() : g_mlp_error.c_str(); }

int64_t spp_mlp_param_count(const spp_mlp* m) { return m ? m->P : -1; }

void spp_mlp_destroy(spp_mlp* m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* bufs[] = {m->params, m->grad, m->mom1, m->mom2, m->p16, m->x, m->t, m->y, m->w, m->loss,
                    m->X, m->T, m->losses, m->idx};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (auto* p : m->Z) (void)hipFree(p);
    for (auto* p : m->dZ32) (void)hipFree(p);
    for (auto* p : m->H) (void)hipFree(p);
    for (auto* p : m->dZ16) (void)hipFree(p);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int spp_mlp_create(const spp_mlp_config* c, spp_mlp** out)
{
    if (!c || !out) return mfail(nullptr, -1, "null argument");
    *out = nullptr;
    if (c->n_in < 1 || c->hidden < 1 || c->n_hidden < 0 || c->n_out < 1 || c->max_batch < 1 ||
        c->n_in > 4096 || c->hidden > 4096 || c->n_out > 4096 || c->max_batch > (1 << 20))
        return mfail(nullptr, -1, "bad network shape");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return mfail(nullptr, -6, "no HIP device visible");
    spp_mlp* m = new spp_mlp();
    m->c = *c;
    m->L = c->n_hidden + 1;
    m->dim.push_back(c->n_in);
    for (int l = 0; l < c->n_hidden; ++l) m->dim.push_back(c->hidden);
    m->dim.push_back(c->n_out);
    for (int l = 0; l < m->L; ++l) {
        m->offW.push_back(m->P);
        m->P += (int64_t)m->dim[l] * m->dim[l + 1];
        m->offb.push_back(m->P);
        m->P += m->dim[l + 1];
    }
    const int64_t B = c->max_batch;
    int rc = 0;
    auto A = [&](int r) {
        if (rc == 0 && r != 0) rc = r;
    };
    A(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess ? 0 : -2);
    A(malloc_dev(m, &m->params, m->P));
    A(malloc_dev(m, &m->grad, m->P));
    A(malloc_dev(m, &m->mom1, m->P));
    A(malloc_dev(m, &m->mom2, m->P));
    A(malloc_dev(m, &m->p16, m->P));
    m->Z.assign(m->L, nullptr);
    m->dZ32.assign(m->L, nullptr);
    m->H.assign(m->L, nullptr);
    m->dZ16.assign(m->L, nullptr);
    for (int l = 0; l < m->L; ++l) {
        A(malloc_dev(m, &m->H[l], B * m->dim[l]));
        A(malloc_dev(m, &m->Z[l], B * m->dim[l + 1]));
        A(malloc_dev(m, &m->dZ32[l], B * m->dim[l + 1]));
        A(malloc_dev(m, &m->dZ16[l], B * m->dim[l + 1]));
    }
    A(malloc_dev(m, &m->x, B * c->n_in));
    A(malloc_dev(m, &m->t, B * c->n_out));
    A(malloc_dev(m, &m->y, B * c->n_out));
    A(malloc_dev(m, &m->w, c->n_out));
    A(malloc_dev(m, &m->loss, 1));
    if (rc == 0 && (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)) rc = -2;
    if (rc != 0) {
        std::string msg = m->err.empty() ? std::string("device allocation failed") : m->err;
        spp_mlp_destroy(m);
        return mfail(nullptr, rc, "%s", msg.c_str());
    }
    // Glorot-uniform weights, zero biases (host generator: deterministic for a seed)
    std::vector<float> p(m->P, 0.0f);
    uint64_t s = c->seed;
    for (int l = 0; l < m->L; ++l) {
        const double a = std::sqrt(6.0 / (m->dim[l] + m->dim[l + 1]));
        for (int64_t k = 0; k < (int64_t)m->dim[l] * m->dim[l + 1]; ++k) {
            const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
            p[m->offW[l] + k] = (float)((2.0 * u - 1.0) * a);
        }
    }
    rc = spp_mlp_set_params(m, p.data(), m->P);
    if (rc != 0) {
        std::string msg = m->err;
        spp_mlp_destroy(m);
        return mfail(nullptr, rc, "%s", msg.c_str());
    }
    *out = m;
    return 0;
}

int spp_mlp_get_params(spp_mlp* m, float* params, int64_t n)
{
    if (!m || !params || n != m->P) return mfail(m, -1, "parameter vector has %lld entries", m ? (long long)m->P : 0LL);
    MHIP(m, hipMemcpyAsync(params, m->params, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int spp_mlp_set_params(spp_mlp* m, const float* params, int64_t n)
{
    if (!m || !params || n != m->P) return mfail(m, -1, "parameter vector has %lld entries", m ? (long long)m->P : 0LL);
    MHIP(m, hipMemcpyAsync(m->params, params, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemsetAsync(m->mom1, 0, n * sizeof(float), m->stream));
    MHIP(m, hipMemsetAsync(m->mom2, 0, n * sizeof(float), m->stream));
    MTRY(mlaunch(m, k_to_bf16, blocks(n), (int)n, (const float*)m->params, m->p16));
    m->adam_t = 0;
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int spp_mlp_forward(spp_mlp* m, const float* x, int32_t batch, float* y)
{
    if (!m || !x || !y) return mfail(m, -1, "null argument");
    MTRY(check_batch(m, batch));
    MHIP(m, hipMemcpyAsync(m->x, x, (size_t)batch * m->dim[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MTRY(forward(m, batch));
    MHIP(m, hipMemcpyAsync(y, m->y, (size_t)batch * m->dim[m->L] * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int spp_mlp_loss_grad(spp_mlp* m, const float* x, const float* t, const float* w, int32_t batch, float* loss,
                      float* grad)
{
    if (!m || !x || !t) return mfail(m, -1, "null argument");
    MTRY(check_batch(m, batch));
    MHIP(m, hipMemcpyAsync(m->x, x, (size_t)batch * m->dim[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->t, t, (size_t)batch * m->dim[m->L] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MTRY(upload_w(m, w));
    MTRY(forward(m, batch));
    MTRY(backward(m, batch, w ? m->w : nullptr, m->loss));
    if (loss) MHIP(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (grad) MHIP(m, hipMemcpyAsync(grad, m->grad, m->P * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int spp_mlp_train_step(spp_mlp* m, const float* x, const float* t, const float* w, int32_t batch, float* loss)
{
    if (!m || !x || !t) return mfail(m, -1, "null argument");
    MTRY(check_batch(m, batch));
    MHIP(m, hipMemcpyAsync(m->x, x, (size_t)batch * m->dim[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->t, t, (size_t)batch * m->dim[m->L] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MTRY(upload_w(m, w));
    MTRY(forward(m, batch));
    MTRY(backward(m, batch, w ? m->w : nullptr, m->loss));
    MTRY(adam(m));
    if (loss) MHIP(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int spp_mlp_train(spp_mlp* m, const float* x, const float* t, const float* w, int32_t n, int32_t epochs,
                  int32_t batch, uint64_t shuffle_seed, float* loss)
{
    if (!m || !x || !t || n <= 0 || epochs <= 0) return mfail(m, -1, "bad arguments");
    MTRY(check_batch(m, std::min<int32_t>(batch, n)));
    const int nx = m->dim[0], nt = m->dim[m->L];
    if (m->data_n < n) {
        for (void* b : {(void*)m->X, (void*)m->T, (void*)m->idx})
            if (b) (void)hipFree(b);
        m->X = m->T = nullptr;
        m->idx = nullptr;
        MTRY(malloc_dev(m, &m->X, (int64_t)n * nx));
        MTRY(malloc_dev(m, &m->T, (int64_t)n * nt));
        MTRY(malloc_dev(m, &m->idx, (int64_t)n));
        m->data_n = n;
    }
    MHIP(m, hipMemcpyAsync(m->X, x, (size_t)n * nx * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->T, t, (size_t)n * nt * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MTRY(upload_w(m, w));
    const int per_epoch = (n + batch - 1) / batch;
    const int64_t steps = (int64_t)per_epoch * epochs;
    if (m->losses_n < steps) {
        if (m->losses) (void)hipFree(m->losses);
        m->losses = nullptr;
        MTRY(malloc_dev(m, &m->losses, steps));
        m->losses_n = steps;
    }
    std::vector<int> order(n);
    uint64_t s = shuffle_seed;
    MHIP(m, hipEventRecord(m->ev0, m->stream));
    int64_t step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = 0; i < n; ++i) order[i] = i;
        for (int i = n - 1; i > 0; --i) std::swap(order[i], order[splitmix64(s) % (uint64_t)(i + 1)]);
        // the order upload must finish before the host vector changes: synchronise per epoch
        MHIP(m, hipMemcpyAsync(m->idx, order.data(), n * sizeof(int), hipMemcpyHostToDevice, m->stream));
        for (int b0 = 0; b0 < n; b0 += batch) {
            const int nb = std::min(batch, n - b0);
            MTRY(mlaunch(m, k_gather_rows, blocks((int64_t)nb * (nx + nt)), nb, nx, nt, (const int*)(m->idx + b0),
                         (const float*)m->X, (const float*)m->T, m->x, m->t));
            MTRY(forward(m, nb));
            MTRY(backward(m, nb, w ? m->w : nullptr, m->losses + step));
            MTRY(adam(m));
            ++step;
        }
        MHIP(m, hipStreamSynchronize(m->stream));
    }
    MHIP(m, hipEventRecord(m->ev1, m->stream));
    MHIP(m, hipEventSynchronize(m->ev1));
    float ms = 0;
    MHIP(m, hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_ms = ms;
    if (loss) {
        std::vector<float> h(per_epoch);
        MHIP(m, hipMemcpy(h.data(), m->losses + (steps - per_epoch), per_epoch * sizeof(float), hipMemcpyDeviceToHost));
        double a = 0;
        for (float v : h) a += v;
        *loss = (float)(a / per_epoch);
    }
    return 0;
}

int spp_mlp_last_train_ms(spp_mlp* m, double* ms)
{
    if (!m || !ms) return mfail(m, -1, "null argument");
    *ms = m->last_ms;
    return 0;
}

}  // extern "C"
