// lyapunov_kernels.hpp -- finite-difference tangent vectors on a batched engine (gfx950).
//
// Included by engine.hip after SysInfo.  A Lyapunov engine is a batch of R replica groups of
// (1 + K) systems with identical parameters and noise seeds: system r*(K+1) is replica r's base
// trajectory, systems r*(K+1)+c (c = 1..K) are its perturbed clones.  All systems of a group hold n
// particles, so particle j of system s has caller index s*n + j.
//
// The reference evolves one perturbed copy per direction next to the unperturbed state and
// re-orthonormalises their differences (parallel/lyapunovbox.h:128-194): differences are taken with
// minimum-image positions and theta wrapped to (-pi, pi] (shared/state-hyper-vector.h:96-106,
// 213-225), the inner product is sum(r.r' + theta theta') (:157-167), and the directions are
// Gram-Schmidt'ed in index order (shared/vector-set.h:118-127).  Here the Gram-Schmidt step is the
// equivalent QR factorisation V = Q R computed as Cholesky QR done twice: a device Gram matrix per
// replica (fixed-order reduction), a per-replica Cholesky in one thread, and an in-place
// V <- V R^-1.  Index order is preserved (R upper triangular), so Q and diag(R) are the
// Gram-Schmidt ones.
#pragma once

namespace lyap {

constexpr int kMaxK = 8;  // tangent vectors per replica

struct Args {
    int R, K, n;                 // replicas, vectors per replica, particles per system
    int nchunk;                  // Gram partial sums per replica
    const SysInfo* sys;          // per-system table (the box of replica r is sys[r*(K+1)].G)
    const int* inv;              // caller index -> slot
    double2* pos;                // current state set: wrapped positions
    double* th;
    double2* U;
    double* dev;                 // [R][K][3][n] deviation / direction components
    double* part;                // [R][nchunk][K(K+1)/2] Gram partials
    double* coef;                // [R][K*K + K]: R^-1 (column-major upper) then the clone scales
    double* rdiag;               // [R][K] diag(R) of the first QR pass
    double* logsum;              // [R][K] accumulated log growth
    double* norm0;               // [R][K] deviation magnitude at the interval start
    int* status;                 // [R] 0 = ok, 1 = Gram matrix not positive definite
};

__device__ __forceinline__ size_t dev_index(const Args& a, int r, int k, int comp, int j)
{
    return (((size_t)r * a.K + k) * 3 + comp) * (size_t)a.n + j;
}

// State_Hyper_Vector::Periodic_Transform of a theta difference (state-hyper-vector.h:223):
// theta - 2 pi ceil((theta - pi) / (2 pi)), into (-pi, pi].
__device__ __forceinline__ double wrap_theta_diff(double t)
{
    return t - 2 * spp::kPi * ceil((t - spp::kPi) / (2 * spp::kPi));
}

// C2DVector::Periodic_Transform of one component (c2dvector.h:67-73).
__device__ __forceinline__ double ref_wrap(double x, double L2)
{
    return x - L2 * (double)(int)floor(x / L2 + 0.5);
}

// caller index -> slot of the current set
__global__ __launch_bounds__(256) void k_inverse(int N, const int* __restrict__ id, int* __restrict__ inv)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N && (unsigned)id[k] < (unsigned)N) inv[id[k]] = k;
}

// Deviations of every clone from its base, thread per (replica, particle):
// dev = clone - base (state-hyper-vector.h:96-106, Periodic_Transform applied to the difference).
__global__ __launch_bounds__(256) void k_deviation(Args a)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.R * a.n) return;
    int r = t / a.n, j = t - r * a.n;
    const Grid& G = a.sys[r * (a.K + 1)].G;
    int sb = a.inv[(r * (a.K + 1)) * a.n + j];
    const double2 b = a.pos[sb];
    double tb = a.th[sb];
    for (int c = 1; c <= a.K; ++c) {
        int sc = a.inv[(r * (a.K + 1) + c) * a.n + j];
        const double2 q = a.pos[sc];
        a.dev[dev_index(a, r, c - 1, 0, j)] = ref_wrap(q.x - b.x, G.Lx2);
        a.dev[dev_index(a, r, c - 1, 1, j)] = ref_wrap(q.y - b.y, G.Ly2);
        a.dev[dev_index(a, r, c - 1, 2, j)] = wrap_theta_diff(a.th[sc] - tb);
    }
}

// Random initial directions: components uniform in [-1, 1) from a Philox stream
// (VectorSet::Rand draws U(+-1e-3), vector-set.h:95-105; the scale drops out at normalisation).
__global__ __launch_bounds__(256) void k_random_dev(Args a, uint64_t seed)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.R * a.n) return;
    int r = t / a.n, j = t - r * a.n;
    for (int k = 0; k < a.K; ++k) {
        uint32_t c[4] = {(uint32_t)j, (uint32_t)r, (uint32_t)k, 0x5350504Cu};
        spp::philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        for (int comp = 0; comp < 3; ++comp)
            a.dev[dev_index(a, r, k, comp, j)] = (double)c[comp] * (2.0 / 4294967296.0) - 1.0;
    }
}

// Gram partial sums, workgroup r*nchunk + chunk: G_ab = sum over the 3n components of v_a v_b.
template <int K>
__global__ __launch_bounds__(256) void k_gram(Args a)
{
    constexpr int NP = K * (K + 1) / 2;
    __shared__ double red[4][NP];
    int r = blockIdx.x / a.nchunk, chunk = blockIdx.x - r * a.nchunk;
    double acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = 0;
    const int M = 3 * a.n;
    const size_t base = (size_t)r * K * M;
    for (int e = chunk * 256 + threadIdx.x; e < M; e += a.nchunk * 256) {
        double v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = a.dev[base + (size_t)k * M + e];
        int q = 0;
#pragma unroll
        for (int x = 0; x < K; ++x)
#pragma unroll
            for (int y = x; y < K; ++y) acc[q++] += v[x] * v[y];
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        double s = acc[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = s;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NP; q += 256)
        a.part[((size_t)r * a.nchunk + chunk) * NP + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
}

// One thread per replica: sum the partials in chunk order, Cholesky G = R^T R, R^-1.
//   pass 0: first QR of the raw deviations; keep diag(R) and Q1 = V R^-1 (no clone update).
//   pass 1: second QR (of Q1); diag of the product R2 R1, the log growth (unless `init`), the clone
//           scales (epsilon, or the grown magnitude when `rescale` is 0: the reference keeps the
//           projected deviation unrescaled, lyapunovbox.h:177-183) and Q = Q1 R2^-1.
__global__ void k_chol(Args a, int pass, int init, int rescale, double eps)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    const int K = a.K, NP = K * (K + 1) / 2;
    double G[kMaxK][kMaxK], L[kMaxK][kMaxK], Ri[kMaxK][kMaxK];
    for (int x = 0; x < K; ++x)
        for (int y = 0; y < K; ++y) G[x][y] = L[x][y] = Ri[x][y] = 0;
    for (int ch = 0; ch < a.nchunk; ++ch) {
        const double* p = a.part + ((size_t)r * a.nchunk + ch) * NP;
        int q = 0;
        for (int x = 0; x < K; ++x)
            for (int y = x; y < K; ++y) {
                G[x][y] += p[q];
                ++q;
            }
    }
    bool ok = true;
    // G = L L^T with L lower triangular; R = L^T
    for (int j = 0; j < K; ++j) {
        double d = G[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if (!(d > 0)) {
            ok = false;
            d = 1;
        }
        L[j][j] = sqrt(d);
        for (int i = j + 1; i < K; ++i) {
            double s = G[j][i];
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            L[i][j] = s / L[j][j];
        }
    }
    if (!ok) a.status[r] = 1;
    // Ri = R^-1 (upper): solve R Ri = I column by column, R[i][k] = L[k][i]
    for (int c = 0; c < K; ++c) {
        for (int i = c; i >= 0; --i) {
            double s = i == c ? 1.0 : 0.0;
            for (int k = i + 1; k <= c; ++k) s -= L[k][i] * Ri[k][c];
            Ri[i][c] = s / L[i][i];
        }
    }
    double* cf = a.coef + (size_t)r * (K * K + K);
    for (int c = 0; c < K; ++c)
        for (int k = 0; k < K; ++k) cf[c * K + k] = Ri[k][c];
    for (int c = 0; c < K; ++c) {
        double* rd = a.rdiag + (size_t)r * K + c;
        if (pass == 0) {
            *rd = L[c][c];
            cf[K * K + c] = 0;  // no clone update
            continue;
        }
        double grown = *rd * L[c][c];
        double scale = eps;
        if (!init) {
            a.logsum[(size_t)r * K + c] += log(grown / a.norm0[(size_t)r * K + c]);
            if (!rescale) scale = grown;
        }
        a.norm0[(size_t)r * K + c] = scale;
        cf[K * K + c] = scale;
    }
}

// Thread per (replica, particle): v <- v R^-1 in place for the three components and, when the
// clone scales are set, clone_c = base + scale_c q_c (LyapunovBox::Add_Deviation,
// lyapunovbox.h:47-67: r += dr then Periodic_Transform; theta += dtheta; v = (cos, sin)).
template <int K>
__global__ __launch_bounds__(256) void k_apply(Args a)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.R * a.n) return;
    int r = t / a.n, j = t - r * a.n;
    const double* cf = a.coef + (size_t)r * (K * K + K);
    double q[3][K];
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
        double v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = a.dev[dev_index(a, r, k, comp, j)];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            double s = 0;
#pragma unroll
            for (int k = 0; k <= c; ++k) s += v[k] * cf[c * K + k];
            q[comp][c] = s;
            a.dev[dev_index(a, r, c, comp, j)] = s;
        }
    }
    if (cf[K * K] == 0) return;
    const Grid& G = a.sys[r * (K + 1)].G;
    int sb = a.inv[(r * (K + 1)) * a.n + j];
    const double2 b = a.pos[sb];
    double tb = a.th[sb];
    double2 ub = a.U[sb];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        double sc = cf[K * K + c];
        double dx = sc * q[0][c], dy = sc * q[1][c], dt = sc * q[2][c];
        int s = a.inv[(r * (K + 1) + c + 1) * a.n + j];
        double th = tb + dt;
        a.pos[s] = make_double2(ref_wrap(b.x + dx, G.Lx2), ref_wrap(b.y + dy, G.Ly2));
        a.th[s] = th;
        a.U[s] = make_double2(ub.x + dx, ub.y + dy);
    }
}

}  // namespace lyap
