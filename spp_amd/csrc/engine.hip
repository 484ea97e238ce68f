// engine.hip -- MI355X (gfx950) engine for the per-time-step SPP particle path, behind the C ABI
// declared in include/spp_engine.h.
//
// Device layout (all arrays in CELL-SORTED slot order; id[slot] is the caller's particle index):
//   pos0  P4[N]      state being advanced: wrapped x, y and theta (reference r, theta)
//   pos1/pos2 P4[N]  RK stage states (pos1 = RK2 midpoint)
//   U     double2[N] unwrapped positions (reference r_original)
//   dth   double[N]  this step's angular noise
//   kacc  P4[N]      RK4 accumulator k1 + 2 k2 + 2 k3 (fx, fy, tau)
//   nbr   int[cap*N] full Verlet list, ELL, slot-major (nbr[k*N + i]) so a wavefront's k-th
//                    neighbour ids are one coalesced 256-B read
//   build float2[N]  positions at the last build, for the displacement trigger
// Cells are keyed in 4x4-cell tiles so consecutive slots cover compact 2-D patches and a
// wavefront's neighbour gathers stay inside a few 128-B lines of L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "spp_device.hpp"

using spp::P4;
using spp::Phys;

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;  // items per thread in the cell-count scan
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kTileBits = 2;   // 4x4 cells per key tile

// ------------------------------------------------------------------------------------------------
// Small device utilities
// ------------------------------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// slots (a contiguous region of space) so its L2 keeps that region's particles.
__device__ __forceinline__ int xcd_block()
{
    int nblk = gridDim.x, b = blockIdx.x;
    if ((nblk & 7) == 0) return (b & 7) * (nblk >> 3) + (b >> 3);
    return b;
}

__device__ __forceinline__ int cell_coord(double x, double L, double L2, int nc, int quirk)
{
    int c;
    if (quirk)  // serial/box.h:116-117: the cast truncates (x + L) before scaling
        c = (int)((double)((int)(x + L) * nc) / L2);
    else
        c = (int)floor((x + L) * nc / L2);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__device__ __forceinline__ int cell_key(int cx, int cy, int ntx)
{
    return (((cy >> kTileBits) * ntx + (cx >> kTileBits)) << (2 * kTileBits)) +
           ((cy & ((1 << kTileBits) - 1)) << kTileBits) + (cx & ((1 << kTileBits) - 1));
}

struct Grid {
    double Lx, Ly, Lx2, Ly2;
    int ncx, ncy, ntx, quirk;
};

template <typename T>
__global__ void k_fill(T* a, T v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// ------------------------------------------------------------------------------------------------
// Rebuild: counting sort of particles by cell, then the Verlet list
// (Box::Update_Cells + Update_Neighbor_List, serial/box.h:107-179)
// ------------------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_bin(int N, Grid G, const P4* __restrict__ pos,
                                                int* __restrict__ key, int* __restrict__ count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    P4 p = pos[i];
    int cx = cell_coord(p.x, G.Lx, G.Lx2, G.ncx, G.quirk);
    int cy = cell_coord(p.y, G.Ly, G.Ly2, G.ncy, G.quirk);
    int k = cell_key(cx, cy, G.ntx);
    key[i] = k;
    atomicAdd(&count[k], 1);
}

// Block-wide exclusive scan of kScanTile ints; returns the tile total.
__device__ int block_exclusive_scan(int (&v)[kScanItems], int* lds_wave)
{
    int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int x = v[k];
        v[k] = local;
        local += x;
    }
    int incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) lds_wave[w] = incl;
    __syncthreads();
    int wave_off = 0, total = 0;
    for (int k = 0; k < kBlock / 64; ++k) {
        if (k < w) wave_off += lds_wave[k];
        total += lds_wave[k];
    }
    int excl = incl - local + wave_off;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] += excl;
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(kBlock) void k_scan_partial(const int* __restrict__ in, int n,
                                                         int* __restrict__ block_sums)
{
    __shared__ int red[kBlock / 64];
    int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < kBlock / 64; ++k) t += red[k];
        block_sums[blockIdx.x] = t;
    }
}

// out[0..n] = exclusive scan of in[0..n), out[n] = total.  Each block first sums the partial
// sums of the blocks before it (there are few: n / 1024).
__global__ __launch_bounds__(kBlock) void k_scan_final(const int* __restrict__ in, int n,
                                                       const int* __restrict__ block_sums,
                                                       int* __restrict__ out)
{
    __shared__ int red[kBlock / 64];
    __shared__ int lds_wave[kBlock / 64];
    int b = blockIdx.x;
    int pre = 0;
    for (int k = threadIdx.x; k < b; k += kBlock) pre += block_sums[k];
    pre = wave_sum(pre);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pre;
    __syncthreads();
    int prefix = 0;
    for (int k = 0; k < kBlock / 64; ++k) prefix += red[k];
    int base = b * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = (base + k < n) ? in[base + k] : 0;
    int total = block_exclusive_scan(v, lds_wave);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] = v[k] + prefix;
    if (b == gridDim.x - 1 && threadIdx.x == 0) out[n] = prefix + total;
}

__global__ __launch_bounds__(kBlock) void k_scatter(int N, const int* __restrict__ key,
                                                    const int* __restrict__ start,
                                                    int* __restrict__ cursor, int* __restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int k = key[i];
    perm[start[k] + atomicAdd(&cursor[k], 1)] = i;
}

// Order each cell's particles by caller index, so the sorted layout is a pure function of the
// positions (the atomic scatter leaves it in arrival order).
__global__ __launch_bounds__(kBlock) void k_cell_sort(int nkeys, const int* __restrict__ start,
                                                      int* __restrict__ perm,
                                                      const int* __restrict__ id)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nkeys) return;
    int s = start[c], e = start[c + 1];
    for (int a = s + 1; a < e; ++a) {
        int v = perm[a], kv = id[v];
        int b = a - 1;
        while (b >= s && id[perm[b]] > kv) {
            perm[b + 1] = perm[b];
            --b;
        }
        perm[b + 1] = v;
    }
}

struct PermuteArgs {
    int N;
    const int* perm;
    const P4* pos_in;
    P4* pos_out;
    const double2* U_in;
    double2* U_out;
    const int* id_in;
    int* id_out;
    // per-stage buffers of the unfused path (null when the permutation happens between steps)
    const P4* a_in[3];
    P4* a_out[3];
    const double* d_in;
    double* d_out;
};

__global__ __launch_bounds__(kBlock) void k_permute(PermuteArgs a)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    int o = a.perm[k];
    a.pos_out[k] = a.pos_in[o];
    a.U_out[k] = a.U_in[o];
    a.id_out[k] = a.id_in[o];
    if (a.d_in) {
        a.d_out[k] = a.d_in[o];
#pragma unroll
        for (int q = 0; q < 3; ++q) a.a_out[q][k] = a.a_in[q][o];
    }
}

struct NlistArgs {
    int N, cap;
    Grid G;
    double rv;
    const P4* pos;
    const int* start;
    int* nbr;
    int* ncount;
    float2* build;
    int* cellxy;
    int* stats;  // [0] max list length, [1] total entries (low 31 bits are enough per build)
};

// Cell::Neighbor_List (shared/cell.h:74-103) over the 3x3 periodic stencil, full list, predicate
// sqrt(d2) < rv exactly as cell.h:84-85.
__global__ __launch_bounds__(kBlock) void k_nlist(NlistArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = i < a.N;
    int c = 0;
    if (live) {
        const Grid& G = a.G;
        P4 me = a.pos[i];
        int cx = cell_coord(me.x, G.Lx, G.Lx2, G.ncx, G.quirk);
        int cy = cell_coord(me.y, G.Ly, G.Ly2, G.ncy, G.quirk);
        for (int oy = -1; oy <= 1; ++oy) {
            int y = cy + oy;
            y = y < 0 ? y + G.ncy : (y >= G.ncy ? y - G.ncy : y);
            for (int ox = -1; ox <= 1; ++ox) {
                int x = cx + ox;
                x = x < 0 ? x + G.ncx : (x >= G.ncx ? x - G.ncx : x);
                int k = cell_key(x, y, G.ntx);
                int s = a.start[k], e = a.start[k + 1];
                for (int j = s; j < e; ++j) {
                    if (j == i) continue;
                    P4 pj = a.pos[j];
                    double dx = spp::min_image(me.x - pj.x, G.Lx, G.Lx2);
                    double dy = spp::min_image(me.y - pj.y, G.Ly, G.Ly2);
                    double d = sqrt(dx * dx + dy * dy);
                    if (d < a.rv) {
                        if (c < a.cap) a.nbr[(size_t)c * a.N + i] = j;
                        ++c;
                    }
                }
            }
        }
        a.ncount[i] = c < a.cap ? c : a.cap;
        a.build[i] = make_float2((float)me.x, (float)me.y);
        a.cellxy[i] = cx | (cy << 16);
    }
    int m = c;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    int t = wave_sum(c);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&a.stats[0], m);
        atomicAdd(&a.stats[1], t);
    }
}

// ------------------------------------------------------------------------------------------------
// Stage kernels: interact + Move fused, one particle per lane.
// ------------------------------------------------------------------------------------------------

struct StageArgs {
    int N;
    int cap;
    Phys P;
    const P4* gather;   // stage input state (also this particle's own stage state)
    P4* out;            // stage output state
    const P4* pos0;     // step-start state (theta_old) for RK stages >= 2
    double2* U;         // unwrapped positions
    double* dth;        // angular noise of this step
    P4* kacc;           // RK4 accumulator
    const int* nbr;
    const int* ncount;
    const int* id;
    const double* inj;  // injected noise, slot order (null = stream)
    const float2* build;
    int64_t step;       // noise-stream step
    int kstep;          // step index inside the current chunk
    double thresh2;     // displacement trigger squared
    int* bad_step;
    unsigned long long* maxdisp;
    int record;         // record the max displacement (last step of a chunk)
};

template <int MODEL, int REP>
__device__ __forceinline__ void gather_forces(const StageArgs& a, int i, const P4& me, double vix,
                                              double viy, double& fx, double& fy, double& tq)
{
    double s1 = 0, s2 = 0;
    fx = vix * a.P.v0;  // Reset: f = v (speed v0), particle.h:559-564
    fy = viy * a.P.v0;
    int cnt = a.ncount[i];
    const int* nb = a.nbr + i;
    for (int k = 0; k < cnt; ++k) {
        int j = nb[(size_t)k * a.N];
        P4 pj = a.gather[j];
        spp::pair_accumulate<MODEL, REP>(a.P, me.x, me.y, me.th, vix, viy, pj, fx, fy, s1, s2);
    }
    tq = spp::drift_torque<MODEL>(a.P, s1, s2);
}

__device__ __forceinline__ void check_disp(const StageArgs& a, int i, bool live, double x,
                                           double y, int mark)
{
    double d2 = 0;
    if (live) {
        float2 b = a.build[i];
        double dx = spp::min_image(x - (double)b.x, a.P.Lx, a.P.Lx2);
        double dy = spp::min_image(y - (double)b.y, a.P.Ly, a.P.Ly2);
        d2 = dx * dx + dy * dy;
        if (d2 > a.thresh2) atomicMin(a.bad_step, mark);
    }
    if (a.record) {
        double m = wave_max(d2);
        if ((threadIdx.x & 63) == 0)
            atomicMax(a.maxdisp, (unsigned long long)__double_as_longlong(m));
    }
}

// INTEG/STAGE: Euler (1,0); RK2 (2,1..2); RK4 (4,1..4).  Reference updates: Move
// (particle.h:571-587), Move_Runge_Kutta2_1/2 (:589-631), Move_Runge_Kutta4_1..4 (:634-725).
template <int MODEL, int REP, int INTEG, int STAGE>
__global__ __launch_bounds__(kBlock) void k_stage(StageArgs a)
{
    int i = xcd_block() * kBlock + threadIdx.x;
    bool live = i < a.N;
    double x = 0, y = 0;
    if (live) {
        const Phys& P = a.P;
        P4 me = a.gather[i];
        double vix = cos(me.th), viy = sin(me.th);
        double fx, fy, tq;
        gather_forces<MODEL, REP>(a, i, me, vix, viy, fx, fy, tq);
        double th;
        if (STAGE <= 1) {
            double dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)a.id[i], a.step);
            if (INTEG == SPP_EULER) {
                th = me.th + P.dt * (tq);
                th += dn;
                double2 u = a.U[i];
                u.x += fx * P.dt;
                u.y += fy * P.dt;
                a.U[i] = u;
                x = spp::periodic_wrap(u.x, P.Lx2);
                y = spp::periodic_wrap(u.y, P.Ly2);
            } else {
                a.dth[i] = dn;
                th = me.th;
                th += P.half_dt * tq;
                th += dn / 2;
                th = spp::wrap_angle(th);
                x = spp::periodic_wrap(me.x + fx * P.half_dt, P.Lx2);
                y = spp::periodic_wrap(me.y + fy * P.half_dt, P.Ly2);
                if (INTEG == SPP_RK4) a.kacc[i] = P4{fx, fy, tq, 0.0};
            }
        } else {
            double thold = a.pos0[i].th;
            double dn = a.dth[i];
            double2 u = a.U[i];
            double h = (INTEG == SPP_RK4 && STAGE == 2) ? P.half_dt : P.dt;
            if (INTEG == SPP_RK4 && STAGE == 4) {
                P4 k = a.kacc[i];
                th = thold + (k.th + tq) * P.dt6;
                th += dn;
                u.x = u.x + (k.x + fx) * P.dt6;
                u.y = u.y + (k.y + fy) * P.dt6;
            } else {
                th = thold + h * tq;
                th += (INTEG == SPP_RK4 && STAGE == 2) ? dn / 2 : dn;
                u.x = u.x + fx * h;
                u.y = u.y + fy * h;
                if (INTEG == SPP_RK4) {
                    P4 k = a.kacc[i];
                    a.kacc[i] = P4{k.x + fx * 2, k.y + fy * 2, k.th + tq * 2, 0.0};
                }
            }
            th = spp::wrap_angle(th);
            bool final_stage = (INTEG == SPP_RK2) || (STAGE == 4);
            if (final_stage) a.U[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly2);
        }
        a.out[i] = P4{x, y, th, 0.0};
    }
    bool final_stage = (INTEG == SPP_EULER) || (INTEG == SPP_RK2 && STAGE == 2) || (STAGE == 4);
    check_disp(a, i, live, x, y, final_stage ? a.kstep + 1 : a.kstep);
}

// ------------------------------------------------------------------------------------------------
// Unfused path: Box::Interact and Move_* as separate launches (drop-in granularity).
// ------------------------------------------------------------------------------------------------

template <int MODEL, int REP>
__global__ __launch_bounds__(kBlock) void k_interact(StageArgs a, P4* force)
{
    int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.N) return;
    P4 me = a.gather[i];
    double vix = cos(me.th), viy = sin(me.th);
    double fx, fy, tq;
    gather_forces<MODEL, REP>(a, i, me, vix, viy, fx, fy, tq);
    force[i] = P4{fx, fy, tq, 0.0};
}

struct MoveArgs {
    int N;
    Phys P;
    int integ, stage;
    P4* pos;            // state, updated in place like the reference's particle array
    double2* U;
    P4* old;            // r_old.x, r_old.y, theta_old
    P4* kacc;
    double* dth;
    const P4* force;
    const int* id;
    const double* inj;
    int64_t step;
};

__global__ __launch_bounds__(kBlock) void k_move(MoveArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const Phys& P = a.P;
    P4 me = a.pos[i];
    P4 f = a.force[i];
    double2 u = a.U[i];
    double th;
    if (a.stage <= 1) {
        double dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)a.id[i], a.step);
        a.dth[i] = dn;
        if (a.integ == SPP_EULER) {
            th = me.th + P.dt * (f.th);
            th += dn;
            u.x += f.x * P.dt;
            u.y += f.y * P.dt;
            a.U[i] = u;
            a.pos[i] = P4{spp::periodic_wrap(u.x, P.Lx2), spp::periodic_wrap(u.y, P.Ly2), th, 0.0};
            return;
        }
        a.old[i] = P4{u.x, u.y, me.th, 0.0};
        th = me.th;
        th += P.half_dt * f.th;
        th += dn / 2;
        th = spp::wrap_angle(th);
        if (a.integ == SPP_RK4) a.kacc[i] = f;
        a.pos[i] = P4{spp::periodic_wrap(me.x + f.x * P.half_dt, P.Lx2),
                      spp::periodic_wrap(me.y + f.y * P.half_dt, P.Ly2), th, 0.0};
        return;
    }
    P4 o = a.old[i];
    double dn = a.dth[i];
    if (a.integ == SPP_RK4 && a.stage == 4) {
        P4 k = a.kacc[i];
        th = o.th + (k.th + f.th) * P.dt6;
        th += dn;
        u.x = o.x + (k.x + f.x) * P.dt6;
        u.y = o.y + (k.y + f.y) * P.dt6;
    } else {
        double h = (a.integ == SPP_RK4 && a.stage == 2) ? P.half_dt : P.dt;
        th = o.th + h * f.th;
        th += (a.integ == SPP_RK4 && a.stage == 2) ? dn / 2 : dn;
        u.x = o.x + f.x * h;
        u.y = o.y + f.y * h;
        if (a.integ == SPP_RK4) {
            P4 k = a.kacc[i];
            a.kacc[i] = P4{k.x + f.x * 2, k.y + f.y * 2, k.th + f.th * 2, 0.0};
        }
    }
    th = spp::wrap_angle(th);
    a.U[i] = u;
    a.pos[i] = P4{spp::periodic_wrap(u.x, P.Lx2), spp::periodic_wrap(u.y, P.Ly2), th, 0.0};
}

// Scatter caller-order values into slot order: dst[slot] = src[id[slot]].
__global__ void k_gather_by_id(int N, const int* __restrict__ id, const double* __restrict__ src,
                               double* __restrict__ dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = src[id[i]];
}

// Polarization partial sums (Box::Compute_Polarization, serial/box.h:531-538), two-pass.
__global__ __launch_bounds__(kBlock) void k_polar(int N, const P4* __restrict__ pos,
                                                  double2* __restrict__ part)
{
    __shared__ double sx[kBlock / 64], sy[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double vx = 0, vy = 0;
    if (i < N) {
        double th = pos[i].th;
        vx = cos(th);
        vy = sin(th);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        vx += __shfl_xor(vx, o);
        vy += __shfl_xor(vy, o);
    }
    if ((threadIdx.x & 63) == 0) {
        sx[threadIdx.x >> 6] = vx;
        sy[threadIdx.x >> 6] = vy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
            a += sx[k];
            b += sy[k];
        }
        part[blockIdx.x] = make_double2(a, b);
    }
}

// ------------------------------------------------------------------------------------------------
// Kernel dispatch tables
// ------------------------------------------------------------------------------------------------

using StageFn = void (*)(StageArgs);
using InteractFn = void (*)(StageArgs, P4*);

template <int MODEL, int REP>
StageFn stage_fn(int integ, int stage)
{
    if (integ == SPP_EULER) return k_stage<MODEL, REP, SPP_EULER, 0>;
    if (integ == SPP_RK2) return stage == 1 ? k_stage<MODEL, REP, SPP_RK2, 1> : k_stage<MODEL, REP, SPP_RK2, 2>;
    switch (stage) {
    case 1: return k_stage<MODEL, REP, SPP_RK4, 1>;
    case 2: return k_stage<MODEL, REP, SPP_RK4, 2>;
    case 3: return k_stage<MODEL, REP, SPP_RK4, 3>;
    default: return k_stage<MODEL, REP, SPP_RK4, 4>;
    }
}

StageFn pick_stage(int model, int rep, int integ, int stage)
{
    if (model == SPP_MODEL_CONTINUOUS) return stage_fn<SPP_MODEL_CONTINUOUS, SPP_REP_NONE>(integ, stage);
    switch (rep) {
    case SPP_REP_NONE: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_NONE>(integ, stage);
    case SPP_REP_R12_TRUNCATED: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_R12_TRUNCATED>(integ, stage);
    case SPP_REP_HARMONIC: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_HARMONIC>(integ, stage);
    case SPP_REP_R12: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_R12>(integ, stage);
    case SPP_REP_YUKAWA_TRUNCATED: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA_TRUNCATED>(integ, stage);
    default: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA>(integ, stage);
    }
}

InteractFn pick_interact(int model, int rep)
{
    if (model == SPP_MODEL_CONTINUOUS) return k_interact<SPP_MODEL_CONTINUOUS, SPP_REP_NONE>;
    switch (rep) {
    case SPP_REP_NONE: return k_interact<SPP_MODEL_REPULSIVE, SPP_REP_NONE>;
    case SPP_REP_R12_TRUNCATED: return k_interact<SPP_MODEL_REPULSIVE, SPP_REP_R12_TRUNCATED>;
    case SPP_REP_HARMONIC: return k_interact<SPP_MODEL_REPULSIVE, SPP_REP_HARMONIC>;
    case SPP_REP_R12: return k_interact<SPP_MODEL_REPULSIVE, SPP_REP_R12>;
    case SPP_REP_YUKAWA_TRUNCATED: return k_interact<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA_TRUNCATED>;
    default: return k_interact<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA>;
    }
}

const char* stage_name(int integ, int stage)
{
    static const char* names[] = {"stage_euler", "stage_rk2_1", "stage_rk2_2", "stage_rk4_1",
                                  "stage_rk4_2", "stage_rk4_3", "stage_rk4_4"};
    if (integ == SPP_EULER) return names[0];
    if (integ == SPP_RK2) return names[stage];
    return names[2 + stage];
}

thread_local std::string g_create_error;

}  // namespace

// ==================================================================================================
// Engine
// ==================================================================================================

struct Snapshot {
    P4* pos = nullptr;
    double2* U = nullptr;
    int* id = nullptr;
    int64_t step = 0;
    double t = 0;
    int chunk_hint = 0;
    double rate_hint = 0;
    bool valid = false;
};

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

struct spp_engine {
    spp_params p{};
    int N = 0;
    Grid G{};
    int nkeys = 0, nscan_blocks = 0;
    int cap = 32;
    Phys P{};
    double r_cut = 0, half_skin = 0, thresh = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    P4 *pos0 = nullptr, *pos1 = nullptr, *pos2 = nullptr, *pos_alt = nullptr;
    double2 *U = nullptr, *U_alt = nullptr;
    int *id = nullptr, *id_alt = nullptr;
    double *dth = nullptr, *dth_alt = nullptr, *inj_host_order = nullptr, *inj = nullptr;
    P4 *kacc = nullptr, *force = nullptr, *old = nullptr;
    P4 *aux_alt[3] = {nullptr, nullptr, nullptr};
    int *nbr = nullptr, *ncount = nullptr, *cellxy = nullptr;
    float2* build = nullptr;
    int *key = nullptr, *perm = nullptr, *cell_count = nullptr, *cell_start = nullptr,
        *cell_cursor = nullptr, *block_sums = nullptr;
    int* dflags = nullptr;              // [0] bad_step, [1] max list length, [2] total entries
    unsigned long long* dmaxdisp = nullptr;
    int* hflags = nullptr;              // pinned mirror: 4 ints + maxdisp (8 B)
    double2* dpolar = nullptr;
    Snapshot rollback;
    std::map<int, Snapshot> slots;

    int64_t step = 0;
    double t = 0;
    int64_t rebuilds = 0, rollbacks = 0, host_syncs = 0;
    int64_t steps_since_build = 0;
    bool have_state = false, list_valid = false, need_rebuild = false;
    bool inj_pending = false;
    int chunk_hint = 8;
    double rate_hint = 0;   // displacement per step seen over the last build interval
    int max_nbrs = 0;
    double mean_nbrs = 0;
    int unfused_stage = 0;  // next expected stage of the unfused path (0 = at a step boundary)

    bool profiling = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> event_pool;

    std::string err;
};

namespace {

int fail(spp_engine* e, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e)
        e->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIPC(e, call)                                                                              \
    do {                                                                                           \
        hipError_t _st = (call);                                                                   \
        if (_st != hipSuccess) return fail((e), SPP_ERR_HIP, "%s: %s", #call, hipGetErrorString(_st)); \
    } while (0)

inline int nblocks(int n) { return (n + kBlock - 1) / kBlock; }

hipEvent_t take_event(spp_engine* e)
{
    if (!e->event_pool.empty()) {
        hipEvent_t ev = e->event_pool.back();
        e->event_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    (void)hipEventCreate(&ev);
    return ev;
}

// Launch helper: records HIP events around the launch on the engine's stream when profiling.
template <typename F, typename... Args>
int launch(spp_engine* e, const char* name, F kernel, int grid, Args... args)
{
    if (grid <= 0) return SPP_OK;
    ProfRec rec{name, nullptr, nullptr};
    if (e->profiling) {
        rec.a = take_event(e);
        rec.b = take_event(e);
        HIPC(e, hipEventRecord(rec.a, e->stream));
    }
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, e->stream, args...);
    HIPC(e, hipGetLastError());
    if (e->profiling) {
        HIPC(e, hipEventRecord(rec.b, e->stream));
        e->prof.push_back(rec);
    }
    return SPP_OK;
}

#define TRY(x)                  \
    do {                        \
        int _rc = (x);          \
        if (_rc != SPP_OK) return _rc; \
    } while (0)

template <typename T>
int dalloc(spp_engine* e, T** p, size_t n)
{
    HIPC(e, hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    return SPP_OK;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

double max_cutoff(const spp_params& p)
{
    double c = p.R_align;
    if (p.model == SPP_MODEL_REPULSIVE && p.repulsion != SPP_REP_NONE && p.rc_rep > c) c = p.rc_rep;
    return c;
}

int alloc_nbr(spp_engine* e)
{
    dfree(e->nbr);
    e->nbr = nullptr;
    return dalloc(e, &e->nbr, (size_t)e->cap * e->N);
}

// ------------------------------------------------------------------------------------------------
// Rebuild: bin, scan, scatter, per-cell order, permute, Verlet list.
// ------------------------------------------------------------------------------------------------
int launch_rebuild(spp_engine* e, bool aux)
{
    int N = e->N, nk = e->nkeys;
    TRY(launch(e, "fill", k_fill<int>, nblocks(nk), e->cell_count, 0, nk));
    TRY(launch(e, "fill", k_fill<int>, nblocks(nk), e->cell_cursor, 0, nk));
    TRY(launch(e, "fill", k_fill<int>, 1, e->dflags + 1, 0, 2));
    TRY(launch(e, "bin", k_bin, nblocks(N), N, e->G, (const P4*)e->pos0, e->key, e->cell_count));
    TRY(launch(e, "scan_partial", k_scan_partial, e->nscan_blocks, (const int*)e->cell_count, nk,
               e->block_sums));
    TRY(launch(e, "scan_final", k_scan_final, e->nscan_blocks, (const int*)e->cell_count, nk,
               (const int*)e->block_sums, e->cell_start));
    TRY(launch(e, "scatter", k_scatter, nblocks(N), N, (const int*)e->key,
               (const int*)e->cell_start, e->cell_cursor, e->perm));
    TRY(launch(e, "cell_sort", k_cell_sort, nblocks(nk), nk, (const int*)e->cell_start, e->perm,
               (const int*)e->id));
    PermuteArgs pa{};
    pa.N = N;
    pa.perm = e->perm;
    pa.pos_in = e->pos0;
    pa.pos_out = e->pos_alt;
    pa.U_in = e->U;
    pa.U_out = e->U_alt;
    pa.id_in = e->id;
    pa.id_out = e->id_alt;
    if (aux) {
        pa.a_in[0] = e->kacc;
        pa.a_in[1] = e->force;
        pa.a_in[2] = e->old;
        for (int q = 0; q < 3; ++q) pa.a_out[q] = e->aux_alt[q];
        pa.d_in = e->dth;
        pa.d_out = e->dth_alt;
    }
    TRY(launch(e, "permute", k_permute, nblocks(N), pa));
    std::swap(e->pos0, e->pos_alt);
    std::swap(e->U, e->U_alt);
    std::swap(e->id, e->id_alt);
    if (aux) {
        std::swap(e->kacc, e->aux_alt[0]);
        std::swap(e->force, e->aux_alt[1]);
        std::swap(e->old, e->aux_alt[2]);
        std::swap(e->dth, e->dth_alt);
    }
    NlistArgs na{N, e->cap, e->G, e->p.r_list, e->pos0, e->cell_start, e->nbr, e->ncount,
                 e->build, e->cellxy, e->dflags + 1};
    TRY(launch(e, "nlist", k_nlist, nblocks(N), na));
    e->rebuilds++;
    e->steps_since_build = 0;
    e->list_valid = true;
    e->need_rebuild = false;
    return SPP_OK;
}

// Re-run only the list kernel (after growing the capacity; cells and order are unchanged).
int launch_nlist_only(spp_engine* e)
{
    TRY(launch(e, "fill", k_fill<int>, 1, e->dflags + 1, 0, 2));
    NlistArgs na{e->N, e->cap, e->G, e->p.r_list, e->pos0, e->cell_start, e->nbr, e->ncount,
                 e->build, e->cellxy, e->dflags + 1};
    return launch(e, "nlist", k_nlist, nblocks(e->N), na);
}

int read_flags(spp_engine* e)
{
    HIPC(e, hipMemcpyAsync(e->hflags, e->dflags, 4 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(e->hflags + 4, e->dmaxdisp, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    e->host_syncs++;
    return SPP_OK;
}

// Synchronously verify the list of the last rebuild fitted; grow and rebuild the list if not.
int ensure_capacity(spp_engine* e)
{
    for (;;) {
        TRY(read_flags(e));
        int need = e->hflags[1];
        e->max_nbrs = need;
        e->mean_nbrs = (double)(unsigned)e->hflags[2] / e->N;
        if (need <= e->cap) return SPP_OK;
        e->cap = ((need + 7) / 8) * 8 + 8;
        TRY(alloc_nbr(e));
        TRY(launch_nlist_only(e));
    }
}

// Injected noise arrives in caller order; put it in slot order with the current permutation.
int take_injected(spp_engine* e)
{
    TRY(launch(e, "gather_noise", k_gather_by_id, nblocks(e->N), e->N, (const int*)e->id,
               (const double*)e->inj_host_order, e->inj));
    return SPP_OK;
}

StageArgs stage_args(spp_engine* e)
{
    StageArgs a{};
    a.N = e->N;
    a.cap = e->cap;
    a.P = e->P;
    a.U = e->U;
    a.dth = e->dth;
    a.kacc = e->kacc;
    a.nbr = e->nbr;
    a.ncount = e->ncount;
    a.id = e->id;
    a.inj = nullptr;
    a.build = e->build;
    a.step = e->step;
    a.thresh2 = e->thresh * e->thresh;
    a.bad_step = e->dflags;
    a.maxdisp = e->dmaxdisp;
    return a;
}

// Launch one fused step (no host synchronisation).
int launch_step(spp_engine* e, int kstep, bool record)
{
    int integ = e->p.integrator;
    StageArgs a = stage_args(e);
    a.kstep = kstep;
    if (e->inj_pending) {
        TRY(take_injected(e));
        a.inj = e->inj;
    }
    int nst = integ == SPP_EULER ? 1 : integ;
    for (int s = 1; s <= nst; ++s) {
        int stage = integ == SPP_EULER ? 0 : s;
        bool last = s == nst;
        a.record = (record && last) ? 1 : 0;
        if (integ == SPP_EULER) {
            a.gather = e->pos0;
            a.out = e->pos1;
        } else if (integ == SPP_RK2) {
            a.gather = s == 1 ? e->pos0 : e->pos1;
            a.out = s == 1 ? e->pos1 : e->pos0;
        } else {
            const P4* g[5] = {nullptr, e->pos0, e->pos1, e->pos2, e->pos1};
            P4* o[5] = {nullptr, e->pos1, e->pos2, e->pos1, e->pos0};
            a.gather = g[s];
            a.out = o[s];
        }
        a.pos0 = e->pos0;
        StageFn fn = pick_stage(e->p.model, e->p.repulsion, integ, stage);
        TRY(launch(e, stage_name(integ, stage), fn, nblocks(e->N), a));
        if (s == 1) a.inj = nullptr;
    }
    if (integ == SPP_EULER) std::swap(e->pos0, e->pos1);
    e->step++;
    e->t += e->p.dt;
    e->steps_since_build++;
    return SPP_OK;
}

int copy_state(spp_engine* e, Snapshot& s, bool save)
{
    size_t N = e->N;
    if (!s.pos) {
        TRY(dalloc(e, &s.pos, N));
        TRY(dalloc(e, &s.U, N));
        TRY(dalloc(e, &s.id, N));
    }
    if (save) {
        HIPC(e, hipMemcpyAsync(s.pos, e->pos0, N * sizeof(P4), hipMemcpyDeviceToDevice, e->stream));
        HIPC(e, hipMemcpyAsync(s.U, e->U, N * sizeof(double2), hipMemcpyDeviceToDevice, e->stream));
        HIPC(e, hipMemcpyAsync(s.id, e->id, N * sizeof(int), hipMemcpyDeviceToDevice, e->stream));
        s.step = e->step;
        s.t = e->t;
        s.chunk_hint = e->chunk_hint;
        s.rate_hint = e->rate_hint;
        s.valid = true;
    } else {
        HIPC(e, hipMemcpyAsync(e->pos0, s.pos, N * sizeof(P4), hipMemcpyDeviceToDevice, e->stream));
        HIPC(e, hipMemcpyAsync(e->U, s.U, N * sizeof(double2), hipMemcpyDeviceToDevice, e->stream));
        HIPC(e, hipMemcpyAsync(e->id, s.id, N * sizeof(int), hipMemcpyDeviceToDevice, e->stream));
        e->step = s.step;
        e->t = s.t;
    }
    return SPP_OK;
}

void free_snapshot(Snapshot& s)
{
    dfree(s.pos);
    dfree(s.U);
    dfree(s.id);
    s = Snapshot{};
}

int check_ready(spp_engine* e)
{
    if (!e) return SPP_ERR_ARG;
    if (!e->have_state) return fail(e, SPP_ERR_STATE, "no particle state: call spp_set_state first");
    return SPP_OK;
}

// Steps under SPP_REBUILD_PERIOD: the reference's Multi_Step(cell_update_period) + Update_Cells.
int multi_step_period(spp_engine* e, int steps)
{
    int P = std::max(1, e->p.rebuild_period);
    for (int k = 0; k < steps; ++k) {
        TRY(launch_step(e, 0, false));
        e->inj_pending = false;
        if (e->steps_since_build >= P) {
            TRY(launch_rebuild(e, false));
            TRY(ensure_capacity(e));
        }
    }
    return SPP_OK;
}

// Steps under SPP_REBUILD_DISPLACEMENT.  Chunks of steps run without host round trips; the stage
// kernels flag the first step whose input state moved farther than the half skin.  At the chunk's
// end the host reads the flag: a late trigger rolls the chunk back to its start and replays only
// the valid steps (the counter-based noise makes the replay exact), then rebuilds.
int multi_step_displacement(spp_engine* e, int steps)
{
    int done = 0;
    while (done < steps) {
        bool rebuilt = false;
        if (e->need_rebuild || !e->list_valid) {
            if (e->steps_since_build > 0 && e->steps_since_build < 1000000)
                e->rate_hint = std::max(e->rate_hint * 0.5, e->thresh / (double)e->steps_since_build);
            TRY(launch_rebuild(e, false));
            rebuilt = true;
        }
        int m = std::min(steps - done, std::max(1, e->chunk_hint));
        TRY(copy_state(e, e->rollback, true));
        TRY(launch(e, "fill", k_fill<int>, 1, e->dflags, INT_MAX, 1));
        TRY(launch(e, "fill", k_fill<unsigned long long>, 1, e->dmaxdisp, 0ull, 1));
        for (int k = 0; k < m; ++k) TRY(launch_step(e, k, k == m - 1));
        TRY(read_flags(e));
        if (rebuilt) {
            e->max_nbrs = e->hflags[1];
            e->mean_nbrs = (double)(unsigned)e->hflags[2] / e->N;
            if (e->hflags[1] > e->cap) {  // list overflowed: restore, grow, redo the chunk
                TRY(copy_state(e, e->rollback, false));
                e->steps_since_build = 0;
                e->cap = ((e->hflags[1] + 7) / 8) * 8 + 8;
                TRY(alloc_nbr(e));
                TRY(launch_nlist_only(e));
                e->rollbacks++;
                continue;
            }
        }
        int bad = e->hflags[0];
        int64_t since0 = e->steps_since_build - m;
        if (bad < m) {
            // steps [0, bad) were computed with a valid list; replay exactly those
            TRY(copy_state(e, e->rollback, false));
            e->steps_since_build = since0;
            for (int k = 0; k < bad; ++k) TRY(launch_step(e, k, false));
            if (bad > 0) e->inj_pending = false;
            done += bad;
            e->rollbacks++;
            e->need_rebuild = true;
            e->chunk_hint = std::max(1, (int)(0.7 * (since0 + bad)));
            continue;
        }
        done += m;
        e->inj_pending = false;
        double disp2;
        std::memcpy(&disp2, e->hflags + 4, sizeof disp2);
        double disp = std::sqrt(disp2);
        double since = (double)e->steps_since_build;
        double rate = std::max(disp / std::max(since, 1.0), e->rate_hint * 0.25);
        double remaining = rate > 0 ? (e->thresh - disp) / rate : 1e9;
        if (remaining < 1.5) {
            e->need_rebuild = true;
            e->chunk_hint = std::max(1, (int)(0.8 * since));
        } else {
            e->chunk_hint = std::max(1, std::min(256, (int)(0.8 * remaining)));
        }
    }
    return SPP_OK;
}

Phys make_phys(const spp_params& p)
{
    Phys P{};
    P.Lx = p.Lx;
    P.Ly = p.Ly;
    P.Lx2 = 2 * p.Lx;
    P.Ly2 = 2 * p.Ly;
    P.dt = p.dt;
    P.half_dt = p.dt / 2;  // shared/parameters.h:78
    P.dt6 = p.dt / 6;      // shared/parameters.h:79
    P.v0 = p.v0;
    P.g = p.g;
    P.Ra = p.R_align;
    P.Ra2 = p.R_align * p.R_align;
    if (p.model == SPP_MODEL_CONTINUOUS) {
        P.c_align = (1 - p.alpha) / spp::kPi;
        P.c_self = p.alpha / spp::kPi;
    } else {
        P.c_align = p.g / (spp::kPi * p.R_align * p.R_align);
        P.c_self = 0;
    }
    P.rc = p.rc_rep;
    P.sigma = p.sigma_rep;
    P.A = p.A_rep;
    double sc = p.sigma_rep / p.rc_rep, sc2 = sc * sc, sc4 = sc2 * sc2, sc8 = sc4 * sc4;
    P.rc13 = sc8 * sc4 * sc;
    P.yuk_c = std::exp(-p.rc_rep / p.sigma_rep) *
              (1. / (p.rc_rep * p.rc_rep) + 1. / (p.sigma_rep * p.rc_rep));
    double c = max_cutoff(p);
    P.guard2 = c * c * (1.0 + 1e-12);
    P.noise_amp = std::sqrt(2 * p.D_rot * p.dt);  // BasicDynamicParticle::Set_Dr, particle.h:106-110
    P.seed = p.seed;
    return P;
}

}  // namespace

// ==================================================================================================
// C ABI
// ==================================================================================================

extern "C" {

int spp_params_default(spp_params* p, int32_t model)
{
    if (!p) return SPP_ERR_ARG;
    std::memset(p, 0, sizeof *p);
    p->abi_version = SPP_ABI_VERSION;
    p->model = model;
    p->integrator = SPP_RK2;
    p->n = 0;
    p->binning = SPP_BIN_EXACT;
    p->rebuild = SPP_REBUILD_DISPLACEMENT;
    p->rebuild_period = 32;  // cell_update_period at dt = 1/1024, shared/parameters.h:80-81
    p->noise_mode = SPP_NOISE_STREAM;
    p->seed = 10;            // shared/parameters.h:60
    p->Lx = p->Ly = 64;      // shared/parameters.h:68-72
    p->dt = 1.0 / 1024;      // shared/parameters.h:77
    p->v0 = 1;
    p->R_align = 1.1;        // RepulsiveParticle::alignment_radius, particle.h:779
    p->g = 1;
    p->alpha = 0;
    p->repulsion = SPP_REP_R12_TRUNCATED;
    p->A_rep = 2;            // particle.h:777
    p->sigma_rep = 1;        // particle.h:778
    p->rc_rep = 1.2;         // particle.h:780
    p->D_rot = 0;
    p->r_list = 1.0625;      // rv, shared/parameters.h:99
    p->mob_par = 0.5;        // ActiveBrownianChain::Set_nb(1), particle.h:861-863
    p->mob_perp = 0.5;
    p->mob_rot = 1;
    p->sigma_mm = 0.5;       // particle.h:1052-1053
    p->rc_mm = 0.55;
    p->A_mm = 1;
    p->k_bond = 100;
    p->l_bond = 1;
    if (model == SPP_MODEL_CONTINUOUS) {
        p->repulsion = SPP_REP_NONE;
        p->R_align = 1;      // d2 < 1, particle.h:305
        p->g = 4;            // ContinuousParticle::g, particle.h:373
    }
    return SPP_OK;
}

int spp_create(const spp_params* p, spp_engine** out)
{
    if (!p || !out) return fail(nullptr, SPP_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->abi_version != SPP_ABI_VERSION)
        return fail(nullptr, SPP_ERR_ARG, "abi_version %d, library is %d", p->abi_version, SPP_ABI_VERSION);
    if (p->model != SPP_MODEL_REPULSIVE && p->model != SPP_MODEL_CONTINUOUS)
        return fail(nullptr, SPP_ERR_UNSUPPORTED, "model %d not available", p->model);
    if (p->integrator != SPP_EULER && p->integrator != SPP_RK2 && p->integrator != SPP_RK4)
        return fail(nullptr, SPP_ERR_ARG, "integrator %d", p->integrator);
    if (p->repulsion < SPP_REP_NONE || p->repulsion > SPP_REP_YUKAWA)
        return fail(nullptr, SPP_ERR_ARG, "repulsion %d", p->repulsion);
    if (p->n <= 0 || !(p->Lx > 0) || !(p->Ly > 0) || !(p->dt > 0) || !(p->r_list > 0))
        return fail(nullptr, SPP_ERR_ARG, "n, Lx, Ly, dt and r_list must be positive");
    if (p->noise_mode != SPP_NOISE_STREAM && p->noise_mode != SPP_NOISE_INJECTED)
        return fail(nullptr, SPP_ERR_ARG, "noise_mode %d", p->noise_mode);
    double rcut = max_cutoff(*p);
    if (p->r_list < rcut) return fail(nullptr, SPP_ERR_ARG, "r_list %g below cutoff %g", p->r_list, rcut);
    int ncx = p->ncell_x > 0 ? p->ncell_x : (int)std::floor(2 * p->Lx / p->r_list);
    int ncy = p->ncell_y > 0 ? p->ncell_y : (int)std::floor(2 * p->Ly / p->r_list);
    // Cell::Cell (shared/cell.h:45-49): cells no narrower than rv.  Three or more per side keep
    // the nine stencil cells distinct.
    if (ncx < 3 || ncy < 3 || 2 * p->Lx / ncx < p->r_list || 2 * p->Ly / ncy < p->r_list)
        return fail(nullptr, SPP_ERR_ARG, "cell grid %dx%d invalid for r_list %g in box %gx%g", ncx,
                    ncy, p->r_list, 2 * p->Lx, 2 * p->Ly);
    if (ncx >= 32768 || ncy >= 32768) return fail(nullptr, SPP_ERR_ARG, "cell grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, SPP_ERR_NO_DEVICE, "no HIP device visible");

    spp_engine* e = new spp_engine();
    e->p = *p;
    e->N = p->n;
    (void)hipGetDevice(&e->device);
    e->G.Lx = p->Lx;
    e->G.Ly = p->Ly;
    e->G.Lx2 = 2 * p->Lx;
    e->G.Ly2 = 2 * p->Ly;
    e->G.ncx = ncx;
    e->G.ncy = ncy;
    e->G.ntx = (ncx + (1 << kTileBits) - 1) >> kTileBits;
    e->G.quirk = p->binning == SPP_BIN_REF_QUIRK;
    int nty = (ncy + (1 << kTileBits) - 1) >> kTileBits;
    e->nkeys = e->G.ntx * nty << (2 * kTileBits);
    e->nscan_blocks = (e->nkeys + kScanTile - 1) / kScanTile;
    e->cap = p->ell_capacity > 0 ? p->ell_capacity : 32;
    e->P = make_phys(*p);
    e->r_cut = rcut;
    e->half_skin = 0.5 * (p->r_list - rcut);
    // float32 build positions: allow for their rounding (|x| <= L) when triggering
    double margin = 4.0 * std::max(p->Lx, p->Ly) * 5.960464477539063e-08;
    e->thresh = std::max(0.0, e->half_skin - margin);
    if (p->rebuild == SPP_REBUILD_DISPLACEMENT && e->thresh <= 0) {
        delete e;
        return fail(nullptr, SPP_ERR_ARG, "r_list leaves no skin for displacement-triggered rebuilds");
    }

    int rc = SPP_OK;
    auto A = [&](int r) {
        if (rc == SPP_OK && r != SPP_OK) rc = r;
    };
    size_t N = e->N;
    A(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess ? SPP_OK : SPP_ERR_HIP);
    A(dalloc(e, &e->pos0, N));
    A(dalloc(e, &e->pos1, N));
    A(dalloc(e, &e->pos2, N));
    A(dalloc(e, &e->pos_alt, N));
    A(dalloc(e, &e->U, N));
    A(dalloc(e, &e->U_alt, N));
    A(dalloc(e, &e->id, N));
    A(dalloc(e, &e->id_alt, N));
    A(dalloc(e, &e->dth, N));
    A(dalloc(e, &e->dth_alt, N));
    A(dalloc(e, &e->inj_host_order, N));
    A(dalloc(e, &e->inj, N));
    A(dalloc(e, &e->kacc, N));
    A(dalloc(e, &e->force, N));
    A(dalloc(e, &e->old, N));
    for (int q = 0; q < 3; ++q) A(dalloc(e, &e->aux_alt[q], N));
    A(dalloc(e, &e->ncount, N));
    A(dalloc(e, &e->cellxy, N));
    A(dalloc(e, &e->build, N));
    A(dalloc(e, &e->key, N));
    A(dalloc(e, &e->perm, N));
    A(dalloc(e, &e->cell_count, (size_t)e->nkeys));
    A(dalloc(e, &e->cell_cursor, (size_t)e->nkeys));
    A(dalloc(e, &e->cell_start, (size_t)e->nkeys + 1));
    A(dalloc(e, &e->block_sums, (size_t)e->nscan_blocks));
    A(dalloc(e, &e->dflags, 4));
    A(dalloc(e, &e->dmaxdisp, 1));
    A(dalloc(e, &e->dpolar, (size_t)nblocks(e->N)));
    A(alloc_nbr(e));
    if (rc == SPP_OK && hipHostMalloc((void**)&e->hflags, 8 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        rc = SPP_ERR_HIP;
    if (rc != SPP_OK) {
        std::string msg = e->err.empty() ? std::string("device allocation failed") : e->err;
        spp_destroy(e);
        return fail(nullptr, rc, "%s", msg.c_str());
    }
    *out = e;
    return SPP_OK;
}

void spp_destroy(spp_engine* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    void* bufs[] = {e->pos0, e->pos1, e->pos2, e->pos_alt, e->U, e->U_alt, e->id, e->id_alt,
                    e->dth, e->dth_alt, e->inj_host_order, e->inj, e->kacc, e->force, e->old,
                    e->aux_alt[0], e->aux_alt[1], e->aux_alt[2], e->nbr, e->ncount, e->cellxy,
                    e->build, e->key, e->perm, e->cell_count, e->cell_cursor, e->cell_start,
                    e->block_sums, e->dflags, e->dmaxdisp, e->dpolar};
    for (void* b : bufs) dfree(b);
    free_snapshot(e->rollback);
    for (auto& kv : e->slots) free_snapshot(kv.second);
    for (auto& r : e->prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
    if (e->hflags) (void)hipHostFree(e->hflags);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* spp_last_error(const spp_engine* e)
{
    return e ? e->err.c_str() : g_create_error.c_str();
}

int spp_get_info(spp_engine* e, spp_info* info)
{
    if (!e || !info) return SPP_ERR_ARG;
    info->steps = e->step;
    info->rebuilds = e->rebuilds;
    info->rollbacks = e->rollbacks;
    info->host_syncs = e->host_syncs;
    info->t = e->t;
    info->ncell_x = e->G.ncx;
    info->ncell_y = e->G.ncy;
    info->ell_capacity = e->cap;
    info->max_neighbors = e->max_nbrs;
    info->mean_neighbors = e->mean_nbrs;
    return SPP_OK;
}

int spp_set_state(spp_engine* e, const double* xw, const double* yw, const double* xu,
                  const double* yu, const double* theta, int32_t n)
{
    if (!e || !xw || !yw || !theta) return e ? fail(e, SPP_ERR_ARG, "null array") : SPP_ERR_ARG;
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n = %d, engine has %d particles", n, e->N);
    std::vector<P4> pos(n);
    std::vector<double2> U(n);
    std::vector<int> id(n);
    for (int i = 0; i < n; ++i) {
        pos[i] = P4{xw[i], yw[i], theta[i], 0.0};
        U[i] = make_double2(xu ? xu[i] : xw[i], yu ? yu[i] : yw[i]);
        id[i] = i;
    }
    HIPC(e, hipMemcpyAsync(e->pos0, pos.data(), n * sizeof(P4), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(e->U, U.data(), n * sizeof(double2), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(e->id, id.data(), n * sizeof(int), hipMemcpyHostToDevice, e->stream));
    e->have_state = true;
    e->unfused_stage = 0;
    TRY(launch_rebuild(e, false));
    TRY(ensure_capacity(e));
    return SPP_OK;
}

int spp_get_state(spp_engine* e, double* xw, double* yw, double* xu, double* yu, double* theta,
                  int32_t n)
{
    TRY(check_ready(e));
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n = %d, engine has %d particles", n, e->N);
    std::vector<P4> pos(n);
    std::vector<double2> U(n);
    std::vector<int> id(n);
    HIPC(e, hipMemcpyAsync(pos.data(), e->pos0, n * sizeof(P4), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(U.data(), e->U, n * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < n; ++k) {
        int i = id[k];
        if (xw) xw[i] = pos[k].x;
        if (yw) yw[i] = pos[k].y;
        if (xu) xu[i] = U[k].x;
        if (yu) yu[i] = U[k].y;
        if (theta) theta[i] = pos[k].th;
    }
    return SPP_OK;
}

int spp_build_neighbors(spp_engine* e)
{
    TRY(check_ready(e));
    TRY(launch_rebuild(e, true));
    return ensure_capacity(e);
}

int spp_get_neighbors(spp_engine* e, int32_t* offsets, int32_t* idx, int64_t cap, int64_t* nnz)
{
    TRY(check_ready(e));
    if (!e->list_valid) TRY(spp_build_neighbors(e));
    int N = e->N, C = e->cap;
    std::vector<int> nb((size_t)C * N), cnt(N), id(N);
    HIPC(e, hipMemcpyAsync(nb.data(), e->nbr, nb.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(cnt.data(), e->ncount, N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->id, N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    std::vector<int64_t> rowlen(N + 1, 0);
    for (int k = 0; k < N; ++k) rowlen[id[k] + 1] = cnt[k];
    for (int i = 0; i < N; ++i) rowlen[i + 1] += rowlen[i];
    int64_t total = rowlen[N];
    if (nnz) *nnz = total;
    if (!offsets || !idx || cap < total)
        return fail(e, SPP_ERR_CAPACITY, "neighbour list has %lld entries", (long long)total);
    for (int i = 0; i <= N; ++i) offsets[i] = (int32_t)rowlen[i];
    for (int k = 0; k < N; ++k) {
        int64_t o = rowlen[id[k]];
        for (int c = 0; c < cnt[k]; ++c) idx[o + c] = id[nb[(size_t)c * N + k]];
    }
    return SPP_OK;
}

int spp_get_cells(spp_engine* e, int32_t* cell_x, int32_t* cell_y, int32_t n)
{
    TRY(check_ready(e));
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n mismatch");
    std::vector<int> cxy(n), id(n);
    HIPC(e, hipMemcpyAsync(cxy.data(), e->cellxy, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < n; ++k) {
        if (cell_x) cell_x[id[k]] = cxy[k] & 0xffff;
        if (cell_y) cell_y[id[k]] = cxy[k] >> 16;
    }
    return SPP_OK;
}

int spp_interact(spp_engine* e)
{
    TRY(check_ready(e));
    StageArgs a = stage_args(e);
    a.gather = e->pos0;
    InteractFn fn = pick_interact(e->p.model, e->p.repulsion);
    return launch(e, "interact", fn, nblocks(e->N), a, e->force);
}

int spp_get_forces(spp_engine* e, double* fx, double* fy, double* tau, int32_t n)
{
    TRY(check_ready(e));
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n mismatch");
    std::vector<P4> f(n);
    std::vector<int> id(n);
    HIPC(e, hipMemcpyAsync(f.data(), e->force, n * sizeof(P4), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < n; ++k) {
        int i = id[k];
        if (fx) fx[i] = f[k].x;
        if (fy) fy[i] = f[k].y;
        if (tau) tau[i] = f[k].th;
    }
    return SPP_OK;
}

int spp_inject_noise(spp_engine* e, const double* dtheta, int32_t n)
{
    TRY(check_ready(e));
    if (e->p.noise_mode != SPP_NOISE_INJECTED)
        return fail(e, SPP_ERR_STATE, "engine was created with noise_mode = SPP_NOISE_STREAM");
    if (!dtheta || n != e->N) return fail(e, SPP_ERR_ARG, "n mismatch");
    HIPC(e, hipMemcpyAsync(e->inj_host_order, dtheta, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    e->inj_pending = true;
    return SPP_OK;
}

int spp_move_stage(spp_engine* e, int32_t stage)
{
    TRY(check_ready(e));
    int integ = e->p.integrator;
    int nst = integ == SPP_EULER ? 1 : integ;
    int expect = integ == SPP_EULER ? 0 : e->unfused_stage + 1;
    if (stage != expect)
        return fail(e, SPP_ERR_STATE, "stage %d called, stage %d expected for integrator %d", stage,
                    expect, integ);
    if (stage <= 1 && e->p.noise_mode == SPP_NOISE_INJECTED && !e->inj_pending)
        return fail(e, SPP_ERR_STATE, "injected-noise engine: call spp_inject_noise before the step");
    MoveArgs m{};
    m.N = e->N;
    m.P = e->P;
    m.integ = integ;
    m.stage = stage;
    m.pos = e->pos0;
    m.U = e->U;
    m.old = e->old;
    m.kacc = e->kacc;
    m.dth = e->dth;
    m.force = e->force;
    m.id = e->id;
    m.inj = nullptr;
    m.step = e->step;
    if (stage <= 1 && e->inj_pending) {
        TRY(take_injected(e));
        m.inj = e->inj;
        e->inj_pending = false;
    }
    TRY(launch(e, "move", k_move, nblocks(e->N), m));
    int done = integ == SPP_EULER ? 1 : stage;
    if (done == nst) {
        e->unfused_stage = 0;
        e->step++;
        e->t += e->p.dt;
        e->steps_since_build++;
    } else {
        e->unfused_stage = done;
    }
    return SPP_OK;
}

int spp_multi_step(spp_engine* e, int32_t steps)
{
    TRY(check_ready(e));
    if (steps < 0) return fail(e, SPP_ERR_ARG, "steps < 0");
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "a split step is in progress");
    if (e->p.noise_mode == SPP_NOISE_INJECTED && (steps > 1 || (steps == 1 && !e->inj_pending)))
        return fail(e, SPP_ERR_STATE, "injected noise covers one step: inject before each step");
    if (steps == 0) return SPP_OK;
    if (e->p.rebuild == SPP_REBUILD_PERIOD) {
        if (!e->list_valid) {
            TRY(launch_rebuild(e, false));
            TRY(ensure_capacity(e));
        }
        return multi_step_period(e, steps);
    }
    return multi_step_displacement(e, steps);
}

int spp_one_step(spp_engine* e) { return spp_multi_step(e, 1); }

int spp_polarization(spp_engine* e, double* p, double* sum_x, double* sum_y)
{
    TRY(check_ready(e));
    int nb = nblocks(e->N);
    TRY(launch(e, "polar", k_polar, nb, e->N, (const P4*)e->pos0, e->dpolar));
    std::vector<double2> part(nb);
    HIPC(e, hipMemcpyAsync(part.data(), e->dpolar, nb * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    double sx = 0, sy = 0;
    for (auto& v : part) {
        sx += v.x;
        sy += v.y;
    }
    if (sum_x) *sum_x = sx;
    if (sum_y) *sum_y = sy;
    double ax = sx / e->N, ay = sy / e->N;
    if (p) *p = std::sqrt(ax * ax + ay * ay);
    return SPP_OK;
}

int spp_snapshot(spp_engine* e, int32_t slot)
{
    TRY(check_ready(e));
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "snapshot inside a split step");
    // Box::Save re-bins so that a later Load re-bins to the same particle order
    // (serial/box.h:313-318); the sorted layout here is a pure function of the positions, so
    // save-rebuild and restore-rebuild give the same order and list, and a replay is bitwise.
    TRY(launch_rebuild(e, false));
    TRY(ensure_capacity(e));
    return copy_state(e, e->slots[slot], true);
}

int spp_restore(spp_engine* e, int32_t slot)
{
    TRY(check_ready(e));
    auto it = e->slots.find(slot);
    if (it == e->slots.end() || !it->second.valid) return fail(e, SPP_ERR_ARG, "no snapshot in slot %d", slot);
    TRY(copy_state(e, it->second, false));
    e->chunk_hint = it->second.chunk_hint;
    e->rate_hint = it->second.rate_hint;
    e->unfused_stage = 0;
    // Box::Load ends in a cell update (parallel/box.h:241); rebuilding from the restored state
    // also makes the replay independent of whatever list was current.
    TRY(launch_rebuild(e, false));
    return ensure_capacity(e);
}

int spp_synchronize(spp_engine* e)
{
    if (!e) return SPP_ERR_ARG;
    HIPC(e, hipStreamSynchronize(e->stream));
    return SPP_OK;
}

int spp_profile_enable(spp_engine* e, int32_t on)
{
    if (!e) return SPP_ERR_ARG;
    HIPC(e, hipStreamSynchronize(e->stream));
    for (auto& r : e->prof) {
        e->event_pool.push_back(r.a);
        e->event_pool.push_back(r.b);
    }
    e->prof.clear();
    e->profiling = on != 0;
    return SPP_OK;
}

int spp_profile_read(spp_engine* e, char* names, int32_t cap, double* ms_avg, int64_t* launches,
                     int32_t max_kernels, int32_t* nk)
{
    if (!e) return SPP_ERR_ARG;
    HIPC(e, hipStreamSynchronize(e->stream));
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int64_t>> acc;
    for (auto& r : e->prof) {
        float ms = 0;
        HIPC(e, hipEventElapsedTime(&ms, r.a, r.b));
        std::string nm(r.name);
        if (!acc.count(nm)) order.push_back(nm);
        acc[nm].first += ms;
        acc[nm].second += 1;
    }
    std::string joined;
    int k = 0;
    for (auto& nm : order) {
        if (k < max_kernels) {
            if (ms_avg) ms_avg[k] = acc[nm].first / acc[nm].second;
            if (launches) launches[k] = acc[nm].second;
        }
        joined += nm;
        joined += '\n';
        ++k;
    }
    if (nk) *nk = k;
    if (names && cap > 0) {
        std::strncpy(names, joined.c_str(), cap - 1);
        names[cap - 1] = 0;
    }
    return SPP_OK;
}

int spp_random_state(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y,
                     double* theta)
{
    if (n < 0 || !x || !y || !theta) return SPP_ERR_ARG;
    for (int32_t i = 0; i < n; ++i) {
        uint32_t c[4] = {(uint32_t)i, 0u, 0u, 0x53505069u};
        spp::philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        double u0 = (double)c[0] * (1.0 / 4294967296.0);
        double u1 = (double)c[1] * (1.0 / 4294967296.0);
        double u2 = (double)c[2] * (1.0 / 4294967296.0);
        x[i] = -Lx + 2 * Lx * u0;
        y[i] = -Ly + 2 * Ly * u1;
        theta[i] = spp::kTwoPi * u2;
    }
    return SPP_OK;
}

double spp_noise_normal(uint64_t seed, int32_t id, int64_t step)
{
    return spp::noise_normal(seed, (uint32_t)id, step);
}

}  // extern "C"
