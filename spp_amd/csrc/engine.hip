// engine.hip -- MI355X (gfx950) engine for the per-time-step SPP particle path, behind the C ABI
// declared in include/spp_engine.h.
//
// Device layout (all per-particle arrays in CELL-SORTED slot order; id[slot] = caller's index):
//   state set (x3)  pos P4[N] (x, y, cos th, sin th), th double[N], U double2[N] (unwrapped
//                   positions, reference r_original), id int[N]
//   pos1/th1, pos2/th2  RK stage states (pos1 = RK2 midpoint)
//   dth    double[N]   this step's angular noise
//   kacc   P4[N]       RK4 accumulator k1 + 2 k2 + 2 k3 (fx, fy, tau)
//   nbr    int[cap*N]  full Verlet list, ELL, slot-major (nbr[k*N + i]): a wavefront's k-th
//                      neighbour ids are one coalesced 256-B read
//   build  float2[N]   positions at the last build, for the displacement trigger
// Cells are row-major (key = cy*ncx + cx), so each stencil row of three cells is one contiguous
// slot range, and each XCD (a contiguous 1/8 of the slots) owns a horizontal band of the box.
//
// Step scheduling (displacement-triggered rebuilds): work is enqueued in intervals, each a
// rebuild followed by m fused steps.  The kernels flag the first step whose input moved farther
// than the half skin; the host checks an interval's flags while the next one already runs, so
// the GPU never waits for the host.  The state at an interval's start is kept untouched (its
// first step writes into another of three state sets), so a late flag rolls back by switching
// sets and replaying only the valid steps -- the counter-based noise makes the replay exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
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

struct Flags {
    int bad_step;  // first step index (within the interval) whose stage input left the skin
    int max_len;   // longest full neighbour list of the interval's rebuild
    int nnz;       // total entries of that list
    int pad;
};

// ------------------------------------------------------------------------------------------------
// Small device utilities
// ------------------------------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// slots (a band of the box) so its L2 keeps that band's particles.
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

struct Grid {
    double Lx, Ly, Lx2, Ly2, invLx2, invLy2;
    int ncx, ncy, quirk;
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
                                                int* __restrict__ key, int* __restrict__ count,
                                                Flags* flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && flags) {
        flags->bad_step = INT_MAX;
        flags->max_len = 0;
        flags->nnz = 0;
    }
    if (i >= N) return;
    P4 p = pos[i];
    int cx = cell_coord(p.x, G.Lx, G.Lx2, G.ncx, G.quirk);
    int cy = cell_coord(p.y, G.Ly, G.Ly2, G.ncy, G.quirk);
    int k = cy * G.ncx + cx;
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
// sums of the blocks before it (few: n / 1024).
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

// Counting-sort scatter.  Counting down returns every cell count to zero, ready for the next
// rebuild without a clearing pass.
__global__ __launch_bounds__(kBlock) void k_scatter(int N, const int* __restrict__ key,
                                                    const int* __restrict__ start,
                                                    int* __restrict__ count, int* __restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int k = key[i];
    perm[start[k] + atomicSub(&count[k], 1) - 1] = i;
}

// Order each cell's particles by caller index, so the sorted layout is a pure function of the
// positions (the atomic scatter leaves arrival order).
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

struct SetPtrs {
    P4* pos;
    double* th;
    double2* U;
    int* id;
};

struct PermuteArgs {
    int N;
    const int* perm;
    SetPtrs in, out;
    // per-stage buffers of the split-call path (null when permuting between steps)
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
    a.out.pos[k] = a.in.pos[o];
    a.out.th[k] = a.in.th[o];
    a.out.U[k] = a.in.U[o];
    a.out.id[k] = a.in.id[o];
    if (a.d_in) {
        a.d_out[k] = a.d_in[o];
#pragma unroll
        for (int q = 0; q < 3; ++q) a.a_out[q][k] = a.a_in[q][o];
    }
}

struct NlistArgs {
    int N, cap;
    Grid G;
    double T_list;  // sqrt(d2) < r_list  <=>  d2 < T_list
    const P4* pos;
    const int* start;
    int* nbr;
    int* ncount;
    float2* build;
    int* cellxy;
    Flags* flags;
};

// Cell::Neighbor_List (shared/cell.h:74-103) over the 3x3 periodic stencil, full list, predicate
// sqrt(d2) < rv (cell.h:84-85) evaluated as d2 < T(rv).  Each stencil row is scanned as one or
// two contiguous slot ranges.
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
        // x ranges of the three stencil columns: [x0, x1] plus an optional wrapped cell xw
        int x0 = cx - 1, x1 = cx + 1, xw = -1;
        if (x0 < 0) {
            xw = G.ncx - 1;
            x0 = 0;
        } else if (x1 >= G.ncx) {
            xw = 0;
            x1 = G.ncx - 1;
        }
        for (int oy = -1; oy <= 1; ++oy) {
            int y = cy + oy;
            y = y < 0 ? y + G.ncy : (y >= G.ncy ? y - G.ncy : y);
            int row = y * G.ncx;
            for (int seg = 0; seg < 2; ++seg) {
                int s, e;
                if (seg == 0) {
                    s = a.start[row + x0];
                    e = a.start[row + x1 + 1];
                } else {
                    if (xw < 0) break;
                    s = a.start[row + xw];
                    e = a.start[row + xw + 1];
                }
                for (int j = s; j < e; ++j) {
                    if (j == i) continue;
                    P4 pj = a.pos[j];
                    double dx = spp::min_image(me.x - pj.x, G.Lx2, G.invLx2);
                    double dy = spp::min_image(me.y - pj.y, G.Ly2, G.invLy2);
                    double d2 = dx * dx + dy * dy;
                    if (d2 < a.T_list) {
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
        atomicMax(&a.flags->max_len, m);
        atomicAdd(&a.flags->nnz, t);
    }
}

// ------------------------------------------------------------------------------------------------
// Stage kernels: interact + Move fused, one particle per lane.
// ------------------------------------------------------------------------------------------------

struct StageArgs {
    int N;
    Phys P;
    const P4* gather;      // stage input: neighbours' (and own) position and heading
    const double* th_in;   // own heading angle of the stage input
    P4* out;               // stage output
    double* th_out;
    const double* th0;     // step-start heading (theta_old) for RK stages >= 2
    const double2* U_in;   // step-start unwrapped position
    double2* U_out;        // written by the final stage
    double* dth;           // angular noise of this step
    P4* kacc;              // RK4 accumulator
    const int* nbr;
    const int* ncount;
    const int* id;
    int* id_out;           // first step of an interval: the output set's copy of id
    const double* inj;     // injected noise, slot order (null = stream)
    const float2* build;
    int64_t step;          // noise-stream step
    int kstep;             // step index inside the current interval
    double thresh2;        // displacement trigger, squared
    Flags* flags;
    double* blockmax;      // per-block max displacement^2 (last step of an interval)
    int record;
};

template <int MODEL, int REP>
__device__ __forceinline__ void gather_forces(const StageArgs& a, int i, const P4& me, double& fx,
                                              double& fy, double& tq)
{
    double s1 = 0, s2 = 0;
    fx = me.c * a.P.v0;  // Reset: f = v (speed v0), particle.h:559-564
    fy = me.s * a.P.v0;
    int cnt = a.ncount[i];
    const int* nb = a.nbr + i;
    for (int k = 0; k < cnt; ++k) {
        int j = nb[(size_t)k * a.N];
        P4 pj = a.gather[j];
        spp::pair_accumulate<MODEL, REP>(a.P, me, pj, fx, fy, s1, s2);
    }
    tq = spp::drift_torque<MODEL>(a.P, s1, s2);
}

__device__ __forceinline__ void check_disp(const StageArgs& a, bool live, int i, double x,
                                           double y, int mark)
{
    __shared__ double red[kBlock / 64];
    double d2 = 0;
    if (live) {
        float2 b = a.build[i];
        double dx = spp::min_image(x - (double)b.x, a.P.Lx2, a.P.invLx2);
        double dy = spp::min_image(y - (double)b.y, a.P.Ly2, a.P.invLy2);
        d2 = dx * dx + dy * dy;
    }
    if (__any(d2 > a.thresh2) && (threadIdx.x & 63) == 0) atomicMin(&a.flags->bad_step, mark);
    if (a.record) {
        double m = wave_max(d2);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int k = 1; k < kBlock / 64; ++k) r = fmax(r, red[k]);
            a.blockmax[blockIdx.x] = r;
        }
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
        double fx, fy, tq;
        gather_forces<MODEL, REP>(a, i, me, fx, fy, tq);
        double th;
        if (STAGE <= 1) {
            int pid = a.id[i];
            if (a.id_out) a.id_out[i] = pid;
            double dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)pid, a.step);
            double th_in = a.th_in[i];
            if (INTEG == SPP_EULER) {
                th = th_in + P.dt * (tq);
                th += dn;
                double2 u = a.U_in[i];
                u.x += fx * P.dt;
                u.y += fy * P.dt;
                a.U_out[i] = u;
                x = spp::periodic_wrap(u.x, P.Lx2);
                y = spp::periodic_wrap(u.y, P.Ly2);
            } else {
                a.dth[i] = dn;
                th = th_in;
                th += P.half_dt * tq;
                th += dn / 2;
                th = spp::wrap_angle(th);
                x = spp::periodic_wrap(me.x + fx * P.half_dt, P.Lx2);
                y = spp::periodic_wrap(me.y + fy * P.half_dt, P.Ly2);
                if (INTEG == SPP_RK4) a.kacc[i] = P4{fx, fy, tq, 0.0};
            }
        } else {
            double thold = a.th0[i];
            double dn = a.dth[i];
            double2 u = a.U_in[i];
            double h = (INTEG == SPP_RK4 && STAGE == 2) ? P.half_dt : P.dt;
            if (INTEG == SPP_RK4 && STAGE == 4) {
                P4 k = a.kacc[i];
                th = thold + (k.c + tq) * P.dt6;
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
                    a.kacc[i] = P4{k.x + fx * 2, k.y + fy * 2, k.c + tq * 2, 0.0};
                }
            }
            th = spp::wrap_angle(th);
            if (INTEG == SPP_RK2 || STAGE == 4) a.U_out[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly2);
        }
        double sn, cs;
        sincos(th, &sn, &cs);
        a.out[i] = P4{x, y, cs, sn};
        a.th_out[i] = th;
    }
    constexpr bool final_stage = (INTEG == SPP_EULER) || (INTEG == SPP_RK2 && STAGE == 2) || (STAGE == 4);
    check_disp(a, live, i, x, y, final_stage ? a.kstep + 1 : a.kstep);
}

// ------------------------------------------------------------------------------------------------
// Split-call path: Box::Interact and Move_* as separate launches (drop-in granularity), on the
// current state set in place like the reference's particle array.
// ------------------------------------------------------------------------------------------------

template <int MODEL, int REP>
__global__ __launch_bounds__(kBlock) void k_interact(StageArgs a, P4* force)
{
    int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.N) return;
    P4 me = a.gather[i];
    double fx, fy, tq;
    gather_forces<MODEL, REP>(a, i, me, fx, fy, tq);
    force[i] = P4{fx, fy, tq, 0.0};
}

struct MoveArgs {
    int N;
    Phys P;
    int integ, stage;
    SetPtrs cur;
    P4* old;     // r_old.x, r_old.y, theta_old
    P4* kacc;
    double* dth;
    const P4* force;
    const double* inj;
    int64_t step;
};

__global__ __launch_bounds__(kBlock) void k_move(MoveArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const Phys& P = a.P;
    P4 me = a.cur.pos[i];
    double th_in = a.cur.th[i];
    P4 f = a.force[i];
    double2 u = a.cur.U[i];
    double th, x, y;
    if (a.stage <= 1) {
        double dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)a.cur.id[i], a.step);
        a.dth[i] = dn;
        if (a.integ == SPP_EULER) {
            th = th_in + P.dt * (f.c);
            th += dn;
            u.x += f.x * P.dt;
            u.y += f.y * P.dt;
            a.cur.U[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly2);
        } else {
            a.old[i] = P4{u.x, u.y, th_in, 0.0};
            th = th_in;
            th += P.half_dt * f.c;
            th += dn / 2;
            th = spp::wrap_angle(th);
            if (a.integ == SPP_RK4) a.kacc[i] = f;
            x = spp::periodic_wrap(me.x + f.x * P.half_dt, P.Lx2);
            y = spp::periodic_wrap(me.y + f.y * P.half_dt, P.Ly2);
        }
    } else {
        P4 o = a.old[i];
        double dn = a.dth[i];
        if (a.integ == SPP_RK4 && a.stage == 4) {
            P4 k = a.kacc[i];
            th = o.c + (k.c + f.c) * P.dt6;
            th += dn;
            u.x = o.x + (k.x + f.x) * P.dt6;
            u.y = o.y + (k.y + f.y) * P.dt6;
        } else {
            double h = (a.integ == SPP_RK4 && a.stage == 2) ? P.half_dt : P.dt;
            th = o.c + h * f.c;
            th += (a.integ == SPP_RK4 && a.stage == 2) ? dn / 2 : dn;
            u.x = o.x + f.x * h;
            u.y = o.y + f.y * h;
            if (a.integ == SPP_RK4) {
                P4 k = a.kacc[i];
                a.kacc[i] = P4{k.x + f.x * 2, k.y + f.y * 2, k.c + f.c * 2, 0.0};
            }
        }
        th = spp::wrap_angle(th);
        a.cur.U[i] = u;  // the reference writes r_original in every stage >= 2
        x = spp::periodic_wrap(u.x, P.Lx2);
        y = spp::periodic_wrap(u.y, P.Ly2);
    }
    double sn, cs;
    sincos(th, &sn, &cs);
    a.cur.pos[i] = P4{x, y, cs, sn};
    a.cur.th[i] = th;
}

// Headings -> (cos, sin) after a host upload, computed by the same device sincos as the stages.
__global__ __launch_bounds__(kBlock) void k_init_heading(int N, P4* __restrict__ pos,
                                                         const double* __restrict__ th)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double sn, cs;
    sincos(th[i], &sn, &cs);
    pos[i].c = cs;
    pos[i].s = sn;
}

// Scatter caller-order values into slot order: dst[slot] = src[id[slot]].
__global__ void k_gather_by_id(int N, const int* __restrict__ id, const double* __restrict__ src,
                               double* __restrict__ dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = src[id[i]];
}

__global__ __launch_bounds__(kBlock) void k_copy_set(int N, SetPtrs src, SetPtrs dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    dst.pos[i] = src.pos[i];
    dst.th[i] = src.th[i];
    dst.U[i] = src.U[i];
    dst.id[i] = src.id[i];
}

// Polarization partial sums (Box::Compute_Polarization, serial/box.h:531-538), two-pass.
__global__ __launch_bounds__(kBlock) void k_polar(int N, const P4* __restrict__ pos,
                                                  double2* __restrict__ part)
{
    __shared__ double sx[kBlock / 64], sy[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double vx = 0, vy = 0;
    if (i < N) {
        vx = pos[i].c;
        vy = pos[i].s;
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

// Smallest double T with fl(sqrt(T)) >= R: then sqrt(d2) < R  <=>  d2 < T for every d2 >= 0
// (correctly rounded sqrt is monotone), so the reference's predicates need no sqrt.
double sqrt_threshold(double R)
{
    if (!(R > 0)) return 0.0;
    double x = R * R;
    if (std::sqrt(x) >= R) {
        for (;;) {
            double p = std::nextafter(x, 0.0);
            if (std::sqrt(p) >= R)
                x = p;
            else
                break;
        }
    } else {
        while (std::sqrt(x) < R) x = std::nextafter(x, INFINITY);
    }
    return x;
}

}  // namespace

// ==================================================================================================
// Engine
// ==================================================================================================

struct Snapshot {
    SetPtrs s{nullptr, nullptr, nullptr, nullptr};
    int64_t step = 0;
    double t = 0;
    bool valid = false;
};

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

// One enqueued interval awaiting verification.
struct Interval {
    int start_set;   // set holding the state at the interval's start (kept untouched)
    int m;           // steps
    int slot;        // flag slot
    int64_t step0;
    int64_t since0;  // steps since the last rebuild at the interval's start (after its rebuild)
    double t0;
    bool rebuilt;
    hipEvent_t done;
};

struct spp_engine {
    spp_params p{};
    int N = 0;
    Grid G{};
    int nkeys = 0, nscan_blocks = 0, nblk = 0;
    int cap = 32;
    Phys P{};
    double r_cut = 0, half_skin = 0, thresh = 0, T_list = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    SetPtrs set[3] = {};
    int cur = 0;
    P4 *pos1 = nullptr, *pos2 = nullptr;
    double *th1 = nullptr, *th2 = nullptr;
    double *dth = nullptr, *dth_alt = nullptr, *inj_host_order = nullptr, *inj = nullptr;
    P4 *kacc = nullptr, *force = nullptr, *old = nullptr;
    P4* aux_alt[3] = {nullptr, nullptr, nullptr};
    int *nbr = nullptr, *ncount = nullptr, *cellxy = nullptr;
    float2* build = nullptr;
    int *key = nullptr, *perm = nullptr, *cell_count = nullptr, *cell_start = nullptr,
        *block_sums = nullptr;
    Flags* dflags = nullptr;     // 2 slots
    double* dblockmax = nullptr; // 2 slots x nblk
    Flags* hflags = nullptr;     // pinned mirrors
    double* hblockmax = nullptr;
    double2* dpolar = nullptr;
    hipEvent_t slot_event[2] = {nullptr, nullptr};
    std::map<int, Snapshot> slots;

    int64_t step = 0;
    double t = 0;
    int64_t rebuilds = 0, rollbacks = 0, host_syncs = 0;
    int64_t steps_since_build = 0;
    int64_t intervals = 0;
    bool have_state = false, list_valid = false;
    bool inj_pending = false;
    double T_cross = 0;          // estimated steps from a rebuild to the displacement trigger
    double T_hist[3] = {0, 0, 0};
    int force_m = 0;             // length of the next interval after a rollback (0 = predict)
    int max_nbrs = 0;
    double mean_nbrs = 0;
    int unfused_stage = 0;       // next expected stage of the split-call path (0 = step boundary)

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

#define TRY(x)                         \
    do {                               \
        int _rc = (x);                 \
        if (_rc != SPP_OK) return _rc; \
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

int alloc_set(spp_engine* e, SetPtrs& s)
{
    TRY(dalloc(e, &s.pos, e->N));
    TRY(dalloc(e, &s.th, e->N));
    TRY(dalloc(e, &s.U, e->N));
    TRY(dalloc(e, &s.id, e->N));
    return SPP_OK;
}

void free_set(SetPtrs& s)
{
    dfree(s.pos);
    dfree(s.th);
    dfree(s.U);
    dfree(s.id);
    s = SetPtrs{nullptr, nullptr, nullptr, nullptr};
}

// ------------------------------------------------------------------------------------------------
// Rebuild: bin, scan, scatter, per-cell order, permute (set src -> set dst), Verlet list.
// ------------------------------------------------------------------------------------------------
int launch_nlist(spp_engine* e, int slot)
{
    NlistArgs na{e->N, e->cap, e->G, e->T_list, e->set[e->cur].pos, e->cell_start, e->nbr,
                 e->ncount, e->build, e->cellxy, e->dflags + slot};
    return launch(e, "nlist", k_nlist, nblocks(e->N), na);
}

int launch_rebuild(spp_engine* e, int dst, int slot, bool aux)
{
    int N = e->N, nk = e->nkeys, src = e->cur;
    TRY(launch(e, "bin", k_bin, nblocks(N), N, e->G, (const P4*)e->set[src].pos, e->key,
               e->cell_count, e->dflags + slot));
    TRY(launch(e, "scan_partial", k_scan_partial, e->nscan_blocks, (const int*)e->cell_count, nk,
               e->block_sums));
    TRY(launch(e, "scan_final", k_scan_final, e->nscan_blocks, (const int*)e->cell_count, nk,
               (const int*)e->block_sums, e->cell_start));
    TRY(launch(e, "scatter", k_scatter, nblocks(N), N, (const int*)e->key,
               (const int*)e->cell_start, e->cell_count, e->perm));
    TRY(launch(e, "cell_sort", k_cell_sort, nblocks(nk), nk, (const int*)e->cell_start, e->perm,
               (const int*)e->set[src].id));
    PermuteArgs pa{};
    pa.N = N;
    pa.perm = e->perm;
    pa.in = e->set[src];
    pa.out = e->set[dst];
    if (aux) {
        pa.a_in[0] = e->kacc;
        pa.a_in[1] = e->force;
        pa.a_in[2] = e->old;
        for (int q = 0; q < 3; ++q) pa.a_out[q] = e->aux_alt[q];
        pa.d_in = e->dth;
        pa.d_out = e->dth_alt;
    }
    TRY(launch(e, "permute", k_permute, nblocks(N), pa));
    if (aux) {
        std::swap(e->kacc, e->aux_alt[0]);
        std::swap(e->force, e->aux_alt[1]);
        std::swap(e->old, e->aux_alt[2]);
        std::swap(e->dth, e->dth_alt);
    }
    e->cur = dst;
    TRY(launch_nlist(e, slot));
    e->rebuilds++;
    e->steps_since_build = 0;
    e->list_valid = true;
    return SPP_OK;
}

int free_set_other_than(int a, int b)
{
    for (int k = 0; k < 3; ++k)
        if (k != a && k != b) return k;
    return -1;
}

// Synchronous rebuild with a capacity check (used outside the interval scheduler).
int rebuild_sync(spp_engine* e, bool aux)
{
    int dst = free_set_other_than(e->cur, -1);
    TRY(launch_rebuild(e, dst, 0, aux));
    for (;;) {
        HIPC(e, hipMemcpyAsync(e->hflags, e->dflags, sizeof(Flags), hipMemcpyDeviceToHost, e->stream));
        HIPC(e, hipStreamSynchronize(e->stream));
        e->host_syncs++;
        int need = e->hflags[0].max_len;
        e->max_nbrs = need;
        e->mean_nbrs = (double)(unsigned)e->hflags[0].nnz / e->N;
        if (need <= e->cap) return SPP_OK;
        e->cap = ((need + 7) / 8) * 8 + 8;
        TRY(alloc_nbr(e));
        TRY(launch(e, "fill", k_fill<int>, 1, (int*)e->dflags, 0, 4));
        TRY(launch_nlist(e, 0));
    }
}

int take_injected(spp_engine* e)
{
    // injected noise arrives in caller order; put it in slot order with the current permutation
    return launch(e, "gather_noise", k_gather_by_id, nblocks(e->N), e->N, (const int*)e->set[e->cur].id,
                  (const double*)e->inj_host_order, e->inj);
}

// Launch one fused step.  S = set holding the step's input state, W = set receiving its output
// (S == W except for the first step of an interval).
int launch_step(spp_engine* e, int S, int W, int kstep, int slot, bool record)
{
    int integ = e->p.integrator;
    StageArgs a{};
    a.N = e->N;
    a.P = e->P;
    a.dth = e->dth;
    a.kacc = e->kacc;
    a.nbr = e->nbr;
    a.ncount = e->ncount;
    a.id = e->set[S].id;
    a.build = e->build;
    a.step = e->step;
    a.kstep = kstep;
    a.thresh2 = e->thresh * e->thresh;
    a.flags = e->dflags + slot;
    a.blockmax = e->dblockmax + (size_t)slot * e->nblk;
    a.th0 = e->set[S].th;
    a.U_in = e->set[S].U;
    a.U_out = e->set[W].U;
    if (e->inj_pending) {
        TRY(take_injected(e));
        a.inj = e->inj;
    }
    // the id array does not change within an interval; the output set gets its own copy
    a.id_out = S != W ? e->set[W].id : nullptr;
    int nst = integ == SPP_EULER ? 1 : integ;
    for (int s = 1; s <= nst; ++s) {
        int stage = integ == SPP_EULER ? 0 : s;
        a.record = (record && s == nst) ? 1 : 0;
        if (integ == SPP_EULER) {
            a.gather = e->set[S].pos;
            a.th_in = e->set[S].th;
            a.out = e->pos1;
            a.th_out = e->th1;
        } else if (integ == SPP_RK2) {
            a.gather = s == 1 ? e->set[S].pos : e->pos1;
            a.th_in = s == 1 ? e->set[S].th : e->th1;
            a.out = s == 1 ? e->pos1 : e->set[W].pos;
            a.th_out = s == 1 ? e->th1 : e->set[W].th;
        } else {
            const P4* g[5] = {nullptr, e->set[S].pos, e->pos1, e->pos2, e->pos1};
            const double* gt[5] = {nullptr, e->set[S].th, e->th1, e->th2, e->th1};
            P4* o[5] = {nullptr, e->pos1, e->pos2, e->pos1, e->set[W].pos};
            double* ot[5] = {nullptr, e->th1, e->th2, e->th1, e->set[W].th};
            a.gather = g[s];
            a.th_in = gt[s];
            a.out = o[s];
            a.th_out = ot[s];
        }
        StageFn fn = pick_stage(e->p.model, e->p.repulsion, integ, stage);
        TRY(launch(e, stage_name(integ, stage), fn, e->nblk, a));
        a.inj = nullptr;
        a.id_out = nullptr;
    }
    if (integ == SPP_EULER) {
        std::swap(e->set[W].pos, e->pos1);
        std::swap(e->set[W].th, e->th1);
    }
    e->cur = W;
    e->step++;
    e->t += e->p.dt;
    e->steps_since_build++;
    return SPP_OK;
}

int copy_set(spp_engine* e, const SetPtrs& src, const SetPtrs& dst)
{
    return launch(e, "copy_set", k_copy_set, nblocks(e->N), e->N, src, dst);
}

int check_ready(spp_engine* e)
{
    if (!e) return SPP_ERR_ARG;
    if (!e->have_state) return fail(e, SPP_ERR_STATE, "no particle state: call spp_set_state first");
    return SPP_OK;
}

// Enqueue one interval: [rebuild] + m steps + flag read-back.  Returns the interval record.
int enqueue_interval(spp_engine* e, int m, bool rebuild, int protect, Interval& iv)
{
    int slot = (int)(e->intervals++ & 1);
    iv.slot = slot;
    iv.m = m;
    iv.rebuilt = rebuild;
    if (rebuild) {
        int dst = free_set_other_than(e->cur, protect);
        TRY(launch_rebuild(e, dst, slot, false));
    } else {
        // flags are reset by k_bin on a rebuild; otherwise reset them here
        TRY(launch(e, "fill", k_fill<int>, 1, (int*)(e->dflags + slot), 0, 4));
        TRY(launch(e, "fill", k_fill<int>, 1, (int*)(e->dflags + slot), INT_MAX, 1));
    }
    iv.start_set = e->cur;
    iv.step0 = e->step;
    iv.since0 = e->steps_since_build;
    iv.t0 = e->t;
    int W = free_set_other_than(e->cur, protect);
    for (int k = 0; k < m; ++k) {
        TRY(launch_step(e, e->cur, k == 0 ? W : e->cur, k, slot, k == m - 1));
        if (k == 0) e->inj_pending = false;
    }
    HIPC(e, hipMemcpyAsync(e->hflags + slot, e->dflags + slot, sizeof(Flags), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(e->hblockmax + (size_t)slot * e->nblk, e->dblockmax + (size_t)slot * e->nblk,
                           e->nblk * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipEventRecord(e->slot_event[slot], e->stream));
    iv.done = e->slot_event[slot];
    return SPP_OK;
}

int predict_m(spp_engine* e, int remaining)
{
    int m;
    if (e->force_m > 0) {
        m = e->force_m;
        e->force_m = 0;
    } else if (e->p.rebuild == SPP_REBUILD_PERIOD) {
        m = std::max(1, e->p.rebuild_period - (int)e->steps_since_build);
    } else if (e->T_cross <= 0) {
        m = 4;  // no history yet: a short interval to measure the displacement rate
    } else {
        m = std::max(1, (int)(0.85 * e->T_cross));
    }
    return std::min(std::min(m, remaining), 4096);
}

// Steps through the interval scheduler (both rebuild policies).
int run_intervals(spp_engine* e, int steps)
{
    std::deque<Interval> q;
    int to_enqueue = steps;
    bool period = e->p.rebuild == SPP_REBUILD_PERIOD;
    while (to_enqueue > 0 || !q.empty()) {
        if (to_enqueue > 0 && q.size() < 2) {
            int m = predict_m(e, to_enqueue);
            bool rebuild = !e->list_valid || e->steps_since_build > 0;
            if (period) rebuild = !e->list_valid || e->steps_since_build >= e->p.rebuild_period;
            int protect = q.empty() ? -1 : q.front().start_set;
            Interval iv{};
            TRY(enqueue_interval(e, m, rebuild, protect, iv));
            q.push_back(iv);
            to_enqueue -= m;
            continue;
        }
        Interval iv = q.front();
        HIPC(e, hipEventSynchronize(iv.done));
        e->host_syncs++;
        const Flags f = e->hflags[iv.slot];
        bool overflow = iv.rebuilt && f.max_len > e->cap;
        int bad = period ? INT_MAX : f.bad_step;
        if (iv.rebuilt) {
            e->max_nbrs = f.max_len;
            e->mean_nbrs = (double)(unsigned)f.nnz / e->N;
        }
        if (!overflow && bad >= iv.m) {
            q.pop_front();
            if (!period && iv.rebuilt) {
                // linear extrapolation of the largest displacement to the trigger
                const double* bm = e->hblockmax + (size_t)iv.slot * e->nblk;
                double d2 = 0;
                for (int b = 0; b < e->nblk; ++b) d2 = std::max(d2, bm[b]);
                double disp = std::sqrt(d2);
                double est = disp > 0 ? e->thresh / disp * iv.m : 4096.0;
                e->T_hist[2] = e->T_hist[1];
                e->T_hist[1] = e->T_hist[0];
                e->T_hist[0] = est;
                double T = est;
                for (double h : e->T_hist)
                    if (h > 0) T = std::min(T, h);
                e->T_cross = T;
            }
            continue;
        }
        // Roll back to the interval's start; it and everything enqueued after it are redone.
        HIPC(e, hipStreamSynchronize(e->stream));
        int redo = 0;
        for (auto& x : q) redo += x.m;
        q.clear();
        e->cur = iv.start_set;
        e->step = iv.step0;
        e->t = iv.t0;
        e->steps_since_build = iv.since0;
        e->rollbacks++;
        if (e->p.noise_mode == SPP_NOISE_INJECTED) e->inj_pending = true;  // replay uses it again
        to_enqueue += redo;
        e->list_valid = false;  // the device list may belong to a later interval
        if (overflow) {
            // same rebuild point (the interval started with one), larger capacity
            e->cap = ((f.max_len + 7) / 8) * 8 + 8;
            TRY(alloc_nbr(e));
            e->force_m = iv.m;
            continue;
        }
        if (bad == 0 && iv.rebuilt)
            return fail(e, SPP_ERR_STATE,
                        "a particle moves farther than the half skin (%g) within one step: "
                        "raise r_list or lower dt", e->half_skin);
        // The rebuild at the interval's start (identical, the state there is already sorted) is
        // redone; the valid prefix [0, bad) is replayed as its own interval and verified again.
        e->force_m = bad > 0 ? bad : 0;
        e->T_cross = std::max(1.0, 0.8 * (iv.since0 + std::max(bad, 1)));
        for (double& h : e->T_hist) h = h > 0 ? std::min(h, e->T_cross) : h;
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
    P.invLx2 = 1.0 / P.Lx2;
    P.invLy2 = 1.0 / P.Ly2;
    P.dt = p.dt;
    P.half_dt = p.dt / 2;  // shared/parameters.h:78
    P.dt6 = p.dt / 6;      // shared/parameters.h:79
    P.v0 = p.v0;
    P.g = p.g;
    P.Ra2 = p.R_align * p.R_align;
    if (p.model == SPP_MODEL_CONTINUOUS) {
        P.c_align = (1 - p.alpha) / spp::kPi;
        P.c_self = p.alpha / spp::kPi;
    } else {
        P.c_align = p.g / (spp::kPi * p.R_align * p.R_align);
        P.c_self = 0;
    }
    P.T_align = sqrt_threshold(p.R_align);
    P.T_rep = p.repulsion != SPP_REP_NONE ? sqrt_threshold(p.rc_rep) : 0.0;
    P.T_max = std::max(P.T_align, P.T_rep);
    P.rc = p.rc_rep;
    P.sigma = p.sigma_rep;
    P.inv_sigma = 1.0 / p.sigma_rep;
    P.A = p.A_rep;
    double sc = p.sigma_rep / p.rc_rep, sc2 = sc * sc, sc4 = sc2 * sc2, sc8 = sc4 * sc4;
    P.rc13 = sc8 * sc4 * sc;
    P.yuk_c = std::exp(-p.rc_rep / p.sigma_rep) *
              (1. / (p.rc_rep * p.rc_rep) + 1. / (p.sigma_rep * p.rc_rep));
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
    if (p->rebuild != SPP_REBUILD_DISPLACEMENT && p->rebuild != SPP_REBUILD_PERIOD)
        return fail(nullptr, SPP_ERR_ARG, "rebuild %d", p->rebuild);
    if (p->rebuild == SPP_REBUILD_PERIOD && p->rebuild_period <= 0)
        return fail(nullptr, SPP_ERR_ARG, "rebuild_period must be positive");
    double rcut = max_cutoff(*p);
    if (p->r_list < rcut) return fail(nullptr, SPP_ERR_ARG, "r_list %g below cutoff %g", p->r_list, rcut);
    int ncx = p->ncell_x > 0 ? p->ncell_x : (int)std::floor(2 * p->Lx / p->r_list);
    int ncy = p->ncell_y > 0 ? p->ncell_y : (int)std::floor(2 * p->Ly / p->r_list);
    // Cell::Cell (shared/cell.h:45-49): cells no narrower than rv.  Three or more per side keep
    // the nine stencil cells distinct.
    if (ncx < 3 || ncy < 3 || 2 * p->Lx / ncx < p->r_list || 2 * p->Ly / ncy < p->r_list)
        return fail(nullptr, SPP_ERR_ARG, "cell grid %dx%d invalid for r_list %g in box %gx%g", ncx,
                    ncy, p->r_list, 2 * p->Lx, 2 * p->Ly);
    if ((int64_t)ncx * ncy >= (1 << 30) || ncx >= 32768 || ncy >= 32768)
        return fail(nullptr, SPP_ERR_ARG, "cell grid too large");
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
    e->G.invLx2 = 1.0 / e->G.Lx2;
    e->G.invLy2 = 1.0 / e->G.Ly2;
    e->G.ncx = ncx;
    e->G.ncy = ncy;
    e->G.quirk = p->binning == SPP_BIN_REF_QUIRK;
    e->nkeys = ncx * ncy;
    e->nscan_blocks = (e->nkeys + kScanTile - 1) / kScanTile;
    e->nblk = nblocks(e->N);
    e->cap = p->ell_capacity > 0 ? p->ell_capacity : 32;
    e->P = make_phys(*p);
    e->T_list = sqrt_threshold(p->r_list);
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
    for (int k = 0; k < 3; ++k) A(alloc_set(e, e->set[k]));
    A(dalloc(e, &e->pos1, N));
    A(dalloc(e, &e->pos2, N));
    A(dalloc(e, &e->th1, N));
    A(dalloc(e, &e->th2, N));
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
    A(dalloc(e, &e->cell_start, (size_t)e->nkeys + 1));
    A(dalloc(e, &e->block_sums, (size_t)e->nscan_blocks));
    A(dalloc(e, &e->dflags, 2));
    A(dalloc(e, &e->dblockmax, 2 * (size_t)e->nblk));
    A(dalloc(e, &e->dpolar, (size_t)e->nblk));
    A(alloc_nbr(e));
    if (rc == SPP_OK && hipHostMalloc((void**)&e->hflags, 2 * sizeof(Flags), hipHostMallocDefault) != hipSuccess)
        rc = SPP_ERR_HIP;
    if (rc == SPP_OK &&
        hipHostMalloc((void**)&e->hblockmax, 2 * (size_t)e->nblk * sizeof(double), hipHostMallocDefault) != hipSuccess)
        rc = SPP_ERR_HIP;
    for (int s = 0; s < 2 && rc == SPP_OK; ++s)
        if (hipEventCreateWithFlags(&e->slot_event[s], hipEventDisableTiming) != hipSuccess) rc = SPP_ERR_HIP;
    if (rc == SPP_OK) {
        // cell counts start at zero and every rebuild's scatter returns them to zero
        if (launch(e, "fill", k_fill<int>, nblocks(e->nkeys), e->cell_count, 0, e->nkeys) != SPP_OK ||
            hipStreamSynchronize(e->stream) != hipSuccess)
            rc = SPP_ERR_HIP;
    }
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
    for (int k = 0; k < 3; ++k) free_set(e->set[k]);
    void* bufs[] = {e->pos1, e->pos2, e->th1, e->th2, e->dth, e->dth_alt, e->inj_host_order,
                    e->inj, e->kacc, e->force, e->old, e->aux_alt[0], e->aux_alt[1], e->aux_alt[2],
                    e->nbr, e->ncount, e->cellxy, e->build, e->key, e->perm, e->cell_count,
                    e->cell_start, e->block_sums, e->dflags, e->dblockmax, e->dpolar};
    for (void* b : bufs) dfree(b);
    for (auto& kv : e->slots) free_set(kv.second.s);
    for (auto& r : e->prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
    for (auto ev : e->slot_event)
        if (ev) (void)hipEventDestroy(ev);
    if (e->hflags) (void)hipHostFree(e->hflags);
    if (e->hblockmax) (void)hipHostFree(e->hblockmax);
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
        pos[i] = P4{xw[i], yw[i], 0.0, 0.0};
        U[i] = make_double2(xu ? xu[i] : xw[i], yu ? yu[i] : yw[i]);
        id[i] = i;
    }
    SetPtrs& s = e->set[e->cur];
    HIPC(e, hipMemcpyAsync(s.pos, pos.data(), n * sizeof(P4), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(s.th, theta, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(s.U, U.data(), n * sizeof(double2), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(s.id, id.data(), n * sizeof(int), hipMemcpyHostToDevice, e->stream));
    TRY(launch(e, "init_heading", k_init_heading, nblocks(n), n, s.pos, (const double*)s.th));
    e->have_state = true;
    e->unfused_stage = 0;
    e->inj_pending = false;
    return rebuild_sync(e, false);
}

int spp_get_state(spp_engine* e, double* xw, double* yw, double* xu, double* yu, double* theta,
                  int32_t n)
{
    TRY(check_ready(e));
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n = %d, engine has %d particles", n, e->N);
    std::vector<P4> pos(n);
    std::vector<double> th(n);
    std::vector<double2> U(n);
    std::vector<int> id(n);
    const SetPtrs& s = e->set[e->cur];
    HIPC(e, hipMemcpyAsync(pos.data(), s.pos, n * sizeof(P4), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(th.data(), s.th, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(U.data(), s.U, n * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), s.id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < n; ++k) {
        int i = id[k];
        if (xw) xw[i] = pos[k].x;
        if (yw) yw[i] = pos[k].y;
        if (xu) xu[i] = U[k].x;
        if (yu) yu[i] = U[k].y;
        if (theta) theta[i] = th[k];
    }
    return SPP_OK;
}

int spp_build_neighbors(spp_engine* e)
{
    TRY(check_ready(e));
    return rebuild_sync(e, true);
}

int spp_get_neighbors(spp_engine* e, int32_t* offsets, int32_t* idx, int64_t cap, int64_t* nnz)
{
    TRY(check_ready(e));
    if (!e->list_valid) TRY(spp_build_neighbors(e));
    int N = e->N, C = e->cap;
    std::vector<int> nb((size_t)C * N), cnt(N), id(N);
    HIPC(e, hipMemcpyAsync(nb.data(), e->nbr, nb.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(cnt.data(), e->ncount, N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->set[e->cur].id, N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
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
    HIPC(e, hipMemcpyAsync(id.data(), e->set[e->cur].id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
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
    StageArgs a{};
    a.N = e->N;
    a.P = e->P;
    a.gather = e->set[e->cur].pos;
    a.nbr = e->nbr;
    a.ncount = e->ncount;
    InteractFn fn = pick_interact(e->p.model, e->p.repulsion);
    return launch(e, "interact", fn, e->nblk, a, e->force);
}

int spp_get_forces(spp_engine* e, double* fx, double* fy, double* tau, int32_t n)
{
    TRY(check_ready(e));
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n mismatch");
    std::vector<P4> f(n);
    std::vector<int> id(n);
    HIPC(e, hipMemcpyAsync(f.data(), e->force, n * sizeof(P4), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->set[e->cur].id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < n; ++k) {
        int i = id[k];
        if (fx) fx[i] = f[k].x;
        if (fy) fy[i] = f[k].y;
        if (tau) tau[i] = f[k].c;
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
    m.cur = e->set[e->cur];
    m.old = e->old;
    m.kacc = e->kacc;
    m.dth = e->dth;
    m.force = e->force;
    m.inj = nullptr;
    m.step = e->step;
    if (stage <= 1 && e->inj_pending) {
        TRY(take_injected(e));
        m.inj = e->inj;
        e->inj_pending = false;
    }
    TRY(launch(e, "move", k_move, e->nblk, m));
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
    return run_intervals(e, steps);
}

int spp_one_step(spp_engine* e) { return spp_multi_step(e, 1); }

int spp_polarization(spp_engine* e, double* p, double* sum_x, double* sum_y)
{
    TRY(check_ready(e));
    TRY(launch(e, "polar", k_polar, e->nblk, e->N, (const P4*)e->set[e->cur].pos, e->dpolar));
    std::vector<double2> part(e->nblk);
    HIPC(e, hipMemcpyAsync(part.data(), e->dpolar, e->nblk * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
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
    TRY(rebuild_sync(e, false));
    Snapshot& s = e->slots[slot];
    if (!s.s.pos) TRY(alloc_set(e, s.s));
    TRY(copy_set(e, e->set[e->cur], s.s));
    s.step = e->step;
    s.t = e->t;
    s.valid = true;
    return SPP_OK;
}

int spp_restore(spp_engine* e, int32_t slot)
{
    TRY(check_ready(e));
    auto it = e->slots.find(slot);
    if (it == e->slots.end() || !it->second.valid) return fail(e, SPP_ERR_ARG, "no snapshot in slot %d", slot);
    TRY(copy_set(e, it->second.s, e->set[e->cur]));
    e->step = it->second.step;
    e->t = it->second.t;
    e->unfused_stage = 0;
    // Box::Load ends in a cell update (parallel/box.h:241)
    return rebuild_sync(e, false);
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
