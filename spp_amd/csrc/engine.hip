// engine.hip -- MI355X (gfx950) engine for the per-time-step SPP particle path, behind the C ABI
// declared in include/spp_engine.h.
//
// Device layout (all per-particle arrays in CELL-SORTED slot order; id[slot] = caller's index):
//   state set (x3-4)  xy double2[N] (wrapped positions) and th double[N] (heading angles): the
//                   record neighbours gather, 24 B in two arrays; U double2[N] (unwrapped
//                   positions, reference r_original); id int[N]
//   xy1/th1, xy2/th2  RK stage states (xy1 = RK2 midpoint); the step's starting angle
//                   theta_old is the input set's th, which no stage overwrites before the last
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
#include <atomic>
#include <climits>
#include <limits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "spp_device.hpp"
#include "binning.hpp"

using spp_bin::cell_coord;

using spp::P4;
using spp::Phys;

namespace {

constexpr int kBlock = 256;
// below this many particles (six wavefronts per SIMD of the device) the stage kernels keep eight
// neighbour records in flight per lane instead of four (SPP_KERNEL_SLOT_DEEP: at every size); at
// 262,144 particles eight give 4 % shorter RK2 stages, at 1,048,576 four are 5 % faster
constexpr int kSmallN = 393216;
// up to this many particles the stage kernels write their outputs with streaming stores: the
// launch is short, and ordinary stores would leave its outputs dirty in the L2s until the
// write-back at its end.  Larger engines keep ordinary stores, which leave the outputs in the
// memory-side cache for the next launch: per step, streaming was 4.6 % faster at 262,144
// particles, 0.8 % slower at 393,216, 9 % slower at 524,288 and 17 % slower at 1,048,576.
constexpr int kStreamMaxN = 262144;
constexpr int kScanItems = 4;  // items per thread in the cell-count scan
constexpr int kSets = 4;       // particle state sets (see spp_engine::set)
constexpr int kScanTile = kBlock * kScanItems;

// One interval's bookkeeping, reduced on the device from per-block partials (no same-address
// atomics across the chip) and written straight into host-mapped memory as tagged words (k_report),
// which the host decodes into this record.
struct Report {
    int bad_step;     // first step index (within the interval) whose stage input left the skin
    int max_len;      // longest full neighbour list of the interval's rebuild
    long long nnz;    // total entries of that list
    double maxdisp2;  // largest squared displacement at the interval's end
    int sort_fail;    // the interval's rebuild took the tile path and could not sort (redo it)
    int pad;
    long long seq;    // the slot's report count
};

// 64-bit words of one slot's report in host memory: the report number in the upper half of each,
// one 32-bit value in the lower half (bad step, longest list, entries low / high, largest squared
// displacement low / high bits, rebuild failure flag).
constexpr int kRepWords = 8;
constexpr int kRepUsed = 7;

struct SlotBufs {
    int* blkbad;      // per-block first bad step (atomicMin inside a block only)
    double* blkmax;   // per-block max squared displacement
    int* blklen;      // per-block longest list
    int* blknnz;      // per-block list entries
};

// One interval's report: its slot's partials, the slot's rebuild failure flag, the slot's report
// count and the host-mapped words it is written to.  An interval's report rides on the next
// interval's first launch (the tile sort or the first stage kernel) as one extra workgroup after the
// launch's own, which saves a kernel launch per interval; the host learns the result while that
// launch runs, and launches the report on its own (k_report) when no such launch follows.
struct ReportArgs {
    int nblk;                  // partials to reduce
    SlotBufs b;
    int with_list;             // the interval rebuilt: list statistics too
    int* sort_fail;            // the slot's tile-sort failure flag (read, then cleared)
    long long* count;          // the slot's report count on the device
    unsigned long long* out;   // the slot's host-mapped words
};

// Reduces one interval's per-block partials into the slot's host-mapped report words (one
// workgroup of kBlock threads) and clears the per-block bad steps for the slot's next interval.
// Each word carries the report number (the slot's count of reports, kept on the device so that
// replays of a captured interval count too) in its upper half and goes out as one system-scope
// store, so the host takes the report once every word's tag matches, with no fence on the device
// between them.  The host polls the words instead of waiting on an event: an event record between
// two intervals idles the GPU for about 5 us on this device.
__device__ void report_body(const ReportArgs& r)
{
    const int nblk = r.nblk;
    const SlotBufs& b = r.b;
    const int with_list = r.with_list;
    int* sort_fail = r.sort_fail;
    long long* count = r.count;
    unsigned long long* out = r.out;
    __shared__ int r_bad[kBlock / 64], r_len[kBlock / 64];
    __shared__ long long r_nnz[kBlock / 64];
    __shared__ double r_max[kBlock / 64];
    int bad = INT_MAX, len = 0;
    long long nnz = 0;
    double mx = 0;
    for (int k = threadIdx.x; k < nblk; k += kBlock) {
        bad = min(bad, b.blkbad[k]);
        b.blkbad[k] = INT_MAX;  // the slot's next interval starts with no bad step
        mx = fmax(mx, b.blkmax[k]);
        if (with_list) {
            len = max(len, b.blklen[k]);
            nnz += b.blknnz[k];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        bad = min(bad, __shfl_xor(bad, o));
        len = max(len, __shfl_xor(len, o));
        nnz += __shfl_xor(nnz, o);
        mx = fmax(mx, __shfl_xor(mx, o));
    }
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        r_bad[w] = bad;
        r_len[w] = len;
        r_nnz[w] = nnz;
        r_max[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            bad = min(bad, r_bad[k]);
            len = max(len, r_len[k]);
            nnz += r_nnz[k];
            mx = fmax(mx, r_max[k]);
        }
        bad = min(bad, r_bad[0]);
        len = max(len, r_len[0]);
        mx = fmax(mx, r_max[0]);
        nnz = 0;
        for (int k = 0; k < kBlock / 64; ++k) nnz += r_nnz[k];
        const int fail = *sort_fail;
        *sort_fail = 0;  // the next interval's rebuild starts clear
        const long long seq = *count + 1;
        *count = seq;
        const unsigned long long tag = (unsigned long long)(unsigned)seq << 32;
        const unsigned long long n = (unsigned long long)nnz;
        const unsigned long long bits = (unsigned long long)__double_as_longlong(mx);
        const unsigned long long w[kRepUsed] = {tag | (unsigned)bad,       tag | (unsigned)len,
                                                tag | (n & 0xffffffffull), tag | (n >> 32),
                                                tag | (bits & 0xffffffffull), tag | (bits >> 32),
                                                tag | (unsigned)fail};
#pragma unroll
        for (int k = 0; k < kRepUsed; ++k) __hip_atomic_store(out + k, w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ------------------------------------------------------------------------------------------------
// Small device utilities
// ------------------------------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// slots (a band of the box) so its L2 keeps that band's particles.
__device__ __forceinline__ int xcd_block_of(int nblk)
{
    const int b = blockIdx.x;
    if ((nblk & 7) == 0) return (b & 7) * (nblk >> 3) + (b >> 3);
    return b;
}

__device__ __forceinline__ int xcd_block() { return xcd_block_of(gridDim.x); }

// Record j of a position array by a 32-bit byte offset from its (uniform) base (engines hold fewer
// than 2^27 particles), which the load takes as a VGPR offset to an SGPR base.
__device__ __forceinline__ double2 load_pos(const double2* base, int j)
{
    return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) + (size_t)((unsigned)j * 16u));
}

struct Grid {
    double Lx, Ly, Lx2, Ly2, invLx2, invLy2;
    int ncx, ncy, quirk;
};

// Per-system constants of a batched engine (one independent periodic box per system).  Systems
// are concatenated; cell keys are system-major, so after every sort system s occupies the slots
// [pid_off, pid_off + n) and its cells the keys [key_off, key_off + ncx*ncy).
struct SysInfo {
    Grid G;
    int key_off, pid_off, n, pad;
    double v0, c_align, g, noise_amp;
    uint64_t seed;
};

// Overrides the per-system fields of the launch constants for a batched engine.
__device__ __forceinline__ void apply_sys(Phys& P, const SysInfo& S)
{
    P.Lx = S.G.Lx;
    P.Ly = S.G.Ly;
    P.Lx2 = S.G.Lx2;
    P.Ly2 = S.G.Ly2;
    P.invLx2 = S.G.invLx2;
    P.invLy2 = S.G.invLy2;
    P.v0 = S.v0;
    P.c_align = S.c_align;
    P.g = S.g;
    P.noise_amp = S.noise_amp;
    P.seed = S.seed;
}

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

__device__ __forceinline__ long long wave_sum_ll(long long v)
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

// Cell key of every particle and its arrival rank among its cell mates, so that the scatter needs
// no atomics.  Slots arrive sorted by the previous cells and particles rarely change cell between
// rebuilds, so equal keys mostly form runs of neighbouring lanes: the first lane of a run reserves
// the whole run with one atomic and its lanes take consecutive ranks (equal keys in separate runs
// reserve separately; any arrival order is fine, the id order inside cells is restored later).
__global__ __launch_bounds__(kBlock) void k_bin(int N, Grid G0, const double2* __restrict__ pos,
                                                int* __restrict__ key, int* __restrict__ count,
                                                int* __restrict__ rank, const SysInfo* __restrict__ sys_tab,
                                                const int* __restrict__ sys_of)
{
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = i < N;  // every lane takes part in the wave's vote and shuffles
    int k = -1;
    if (live) {
        Grid G = G0;
        int koff = 0;
        if (sys_of) {
            const SysInfo& S = sys_tab[sys_of[i]];
            G = S.G;
            koff = S.key_off;
        }
        const double2 p = pos[i];
        int cx = cell_coord(p.x, G.Lx, G.Lx2, G.invLx2, G.ncx, G.quirk);
        int cy = cell_coord(p.y, G.Ly, G.Ly2, G.invLy2, G.ncy, G.quirk);
        k = koff + cy * G.ncx + cx;
    }
    const int prev = __shfl_up(k, 1, 64);
    const bool head = !live || lane == 0 || prev != k;  // dead lanes end runs and reserve nothing
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
    const int next = above ? __builtin_ctzll(above) : 64;  // the next run's first lane
    int base = 0;
    if (live && head) base = atomicAdd(&count[k], next - lane);
    const unsigned long long upto = lane == 63 ? heads : heads & ((2ull << lane) - 1);
    const int first = 63 - __builtin_clzll(upto);  // lane 0 is always a head, so upto != 0
    base = __shfl(base, first, 64);
    if (live) {
        key[i] = k;
        rank[i] = base + (lane - first);
    }
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

// out[0..n] = exclusive scan of in[0..n), out[n] = total; in[] is left zero.  Each block first
// sums the partial sums of the blocks before it (few: n / 1024).
__global__ __launch_bounds__(kBlock) void k_scan_final(int* __restrict__ in, int n,
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
        if (base + k < n) {
            out[base + k] = v[k] + prefix;
            in[base + k] = 0;  // the counts' last reader: zero for the next rebuild
        }
    if (b == gridDim.x - 1 && threadIdx.x == 0) out[n] = prefix + total;
}

// Counting-sort scatter to the cell start plus the arrival rank recorded by the binning.
__global__ __launch_bounds__(kBlock) void k_scatter(int N, const int* __restrict__ key,
                                                    const int* __restrict__ start,
                                                    const int* __restrict__ rank, int* __restrict__ perm)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= N) return;
    perm[start[key[i]] + rank[i]] = i;
}

// One state set in slot order: the record neighbours gather is (xy, th), two arrays so that a
// gather touches 24 B per neighbour (16 B + 8 B) where an interleaved 32-B record would touch 32.
struct SetPtrs {
    double2* xy;  // wrapped position (reference r)
    double* th;   // heading angle (reference theta)
    double2* U;   // unwrapped position (reference r_original)
    int* id;      // caller index
};

struct PermuteArgs {
    int N;
    const int* perm;
    SetPtrs in, out;
    int* inv;  // caller index -> slot after the permutation
    // per-stage buffers of the split-call path (null when permuting between steps)
    const P4* a_in[3];
    P4* a_out[3];
    const double* d_in;
    double* d_out;
};

// Canonical order inside each cell (the scatter leaves arrival order) and the permutation in one
// pass: a particle's slot is its cell's start plus the number of cell mates with a smaller id (one
// pass over the cell per particle, so a crowded cell costs each of its particles one pass), and its
// state is copied there.  perm2 keeps the permutation.
__global__ __launch_bounds__(kBlock) void k_rank_permute(PermuteArgs a, const int* __restrict__ key,
                                                         const int* __restrict__ start, const int* __restrict__ perm,
                                                         int* __restrict__ perm2)
{
    int j = xcd_block() * blockDim.x + threadIdx.x;
    if (j >= a.N) return;
    const int me = perm[j], c = key[me], mine = a.in.id[me];
    const int s = start[c], e = start[c + 1];
    int r = 0;
#pragma unroll 4
    for (int q = s; q < e; ++q) r += a.in.id[perm[q]] < mine;
    const int k = s + r;
    perm2[k] = me;
    a.out.xy[k] = a.in.xy[me];
    a.out.th[k] = a.in.th[me];
    a.out.U[k] = a.in.U[me];
    a.out.id[k] = mine;
    if ((unsigned)mine < (unsigned)a.N) a.inv[mine] = k;
    if (a.d_in) {
        a.d_out[k] = a.d_in[me];
#pragma unroll
        for (int q = 0; q < 3; ++q) a.a_out[q][k] = a.a_in[q][me];
    }
}

// ------------------------------------------------------------------------------------------------
// Rebuild, tile path: binning, counting sort, canonical order inside cells and the permutation in
// ONE launch (the general path above takes five).
//
// Between two builds of the displacement policy no particle moves farther than the half skin,
// which is less than a cell, so a particle now in cell (cx, cy) sat at the last build in one of
// the cells (cx + a, cy + b), |a|, |b| <= 1, whose slots that build's cell starts still describe
// (the slot order has not changed since).  A tile -- cells [c0, c1) of row y -- therefore collects
// its particles from at most six runs of the current order (rows y-1..y+1, columns c0-1..c1,
// periodic), sorts them in LDS by (cell, caller index) -- the general path's canonical order, so
// both paths give the same slots bit for bit -- and learns its first slot from the tiles before it:
// each tile publishes its count at once, then its 256 threads read the counts of all earlier tiles
// together and sum them (a few hundred tiles, so one parallel read replaces a chained look-back,
// whose memory-side loads would line up one after another).  Tile b is workgroup b, in row-major
// order: workgroups are dispatched in index order and all of them fit on the device at once (a few
// hundred of 256 threads and 33 KB of LDS), so a tile waits only for running ones.  The counts
// carry the launch's epoch, so stale entries of earlier launches read as unpublished and nothing
// has to be cleared between launches -- and no counter is shared by every workgroup (same-address
// atomics from all XCDs serialise).
//
// A particle that moved farther than one cell is collected by no tile, so the total falls short of
// N; that, a tile overflowing its LDS or a look-back that does not settle sets *fail, and the host
// rolls the interval back and redoes the rebuild on the general path.  Every index a failed launch
// computes is clamped into its array, so even a garbage input order stays in bounds.
// ------------------------------------------------------------------------------------------------
constexpr int kSortCap = 2048;     // particles one tile can hold (LDS)
constexpr int kSortMaxW = 1024;    // cells per tile (LDS histogram; kBlock * kScanItems)
#ifndef SPP_SORT_TARGET
#define SPP_SORT_TARGET 384
#endif
constexpr int kSortTarget = SPP_SORT_TARGET;  // particles per tile the host aims for
constexpr int kSpinLimit = 1 << 20;

struct SortArgs {
    int N;
    Grid G;
    int seg, segw, ntiles;         // tiles per row, cells per tile (the row's last may be narrower)
    unsigned epoch;                // this launch's tag in the tile states (1 .. 2^30 - 1)
    SetPtrs in, out;
    int* inv;
    const int* cs_old;             // cell starts of the build the input order follows
    int* cs_new;
    unsigned long long* tstate;    // per tile: (epoch << 2 | 1) << 32 | particle count
    int* fail;
    int rep_on;                    // workgroup ntiles reduces the previous interval's report
    ReportArgs rep;
};

__device__ __forceinline__ unsigned long long tile_load(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void tile_store(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int clamp_slot(int v, int N) { return v < 0 ? 0 : (v > N ? N : v); }

// Slot of tile entry q inside the tile: its cell's start + the number of cell mates with a smaller
// caller index (the general path's canonical order); src = its slot in the input order.
__device__ __forceinline__ void tile_rank(int q, const uint16_t* order, const uint16_t* e_col, const int* e_id,
                                          const int* e_old, const int* hist, int& lslot, int& mine, int& src)
{
    const int e = order[q];
    const int c = e_col[e], c_s = hist[c], c_e = hist[c + 1];
    mine = e_id[e];
    int r = 0;
    for (int k = c_s; k < c_e; ++k) r += e_id[order[k]] < mine;
    lslot = c_s + r;
    src = e_old[e];
}

// Moves one particle's state to its slot; returns 1 when the slot falls outside the engine (a
// failed sort).
__device__ __forceinline__ int tile_put(const SortArgs& a, int slot, int mine, double2 xy, double th, double2 u)
{
    if (slot >= a.N) return 1;
    a.out.xy[slot] = xy;
    a.out.th[slot] = th;
    a.out.U[slot] = u;
    a.out.id[slot] = mine;
    if ((unsigned)mine < (unsigned)a.N) a.inv[mine] = slot;
    return 0;
}

__global__ __launch_bounds__(kBlock) void k_sort_tiles(SortArgs a)
{
    if (a.rep_on && (int)blockIdx.x == a.ntiles) {  // the previous interval's report (dispatched last)
        report_body(a.rep);
        return;
    }
    __shared__ int e_old[kSortCap], e_id[kSortCap];
    __shared__ uint16_t e_col[kSortCap], order[kSortCap];
    __shared__ int hist[kSortMaxW + 1], fillc[kSortMaxW];
    __shared__ int lds_wave[kBlock / 64];
    __shared__ long long red[kBlock / 64];
    __shared__ int s_n, s_base, s_bad;
    const int t = threadIdx.x, N = a.N;
    const int W = a.G.ncx, H = a.G.ncy;
    const int tile = blockIdx.x;
    const unsigned tag = a.epoch << 2;
    if (t == 0) {
        s_n = 0;
        s_bad = 0;
    }
    for (int k = t; k <= a.segw; k += kBlock) hist[k] = 0;
    for (int k = t; k < a.segw; k += kBlock) fillc[k] = 0;
    __syncthreads();
    const int y = tile / a.seg, c0 = (tile - y * a.seg) * a.segw;
    const int c1 = min(W, c0 + a.segw), w = c1 - c0;

    // The runs of the current order that hold every candidate: columns [c0-1, c1+1) of rows
    // y-1..y+1, with the periodic column on the far side when the tile touches a row end.  Run k
    // covers flattened candidates [p[k], p[k+1]); candidate q sits at slot q + off[k].
    int p[7], off[6];
    {
        const int lo = c0 - 1, hi = c1 + 1;
        int ra[6], rb[6];  // column range of each run (empty when ra == rb)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int r = y + d - 1;
            r = r < 0 ? r + H : (r >= H ? r - H : r);
            const int ca = max(lo, 0), cb = min(hi, W);
            ra[2 * d] = r * W + ca;
            rb[2 * d] = r * W + cb;
            // the periodic column beyond a row end, unless the main run already holds it
            const bool wl = lo < 0 && cb < W, wr = hi > W && ca > 0;
            ra[2 * d + 1] = wl ? r * W + W - 1 : (wr ? r * W : 0);
            rb[2 * d + 1] = wl ? r * W + W : (wr ? r * W + 1 : 0);
        }
        p[0] = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int s0 = clamp_slot(a.cs_old[ra[k]], N), s1 = clamp_slot(a.cs_old[rb[k]], N);
            const int len = ra[k] < rb[k] && s1 > s0 ? s1 - s0 : 0;
            p[k + 1] = p[k] + len;
            off[k] = s0 - p[k];
        }
    }
    const int total = p[6];

    // Collect the particles whose new cell lies in the tile.  Each round requests all of a thread's
    // candidates (kCollect) before looking at any, so a tile pays one memory round trip per round,
    // not one per candidate.
    constexpr int kCollect = 8;
    for (int q0 = t; q0 < total; q0 += kBlock * kCollect) {
        int sl[kCollect], idv[kCollect];
        double2 xy[kCollect];
#pragma unroll
        for (int u = 0; u < kCollect; ++u) {
            const int q = q0 + u * kBlock;
            int o = off[0];
#pragma unroll
            for (int k = 1; k < 6; ++k) o = q >= p[k] ? off[k] : o;
            sl[u] = q < total ? q + o : -1;
            if (sl[u] >= 0) {
                xy[u] = load_pos(a.in.xy, sl[u]);
                idv[u] = a.in.id[sl[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kCollect; ++u) {
            if (sl[u] < 0) continue;
            const int cy = cell_coord(xy[u].y, a.G.Ly, a.G.Ly2, a.G.invLy2, H, a.G.quirk);
            if (cy != y) continue;
            const int cx = cell_coord(xy[u].x, a.G.Lx, a.G.Lx2, a.G.invLx2, W, a.G.quirk);
            if (cx < c0 || cx >= c1) continue;
            const int e = atomicAdd(&s_n, 1);
            if (e < kSortCap) {
                e_old[e] = sl[u];
                e_col[e] = (uint16_t)(cx - c0);
                e_id[e] = idv[u];
                atomicAdd(&hist[cx - c0], 1);
            }
        }
    }
    __syncthreads();
    const int n = min(s_n, kSortCap);
    if (t == 0) {
        if (s_n > kSortCap) s_bad = 1;
        // publish the count at once, so later tiles can sum it
        tile_store(&a.tstate[tile], ((unsigned long long)(tag | 1u) << 32) | (unsigned)n);
    }

    // column starts: exclusive scan of the column counts (w <= kSortMaxW = kBlock * kScanItems)
    {
        int v[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int c = t * kScanItems + k;
            v[k] = c < w ? hist[c] : 0;
        }
        __syncthreads();
        block_exclusive_scan(v, lds_wave);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int c = t * kScanItems + k;
            if (c < w) hist[c] = v[k];
        }
        if (t == 0) hist[w] = n;
    }
    __syncthreads();
    for (int e = t; e < n; e += kBlock) {
        const int c = e_col[e];
        order[hist[c] + atomicAdd(&fillc[c], 1)] = (uint16_t)e;
    }
    __syncthreads();

    // Each particle's slot inside the tile: its cell's start + the number of cell mates with a
    // smaller caller index (the general path's canonical order).  The first kMove particles of each
    // thread have their state requested now, so those loads overlap the wait for the tile's base.
    // The first two particles of each thread: slots inside the tile and their state, requested now
    // so the loads overlap the wait for the tile's first slot (lanes without a particle read slot 0
    // and store nothing: unconditional loads keep the state in registers).
    int ls0 = -1, id0 = 0, sr0 = 0, ls1 = -1, id1 = 0, sr1 = 0;
    if (t < n) tile_rank(t, order, e_col, e_id, e_old, hist, ls0, id0, sr0);
    if (t + kBlock < n) tile_rank(t + kBlock, order, e_col, e_id, e_old, hist, ls1, id1, sr1);
    const double2 xy0 = a.in.xy[sr0];
    const double th0 = a.in.th[sr0];
    const double2 u0 = a.in.U[sr0];
    const double2 xy1 = a.in.xy[sr1];
    const double th1 = a.in.th[sr1];
    const double2 u1 = a.in.U[sr1];

    // The tile's first slot: the sum of the counts of all earlier tiles, all of a thread's reads
    // issued together; tiles whose count is not published yet are read again, a bounded number of
    // times.
    {
        constexpr int kPre = 4;
        long long part = 0;
        bool bad = false;
        for (int j0 = t; j0 < tile && !bad; j0 += kBlock * kPre) {
            unsigned long long st[kPre];
            bool ready[kPre];
#pragma unroll
            for (int u = 0; u < kPre; ++u) ready[u] = j0 + u * kBlock >= tile;
            for (int spins = 0;; ++spins) {
#pragma unroll
                for (int u = 0; u < kPre; ++u)
                    if (!ready[u]) st[u] = tile_load(&a.tstate[j0 + u * kBlock]);
                bool all = true;
#pragma unroll
                for (int u = 0; u < kPre; ++u) {
                    if (!ready[u] && (unsigned)(st[u] >> 32) == (tag | 1u)) {
                        ready[u] = true;
                        part += (unsigned)st[u];
                    }
                    all = all && ready[u];
                }
                if (all) break;
                if (spins >= kSpinLimit) {
                    bad = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (bad) s_bad = 1;
        part = wave_sum_ll(part);
        if ((t & 63) == 0) red[t >> 6] = part;
    }
    __syncthreads();
    if (t == 0) {
        long long prefix = 0;
        for (int k = 0; k < kBlock / 64; ++k) prefix += red[k];
        if (prefix > N) {
            s_bad = 1;
            prefix = N;
        }
        s_base = (int)prefix;
        if (tile == a.ntiles - 1 && prefix + n != N) s_bad = 1;  // a particle went missing
    }
    __syncthreads();
    const int base = s_base;

    for (int c = t; c < w; c += kBlock) a.cs_new[y * W + c0 + c] = clamp_slot(base + hist[c], N);
    if (tile == a.ntiles - 1 && t == 0) a.cs_new[H * W] = clamp_slot(base + n, N);
    int put_bad = 0;
    if (ls0 >= 0) put_bad |= tile_put(a, base + ls0, id0, xy0, th0, u0);
    if (ls1 >= 0) put_bad |= tile_put(a, base + ls1, id1, xy1, th1, u1);
    if (put_bad) s_bad = 1;
    // particles beyond two per thread (crowded tiles)
    for (int q = t + 2 * kBlock; q < n; q += kBlock) {
        int ls, me, sr;
        tile_rank(q, order, e_col, e_id, e_old, hist, ls, me, sr);
        if (tile_put(a, base + ls, me, a.in.xy[sr], a.in.th[sr], a.in.U[sr])) s_bad = 1;
    }
    __syncthreads();
    if (t == 0 && s_bad) atomicOr(a.fail, 1);
}

// ------------------------------------------------------------------------------------------------
// Stage kernels: interact + Move fused, one particle per lane.
// ------------------------------------------------------------------------------------------------

// Constants of the single-bead active Brownian model (zero for the Vicsek models).
struct AbpConst {
    int n_membrane;     // ring beads: caller indices [0, n_membrane)
    int nb;             // beads per swimmer (1: single-bead swimmers)
    double F0;          // swimmers' self-propulsion force
    double mob_par, mob_perp, mob_rot, torque0;  // ring beads (and single-bead swimmers)
    double c_par, c_perp, c_rot;                 // chains of nb beads (Set_nb, particle.h:856-876)
    double k_bond, l_bond, kappa;
    // bead-pair law of chains: R12_Repulsive_Truncated(sigma_b, rc_b, A_b), predicate d2 < T_b
    double sep, sigma_b, A_b, rc13_b, T_b;
};

struct StageArgs {
    int N;
    Phys P;
    AbpConst abp;
    const int* inv;        // caller index -> slot (ring neighbours of the active Brownian model)
    const SysInfo* sys_tab;  // batched engine (null: single system, constants in P)
    const int* sys_of;
    const double2* gxy;    // stage input positions (neighbours' and own)
    const double* gth;     // stage input headings (neighbours' and own)
    double2* oxy;          // stage output positions
    double* oth;           // stage output headings
    const double* th0;     // step-start heading (theta_old) for RK stages >= 2: the input set's,
                           // which no stage overwrites before the step's last
    const double2* U_in;   // step-start unwrapped position
    double2* U_out;        // written by the final stage
    double* dth;           // angular noise of this step
    P4* kacc;              // RK4 accumulator
    const int* nbr32;
    const int* ncount;
    const int* id;
    int* id_out;           // first step of an interval: the output set's copy of id
    const double* inj;     // injected noise, slot order (null = stream)
    const float2* build;
    int64_t step;          // noise-stream step (relative to *clock when clock is set)
    const long long* clock;  // graph replays: the interval's first step, set before each replay
    int kstep;             // step index inside the current interval
    double thresh2;        // displacement trigger, squared
    int* blkbad;           // per-workgroup first bad step
    double* blockmax;      // per-workgroup max displacement^2 (last step of an interval)
    int record;
    bool stream;           // write the stage outputs with streaming stores (kStreamMaxN)
    int grid;              // the launch's workgroups for the particles (one more carries rep)
    int rep_on;            // workgroup grid reduces the previous interval's report
    ReportArgs rep;
};

// The noise stream's step of a stage launch: absolute, or relative to the interval's first step
// held on the device when the launch belongs to a captured (replayed) interval.
__device__ __forceinline__ int64_t noise_step(const StageArgs& a)
{
    return a.clock ? a.step + *a.clock : a.step;
}

// Slot j's record of the stage input: (x, y, theta) in a P4 (w unused).  Addressed by 32-bit byte
// offsets from the arrays' (uniform) bases, which the loads take as a VGPR offset to an SGPR base
// (engines hold fewer than 2^27 particles).
__device__ __forceinline__ P4 stage_record(const StageArgs& a, int j)
{
    const unsigned u = (unsigned)j;
    const double2 p = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(a.gxy) + (size_t)(u * 16u));
    const double t = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(a.gth) + (size_t)(u * 8u));
    return P4{p.x, p.y, t, 0.0};
}

// Whether any live lane of the wavefront sits within img_m of a periodic edge (then every lane
// takes the minimum image of its pair differences; otherwise none needs it).  Wave-uniform.
__device__ __forceinline__ bool wave_needs_image(const Phys& P, bool live, double x, double y)
{
    const bool edge = live && !(fabs(x) <= P.Lx - P.img_m && fabs(y) <= P.Ly - P.img_m);
    return __any(edge) != 0;
}

// Displacement trigger: marks the step (per workgroup, atomics stay inside one workgroup's slot)
// and, on the last step of an interval, records the workgroup's largest squared displacement.
__device__ __forceinline__ void check_disp(const StageArgs& a, const Phys& P, int s, bool live, int i,
                                           double x, double y, int mark, double& dmax)
{
    double d2 = 0;
    if (live) {
        float2 b = a.build[i];
        double dx = spp::min_image(x - (double)b.x, P.Lx2, P.invLx2);
        double dy = spp::min_image(y - (double)b.y, P.Ly2, P.invLy2);
        d2 = dx * dx + dy * dy;
    }
    if (__any(d2 > a.thresh2) && (threadIdx.x & 63) == 0) atomicMin(&a.blkbad[s], mark);
    dmax = fmax(dmax, d2);
}

__device__ __forceinline__ void finish_disp(const StageArgs& a, int s, double dmax)
{
    __shared__ double red[kBlock / 64];
    if (!a.record) return;
    double m = wave_max(dmax);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < kBlock / 64; ++k) r = fmax(r, red[k]);
        a.blockmax[s] = r;
    }
}

// ------------------------------------------------------------------------------------------------
// Slot-parallel variant: one lane per particle over the whole slot range, neighbour records
// gathered from L2 through a 32-bit slot-index list.  Loads of lanes past their list length are
// masked off (they would cost texture-address cycles for nothing).
// ------------------------------------------------------------------------------------------------

struct NlistFlatArgs {
    int N, cap;
    Grid G;
    const SysInfo* sys_tab;  // batched engine (null: single system)
    const int* sys_of;
    double T_list;
    int wide;                // 1: the 5x5 stencil (the wide list after a deep overlap)
    const double2* pos;
    const int* start;
    int* nbr32;
    int* ncount;
    float2* build;
    int* cellxy;
    int* blklen;
    int* blknnz;
};

// The candidates of slot i in flattened order (range k covers [p[k], p[k+1]), candidate t at slot
// t + off[k]), eight requested per round; writes the list and returns its length before the cut.
template <int kSeg, int kBatch, bool IMG>
__device__ __forceinline__ int nlist_scan(const NlistFlatArgs& a, int i, double2 me, const Grid& G, const int* p,
                                          const int* off)
{
    const int total = p[kSeg];
    int c = 0;
    int* w = a.nbr32 + i;  // the next list entry's address: row c of the slot-major list
    for (int t0 = 0; t0 < total; t0 += kBatch) {
        int jj[kBatch];
        double2 q[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int t = t0 + u;
            int o = off[0];
#pragma unroll
            for (int k = 1; k < kSeg; ++k) o = t >= p[k] ? off[k] : o;
            jj[u] = t + o;
            if (t < total) q[u] = load_pos(a.pos, jj[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (t0 + u < total && jj[u] != i) {
                double dx = spp::image_diff<IMG>(me.x - q[u].x, G.Lx2, G.invLx2);
                double dy = spp::image_diff<IMG>(me.y - q[u].y, G.Ly2, G.invLy2);
                double d2 = dx * dx + dy * dy;
                if (d2 < a.T_list) {
                    if (c < a.cap) {
                        *w = jj[u];
                        w += a.N;
                    }
                    ++c;
                }
            }
        }
    }
    return c;
}

// Cell::Neighbor_List (shared/cell.h:74-103) for slot i over the 3x3 periodic stencil of grid G
// (R = 1; R = 2: the 5x5 stencil of the wide list, for pairs up to two cell widths apart)
// (cell keys from koff).  Each stencil row is one or two contiguous slot ranges (the columns inside
// the grid, then the column across the periodic edge), so the stencil is at most six ranges; their
// bounds are read at once and the candidates are streamed eight at a time across range ends, in
// range order, so the list keeps the row-by-row candidate order while a wave waits on a few rounds
// of loads instead of one per range.  Returns the list length before the capacity cut.
template <int R = 1>
__device__ __forceinline__ int nlist_particle(const NlistFlatArgs& a, int i, const Grid& G, int koff)
{
    constexpr int kRows = 2 * R + 1, kSeg = 2 * kRows, kBatch = 8;
    const double2 me = a.pos[i];
    const int cx = cell_coord(me.x, G.Lx, G.Lx2, G.invLx2, G.ncx, G.quirk);
    const int cy = cell_coord(me.y, G.Ly, G.Ly2, G.invLy2, G.ncy, G.quirk);
    // columns inside the grid [x0, x1], then the columns across the periodic edge [w0, w1)
    int x0 = cx - R, x1 = cx + R, w0 = 0, w1 = 0;
    if (x0 < 0) {
        w0 = G.ncx + x0;
        w1 = G.ncx;
        x0 = 0;
    } else if (x1 >= G.ncx) {
        w0 = 0;
        w1 = x1 - G.ncx + 1;
        x1 = G.ncx - 1;
    }
    // range k covers flattened candidates [p[k], p[k+1]); candidate t sits at slot t + off[k]
    int p[kSeg + 1], off[kSeg];
    {
        int s[kSeg], e[kSeg];
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            int y = cy + r - R;
            y = y < 0 ? y + G.ncy : (y >= G.ncy ? y - G.ncy : y);
            const int row = koff + y * G.ncx;
            s[2 * r] = a.start[row + x0];
            e[2 * r] = a.start[row + x1 + 1];
            s[2 * r + 1] = w1 > w0 ? a.start[row + w0] : 0;
            e[2 * r + 1] = w1 > w0 ? a.start[row + w1] : 0;
        }
        p[0] = 0;
#pragma unroll
        for (int k = 0; k < kSeg; ++k) {
            p[k + 1] = p[k] + (e[k] - s[k]);
            off[k] = s[k] - p[k];
        }
    }
    // Cells away from the grid's edges (exact binning, five or more cells per side): every candidate
    // sits in an adjacent cell on the same side of the periodic edges, closer than L, so the
    // differences need no minimum image (image_diff).  Decided for the whole wavefront.
    const bool inner = R == 1 && !G.quirk && G.ncx >= 5 && G.ncy >= 5 && cx >= 1 && cx <= G.ncx - 2 && cy >= 1 &&
                       cy <= G.ncy - 2;
    int c;
    if (__all(inner)) {
        // no periodic column: the stencil is the three row ranges 0, 2, 4
        const int p3[4] = {p[0], p[2], p[4], p[kSeg]}, o3[3] = {off[0], off[2], off[4]};
        c = nlist_scan<3, kBatch, false>(a, i, me, G, p3, o3);
    } else {
        c = nlist_scan<kSeg, kBatch, true>(a, i, me, G, p, off);
    }
    a.ncount[i] = c < a.cap ? c : a.cap;
    a.build[i] = make_float2((float)me.x, (float)me.y);
    a.cellxy[i] = cx | (cy << 16);
    return c;
}

// Grid and first cell key of slot i's system.
__device__ __forceinline__ void system_grid(const NlistFlatArgs& a, int i, Grid& G, int& koff)
{
    G = a.G;
    koff = 0;
    if (a.sys_of) {
        const SysInfo& S = a.sys_tab[a.sys_of[i]];
        G = S.G;
        koff = S.key_off;
    }
}

// Longest list and entry count of workgroup b (list lengths before the capacity cut), for the
// interval report.
__device__ __forceinline__ void list_stats(const NlistFlatArgs& a, int b, int c)
{
    __shared__ int red_len[kBlock / 64], red_nnz[kBlock / 64];
    int m = c;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    int t = wave_sum(c);
    if ((threadIdx.x & 63) == 0) {
        red_len[threadIdx.x >> 6] = m;
        red_nnz[threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int mm = 0, tt = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
            mm = max(mm, red_len[k]);
            tt += red_nnz[k];
        }
        a.blklen[b] = mm;
        a.blknnz[b] = tt;
    }
}

__global__ __launch_bounds__(kBlock) void k_nlist_flat(NlistFlatArgs a)
{
    const int b = xcd_block();
    int i = b * blockDim.x + threadIdx.x;
    bool live = i < a.N;
    int c = 0;
    if (live) {
        Grid G;
        int koff;
        system_grid(a, i, G, koff);
        c = a.wide ? nlist_particle<2>(a, i, G, koff) : nlist_particle<1>(a, i, G, koff);
    }
    list_stats(a, b, c);
}

// GU: neighbour records in flight per lane (a block of the list loop).  Small systems, whose stage
// time is one wavefront's chain of loads, take 8; large ones 4 (fewer VGPRs, same speed or better).
// ONE: one cutoff for both pair laws (pair_accumulate).  RAW: return the two torque sums in raw[]
// instead of the torque (the caller forms it from constants it re-reads after the loop).
template <int MODEL, int REP, int GU = 4, bool IMG = true, bool ONE = false, bool RAW = false>
__device__ __forceinline__ void gather_forces_flat(const StageArgs& a, const Phys& P, int i, const P4& me,
                                                   double mc, double ms, double& fx, double& fy, double& tq,
                                                   double* raw = nullptr)
{
    const int cnt = a.ncount[i];
    const int* nb = a.nbr32 + i;
    double s1 = 0, s2 = 0;
    fx = mc * P.v0;  // Reset: f = v (speed v0), particle.h:559-564
    fy = ms * P.v0;
    for (int k = 0; k < cnt; k += GU) {
        int j[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u)
            if (k + u < cnt) j[u] = nb[(size_t)(k + u) * a.N];
        P4 q[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u)
            if (k + u < cnt) q[u] = stage_record(a, j[u]);
#pragma unroll
        for (int u = 0; u < GU; ++u)
            if (k + u < cnt) spp::pair_accumulate<MODEL, REP, IMG, ONE>(P, me, mc, ms, q[u], fx, fy, s1, s2);
    }
    if (RAW) {
        raw[0] = s1;
        raw[1] = s2;
    } else {
        tq = spp::drift_torque<MODEL>(P, s1, s2);
    }
}

// One particle's fused stage (interaction at the stage input + the move), for slot i of a system
// whose constants P already hold and whose particle indices start at pid_off.  Returns the wrapped
// output position for the displacement check.
//
// Stages after the first of large engines (four records in flight: many waves per SIMD, so
// occupancy sets the speed) load their own inputs (theta_old, the noise, the unwrapped position,
// RK4's accumulator) after the pair loop, and batched ones re-read their system's constants there,
// instead of holding both in registers across the loop: the later stages' kernels fit one more
// wave per SIMD (e.g. 102 -> 88 VGPRs for the batched continuous RK2 stage 2).  Same arithmetic.
template <int MODEL, int REP, int INTEG, int STAGE, int GU = 4, bool ONE = false, bool BATCH = false>
__device__ __forceinline__ void stage_particle(const StageArgs& a, const Phys& Pin, int i, int pid_off, double& x,
                                               double& y)
{
    constexpr bool kLate = STAGE > 1 && GU != 8;
    Phys P = Pin;
    P4 me = stage_record(a, i);
    const bool img = wave_needs_image(P, true, me.x, me.y);  // over the lanes that run this
    double mc, ms;  // v = (cos, sin) of the stage input's heading
    spp::heading_unit(me.z, mc, ms);
    double dn = 0, th_in = 0, thold = 0;
    double2 u = make_double2(0, 0);
    P4 kk{0, 0, 0, 0};
    if (STAGE <= 1) {
        int pid = a.id[i];
        if (a.id_out) a.id_out[i] = pid;
        th_in = me.z;
        if (INTEG == SPP_EULER) u = a.U_in[i];
        // the stream is addressed by the particle's index within its system, so systems with
        // equal seeds (Lyapunov clones) draw identical noise
        dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)(pid - pid_off), noise_step(a));
    } else if (!kLate) {
        thold = a.th0[i];  // theta_old: the step's input set, which no stage has overwritten yet
        dn = a.dth[i];
        u = a.U_in[i];
        if (INTEG == SPP_RK4) kk = a.kacc[i];
    }
    double fx, fy, tq;
    if (BATCH && kLate) {
        double raw[2];
        if (img)
            gather_forces_flat<MODEL, REP, GU, true, ONE, true>(a, P, i, me, mc, ms, fx, fy, tq, raw);
        else
            gather_forces_flat<MODEL, REP, GU, false, ONE, true>(a, P, i, me, mc, ms, fx, fy, tq, raw);
        asm volatile("" ::: "memory");  // re-read from memory, not kept from before the loop
        apply_sys(P, a.sys_tab[a.sys_of[i]]);
        tq = spp::drift_torque<MODEL>(P, raw[0], raw[1]);
    } else {
        if (img)
            gather_forces_flat<MODEL, REP, GU, true, ONE>(a, P, i, me, mc, ms, fx, fy, tq);
        else
            gather_forces_flat<MODEL, REP, GU, false, ONE>(a, P, i, me, mc, ms, fx, fy, tq);
    }
    if (kLate) {
        asm volatile("" ::: "memory");  // issued here, after the pair loop
        thold = a.th0[i];
        dn = a.dth[i];
        u = a.U_in[i];
        if (INTEG == SPP_RK4) kk = a.kacc[i];
    }
    double th;
    if (STAGE <= 1) {
        if (INTEG == SPP_EULER) {
            th = th_in + P.dt * (tq);
            th += dn;
            u.x += fx * P.dt;
            u.y += fy * P.dt;
            spp::store_out(a.stream, &a.U_out[i], u);
            x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
        } else {
            spp::store_out(a.stream, &a.dth[i], dn);
            th = th_in;
            th += P.half_dt * tq;
            th += dn / 2;
            th = spp::wrap_angle(th);
            x = spp::periodic_wrap(me.x + fx * P.half_dt, P.Lx, P.Lx2);
            y = spp::periodic_wrap(me.y + fy * P.half_dt, P.Ly, P.Ly2);
            if (INTEG == SPP_RK4) spp::store_out(a.stream, &a.kacc[i], fx, fy, tq, 0.0);
        }
    } else {
        double h = (INTEG == SPP_RK4 && STAGE == 2) ? P.half_dt : P.dt;
        if (INTEG == SPP_RK4 && STAGE == 4) {
            th = thold + (kk.z + tq) * P.dt6;
            th += dn;
            u.x = u.x + (kk.x + fx) * P.dt6;
            u.y = u.y + (kk.y + fy) * P.dt6;
        } else {
            th = thold + h * tq;
            th += (INTEG == SPP_RK4 && STAGE == 2) ? dn / 2 : dn;
            u.x = u.x + fx * h;
            u.y = u.y + fy * h;
            if (INTEG == SPP_RK4) spp::store_out(a.stream, &a.kacc[i], kk.x + fx * 2, kk.y + fy * 2, kk.z + tq * 2, 0.0);
        }
        th = spp::wrap_angle(th);
        if (INTEG == SPP_RK2 || STAGE == 4) spp::store_out(a.stream, &a.U_out[i], u);
        x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
        y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
    }
    spp::store_out(a.stream, &a.oxy[i], make_double2(x, y));
    spp::store_out(a.stream, &a.oth[i], th);
}

template <int MODEL, int REP, int INTEG, int STAGE, bool BATCH, int GU = 4, bool ONE = false>
__global__ __launch_bounds__(kBlock) void k_stage_flat(StageArgs a)
{
    if (a.rep_on && (int)blockIdx.x == a.grid) {  // the previous interval's report (dispatched last)
        report_body(a.rep);
        return;
    }
    int b = xcd_block_of(a.grid);
    int i = b * kBlock + threadIdx.x;
    bool live = i < a.N;
    Phys P = a.P;
    int pid_off = 0;
    if (BATCH && live) {
        const SysInfo& S = a.sys_tab[a.sys_of[i]];
        apply_sys(P, S);
        pid_off = S.pid_off;
    }
    constexpr bool final_stage = (INTEG == SPP_EULER) || (INTEG == SPP_RK2 && STAGE == 2) || (STAGE == 4);
    double x = 0, y = 0, dmax = 0;
    if (live) stage_particle<MODEL, REP, INTEG, STAGE, GU, ONE, BATCH>(a, P, i, pid_off, x, y);
    check_disp(a, P, b, live, i, x, y, final_stage ? a.kstep + 1 : a.kstep, dmax);
    finish_disp(a, b, dmax);
}

template <int MODEL, int REP, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_interact_flat(StageArgs a, P4* force)
{
    int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.N) return;
    Phys P = a.P;
    if (BATCH) apply_sys(P, a.sys_tab[a.sys_of[i]]);
    P4 me = stage_record(a, i);
    double mc, ms, fx, fy, tq;
    spp::heading_unit(me.z, mc, ms);
    gather_forces_flat<MODEL, REP>(a, P, i, me, mc, ms, fx, fy, tq);
    force[i] = P4{fx, fy, tq, 0.0};
}

// ------------------------------------------------------------------------------------------------
// Split-call path: Box::Interact and Move_* as separate launches (drop-in granularity), on the
// current state set in place like the reference's particle array.
// ------------------------------------------------------------------------------------------------


struct MoveArgs {
    int N;
    Phys P;
    AbpConst abp;
    const SysInfo* sys_tab;
    const int* sys_of;
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
    Phys P = a.P;
    int pid_off = 0;
    if (a.sys_of) {
        const SysInfo& S = a.sys_tab[a.sys_of[i]];
        apply_sys(P, S);
        pid_off = S.pid_off;
    }
    const double2 me = a.cur.xy[i];
    double th_in = a.cur.th[i];
    P4 f = a.force[i];
    double2 u = a.cur.U[i];
    double th, x, y;
    if (a.stage <= 1) {
        double dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)(a.cur.id[i] - pid_off), a.step);
        a.dth[i] = dn;
        if (a.integ == SPP_EULER) {
            th = th_in + P.dt * (f.z);
            th += dn;
            u.x += f.x * P.dt;
            u.y += f.y * P.dt;
            a.cur.U[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
        } else {
            a.old[i] = P4{u.x, u.y, th_in, 0.0};
            th = th_in;
            th += P.half_dt * f.z;
            th += dn / 2;
            th = spp::wrap_angle(th);
            if (a.integ == SPP_RK4) a.kacc[i] = f;
            x = spp::periodic_wrap(me.x + f.x * P.half_dt, P.Lx, P.Lx2);
            y = spp::periodic_wrap(me.y + f.y * P.half_dt, P.Ly, P.Ly2);
        }
    } else {
        P4 o = a.old[i];
        double dn = a.dth[i];
        if (a.integ == SPP_RK4 && a.stage == 4) {
            P4 k = a.kacc[i];
            th = o.z + (k.z + f.z) * P.dt6;
            th += dn;
            u.x = o.x + (k.x + f.x) * P.dt6;
            u.y = o.y + (k.y + f.y) * P.dt6;
        } else {
            double h = (a.integ == SPP_RK4 && a.stage == 2) ? P.half_dt : P.dt;
            th = o.z + h * f.z;
            th += (a.integ == SPP_RK4 && a.stage == 2) ? dn / 2 : dn;
            u.x = o.x + f.x * h;
            u.y = o.y + f.y * h;
            if (a.integ == SPP_RK4) {
                P4 k = a.kacc[i];
                a.kacc[i] = P4{k.x + f.x * 2, k.y + f.y * 2, k.z + f.z * 2, 0.0};
            }
        }
        th = spp::wrap_angle(th);
        a.cur.U[i] = u;  // the reference writes r_original in every stage >= 2
        x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
        y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
    }
    a.cur.xy[i] = make_double2(x, y);
    a.cur.th[i] = th;
}

#include "abp_kernels.hpp"

// Scatter caller-order values into slot order: dst[slot] = src[id[slot]].
__global__ void k_gather_by_id(int N, const int* __restrict__ id, const double* __restrict__ src,
                               double* __restrict__ dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const int j = id[i];
        dst[i] = (unsigned)j < (unsigned)N ? src[j] : 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_copy_set(int N, SetPtrs src, SetPtrs dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    dst.xy[i] = src.xy[i];
    dst.th[i] = src.th[i];
    dst.U[i] = src.U[i];
    dst.id[i] = src.id[i];
}

// Polarization partial sums (Box::Compute_Polarization, serial/box.h:531-538), two-pass.
__global__ __launch_bounds__(kBlock) void k_polar(int N, const double* __restrict__ th,
                                                  double2* __restrict__ part)
{
    __shared__ double sx[kBlock / 64], sy[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double vx = 0, vy = 0;
    if (i < N) spp::heading_unit(th[i], vx, vy);
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

// Per-system heading sums of a batched engine: one workgroup per system over its slot range.
__global__ __launch_bounds__(kBlock) void k_polar_sys(const SysInfo* __restrict__ sys, const double* __restrict__ th,
                                                      double2* __restrict__ out)
{
    __shared__ double sx[kBlock / 64], sy[kBlock / 64];
    const SysInfo& S = sys[blockIdx.x];
    double vx = 0, vy = 0;
    for (int i = S.pid_off + threadIdx.x; i < S.pid_off + S.n; i += kBlock) {
        double c, s;
        spp::heading_unit(th[i], c, s);
        vx += c;
        vy += s;
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
        out[blockIdx.x] = make_double2(a, b);
    }
}

// Pair-distance histogram over the full Verlet list (each pair counted from both ends), integer
// counts so the result does not depend on the order of the atomic additions.
__global__ __launch_bounds__(kBlock) void k_pair_hist(int N, const double2* __restrict__ pos, const int* __restrict__ nbr,
                                                      const int* __restrict__ ncount, Grid G, double inv_dr,
                                                      int nbins, unsigned long long* __restrict__ hist)
{
    extern __shared__ unsigned int sh[];
    for (int k = threadIdx.x; k < nbins; k += kBlock) sh[k] = 0;
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const double2 me = pos[i];
        int c = ncount[i];
        for (int k = 0; k < c; ++k) {
            const double2 q = pos[nbr[(size_t)k * N + i]];
            double dx = spp::min_image(me.x - q.x, G.Lx2, G.invLx2);
            double dy = spp::min_image(me.y - q.y, G.Ly2, G.invLy2);
            int b = (int)(sqrt(dx * dx + dy * dy) * inv_dr);
            if (b < nbins) atomicAdd(&sh[b], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nbins; k += kBlock)
        if (sh[k]) atomicAdd(&hist[k], (unsigned long long)sh[k]);
}

// One interval's report as a kernel of its own (the interval's report could not ride on the next
// interval's first launch, see ReportArgs).
__global__ __launch_bounds__(kBlock) void k_report(ReportArgs r)
{
    report_body(r);
}

// ------------------------------------------------------------------------------------------------
// Kernel dispatch tables
// ------------------------------------------------------------------------------------------------

using StageFn = void (*)(StageArgs);
constexpr int kVariantBatch = -1;  // internal: slot-parallel kernels with per-system constants
using InteractFn = void (*)(StageArgs, P4*);

// one: the engine's repulsion and alignment cutoffs coincide (single engines of the repulsive model
// with a repulsion law only; batched engines keep the general pair path).
template <int MODEL, int REP, int I, int ST>
StageFn variant_fn(int variant, bool small, bool one)
{
    if (variant == kVariantBatch)
        return small ? k_stage_flat<MODEL, REP, I, ST, true, 8> : k_stage_flat<MODEL, REP, I, ST, true, 4>;
    if constexpr (MODEL == SPP_MODEL_REPULSIVE && REP != SPP_REP_NONE) {
        if (one)
            return small ? k_stage_flat<MODEL, REP, I, ST, false, 8, true> : k_stage_flat<MODEL, REP, I, ST, false, 4, true>;
    }
    return small ? k_stage_flat<MODEL, REP, I, ST, false, 8> : k_stage_flat<MODEL, REP, I, ST, false, 4>;
}

template <int MODEL, int REP>
StageFn stage_fn(int integ, int stage, int variant, bool small, bool one)
{
    if (integ == SPP_EULER) return variant_fn<MODEL, REP, SPP_EULER, 0>(variant, small, one);
    if (integ == SPP_RK2)
        return stage == 1 ? variant_fn<MODEL, REP, SPP_RK2, 1>(variant, small, one)
                          : variant_fn<MODEL, REP, SPP_RK2, 2>(variant, small, one);
    switch (stage) {
    case 1: return variant_fn<MODEL, REP, SPP_RK4, 1>(variant, small, one);
    case 2: return variant_fn<MODEL, REP, SPP_RK4, 2>(variant, small, one);
    case 3: return variant_fn<MODEL, REP, SPP_RK4, 3>(variant, small, one);
    default: return variant_fn<MODEL, REP, SPP_RK4, 4>(variant, small, one);
    }
}

template <int REP>
StageFn abp_stage_fn(int integ, int stage)
{
    if (integ == SPP_EULER) return k_stage_abp<REP, SPP_EULER, 0>;
    return stage <= 1 ? k_stage_abp<REP, SPP_RK2, 1> : k_stage_abp<REP, SPP_RK2, 2>;
}

StageFn pick_stage(int model, int rep, int integ, int stage, int variant, bool small, bool one)
{
    if (model == SPP_MODEL_ABP) {
        switch (rep) {
        case SPP_REP_NONE: return abp_stage_fn<SPP_REP_NONE>(integ, stage);
        case SPP_REP_R12_TRUNCATED: return abp_stage_fn<SPP_REP_R12_TRUNCATED>(integ, stage);
        case SPP_REP_HARMONIC: return abp_stage_fn<SPP_REP_HARMONIC>(integ, stage);
        case SPP_REP_R12: return abp_stage_fn<SPP_REP_R12>(integ, stage);
        case SPP_REP_YUKAWA_TRUNCATED: return abp_stage_fn<SPP_REP_YUKAWA_TRUNCATED>(integ, stage);
        default: return abp_stage_fn<SPP_REP_YUKAWA>(integ, stage);
        }
    }
    if (model == SPP_MODEL_CONTINUOUS)
        return stage_fn<SPP_MODEL_CONTINUOUS, SPP_REP_NONE>(integ, stage, variant, small, false);
    switch (rep) {
    case SPP_REP_NONE: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_NONE>(integ, stage, variant, small, false);
    case SPP_REP_R12_TRUNCATED:
        return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_R12_TRUNCATED>(integ, stage, variant, small, one);
    case SPP_REP_HARMONIC: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_HARMONIC>(integ, stage, variant, small, one);
    case SPP_REP_R12: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_R12>(integ, stage, variant, small, one);
    case SPP_REP_YUKAWA_TRUNCATED:
        return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA_TRUNCATED>(integ, stage, variant, small, one);
    default: return stage_fn<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA>(integ, stage, variant, small, one);
    }
}

template <int MODEL, int REP>
InteractFn interact_fn(int variant)
{
    if (variant == kVariantBatch) return k_interact_flat<MODEL, REP, true>;
    return k_interact_flat<MODEL, REP, false>;
}

// Particle::Reset for every particle (particle.h:559-564, 878-883): f = v0 v (F0 v for swimmers,
// 0 for ring beads), torque 0 (torque0 for the active Brownian model), into the force buffer that
// the split-call moves apply.
__global__ __launch_bounds__(kBlock) void k_reset_forces(StageArgs a, int abp, P4* force)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    Phys P = a.P;
    if (a.sys_of) apply_sys(P, a.sys_tab[a.sys_of[i]]);
    double mc, ms;
    spp::heading_unit(a.gth[i], mc, ms);
    double v0 = P.v0, tq = 0.0;
    if (abp) {
        const bool bead = (unsigned)a.id[i] < (unsigned)a.abp.n_membrane;
        v0 = bead ? 0.0 : a.abp.F0;
        tq = a.abp.torque0;
    }
    force[i] = P4{mc * v0, ms * v0, tq, 0.0};
}

// Particle::Interact(Particle&) for one pair of slots (particle.h:728-757, 296-338, 964-1012): each
// side's force and torque are added to its accumulated ones (one thread, so no other update races).
template <int MODEL, int REP>
__device__ __forceinline__ void pair_side(const StageArgs& a, const Phys& P, int si, int sj, double& fx,
                                          double& fy, double& tq)
{
    if (MODEL == SPP_MODEL_ABP) {
        abp_pair_side<REP>(a, P, si, sj, fx, fy, tq);
        return;
    }
    const P4 me = stage_record(a, si);
    double mc, ms, s1 = 0, s2 = 0;
    spp::heading_unit(me.z, mc, ms);
    spp::pair_accumulate<MODEL, REP>(P, me, mc, ms, stage_record(a, sj), fx, fy, s1, s2);
    tq += spp::drift_torque<MODEL>(P, s1, s2);
}

template <int MODEL, int REP>
__global__ void k_pair_interact(StageArgs a, int ci, int cj, P4* force)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int si = a.inv[ci], sj = a.inv[cj];
    Phys P = a.P;
    if (a.sys_of) apply_sys(P, a.sys_tab[a.sys_of[si]]);
    double fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0};
    pair_side<MODEL, REP>(a, P, si, sj, fi[0], fi[1], fi[2]);
    pair_side<MODEL, REP>(a, P, sj, si, fj[0], fj[1], fj[2]);
    force[si].x += fi[0];
    force[si].y += fi[1];
    force[si].z += fi[2];
    force[sj].x += fj[0];
    force[sj].y += fj[1];
    force[sj].z += fj[2];
}

using PairFn = void (*)(StageArgs, int, int, P4*);

template <int MODEL>
PairFn pair_fn(int rep)
{
    switch (rep) {
    case SPP_REP_NONE: return k_pair_interact<MODEL, SPP_REP_NONE>;
    case SPP_REP_R12_TRUNCATED: return k_pair_interact<MODEL, SPP_REP_R12_TRUNCATED>;
    case SPP_REP_HARMONIC: return k_pair_interact<MODEL, SPP_REP_HARMONIC>;
    case SPP_REP_R12: return k_pair_interact<MODEL, SPP_REP_R12>;
    case SPP_REP_YUKAWA_TRUNCATED: return k_pair_interact<MODEL, SPP_REP_YUKAWA_TRUNCATED>;
    default: return k_pair_interact<MODEL, SPP_REP_YUKAWA>;
    }
}

PairFn pick_pair(int model, int rep)
{
    if (model == SPP_MODEL_ABP) return pair_fn<SPP_MODEL_ABP>(rep);
    if (model == SPP_MODEL_CONTINUOUS) return pair_fn<SPP_MODEL_CONTINUOUS>(SPP_REP_NONE);
    return pair_fn<SPP_MODEL_REPULSIVE>(rep);
}

InteractFn pick_interact(int model, int rep, int variant)
{
    if (model == SPP_MODEL_ABP) {
        switch (rep) {
        case SPP_REP_NONE: return k_interact_abp<SPP_REP_NONE>;
        case SPP_REP_R12_TRUNCATED: return k_interact_abp<SPP_REP_R12_TRUNCATED>;
        case SPP_REP_HARMONIC: return k_interact_abp<SPP_REP_HARMONIC>;
        case SPP_REP_R12: return k_interact_abp<SPP_REP_R12>;
        case SPP_REP_YUKAWA_TRUNCATED: return k_interact_abp<SPP_REP_YUKAWA_TRUNCATED>;
        default: return k_interact_abp<SPP_REP_YUKAWA>;
        }
    }
    if (model == SPP_MODEL_CONTINUOUS) return interact_fn<SPP_MODEL_CONTINUOUS, SPP_REP_NONE>(variant);
    switch (rep) {
    case SPP_REP_NONE: return interact_fn<SPP_MODEL_REPULSIVE, SPP_REP_NONE>(variant);
    case SPP_REP_R12_TRUNCATED: return interact_fn<SPP_MODEL_REPULSIVE, SPP_REP_R12_TRUNCATED>(variant);
    case SPP_REP_HARMONIC: return interact_fn<SPP_MODEL_REPULSIVE, SPP_REP_HARMONIC>(variant);
    case SPP_REP_R12: return interact_fn<SPP_MODEL_REPULSIVE, SPP_REP_R12>(variant);
    case SPP_REP_YUKAWA_TRUNCATED: return interact_fn<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA_TRUNCATED>(variant);
    default: return interact_fn<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA>(variant);
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

// Device scratch of the allocation-free entry point (spp_pair_distribution_noalloc): a fixed
// array in the library's code object, one copy per GPU, held until the process ends.  Host code
// uses its per-device address from hipGetSymbolAddress (resolved once per device when the first
// engine on that device is created), never the array's name.
constexpr int kScratchBins = 8192;
static __device__ __align__(256) unsigned long long scratch[kScratchBins];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// The calling thread's current device must be `dev`.
int resolve_scratch(int dev, unsigned long long** out)
{
    if (dev < 0 || dev >= kMaxDevices) return SPP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return SPP_ERR_HIP;
        g_scratch_ptr[dev] = static_cast<unsigned long long*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return SPP_OK;
}

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

#include "lyapunov_kernels.hpp"
#include "persist_kernels.hpp"

// ==================================================================================================
// Engine
// ==================================================================================================

struct Snapshot {
    SetPtrs s{nullptr, nullptr, nullptr, nullptr};
    int64_t step = 0;
    double t = 0;
    double T_hist[3] = {0, 0, 0};  // scheduler state: a replay rebuilds at the same steps
    bool valid = false;
};

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

// One enqueued interval awaiting verification.
struct Interval {
    int anchor_set;  // set a rollback returns to, kept untouched until the interval is verified
    int anchor_since;// steps since the last build there
    int start_set;   // set holding the state the interval's first step reads (after its rebuild)
    int m;           // steps
    int slot;        // flag slot
    int64_t step0;
    int64_t since0;  // steps since the last rebuild at the interval's start (after its rebuild)
    double t0;
    bool rebuilt;
    bool wide;       // its steps ran on the wide list
    long long seq;   // the report count of its slot once the interval's report has been written
};

// Tangent-vector state of a Lyapunov engine (device arrays described in lyapunov_kernels.hpp).
struct LyapState {
    bool on = false;
    int R = 0, K = 0, n = 0, nchunk = 0, rescale = 1;
    double eps = 0, t_start = 0, t_last = 0;
    int64_t renorms = 0;
    int* inv = nullptr;
    int* status = nullptr;
    double *dev = nullptr, *part = nullptr, *coef = nullptr, *rdiag = nullptr, *logsum = nullptr,
           *norm0 = nullptr;
};

struct spp_engine {
    spp_params p{};
    int N = 0;
    Grid G{};
    int nkeys = 0, nscan_blocks = 0, nblk = 0;
    int cap = 32;
    Phys P{};
    double r_cut = 0, half_skin = 0, thresh = 0, T_list = 0;
    // The wide list: a 5x5-cell stencil with radius r_wide (two cell widths), built instead of the
    // ordinary one after a step moved a particle beyond the half skin right after a rebuild (a deep
    // overlap), until the particles slow down.  wide: the current list is wide; wide_next: the next
    // rebuild builds one.
    bool wide_ok = false, wide = false, wide_next = false;
    double r_wide = 0, T_wide = 0, thresh_wide = 0, img_m_wide = 0;
    int64_t wide_rebuilds = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    // State sets.  A rollback returns to an interval's anchor: the state its first step reads (three
    // sets), or, for engines with the tile sort, the state before its rebuild -- a failed tile sort
    // leaves no usable rebuilt state.  Two intervals may be in flight, each keeping its anchor
    // untouched until the host has verified it, so the tile engines hold a fourth set.
    SetPtrs set[kSets] = {};
    int nsets = 3;
    int cur = 0;
    double2 *xy1 = nullptr, *xy2 = nullptr;
    double *th1 = nullptr, *th2 = nullptr;
    double *dth = nullptr, *dth_alt = nullptr, *inj_host_order = nullptr, *inj = nullptr;
    P4 *kacc = nullptr, *force = nullptr, *old = nullptr;
    P4* aux_alt[3] = {nullptr, nullptr, nullptr};
    int* nbr32 = nullptr;        // list of slot indices (always complete)
    int *ncount = nullptr, *cellxy = nullptr;
    int variant = SPP_KERNEL_SLOT;
    bool deep = false;           // SPP_KERNEL_SLOT_DEEP: eight records in flight at every size
    bool one_cut = false;        // repulsion and alignment share one cutoff (single repulsive engines)
    int nstat = 0;               // entries of the per-workgroup statistics arrays
    float2* build = nullptr;
    int *key = nullptr, *perm = nullptr, *cell_count = nullptr, *cell_start = nullptr,
        *block_sums = nullptr;
    int* inv = nullptr;          // caller index -> slot of the current set
    int* perm2 = nullptr;        // the scatter's order, canonicalised within cells
    // tile path of the rebuild (k_sort_tiles): eligible engines, the ping-pong cell starts (the
    // launch reads the previous build's while it writes the new ones), the look-back states and
    // their epoch
    bool sort_ok = false;        // single box, slot kernels, >= 3 cells per side, no graphs
    bool order_valid = false;    // the current set's slot order is the last build's (cell_start)
    bool force_general = false;  // the next rebuild takes the general path (after a tile failure)
    int sort_seg = 0, sort_segw = 0, sort_ntiles = 0;
    int* cell_start_alt = nullptr;
    unsigned long long* tile_state = nullptr;
    unsigned sort_epoch = 0;
    int* sort_fail = nullptr;
    int64_t tile_rebuilds = 0, tile_fallbacks = 0;
    AbpConst abp{};
    SlotBufs slotb[2] = {};      // per-block partials of the (at most two) intervals in flight
    Report hrep[2] = {};         // the slots' last reports, decoded by wait_report
    unsigned long long* hwords = nullptr;  // host-mapped report words, 2 slots x kRepWords
    unsigned long long* dwords = nullptr;  // device view of hwords
    long long* rep_count = nullptr;  // device: reports written per slot
    long long rep_issued[2] = {0, 0};  // host: reports enqueued per slot (launched or replayed)
    // the last enqueued interval's report, not launched yet: it rides on the next interval's first
    // launch, or is launched on its own before the host waits for it
    bool rep_pending = false;
    int rep_pending_slot = 0;
    bool rep_pending_list = false;
    double2* dpolar = nullptr;
    // batched engine: per-system constants and the (fixed) slot -> system map; null for one box
    int nsys = 1;
    std::vector<SysInfo> hsys;
    SysInfo* dsys = nullptr;
    int* sys_of = nullptr;
    double2* dpolar_sys = nullptr;
    SysInfo* dsys1 = nullptr;    // one-entry table describing a single-box engine (sampling)
    double2* dsamples = nullptr; // heading sums recorded during spp_multi_step_sampled
    int64_t samples_alloc = 0, sample_cap = 0, sample_base = 0;
    int sample_every = 0;
    hipEvent_t dev_t0 = nullptr, dev_t1 = nullptr;  // device time of the outermost stepping call
    int dev_depth = 0;
    double device_ms = 0;
    std::map<int, Snapshot> slots;
    LyapState ly;

    int64_t step = 0;
    double t = 0;
    int64_t rebuilds = 0, rollbacks = 0, host_syncs = 0;
    int64_t steps_since_build = 0;
    int64_t intervals = 0;
    bool have_state = false, list_valid = false;
    bool inj_pending = false;
    double T_cross = 0;          // estimated steps from a rebuild to the displacement trigger
    double T_hist[3] = {0, 0, 0};  // the last three estimates, newest first (T_cross derives from them)
    double T_margin = 0.85;      // fraction of T_cross an interval may run on one list
    int force_m = 0;             // length of the next interval after a rollback (0 = predict)
    int max_nbrs = 0;
    double mean_nbrs = 0;
    int unfused_stage = 0;       // next expected stage of the split-call path (0 = step boundary)
    // kernel_variant SPP_KERNEL_PERSIST: one workgroup per system runs whole intervals
    bool persist = false;
    bool persist_valid = false;  // the device statuses describe the current state and lists
    persist::Status* pstatus = nullptr;
    int64_t persist_rebuild_base = 0;
    // graph_intervals: captured fixed-period intervals, keyed by what their launches embed
    struct GraphEntry {
        int cur, protect, slot, m;
        uint64_t gen;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    uint64_t alloc_gen = 0;      // bumped whenever a buffer an interval launch embeds is replaced
    bool dry = false;            // bookkeeping only: launch() enqueues nothing (graph replays)
    bool graph_mode = false;     // launches being captured: stage steps relative to *dclock
    int64_t graph_base = 0;
    long long* dclock = nullptr; // the replayed interval's first step
    int64_t graph_replays = 0, graph_captures = 0;

    unsigned long long* scratch = nullptr;  // this device's static scratch (spp_pair_distribution_noalloc)

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
    if (grid <= 0 || e->dry) return SPP_OK;
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

// ActiveBrownianChain::Set_nb (shared/particle.h:856-876): mobilities of a chain of nb beads
// (single beads 0.5 / 0.5 / 1; longer chains from the aspect ratio of a prolate spheroid).
void chain_mobilities(int nb, double sep, double sigma, double& mpar, double& mperp, double& kperp)
{
    if (nb <= 1) {
        mpar = 0.5;
        mperp = 0.5;
        kperp = 1;
        return;
    }
    const double aspect_ratio = 1 / ((sep * (nb - 1) / sigma) + 1);
    const double e2 = 1 - aspect_ratio * aspect_ratio;
    const double ecc = std::sqrt(e2);
    const double L = std::log((1 + ecc) / (1 - ecc));
    mpar = 1.0;
    mperp = 0.5 * (2 * ecc + (3 * e2 - 1) * L) / ((1 + e2) * L - 2 * ecc);
    kperp = 6 / (1 + aspect_ratio * aspect_ratio) / sigma / sigma;
}

double max_cutoff(const spp_params& p)
{
    if (p.model == SPP_MODEL_ABP) {
        double c = p.repulsion != SPP_REP_NONE ? p.rc_rep : 0.0;
        // two chains interact while bead pairs, up to bead_sep (nb - 1) / 2 from either centre,
        // are closer than rc_bead: the list covers centre distances up to rc_bead + bead_sep (nb - 1)
        if (p.nb_swimmer >= 2) c = std::max(c, p.rc_bead + p.bead_sep * (p.nb_swimmer - 1));
        return c;
    }
    double c = p.R_align;
    if (p.model == SPP_MODEL_REPULSIVE && p.repulsion != SPP_REP_NONE && p.rc_rep > c) c = p.rc_rep;
    return c;
}

void drop_graphs(spp_engine* e)
{
    for (auto& g : e->graphs) (void)hipGraphExecDestroy(g.exec);
    e->graphs.clear();
}

int alloc_nbr(spp_engine* e)
{
    e->alloc_gen++;
    drop_graphs(e);
    dfree(e->nbr32);
    e->nbr32 = nullptr;
    return dalloc(e, &e->nbr32, (size_t)e->cap * e->N);
}

int alloc_set(spp_engine* e, SetPtrs& s)
{
    TRY(dalloc(e, &s.xy, e->N));
    TRY(dalloc(e, &s.th, e->N));
    TRY(dalloc(e, &s.U, e->N));
    TRY(dalloc(e, &s.id, e->N));
    return SPP_OK;
}

void free_set(SetPtrs& s)
{
    dfree(s.xy);
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
    e->wide = e->wide_next && e->wide_ok;
    if (e->wide) e->wide_rebuilds++;
    NlistFlatArgs fa{e->N, e->cap, e->G, e->dsys, e->sys_of, e->wide ? e->T_wide : e->T_list, e->wide ? 1 : 0,
                     e->set[e->cur].xy, e->cell_start, e->nbr32, e->ncount, e->build, e->cellxy,
                     e->slotb[slot].blklen, e->slotb[slot].blknnz};
    return launch(e, "nlist", k_nlist_flat, e->nblk, fa);
}

// The report of the slot's interval, counted as issued (also in dry mode: a dry interval is replayed
// from a captured graph, which runs it).
ReportArgs report_args(spp_engine* e, int slot, bool with_list)
{
    e->rep_issued[slot]++;
    return ReportArgs{e->nstat, e->slotb[slot], with_list ? 1 : 0, e->sort_fail + slot, e->rep_count + slot,
                      e->dwords + slot * kRepWords};
}

int launch_report(spp_engine* e, int slot, bool with_list)
{
    return launch(e, "report", k_report, 1, report_args(e, slot, with_list));
}

// The pending report for a launch that can carry it as its extra workgroup (false: none pending).
bool take_pending_report(spp_engine* e, ReportArgs* r)
{
    if (!e->rep_pending) return false;
    e->rep_pending = false;
    *r = report_args(e, e->rep_pending_slot, e->rep_pending_list);
    return true;
}

// Launches the pending report on its own (before a launch that cannot carry it, or before the
// host waits for it).
int flush_pending_report(spp_engine* e)
{
    if (!e->rep_pending) return SPP_OK;
    e->rep_pending = false;
    return launch_report(e, e->rep_pending_slot, e->rep_pending_list);
}

// Waits until the slot's report number `want` has reached host memory, then decodes it into
// e->hrep[slot].  The host spins until every word of the slot carries the number's tag (reports of
// one slot are numbered 1, 2, ... and complete in order); every ~1024 spins it asks the stream
// whether it failed or ran dry without the report (an error either way), so the wait always ends.
int wait_report(spp_engine* e, int slot, long long want)
{
    const volatile unsigned long long* w = e->hwords + slot * kRepWords;
    const unsigned long long tag = (unsigned long long)(unsigned)want << 32;
    auto done = [&]() {
        for (int k = 0; k < kRepUsed; ++k)
            if ((w[k] & 0xffffffff00000000ull) != tag) return false;
        return true;
    };
    for (long spins = 1;; ++spins) {
        if (done()) break;
        if ((spins & 1023) == 0) {
            const hipError_t q = hipStreamQuery(e->stream);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(e, SPP_ERR_HIP, "interval: %s", hipGetErrorString(q));
            if (q == hipSuccess) {
                // the stream has finished; the words may still be on their way to host memory, so
                // allow them a grace period (about a millisecond) before calling the report lost
                bool ok = false;
                for (long g = 0; g < (1L << 14) && !(ok = done()); ++g) __builtin_ia32_pause();
                if (ok) break;
                return fail(e, SPP_ERR_HIP, "interval report %lld of slot %d missing after the stream drained",
                            want, slot);
            }
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    unsigned long long v[kRepUsed];
    for (int k = 0; k < kRepUsed; ++k) v[k] = w[k] & 0xffffffffull;
    Report r{};
    r.bad_step = (int)(unsigned)v[0];
    r.max_len = (int)(unsigned)v[1];
    r.nnz = (long long)(v[2] | v[3] << 32);
    const unsigned long long bits = v[4] | v[5] << 32;
    std::memcpy(&r.maxdisp2, &bits, sizeof r.maxdisp2);
    r.sort_fail = (int)(unsigned)v[6];
    r.seq = want;
    e->hrep[slot] = r;
    return SPP_OK;
}

// tile: the caller reads the rebuild's failure flag from the interval report (only the interval
// scheduler does); every other caller takes the general sort, which needs no premise on the order.
int launch_rebuild(spp_engine* e, int dst, int slot, bool aux, bool tile)
{
    int N = e->N, nk = e->nkeys, src = e->cur;
    if (tile && !aux && e->sort_ok && e->order_valid && !e->force_general) {
        // tile path: one launch from the current order to the new sorted set
        SortArgs sa{};
        sa.N = N;
        sa.G = e->G;
        sa.seg = e->sort_seg;
        sa.segw = e->sort_segw;
        sa.ntiles = e->sort_ntiles;
        sa.in = e->set[src];
        sa.out = e->set[dst];
        sa.inv = e->inv;
        sa.cs_old = e->cell_start;
        sa.cs_new = e->cell_start_alt;
        sa.tstate = e->tile_state;
        sa.fail = e->sort_fail + slot;
        e->sort_epoch = e->sort_epoch % ((1u << 30) - 1) + 1;  // 1 .. 2^30 - 1, never the cleared 0
        sa.epoch = e->sort_epoch;
        sa.rep_on = take_pending_report(e, &sa.rep) ? 1 : 0;
        TRY(launch(e, "sort_tiles", k_sort_tiles, e->sort_ntiles + sa.rep_on, sa));
        std::swap(e->cell_start, e->cell_start_alt);
        e->tile_rebuilds++;
        e->cur = dst;
        TRY(launch_nlist(e, slot));
        e->persist_valid = false;
        e->rebuilds++;
        e->steps_since_build = 0;
        e->list_valid = true;
        return SPP_OK;
    }
    TRY(flush_pending_report(e));
    e->force_general = false;
    e->order_valid = true;  // the general path sorts any order
    // the arrival ranks live in perm2 until the per-cell ordering overwrites it
    TRY(launch(e, "bin", k_bin, nblocks(N), N, e->G, (const double2*)e->set[src].xy, e->key,
               e->cell_count, e->perm2, (const SysInfo*)e->dsys, (const int*)e->sys_of));
    TRY(launch(e, "scan_partial", k_scan_partial, e->nscan_blocks, (const int*)e->cell_count, nk,
               e->block_sums));
    TRY(launch(e, "scan_final", k_scan_final, e->nscan_blocks, e->cell_count, nk, (const int*)e->block_sums,
               e->cell_start));
    TRY(launch(e, "scatter", k_scatter, nblocks(N), N, (const int*)e->key,
               (const int*)e->cell_start, (const int*)e->perm2, e->perm));
    PermuteArgs pa{};
    pa.N = N;
    pa.perm = e->perm2;
    pa.in = e->set[src];
    pa.out = e->set[dst];
    pa.inv = e->inv;
    if (aux) {
        pa.a_in[0] = e->kacc;
        pa.a_in[1] = e->force;
        pa.a_in[2] = e->old;
        for (int q = 0; q < 3; ++q) pa.a_out[q] = e->aux_alt[q];
        pa.d_in = e->dth;
        pa.d_out = e->dth_alt;
    }
    TRY(launch(e, "rank_permute", k_rank_permute, nblocks(N), pa, (const int*)e->key, (const int*)e->cell_start,
               (const int*)e->perm, e->perm2));
    if (aux) {
        e->alloc_gen++;
        drop_graphs(e);
        std::swap(e->kacc, e->aux_alt[0]);
        std::swap(e->force, e->aux_alt[1]);
        std::swap(e->old, e->aux_alt[2]);
        std::swap(e->dth, e->dth_alt);
    }
    e->cur = dst;
    TRY(launch_nlist(e, slot));
    e->persist_valid = false;
    e->rebuilds++;
    e->steps_since_build = 0;
    e->list_valid = true;
    return SPP_OK;
}

// Host-side guard: device ids index caller arrays in the getters, so they are checked first.
int check_ids(spp_engine* e, const std::vector<int>& id)
{
    for (int v : id)
        if ((unsigned)v >= (unsigned)e->N) return fail(e, SPP_ERR_STATE, "device particle id %d out of range", v);
    return SPP_OK;
}

int free_set_other_than(const spp_engine* e, int a, int b, int c = -1)
{
    for (int k = 0; k < e->nsets; ++k)
        if (k != a && k != b && k != c) return k;
    return -1;
}

// Synchronous rebuild with a capacity check (used outside the interval scheduler).
int rebuild_sync(spp_engine* e, bool aux)
{
    int dst = free_set_other_than(e, e->cur, -1);
    TRY(launch_rebuild(e, dst, 0, aux, false));
    for (;;) {
        TRY(launch(e, "fill", k_fill<double>, nblocks(e->nstat), e->slotb[0].blkmax, 0.0, e->nstat));
        TRY(launch_report(e, 0, true));
        HIPC(e, hipStreamSynchronize(e->stream));
        TRY(wait_report(e, 0, e->rep_issued[0]));
        e->host_syncs++;
        int need = e->hrep[0].max_len;
        e->max_nbrs = need;
        e->mean_nbrs = (double)e->hrep[0].nnz / e->N;
        if (need <= e->cap) return SPP_OK;
        e->cap = ((need + 7) / 8) * 8 + 8;
        TRY(alloc_nbr(e));
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
// The displacement trigger of the list in place (ordinary or wide).
double cur_thresh(const spp_engine* e) { return e->wide ? e->thresh_wide : e->thresh; }

int launch_step(spp_engine* e, int S, int W, int kstep, int slot, bool record)
{
    int integ = e->p.integrator;
    StageArgs a{};
    a.N = e->N;
    a.P = e->P;
    a.sys_tab = e->dsys;
    a.sys_of = e->sys_of;
    a.abp = e->abp;
    a.inv = e->inv;
    a.dth = e->dth;
    a.kacc = e->kacc;
    a.nbr32 = e->nbr32;
    a.ncount = e->ncount;
    a.id = e->set[S].id;
    a.build = e->build;
    a.step = e->graph_mode ? e->step - e->graph_base : e->step;
    a.clock = e->graph_mode ? e->dclock : nullptr;
    a.kstep = kstep;
    a.thresh2 = cur_thresh(e) * cur_thresh(e);
    if (e->wide) a.P.img_m = e->img_m_wide;
    a.blkbad = e->slotb[slot].blkbad;
    a.blockmax = e->slotb[slot].blkmax;
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
            a.gxy = e->set[S].xy;
            a.gth = e->set[S].th;
            a.oxy = e->xy1;
            a.oth = e->th1;
        } else if (integ == SPP_RK2) {
            a.gxy = s == 1 ? e->set[S].xy : e->xy1;
            a.gth = s == 1 ? e->set[S].th : e->th1;
            a.oxy = s == 1 ? e->xy1 : e->set[W].xy;
            a.oth = s == 1 ? e->th1 : e->set[W].th;
        } else {
            const double2* g[5] = {nullptr, e->set[S].xy, e->xy1, e->xy2, e->xy1};
            const double* gt[5] = {nullptr, e->set[S].th, e->th1, e->th2, e->th1};
            double2* o[5] = {nullptr, e->xy1, e->xy2, e->xy1, e->set[W].xy};
            double* ot[5] = {nullptr, e->th1, e->th2, e->th1, e->set[W].th};
            a.gxy = g[s];
            a.gth = gt[s];
            a.oxy = o[s];
            a.oth = ot[s];
        }
        a.stream = e->N <= kStreamMaxN;
        StageFn fn = pick_stage(e->p.model, e->p.repulsion, integ, stage, e->variant, e->deep || e->N < kSmallN,
                                e->one_cut);
        a.grid = e->nblk;
        a.rep_on = 0;
        if (e->rep_pending) {
            // the previous interval's report rides on this launch (the stage kernels of the slot
            // family carry it; the active-Brownian stages do not)
            if (e->p.model == SPP_MODEL_ABP)
                TRY(flush_pending_report(e));
            else
                a.rep_on = take_pending_report(e, &a.rep) ? 1 : 0;
        }
        TRY(launch(e, stage_name(integ, stage), fn, e->nblk + a.rep_on, a));
        a.rep_on = 0;
        a.inj = nullptr;
        a.id_out = nullptr;
    }
    if (integ == SPP_EULER) {
        std::swap(e->set[W].xy, e->xy1);
        std::swap(e->set[W].th, e->th1);
    }
    e->cur = W;
    e->step++;
    e->t += e->p.dt;
    e->steps_since_build++;
    if (e->sample_every > 0 && (e->step - e->sample_base) % e->sample_every == 0) {
        // heading sums of every system after this step (a replay after a rollback rewrites them)
        int64_t k = (e->step - e->sample_base) / e->sample_every - 1;
        if (k >= 0 && k < e->sample_cap) {
            const int ns = e->dsys ? e->nsys : 1;
            TRY(launch(e, "polar_sample", k_polar_sys, ns, (const SysInfo*)(e->dsys ? e->dsys : e->dsys1),
                       (const double*)e->set[W].th, e->dsamples + k * ns));
        }
    }
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
// The launches of one interval: [rebuild] + m steps + report.  Host bookkeeping only when e->dry.
int interval_body(spp_engine* e, int m, bool rebuild, int protect, int slot, Interval& iv)
{
    // slot's per-block bad steps: INT_MAX from creation or from the report that closed its last use
    // The anchor: the state before the rebuild on tile engines (a failed tile sort is redone from
    // there), the state after it otherwise (a rollback then rebuilds that already sorted state).
    const int origin = e->cur;
    const int64_t origin_since = e->steps_since_build;
    if (rebuild) {
        int dst = free_set_other_than(e, e->cur, protect);
        TRY(launch_rebuild(e, dst, slot, false, true));
    }
    iv.start_set = e->cur;
    iv.wide = e->wide;
    iv.step0 = e->step;
    iv.since0 = e->steps_since_build;
    iv.t0 = e->t;
    iv.anchor_set = e->sort_ok ? origin : e->cur;
    iv.anchor_since = (int)std::min<int64_t>(e->sort_ok ? origin_since : e->steps_since_build, INT_MAX);
    e->graph_base = e->step;
    // the output set spares the anchor and the front interval's anchor
    int W = free_set_other_than(e, e->cur, protect, iv.anchor_set);
    for (int k = 0; k < m; ++k) {
        TRY(launch_step(e, e->cur, k == 0 ? W : e->cur, k, slot, k == m - 1));
        if (k == 0) e->inj_pending = false;
    }
    if (e->graph_mode || e->dry) return launch_report(e, slot, rebuild);
    TRY(flush_pending_report(e));  // (this interval's first launch has taken the previous one)
    // the report rides on the next interval's first launch (or is launched before the host waits)
    e->rep_pending = true;
    e->rep_pending_slot = slot;
    e->rep_pending_list = rebuild;
    return SPP_OK;
}

__global__ void k_set_i64(long long* p, long long v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// A full fixed-period interval replayed from a captured HIP graph.  Everything its launches embed
// is fixed by the key: the state set it starts from and the one kept for rollback (which fix the
// rebuild's destination and the steps' output set), the report slot, the length and the buffer
// generation; the noise step comes from the device clock set just before the replay.
bool graphable(const spp_engine* e, int m, bool rebuild)
{
    return e->p.graph_intervals && e->p.rebuild == SPP_REBUILD_PERIOD && rebuild && m == e->p.rebuild_period &&
           e->p.integrator != SPP_EULER && !e->inj_pending && !e->profiling && e->sample_every == 0;
}

int enqueue_interval(spp_engine* e, int m, bool rebuild, int protect, Interval& iv)
{
    int slot = (int)(e->intervals++ & 1);
    iv.slot = slot;
    iv.m = m;
    iv.rebuilt = rebuild;
    if (!graphable(e, m, rebuild)) {
        TRY(interval_body(e, m, rebuild, protect, slot, iv));
    } else {
        TRY(flush_pending_report(e));  // a captured launch must not carry another interval's report
        if (!e->dclock) TRY(dalloc(e, &e->dclock, 1));
        spp_engine::GraphEntry* g = nullptr;
        for (auto& x : e->graphs)
            if (x.cur == e->cur && x.protect == protect && x.slot == slot && x.m == m && x.gen == e->alloc_gen) g = &x;
        const int cur0 = e->cur;
        if (!g) {
            // capture: the launches are recorded, not run; the bookkeeping advances as for a launch
            hipGraph_t graph = nullptr;
            const long long issued0 = e->rep_issued[slot];  // a capture that fails runs no report
            HIPC(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
            e->graph_mode = true;
            int rc = interval_body(e, m, rebuild, protect, slot, iv);
            e->graph_mode = false;
            hipError_t ce = hipStreamEndCapture(e->stream, &graph);
            if (rc != SPP_OK || ce != hipSuccess) e->rep_issued[slot] = issued0;
            if (rc != SPP_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ce != hipSuccess) return fail(e, SPP_ERR_HIP, "interval capture: %s", hipGetErrorString(ce));
            hipGraphExec_t exec = nullptr;
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                e->rep_issued[slot] = issued0;
                return fail(e, SPP_ERR_HIP, "interval graph: %s", hipGetErrorString(ie));
            }
            e->graphs.push_back(spp_engine::GraphEntry{cur0, protect, slot, m, e->alloc_gen, exec});
            g = &e->graphs.back();
            e->graph_captures++;
        } else {
            e->dry = true;
            int rc = interval_body(e, m, rebuild, protect, slot, iv);
            e->dry = false;
            TRY(rc);
        }
        hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(64), 0, e->stream, e->dclock, (long long)iv.step0);
        HIPC(e, hipGetLastError());
        HIPC(e, hipGraphLaunch(g->exec, e->stream));
        e->graph_replays++;
    }
    iv.seq = e->rep_issued[slot] + (e->rep_pending && e->rep_pending_slot == slot ? 1 : 0);
    return SPP_OK;
}

// Whether the next interval starts with a rebuild.  Displacement policy: always when the list does
// not describe the current state; otherwise only once the steps since the last build reach the
// predicted safe length (0.85 of the estimated crossing).  An interval that continues on the old
// list is still checked step by step against the build positions, and a crossing rolls it back to
// a rebuild (run_intervals), so continuing is safe; it spares a rebuild at the start of every call
// of a short-call loop (Box::One_Step).
bool rebuild_due(const spp_engine* e)
{
    if (!e->list_valid) return true;
    if (e->p.rebuild == SPP_REBUILD_PERIOD) return e->steps_since_build >= e->p.rebuild_period;
    if (e->steps_since_build == 0) return false;
    if (e->T_cross <= 0) return true;  // no displacement history yet
    return e->T_margin * e->T_cross - (double)e->steps_since_build < 1.0;
}

// The displacement policy's estimate of the steps from a rebuild to the trigger, from the last
// three per-interval estimates (each a linear extrapolation of that interval's largest
// displacement).  Normally the smallest of them with a 15 % margin, which covers fluctuations and
// acceleration.  While the estimates grow by more than 5 % twice in a row and stay short (the
// particles slow down, as after a random start whose overlaps push particles apart), the newest one
// without margin: linear extrapolation of a decelerating displacement already overestimates it, and
// a mistake costs at most a short interval (the crossing is still caught step by step).
void derive_schedule(spp_engine* e)
{
    const double* h = e->T_hist;
    double T = 0.0;
    for (int k = 0; k < 3; ++k)
        if (h[k] > 0 && (T == 0.0 || h[k] < T)) T = h[k];
    const bool slowing = h[2] > 0 && h[0] > 1.05 * h[1] && h[1] > 1.05 * h[2] && h[0] < 8.0;
    e->T_cross = slowing ? h[0] : T;
    e->T_margin = slowing ? 1.0 : 0.85;
}

int predict_m(spp_engine* e, int remaining, bool rebuild)
{
    int m;
    if (e->force_m > 0) {
        m = e->force_m;
        e->force_m = 0;
    } else if (e->p.rebuild == SPP_REBUILD_PERIOD) {
        // an interval that starts with a rebuild runs a whole period
        m = rebuild ? e->p.rebuild_period : std::max(1, e->p.rebuild_period - (int)e->steps_since_build);
    } else if (e->T_cross <= 0) {
        m = 4;  // no history yet: a short interval to measure the displacement rate
    } else if (rebuild) {
        m = std::max(1, (int)(e->T_margin * e->T_cross));
    } else {
        // the rest of the current list's predicted safe length
        m = std::max(1, (int)(e->T_margin * e->T_cross - (double)e->steps_since_build));
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
            // a replay after a rollback (force_m) restarts from the rebuild it rolled back to
            bool rebuild = rebuild_due(e);
            int m = predict_m(e, to_enqueue, rebuild);
            int protect = q.empty() ? -1 : q.front().anchor_set;
            Interval iv{};
            TRY(enqueue_interval(e, m, rebuild, protect, iv));
            q.push_back(iv);
            to_enqueue -= m;
            continue;
        }
        Interval iv = q.front();
        if (e->rep_pending && e->rep_pending_slot == iv.slot) TRY(flush_pending_report(e));
        TRY(wait_report(e, iv.slot, iv.seq));
        e->host_syncs++;
        const Report f = e->hrep[iv.slot];
        const bool sort_fail = iv.rebuilt && f.sort_fail != 0;
        bool overflow = !sort_fail && iv.rebuilt && f.max_len > e->cap;
        int bad = period ? INT_MAX : f.bad_step;
        if (iv.rebuilt && !sort_fail) {
            e->max_nbrs = f.max_len;
            e->mean_nbrs = (double)f.nnz / e->N;
        }
        if (!sort_fail && !overflow && bad >= iv.m) {
            q.pop_front();
            if (!period && iv.rebuilt) {
                // linear extrapolation of the largest displacement to the trigger
                double disp = std::sqrt(f.maxdisp2);
                double est = disp > 0 ? (iv.wide ? e->thresh_wide : e->thresh) / disp * iv.m : 4096.0;
                e->T_hist[2] = e->T_hist[1];
                e->T_hist[1] = e->T_hist[0];
                e->T_hist[0] = est;
                derive_schedule(e);
                // back to the ordinary list once the fastest particle covers less than half its
                // trigger per step (the next rebuild builds it; the estimates start afresh)
                if (iv.wide && e->wide_next && disp < 0.5 * e->thresh * iv.m) {
                    e->wide_next = false;
                    for (double& h : e->T_hist) h = 0;
                    derive_schedule(e);
                }
            }
            continue;
        }
        // Roll back to the interval's start; it and everything enqueued after it are redone.
        HIPC(e, hipStreamSynchronize(e->stream));
        // a discarded interval's report still runs: it clears its slot's per-block bad steps and
        // failure flag for the slot's next interval (its result is not read)
        TRY(flush_pending_report(e));
        int redo = 0;
        for (auto& x : q) redo += x.m;
        q.clear();
        // back to the interval's anchor; the redo rebuilds from there
        e->cur = iv.anchor_set;
        e->step = iv.step0;
        e->t = iv.t0;
        e->steps_since_build = iv.anchor_since;
        e->rollbacks++;
        if (e->p.noise_mode == SPP_NOISE_INJECTED) e->inj_pending = true;  // replay uses it again
        to_enqueue += redo;
        e->list_valid = false;  // the device list may belong to a later interval
        e->order_valid = false; // and the cell starts to a later build
        if (sort_fail) {
            // the tile path could not sort (a particle moved farther than a cell, or a tile
            // overflowed): same rebuild point on the general path; after three, general for good
            e->force_general = true;
            if (++e->tile_fallbacks >= 3) e->sort_ok = false;
            e->force_m = iv.m;
            continue;
        }
        if (overflow) {
            // same rebuild point (the interval started with one), larger capacity
            e->cap = ((f.max_len + 7) / 8) * 8 + 8;
            TRY(alloc_nbr(e));
            e->force_m = iv.m;
            continue;
        }
        if (bad == 0 && iv.rebuilt) {
            // A step right after a rebuild moved a particle beyond the trigger (a deep overlap
            // pushed apart): redo it on the wide list, whose skin is several times wider.
            if (e->wide_ok && !iv.wide) {
                e->wide_next = true;
                e->force_m = 1;
                for (double& h : e->T_hist) h = 0;
                derive_schedule(e);
                continue;
            }
            return fail(e, SPP_ERR_STATE,
                        "a particle moves farther than %g within one step%s: raise r_list or lower dt",
                        iv.wide ? e->thresh_wide : e->half_skin, iv.wide ? " (on the wide list)" : "");
        }
        // The rebuild at the interval's start (identical, the state there is already sorted) is
        // redone; the valid prefix [0, bad) is replayed as its own interval and verified again.
        e->force_m = bad > 0 ? bad : 0;
        // the crossing came earlier than predicted: every estimate becomes 0.8 of the steps it took
        for (double& h : e->T_hist) h = std::max(1.0, 0.8 * (iv.since0 + std::max(bad, 1)));
        derive_schedule(e);
    }
    return SPP_OK;
}

using PersistFn = void (*)(persist::Args);

template <int MODEL, int REP>
PersistFn persist_integ(int integ)
{
    if (integ == SPP_EULER) return persist::k_persist<MODEL, REP, SPP_EULER>;
    if (integ == SPP_RK2) return persist::k_persist<MODEL, REP, SPP_RK2>;
    return persist::k_persist<MODEL, REP, SPP_RK4>;
}

PersistFn pick_persist(int model, int rep, int integ)
{
    if (model == SPP_MODEL_CONTINUOUS) return persist_integ<SPP_MODEL_CONTINUOUS, SPP_REP_NONE>(integ);
    switch (rep) {
    case SPP_REP_NONE: return persist_integ<SPP_MODEL_REPULSIVE, SPP_REP_NONE>(integ);
    case SPP_REP_R12_TRUNCATED: return persist_integ<SPP_MODEL_REPULSIVE, SPP_REP_R12_TRUNCATED>(integ);
    case SPP_REP_HARMONIC: return persist_integ<SPP_MODEL_REPULSIVE, SPP_REP_HARMONIC>(integ);
    case SPP_REP_R12: return persist_integ<SPP_MODEL_REPULSIVE, SPP_REP_R12>(integ);
    case SPP_REP_YUKAWA_TRUNCATED: return persist_integ<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA_TRUNCATED>(integ);
    default: return persist_integ<SPP_MODEL_REPULSIVE, SPP_REP_YUKAWA>(integ);
    }
}

// Steps through the persistent kernel: every system runs to the common target step in one launch;
// a list longer than the capacity stops that system at the step it was building for, the lists are
// reallocated and the launch repeated (finished systems return at once).
int run_persist(spp_engine* e, int steps)
{
    const int ns = e->dsys ? e->nsys : 1;
    std::vector<persist::Status> h(ns);
    if (!e->persist_valid) {
        for (auto& x : h) {
            x = persist::Status{};
            x.step = e->step;
            x.since = -1;
        }
        HIPC(e, hipMemcpy(e->pstatus, h.data(), ns * sizeof(persist::Status), hipMemcpyHostToDevice));
        e->persist_valid = true;
        e->persist_rebuild_base = e->rebuilds;
    }
    const long long target = e->step + steps;
    const int spare = free_set_other_than(e, e->cur, -1);
    PersistFn fn = pick_persist(e->p.model, e->p.repulsion, e->p.integrator);
    for (;;) {
        persist::Args a{};
        StageArgs& st = a.st;
        st.N = e->N;
        st.P = e->P;
        st.abp = e->abp;
        st.inv = e->inv;
        st.sys_tab = e->dsys;
        st.sys_of = e->sys_of;
        st.dth = e->dth;
        st.kacc = e->kacc;
        st.nbr32 = e->nbr32;
        st.ncount = e->ncount;
        st.build = e->build;
        a.nl = NlistFlatArgs{e->N, e->cap, e->G, e->dsys, e->sys_of, e->T_list, 0, nullptr, e->cell_start,
                             e->nbr32, e->ncount, e->build, e->cellxy, nullptr, nullptr};
        a.sys_tab = e->dsys ? e->dsys : e->dsys1;
        a.cur = e->set[e->cur];
        a.spare = e->set[spare];
        a.xy1 = e->xy1;
        a.xy2 = e->xy2;
        a.th1 = e->th1;
        a.th2 = e->th2;
        a.key = e->key;
        a.count = e->cell_count;
        a.start = e->cell_start;
        a.perm = e->perm;
        a.perm2 = e->perm2;
        a.inv = e->inv;
        a.target = target;
        a.period = e->p.rebuild == SPP_REBUILD_PERIOD ? e->p.rebuild_period : 0;
        a.nsys = ns;
        a.thresh2 = e->thresh * e->thresh;
        a.samples = e->sample_every > 0 ? e->dsamples : nullptr;
        a.sample_every = e->sample_every;
        a.sample_base = e->sample_base;
        a.sample_cap = e->sample_cap;
        a.status = e->pstatus;
        ProfRec rec{"persist", nullptr, nullptr};
        if (e->profiling) {
            rec.a = take_event(e);
            rec.b = take_event(e);
            HIPC(e, hipEventRecord(rec.a, e->stream));
        }
        hipLaunchKernelGGL(fn, dim3(ns), dim3(persist::kBlock1024), 0, e->stream, a);
        HIPC(e, hipGetLastError());
        if (e->profiling) {
            HIPC(e, hipEventRecord(rec.b, e->stream));
            e->prof.push_back(rec);
        }
        HIPC(e, hipMemcpyAsync(h.data(), e->pstatus, ns * sizeof(persist::Status), hipMemcpyDeviceToHost, e->stream));
        HIPC(e, hipStreamSynchronize(e->stream));
        e->host_syncs++;
        int need = 0;
        long long nnz = 0;
        for (int s = 0; s < ns; ++s) {
            if (h[s].code == persist::kSkinJump)
                return fail(e, SPP_ERR_STATE,
                            "a particle moves farther than the half skin (%g) within one step: "
                            "raise r_list or lower dt", e->half_skin);
            if (h[s].code == persist::kNeedCap) need = std::max(need, h[s].need_cap);
            e->max_nbrs = std::max(e->max_nbrs, h[s].max_len);
            nnz += h[s].nnz;
        }
        e->mean_nbrs = (double)nnz / e->N;
        if (need == 0) break;
        // Longer lists.  Row k of slot i sits at k N + i, so the old array is a prefix of the new
        // one: the other systems keep their lists (and so their rebuild steps); the systems that
        // overflowed rebuild where they stopped.
        const int old_cap = e->cap;
        int* old32 = e->nbr32;
        e->nbr32 = nullptr;
        e->cap = ((need + 7) / 8) * 8 + 8;
        TRY(dalloc(e, &e->nbr32, (size_t)e->cap * e->N));
        HIPC(e, hipMemcpyAsync(e->nbr32, old32, (size_t)old_cap * e->N * sizeof(int), hipMemcpyDeviceToDevice,
                               e->stream));
        HIPC(e, hipStreamSynchronize(e->stream));
        dfree(old32);
        for (auto& x : h)
            if (x.code == persist::kNeedCap) {
                x.code = persist::kRunning;
                x.since = -1;
            }
        HIPC(e, hipMemcpy(e->pstatus, h.data(), ns * sizeof(persist::Status), hipMemcpyHostToDevice));
    }
    int64_t rb = 0;
    bool lists = true;
    for (auto& x : h) {
        rb += x.rebuilds;
        lists = lists && x.since >= 0;
    }
    e->rebuilds = e->persist_rebuild_base + rb;
    for (int k = 0; k < steps; ++k) e->t += e->p.dt;  // the per-step sum the multi-launch path forms
    e->step = target;
    // every system's list covers its state unless one still has to rebuild
    e->list_valid = lists;
    e->steps_since_build = 0;
    return SPP_OK;
}

void lyap_free(LyapState& L)
{
    void* bufs[] = {L.inv, L.status, L.dev, L.part, L.coef, L.rdiag, L.logsum, L.norm0};
    for (void* b : bufs) dfree(b);
    L = LyapState{};
}

lyap::Args lyap_args(spp_engine* e)
{
    const LyapState& L = e->ly;
    SetPtrs& s = e->set[e->cur];
    lyap::Args a{};
    a.R = L.R;
    a.K = L.K;
    a.n = L.n;
    a.nchunk = L.nchunk;
    a.sys = e->dsys;
    a.inv = L.inv;
    a.pos = s.xy;
    a.th = s.th;
    a.U = s.U;
    a.dev = L.dev;
    a.part = L.part;
    a.coef = L.coef;
    a.rdiag = L.rdiag;
    a.logsum = L.logsum;
    a.norm0 = L.norm0;
    a.status = L.status;
    return a;
}

template <int K>
int lyap_qr_k(spp_engine* e, const lyap::Args& a, bool init)
{
    const LyapState& L = e->ly;
    for (int pass = 0; pass < 2; ++pass) {
        TRY(launch(e, "lyap_gram", lyap::k_gram<K>, L.R * L.nchunk, a));
        TRY(launch(e, "lyap_chol", lyap::k_chol, nblocks(L.R), a, pass, init ? 1 : 0, L.rescale, L.eps));
        TRY(launch(e, "lyap_apply", lyap::k_apply<K>, nblocks(L.R * L.n), a));
    }
    return SPP_OK;
}

// Orthonormalise the deviations in e->ly.dev (Cholesky QR, twice), accumulate the growth and put
// every clone at base + scale * direction.
int lyap_qr(spp_engine* e, bool init)
{
    lyap::Args a = lyap_args(e);
    switch (e->ly.K) {
    case 1: return lyap_qr_k<1>(e, a, init);
    case 2: return lyap_qr_k<2>(e, a, init);
    case 3: return lyap_qr_k<3>(e, a, init);
    case 4: return lyap_qr_k<4>(e, a, init);
    case 5: return lyap_qr_k<5>(e, a, init);
    case 6: return lyap_qr_k<6>(e, a, init);
    case 7: return lyap_qr_k<7>(e, a, init);
    case 8: return lyap_qr_k<8>(e, a, init);
    }
    return fail(e, SPP_ERR_ARG, "vectors must be 1..%d", lyap::kMaxK);
}

int lyap_check_status(spp_engine* e)
{
    std::vector<int> st(e->ly.R);
    HIPC(e, hipMemcpyAsync(st.data(), e->ly.status, st.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int r = 0; r < e->ly.R; ++r)
        if (st[r])
            return fail(e, SPP_ERR_STATE, "replica %d: tangent vectors became linearly dependent "
                        "(Gram matrix not positive definite)", r);
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
    } else if (p.model == SPP_MODEL_ABP) {
        P.c_align = 0;  // no alignment: the pair kernel sees T_align = 0
        P.c_self = 0;
    } else {
        P.c_align = p.g / (spp::kPi * p.R_align * p.R_align);
        P.c_self = 0;
    }
    P.T_align = p.model == SPP_MODEL_ABP ? 0.0 : sqrt_threshold(p.R_align);
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
    P.img_m = std::numeric_limits<double>::infinity();  // set by the engine for displacement lists
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
    p->k_bond = 100;
    p->l_bond = 1;
    if (model == SPP_MODEL_ABP) {
        // single-bead ActiveBrownianChain: every pair uses A_p with the membrane-bead sigma_m and
        // repulsion_radius_m (particle.h:968-981, 1048-1053); F0 = 1 (membrane.cpp:150); ring rest
        // length repulsion_radius_m (serial/beadbox.h:292)
        p->repulsion = SPP_REP_R12_TRUNCATED;
        p->A_rep = 1;
        p->sigma_rep = 0.5;
        p->rc_rep = 0.55;
        p->l_bond = 0.55;
        p->v0 = 1;
        p->r_list = 0.8;
    }
    // multi-bead swimmers (particle.h:1048-1056): single beads unless nb_swimmer >= 2
    p->nb_swimmer = 1;
    p->bead_sep = 1;     // ActiveBrownianChain::separation
    p->sigma_bead = 1;   // sigma_p
    p->rc_bead = 1.1;    // repulsion_radius
    p->A_bead = 1;       // A_p
    if (model == SPP_MODEL_CONTINUOUS) {
        p->repulsion = SPP_REP_NONE;
        p->R_align = 1;      // d2 < 1, particle.h:305
        p->g = 4;            // ContinuousParticle::g, particle.h:373
    }
    return SPP_OK;
}

}  // extern "C"

namespace {

// One box of an engine before validation: spp_create describes one, spp_create_batch many.
struct SysDesc {
    int n, ncx, ncy;
    double Lx, Ly, v0, g, D;
    uint64_t seed;
};

int create_engine(const spp_params* p, const spp_system* systems, int nsys, spp_engine** out)
{
    if (!p || !out) return fail(nullptr, SPP_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->abi_version != SPP_ABI_VERSION)
        return fail(nullptr, SPP_ERR_ARG, "abi_version %d, library is %d", p->abi_version, SPP_ABI_VERSION);
    if (p->model != SPP_MODEL_REPULSIVE && p->model != SPP_MODEL_CONTINUOUS && p->model != SPP_MODEL_ABP)
        return fail(nullptr, SPP_ERR_UNSUPPORTED, "model %d not available", p->model);
    if (p->integrator != SPP_EULER && p->integrator != SPP_RK2 && p->integrator != SPP_RK4)
        return fail(nullptr, SPP_ERR_ARG, "integrator %d", p->integrator);
    if (p->model == SPP_MODEL_ABP) {
        // ActiveBrownianChain has Move and Move_Runge_Kutta2_1/2 only (particle.h:898-962)
        if (p->integrator == SPP_RK4)
            return fail(nullptr, SPP_ERR_UNSUPPORTED, "the active Brownian model integrates with Euler or RK2");
        if (systems || p->kernel_variant != SPP_KERNEL_SLOT)
            return fail(nullptr, SPP_ERR_UNSUPPORTED, "the active Brownian model runs one box on the slot kernels");
        if (p->n_membrane < 0 || p->n_membrane > p->n || (p->n_membrane > 0 && p->n_membrane < 3))
            return fail(nullptr, SPP_ERR_ARG, "n_membrane must be 0 or in [3, n]");
        if (p->n_membrane > 0 && !(p->l_bond > 0))
            return fail(nullptr, SPP_ERR_ARG, "l_bond must be positive");
        if (p->nb_swimmer >= 2 &&
            (p->nb_swimmer > 16 || !(p->bead_sep > 0) || !(p->sigma_bead > 0) || !(p->rc_bead > 0)))
            return fail(nullptr, SPP_ERR_ARG, "chains need 2 <= nb_swimmer <= 16 and positive bead_sep, sigma_bead, rc_bead");
    }
    if (p->repulsion < SPP_REP_NONE || p->repulsion > SPP_REP_YUKAWA)
        return fail(nullptr, SPP_ERR_ARG, "repulsion %d", p->repulsion);
    if (!(p->dt > 0) || !(p->r_list > 0)) return fail(nullptr, SPP_ERR_ARG, "dt and r_list must be positive");
    if (p->noise_mode != SPP_NOISE_STREAM && p->noise_mode != SPP_NOISE_INJECTED)
        return fail(nullptr, SPP_ERR_ARG, "noise_mode %d", p->noise_mode);
    if (p->rebuild != SPP_REBUILD_DISPLACEMENT && p->rebuild != SPP_REBUILD_PERIOD)
        return fail(nullptr, SPP_ERR_ARG, "rebuild %d", p->rebuild);
    if (p->kernel_variant < SPP_KERNEL_SLOT || p->kernel_variant > SPP_KERNEL_SLOT_DEEP)
        return fail(nullptr, SPP_ERR_ARG, "kernel_variant %d", p->kernel_variant);
    if (p->kernel_variant == 1 || p->kernel_variant == 2)
        return fail(nullptr, SPP_ERR_UNSUPPORTED, "kernel_variant %d is retired (slower than 0 everywhere)",
                    p->kernel_variant);
    if (systems && p->kernel_variant != SPP_KERNEL_SLOT && p->kernel_variant != SPP_KERNEL_PERSIST &&
        p->kernel_variant != SPP_KERNEL_SLOT_DEEP)
        return fail(nullptr, SPP_ERR_UNSUPPORTED, "batched engines run the slot kernels (kernel_variant 0, 3 or 4)");
    const bool persist = p->kernel_variant == SPP_KERNEL_PERSIST;
    if (persist && p->noise_mode != SPP_NOISE_STREAM)
        return fail(nullptr, SPP_ERR_UNSUPPORTED, "the persistent kernel draws its noise from the stream");
    if (p->rebuild == SPP_REBUILD_PERIOD && p->rebuild_period <= 0)
        return fail(nullptr, SPP_ERR_ARG, "rebuild_period must be positive");
    if (p->rebuild_sort != SPP_REBUILD_SORT_AUTO && p->rebuild_sort != SPP_REBUILD_SORT_GENERAL)
        return fail(nullptr, SPP_ERR_ARG, "rebuild_sort %d", p->rebuild_sort);
    double rcut = max_cutoff(*p);
    if (p->r_list < rcut) return fail(nullptr, SPP_ERR_ARG, "r_list %g below cutoff %g", p->r_list, rcut);

    std::vector<SysDesc> sd(nsys);
    for (int s = 0; s < nsys; ++s) {
        if (systems) {
            const spp_system& q = systems[s];
            sd[s] = SysDesc{q.n, q.ncell_x, q.ncell_y, q.Lx, q.Ly, q.v0, q.g, q.D_rot, q.seed};
        } else {
            sd[s] = SysDesc{p->n, p->ncell_x, p->ncell_y, p->Lx, p->Ly, p->v0, p->g, p->D_rot, p->seed};
        }
    }
    int64_t ntot = 0, keys = 0;
    double Lmax = 0;
    for (int s = 0; s < nsys; ++s) {
        SysDesc& d = sd[s];
        if (d.n <= 0 || !(d.Lx > 0) || !(d.Ly > 0) || !(d.D >= 0))
            return fail(nullptr, SPP_ERR_ARG, "system %d: n, Lx and Ly must be positive and D_rot >= 0", s);
        if (d.ncx <= 0) d.ncx = (int)std::min(1e9, std::floor(2 * d.Lx / p->r_list));
        if (d.ncy <= 0) d.ncy = (int)std::min(1e9, std::floor(2 * d.Ly / p->r_list));
        // Cell::Cell (shared/cell.h:45-49): cells no narrower than rv.  Three or more per side keep
        // the nine stencil cells distinct.
        if (d.ncx < 3 || d.ncy < 3 || 2 * d.Lx / d.ncx < p->r_list || 2 * d.Ly / d.ncy < p->r_list)
            return fail(nullptr, SPP_ERR_ARG, "system %d: cell grid %dx%d invalid for r_list %g in box %gx%g",
                        s, d.ncx, d.ncy, p->r_list, 2 * d.Lx, 2 * d.Ly);
        if (d.ncx >= 32768 || d.ncy >= 32768) return fail(nullptr, SPP_ERR_ARG, "cell grid too large");
        if (persist && d.n > persist::kMaxN)
            return fail(nullptr, SPP_ERR_UNSUPPORTED, "system %d: %d particles, the persistent kernel takes up to %d",
                        s, d.n, persist::kMaxN);
        ntot += d.n;
        keys += (int64_t)d.ncx * d.ncy;
        Lmax = std::max(Lmax, std::max(d.Lx, d.Ly));
    }
    // (neighbour records are addressed by 32-bit byte offsets: 16 B per particle)
    if (ntot >= (int64_t)1 << 27 || keys >= (int64_t)1 << 30)
        return fail(nullptr, SPP_ERR_ARG, "%lld particles / %lld cells: too many for one engine",
                    (long long)ntot, (long long)keys);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, SPP_ERR_NO_DEVICE, "no HIP device visible");

    spp_engine* e = new spp_engine();
    e->p = *p;
    e->N = (int)ntot;
    (void)hipGetDevice(&e->device);
    if (resolve_scratch(e->device, &e->scratch) != SPP_OK) {
        delete e;
        return fail(nullptr, SPP_ERR_HIP, "device scratch of the allocation-free entry points not resolvable");
    }
    auto grid_of = [&](const SysDesc& d) {
        Grid G{};
        G.Lx = d.Lx;
        G.Ly = d.Ly;
        G.Lx2 = 2 * d.Lx;
        G.Ly2 = 2 * d.Ly;
        G.invLx2 = 1.0 / G.Lx2;
        G.invLy2 = 1.0 / G.Ly2;
        G.ncx = d.ncx;
        G.ncy = d.ncy;
        G.quirk = p->binning == SPP_BIN_REF_QUIRK;
        return G;
    };
    e->G = grid_of(sd[0]);
    e->nkeys = (int)keys;
    e->nscan_blocks = (e->nkeys + kScanTile - 1) / kScanTile;
    e->nblk = nblocks(e->N);
    const bool deep = p->kernel_variant == SPP_KERNEL_SLOT_DEEP;
    e->variant = systems ? kVariantBatch : ((persist || deep) ? SPP_KERNEL_SLOT : p->kernel_variant);
    e->deep = deep;
    e->persist = persist;
    e->nstat = e->nblk;
    // Tile path of the rebuild: single boxes on the slot kernels (captured graphs would freeze the
    // ping-pong of the cell starts).  A tile is about kSortTarget particles of one cell row, at most
    // kSortMaxW cells.
    e->sort_ok = !systems && !persist && p->rebuild_sort == SPP_REBUILD_SORT_AUTO && !p->graph_intervals;
    if (e->sort_ok) {
        const int W = e->G.ncx;
        const double per_cell = (double)e->N / ((double)W * e->G.ncy);
        int segw = (int)std::lround(kSortTarget / std::max(per_cell, 1e-9));
        segw = std::max(1, std::min(std::min(segw, W), kSortMaxW));
        static_assert(kSortCap <= 65536 && kSortMaxW <= 65536, "16-bit tile indices");
        e->sort_segw = segw;
        e->sort_seg = (W + segw - 1) / segw;
        e->sort_ntiles = e->sort_seg * e->G.ncy;
    }
    e->cap = p->ell_capacity > 0 ? p->ell_capacity : 32;
    {
        spp_params q = *p;
        q.Lx = sd[0].Lx;
        q.Ly = sd[0].Ly;
        q.v0 = sd[0].v0;
        q.g = sd[0].g;
        q.D_rot = sd[0].D;
        q.seed = sd[0].seed;
        q.n = sd[0].n;
        e->P = make_phys(q);
        // the pair kernel's one-cutoff path: both laws end at the same exact threshold
        e->one_cut = !systems && p->model == SPP_MODEL_REPULSIVE && p->repulsion != SPP_REP_NONE &&
                     e->P.T_rep == e->P.T_align && e->P.T_align == e->P.T_max;
        if (systems) {
            e->nsys = nsys;
            e->hsys.resize(nsys);
            int koff = 0, poff = 0;
            for (int s = 0; s < nsys; ++s) {
                q.Lx = sd[s].Lx;
                q.Ly = sd[s].Ly;
                q.v0 = sd[s].v0;
                q.g = sd[s].g;
                q.D_rot = sd[s].D;
                q.seed = sd[s].seed;
                Phys Ps = make_phys(q);
                SysInfo& S = e->hsys[s];
                S.G = grid_of(sd[s]);
                S.key_off = koff;
                S.pid_off = poff;
                S.n = sd[s].n;
                S.pad = 0;
                S.v0 = Ps.v0;
                S.c_align = Ps.c_align;
                S.g = Ps.g;
                S.noise_amp = Ps.noise_amp;
                S.seed = Ps.seed;
                koff += sd[s].ncx * sd[s].ncy;
                poff += sd[s].n;
            }
        }
    }
    if (p->model == SPP_MODEL_ABP) {
        e->abp.n_membrane = p->n_membrane;
        e->abp.nb = p->nb_swimmer >= 2 ? p->nb_swimmer : 1;
        e->abp.F0 = p->v0;
        e->abp.mob_par = p->mob_par;
        e->abp.mob_perp = p->mob_perp;
        e->abp.mob_rot = p->mob_rot;
        e->abp.torque0 = p->torque0;
        e->abp.k_bond = p->k_bond;
        e->abp.l_bond = p->l_bond;
        e->abp.kappa = p->kappa_bend;
        chain_mobilities(e->abp.nb, p->bead_sep, p->sigma_bead, e->abp.c_par, e->abp.c_perp, e->abp.c_rot);
        e->abp.sep = p->bead_sep;
        e->abp.sigma_b = p->sigma_bead;
        e->abp.A_b = p->A_bead;
        const double sc = p->sigma_bead / p->rc_bead, sc2 = sc * sc, sc4 = sc2 * sc2, sc8 = sc4 * sc4;
        e->abp.rc13_b = sc8 * sc4 * sc;
        e->abp.T_b = sqrt_threshold(p->rc_bead);
    }
    e->T_list = sqrt_threshold(p->r_list);
    e->r_cut = rcut;
    e->half_skin = 0.5 * (p->r_list - rcut);
    // float32 build positions: allow for their rounding (|x| <= L) when triggering
    double margin = 4.0 * Lmax * 5.960464477539063e-08;
    e->thresh = std::max(0.0, e->half_skin - margin);
    // Displacement-triggered lists keep every listed neighbour of a kept step within r_list plus
    // twice the trigger; a particle that far from the periodic edges has its neighbours on the same
    // side (its differences need no minimum image).  A step whose check fails is discarded whole, so
    // its pairs do not matter.  The fixed-period policy gives no such bound.
    if (p->rebuild == SPP_REBUILD_DISPLACEMENT) e->P.img_m = p->r_list + 2 * e->half_skin + 1e-6;
    // The wide list: five or more cells per side in every system (the 5x5 stencil then holds every
    // pair up to two cell widths apart, and two widths stay below L), exact binning, displacement
    // policy, multi-launch kernels.
    {
        bool ok = p->rebuild == SPP_REBUILD_DISPLACEMENT && !persist && p->binning != SPP_BIN_REF_QUIRK;
        double wmin = 1e300;
        for (int s = 0; s < nsys; ++s) {
            ok = ok && sd[s].ncx >= 5 && sd[s].ncy >= 5;
            wmin = std::min(wmin, std::min(2 * sd[s].Lx / sd[s].ncx, 2 * sd[s].Ly / sd[s].ncy));
        }
        if (ok) {
            e->wide_ok = true;
            e->r_wide = 2 * wmin;
            e->T_wide = sqrt_threshold(e->r_wide);
            e->thresh_wide = std::max(0.0, 0.5 * (e->r_wide - rcut) - margin);
            e->img_m_wide = e->r_wide + 2 * 0.5 * (e->r_wide - rcut) + 1e-6;
        }
    }
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
    e->nsets = e->sort_ok ? 4 : 3;
    for (int k = 0; k < e->nsets; ++k) A(alloc_set(e, e->set[k]));
    A(dalloc(e, &e->xy1, N));
    A(dalloc(e, &e->xy2, N));
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
    A(dalloc(e, &e->perm2, N));
    A(dalloc(e, &e->inv, N));
    A(dalloc(e, &e->cell_count, (size_t)e->nkeys));
    A(dalloc(e, &e->cell_start, (size_t)e->nkeys + 1));
    if (e->sort_ok) {
        A(dalloc(e, &e->cell_start_alt, (size_t)e->nkeys + 1));
        A(dalloc(e, &e->tile_state, (size_t)e->sort_ntiles));
    }
    A(dalloc(e, &e->sort_fail, 2));  // one per report slot
    A(dalloc(e, &e->block_sums, (size_t)e->nscan_blocks));
    for (int sl = 0; sl < 2; ++sl) {
        A(dalloc(e, &e->slotb[sl].blkbad, (size_t)e->nstat));
        A(dalloc(e, &e->slotb[sl].blkmax, (size_t)e->nstat));
        A(dalloc(e, &e->slotb[sl].blklen, (size_t)e->nstat));
        A(dalloc(e, &e->slotb[sl].blknnz, (size_t)e->nstat));
    }
    A(dalloc(e, &e->dpolar, (size_t)e->nblk));
    A(alloc_nbr(e));
    if (systems) {
        A(dalloc(e, &e->dsys, (size_t)nsys));
        A(dalloc(e, &e->sys_of, N));
        A(dalloc(e, &e->dpolar_sys, (size_t)nsys));
    }
    if (persist) {
        A(dalloc(e, &e->pstatus, (size_t)nsys));
        if (!systems) {  // the persistent kernel reads every constant from a system table
            SysInfo S{};
            S.G = e->G;
            S.n = e->N;
            S.v0 = e->P.v0;
            S.c_align = e->P.c_align;
            S.g = e->P.g;
            S.noise_amp = e->P.noise_amp;
            S.seed = e->P.seed;
            A(dalloc(e, &e->dsys1, 1));
            if (rc == SPP_OK && hipMemcpy(e->dsys1, &S, sizeof S, hipMemcpyHostToDevice) != hipSuccess) rc = SPP_ERR_HIP;
        }
    }
    // coherent host memory: the host polls the report words while kernels write them (tag 0: no
    // report yet)
    const size_t wbytes = 2 * kRepWords * sizeof(unsigned long long);
    if (rc == SPP_OK && hipHostMalloc((void**)&e->hwords, wbytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        rc = SPP_ERR_HIP;
    if (rc == SPP_OK) std::memset((void*)e->hwords, 0, wbytes);
    if (rc == SPP_OK && hipHostGetDevicePointer((void**)&e->dwords, e->hwords, 0) != hipSuccess) rc = SPP_ERR_HIP;
    A(dalloc(e, &e->rep_count, 2));
    if (rc == SPP_OK && hipMemset(e->rep_count, 0, 2 * sizeof(long long)) != hipSuccess) rc = SPP_ERR_HIP;
    if (p->device_timing &&
        (hipEventCreate(&e->dev_t0) != hipSuccess || hipEventCreate(&e->dev_t1) != hipSuccess))
        rc = SPP_ERR_HIP;
    if (rc == SPP_OK && systems) {
        // systems occupy fixed slot ranges: cell keys are system-major, so every sort keeps them
        std::vector<int> so(N);
        for (int s = 0; s < nsys; ++s)
            for (int k = 0; k < e->hsys[s].n; ++k) so[e->hsys[s].pid_off + k] = s;
        if (hipMemcpy(e->dsys, e->hsys.data(), nsys * sizeof(SysInfo), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(e->sys_of, so.data(), N * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            rc = SPP_ERR_HIP;
    }
    if (rc == SPP_OK) {
        // cell counts start at zero and every rebuild's scatter returns them to zero; the tile
        // path's ticket counters and look-back states start at zero and every launch clears them
        if (launch(e, "fill", k_fill<int>, nblocks(e->nkeys), e->cell_count, 0, e->nkeys) != SPP_OK ||
            launch(e, "fill", k_fill<int>, 1, e->sort_fail, 0, 2) != SPP_OK)
            rc = SPP_ERR_HIP;
        if (rc == SPP_OK && e->sort_ok &&
            launch(e, "fill", k_fill<unsigned long long>, (e->sort_ntiles + 255) / 256, e->tile_state, 0ull,
                   e->sort_ntiles) != SPP_OK)
            rc = SPP_ERR_HIP;
        for (int sl = 0; sl < 2 && rc == SPP_OK; ++sl) {
            const SlotBufs& b = e->slotb[sl];
            int G8 = e->nstat;
            if (launch(e, "fill", k_fill<int>, nblocks(G8), b.blkbad, INT_MAX, G8) != SPP_OK ||
                launch(e, "fill", k_fill<double>, nblocks(G8), b.blkmax, 0.0, G8) != SPP_OK ||
                launch(e, "fill", k_fill<int>, nblocks(G8), b.blklen, 0, G8) != SPP_OK ||
                launch(e, "fill", k_fill<int>, nblocks(G8), b.blknnz, 0, G8) != SPP_OK)
                rc = SPP_ERR_HIP;
        }
        if (rc == SPP_OK && hipStreamSynchronize(e->stream) != hipSuccess) rc = SPP_ERR_HIP;
    }
    if (rc != SPP_OK) {
        std::string msg = e->err.empty() ? std::string("device allocation failed") : e->err;
        spp_destroy(e);
        return fail(nullptr, rc, "%s", msg.c_str());
    }
    *out = e;
    return SPP_OK;
}

}  // namespace

extern "C" {

int spp_create(const spp_params* p, spp_engine** out) { return create_engine(p, nullptr, 1, out); }

int spp_create_batch(const spp_params* p, const spp_system* systems, int32_t nsys, spp_engine** out)
{
    if (!systems || nsys <= 0) {
        if (out) *out = nullptr;
        return fail(nullptr, SPP_ERR_ARG, "spp_create_batch needs at least one system");
    }
    return create_engine(p, systems, nsys, out);
}

void spp_destroy(spp_engine* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (int k = 0; k < kSets; ++k) free_set(e->set[k]);
    void* bufs[] = {e->xy1, e->xy2, e->th1, e->th2, e->dth, e->dth_alt, e->inj_host_order,
                    e->inj, e->kacc, e->force, e->old, e->aux_alt[0], e->aux_alt[1], e->aux_alt[2],
                    e->nbr32, e->ncount, e->cellxy, e->build, e->key, e->perm, e->perm2, e->inv, e->cell_count,
                    e->cell_start, e->block_sums, e->dpolar, e->dsys, e->sys_of, e->dpolar_sys, e->dsys1, e->dsamples,
                    e->pstatus, e->cell_start_alt, e->tile_state, e->sort_fail, e->rep_count};
    for (void* b : bufs) dfree(b);
    for (auto& b : e->slotb) {
        dfree(b.blkbad);
        dfree(b.blkmax);
        dfree(b.blklen);
        dfree(b.blknnz);
    }
    for (auto& kv : e->slots) free_set(kv.second.s);
    drop_graphs(e);
    dfree(e->dclock);
    lyap_free(e->ly);
    for (auto& r : e->prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
    for (auto ev : {e->dev_t0, e->dev_t1})
        if (ev) (void)hipEventDestroy(ev);
    if (e->hwords) (void)hipHostFree(e->hwords);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* spp_last_error(const spp_engine* e)
{
    return e ? e->err.c_str() : g_create_error.c_str();
}

// Device time of a public call that steps or measures: HIP events on the engine's stream around
// the outermost such call, accumulated into spp_info.device_ms (the calls end synchronised).  Only
// engines created with spp_params.device_timing hold the events: each record idles the GPU for a
// few microseconds, which a short stepping call would otherwise pay twice.
struct DeviceClock {
    spp_engine* e;
    bool on;
    explicit DeviceClock(spp_engine* eng) : e(eng), on(eng && eng->dev_t0)
    {
        if (on && e->dev_depth++ == 0) (void)hipEventRecord(e->dev_t0, e->stream);
    }
    ~DeviceClock()
    {
        if (!on || --e->dev_depth != 0) return;
        float ms = 0;
        if (hipEventRecord(e->dev_t1, e->stream) == hipSuccess && hipEventSynchronize(e->dev_t1) == hipSuccess &&
            hipEventElapsedTime(&ms, e->dev_t0, e->dev_t1) == hipSuccess)
            e->device_ms += ms;
    }
};

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
    info->tile_rebuilds = e->tile_rebuilds;
    info->tile_fallbacks = e->tile_fallbacks;
    info->device_ms = e->device_ms;
    info->wide_rebuilds = e->wide_rebuilds;
    return SPP_OK;
}

int spp_set_state(spp_engine* e, const double* xw, const double* yw, const double* xu,
                  const double* yu, const double* theta, int32_t n)
{
    if (!e || !xw || !yw || !theta) return e ? fail(e, SPP_ERR_ARG, "null array") : SPP_ERR_ARG;
    DeviceClock dc(e);
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n = %d, engine has %d particles", n, e->N);
    // Wrapped positions must lie in their box [-L, L] (the cell binning clamps, so a particle
    // outside would be binned into a wrong cell and miss neighbours); no value may be NaN or inf.
    const int ns = e->dsys ? e->nsys : 1;
    for (int s = 0; s < ns; ++s) {
        const int lo = e->dsys ? e->hsys[s].pid_off : 0, hi = e->dsys ? lo + e->hsys[s].n : n;
        const double Lx = e->dsys ? e->hsys[s].G.Lx : e->G.Lx, Ly = e->dsys ? e->hsys[s].G.Ly : e->G.Ly;
        for (int i = lo; i < hi; ++i) {
            if (!(std::fabs(xw[i]) <= Lx) || !(std::fabs(yw[i]) <= Ly))
                return fail(e, SPP_ERR_ARG, "particle %d at (%g, %g) is outside its box [-%g, %g] x [-%g, %g]", i,
                            xw[i], yw[i], Lx, Lx, Ly, Ly);
            if (!std::isfinite(theta[i]) || (xu && !std::isfinite(xu[i])) || (yu && !std::isfinite(yu[i])))
                return fail(e, SPP_ERR_ARG, "particle %d has a non-finite heading or unwrapped position", i);
        }
    }
    std::vector<double2> pos(n);
    std::vector<double2> U(n);
    std::vector<int> id(n);
    for (int i = 0; i < n; ++i) {
        pos[i] = make_double2(xw[i], yw[i]);
        U[i] = make_double2(xu ? xu[i] : xw[i], yu ? yu[i] : yw[i]);
        id[i] = i;
    }
    SetPtrs& s = e->set[e->cur];
    HIPC(e, hipMemcpyAsync(s.xy, pos.data(), n * sizeof(double2), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(s.th, theta, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(s.U, U.data(), n * sizeof(double2), hipMemcpyHostToDevice, e->stream));
    HIPC(e, hipMemcpyAsync(s.id, id.data(), n * sizeof(int), hipMemcpyHostToDevice, e->stream));
    e->have_state = true;
    e->unfused_stage = 0;
    e->inj_pending = false;
    e->wide_next = false;  // a new state starts on the ordinary list
    return rebuild_sync(e, false);
}

int spp_get_state(spp_engine* e, double* xw, double* yw, double* xu, double* yu, double* theta,
                  int32_t n)
{
    TRY(check_ready(e));
    if (n != e->N) return fail(e, SPP_ERR_ARG, "n = %d, engine has %d particles", n, e->N);
    std::vector<double2> pos(n);
    std::vector<double> th(n);
    std::vector<double2> U(n);
    std::vector<int> id(n);
    const SetPtrs& s = e->set[e->cur];
    HIPC(e, hipMemcpyAsync(pos.data(), s.xy, n * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(th.data(), s.th, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(U.data(), s.U, n * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), s.id, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    TRY(check_ids(e, id));
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
    std::vector<int> n32((size_t)C * N), cnt(N), id(N);
    HIPC(e, hipMemcpyAsync(n32.data(), e->nbr32, n32.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(cnt.data(), e->ncount, N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipMemcpyAsync(id.data(), e->set[e->cur].id, N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    TRY(check_ids(e, id));
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
        for (int c = 0; c < cnt[k]; ++c) {
            int slot = n32[(size_t)c * N + k];
            idx[o + c] = id[slot];
        }
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
    TRY(check_ids(e, id));
    for (int k = 0; k < n; ++k) {
        if (cell_x) cell_x[id[k]] = cxy[k] & 0xffff;
        if (cell_y) cell_y[id[k]] = cxy[k] >> 16;
    }
    return SPP_OK;
}

int spp_interact(spp_engine* e)
{
    TRY(check_ready(e));
    if (!e->list_valid) TRY(rebuild_sync(e, true));
    StageArgs a{};
    a.N = e->N;
    a.P = e->P;
    a.sys_tab = e->dsys;
    a.sys_of = e->sys_of;
    a.abp = e->abp;
    a.inv = e->inv;
    a.gxy = e->set[e->cur].xy;
    a.gth = e->set[e->cur].th;
    a.id = e->set[e->cur].id;
    a.nbr32 = e->nbr32;
    a.ncount = e->ncount;
    InteractFn fn = pick_interact(e->p.model, e->p.repulsion, e->variant);
    return launch(e, "interact", fn, e->nblk, a, e->force);
}

StageArgs pair_args(spp_engine* e)
{
    StageArgs a{};
    a.N = e->N;
    a.P = e->P;
    a.sys_tab = e->dsys;
    a.sys_of = e->sys_of;
    a.abp = e->abp;
    a.inv = e->inv;
    a.gxy = e->set[e->cur].xy;
    a.gth = e->set[e->cur].th;
    a.id = e->set[e->cur].id;
    return a;
}

int spp_reset_forces(spp_engine* e)
{
    TRY(check_ready(e));
    return launch(e, "reset_forces", k_reset_forces, e->nblk, pair_args(e), e->p.model == SPP_MODEL_ABP ? 1 : 0,
                  e->force);
}

int spp_pair_interact(spp_engine* e, int32_t i, int32_t j)
{
    TRY(check_ready(e));
    if (i < 0 || j < 0 || i >= e->N || j >= e->N || i == j)
        return fail(e, SPP_ERR_ARG, "particles %d and %d: two distinct indices in [0, %d) needed", i, j, e->N);
    if (e->dsys) {
        int si = 0, sj = 0;
        for (int s = 0; s < e->nsys; ++s) {
            if (i >= e->hsys[s].pid_off && i < e->hsys[s].pid_off + e->hsys[s].n) si = s;
            if (j >= e->hsys[s].pid_off && j < e->hsys[s].pid_off + e->hsys[s].n) sj = s;
        }
        if (si != sj) return fail(e, SPP_ERR_ARG, "particles %d and %d belong to different systems", i, j);
    }
    PairFn fn = pick_pair(e->p.model, e->p.repulsion);
    if (e->dry) return SPP_OK;
    hipLaunchKernelGGL(fn, dim3(1), dim3(64), 0, e->stream, pair_args(e), i, j, e->force);
    HIPC(e, hipGetLastError());
    return SPP_OK;
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
    TRY(check_ids(e, id));
    for (int k = 0; k < n; ++k) {
        int i = id[k];
        if (fx) fx[i] = f[k].x;
        if (fy) fy[i] = f[k].y;
        if (tau) tau[i] = f[k].z;
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
    m.sys_tab = e->dsys;
    m.sys_of = e->sys_of;
    m.abp = e->abp;
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
    if (e->p.model == SPP_MODEL_ABP)
        TRY(launch(e, "move", k_move_abp, e->nblk, m));
    else
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
    DeviceClock dc(e);
    if (steps < 0) return fail(e, SPP_ERR_ARG, "steps < 0");
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "a split step is in progress");
    if (e->p.noise_mode == SPP_NOISE_INJECTED && (steps > 1 || (steps == 1 && !e->inj_pending)))
        return fail(e, SPP_ERR_STATE, "injected noise covers one step: inject before each step");
    if (steps == 0) return SPP_OK;
    if (e->persist) return run_persist(e, steps);
    return run_intervals(e, steps);
}

int spp_one_step(spp_engine* e) { return spp_multi_step(e, 1); }

int spp_multi_step_sampled(spp_engine* e, int32_t steps, int32_t every, double* polar, int64_t len)
{
    TRY(check_ready(e));
    DeviceClock dc(e);
    if (every <= 0 || steps < 0 || steps % every) return fail(e, SPP_ERR_ARG, "steps must be a multiple of every > 0");
    const int ns = e->dsys ? e->nsys : 1;
    const int64_t nsamp = steps / every;
    if (!polar || len != nsamp * ns)
        return fail(e, SPP_ERR_ARG, "polar needs %lld entries", (long long)(nsamp * ns));
    if (!e->dsys && !e->dsys1) {
        SysInfo S{};
        S.G = e->G;
        S.n = e->N;
        TRY(dalloc(e, &e->dsys1, 1));
        HIPC(e, hipMemcpy(e->dsys1, &S, sizeof S, hipMemcpyHostToDevice));
    }
    if (e->samples_alloc < nsamp * ns) {
        dfree(e->dsamples);
        e->dsamples = nullptr;
        e->samples_alloc = 0;
        TRY(dalloc(e, &e->dsamples, (size_t)(nsamp * ns)));
        e->samples_alloc = nsamp * ns;
    }
    e->sample_every = every;
    e->sample_base = e->step;
    e->sample_cap = nsamp;
    int rc = spp_multi_step(e, steps);
    e->sample_every = 0;
    if (rc != SPP_OK) return rc;
    std::vector<double2> h(nsamp * ns);
    HIPC(e, hipMemcpyAsync(h.data(), e->dsamples, h.size() * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int64_t k = 0; k < nsamp; ++k)
        for (int s = 0; s < ns; ++s) {
            const double n = e->dsys ? e->hsys[s].n : e->N;
            const double ax = h[k * ns + s].x / n, ay = h[k * ns + s].y / n;
            polar[k * ns + s] = std::sqrt(ax * ax + ay * ay);
        }
    return SPP_OK;
}

int spp_polarization(spp_engine* e, double* p, double* sum_x, double* sum_y)
{
    TRY(check_ready(e));
    TRY(launch(e, "polar", k_polar, e->nblk, e->N, (const double*)e->set[e->cur].th, e->dpolar));
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

int spp_batch_polarization(spp_engine* e, double* p, int32_t nsys)
{
    TRY(check_ready(e));
    DeviceClock dc(e);
    if (!p) return fail(e, SPP_ERR_ARG, "null array");
    if (!e->dsys) {
        if (nsys != 1) return fail(e, SPP_ERR_ARG, "single-system engine: nsys must be 1");
        return spp_polarization(e, p, nullptr, nullptr);
    }
    if (nsys != e->nsys) return fail(e, SPP_ERR_ARG, "nsys = %d, engine has %d systems", nsys, e->nsys);
    TRY(launch(e, "polar_sys", k_polar_sys, nsys, (const SysInfo*)e->dsys, (const double*)e->set[e->cur].th,
               e->dpolar_sys));
    std::vector<double2> v(nsys);
    HIPC(e, hipMemcpyAsync(v.data(), e->dpolar_sys, nsys * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    for (int s = 0; s < nsys; ++s) {
        double ax = v[s].x / e->hsys[s].n, ay = v[s].y / e->hsys[s].n;
        p[s] = std::sqrt(ax * ax + ay * ay);
    }
    return SPP_OK;
}

int spp_pair_distribution(spp_engine* e, double r_max, int32_t nbins, int64_t* counts, double* g)
{
    TRY(check_ready(e));
    if (e->dsys) return fail(e, SPP_ERR_UNSUPPORTED, "pair distribution of a batched engine");
    if (nbins <= 0 || nbins > 8192 || !(r_max > 0)) return fail(e, SPP_ERR_ARG, "need 0 < nbins <= 8192, r_max > 0");
    if (r_max > e->p.r_list) return fail(e, SPP_ERR_ARG, "r_max %g beyond the list radius %g", r_max, e->p.r_list);
    if (!e->list_valid) TRY(rebuild_sync(e, false));
    unsigned long long* dh = nullptr;
    TRY(dalloc(e, &dh, (size_t)nbins));
    int rc = launch(e, "fill", k_fill<unsigned long long>, nblocks(nbins), dh, 0ull, nbins);
    const double dr = r_max / nbins;
    if (rc == SPP_OK) {
        hipLaunchKernelGGL(k_pair_hist, dim3(e->nblk), dim3(kBlock), nbins * sizeof(unsigned int), e->stream, e->N,
                           (const double2*)e->set[e->cur].xy, (const int*)e->nbr32, (const int*)e->ncount, e->G,
                           1.0 / dr, nbins, dh);
        if (hipGetLastError() != hipSuccess) rc = fail(e, SPP_ERR_HIP, "pair histogram launch failed");
    }
    std::vector<unsigned long long> h(nbins);
    if (rc == SPP_OK && (hipMemcpyAsync(h.data(), dh, nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                        e->stream) != hipSuccess ||
                         hipStreamSynchronize(e->stream) != hipSuccess))
        rc = fail(e, SPP_ERR_HIP, "pair histogram read-back failed");
    dfree(dh);
    if (rc != SPP_OK) return rc;
    // g(r_k) = C_k / (N rho pi (r_{k+1}^2 - r_k^2)), C_k counting each pair from both ends
    const double rho = e->N / (4.0 * e->p.Lx * e->p.Ly);
    for (int k = 0; k < nbins; ++k) {
        if (counts) counts[k] = (int64_t)h[k];
        const double a = spp::kPi * ((k + 1) * dr * (k + 1) * dr - k * dr * k * dr);
        if (g) g[k] = (double)h[k] / (e->N * rho * a);
    }
    return SPP_OK;
}

int spp_pair_distribution_noalloc(spp_engine* e, double r_max, int32_t nbins, int64_t* counts, double* g)
{
    TRY(check_ready(e));
    if (e->dsys) return fail(e, SPP_ERR_UNSUPPORTED, "pair distribution of a batched engine");
    if (nbins <= 0 || nbins > kScratchBins || !(r_max > 0))
        return fail(e, SPP_ERR_ARG, "need 0 < nbins <= %d (the static scratch), r_max > 0", kScratchBins);
    if (r_max > e->p.r_list) return fail(e, SPP_ERR_ARG, "r_max %g beyond the list radius %g", r_max, e->p.r_list);
    // a rebuild may have to grow the list (an allocation): the caller builds it beforehand
    if (!e->list_valid) return fail(e, SPP_ERR_STATE, "no current neighbour list: call spp_build_neighbors first");
    if (!counts && !g) return SPP_OK;
    if (e->profiling) return fail(e, SPP_ERR_STATE, "profiling records events (allocations): disable it first");
    // the scratch keeps the previous call's contents: clear the bins this call uses (a fill
    // kernel, so the call may also be captured into a graph)
    unsigned long long* dh = e->scratch;
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(nblocks(nbins)), dim3(kBlock), 0, e->stream, dh, 0ull, nbins);
    const double dr = r_max / nbins;
    hipLaunchKernelGGL(k_pair_hist, dim3(e->nblk), dim3(kBlock), nbins * sizeof(unsigned int), e->stream, e->N,
                       (const double2*)e->set[e->cur].xy, (const int*)e->nbr32, (const int*)e->ncount, e->G, 1.0 / dr,
                       nbins, dh);
    HIPC(e, hipGetLastError());
    // the counts land in the caller's array (or, without one, in g, converted in place): 8 B per bin
    // either way, so no host buffer is needed
    void* dst = counts ? static_cast<void*>(counts) : static_cast<void*>(g);
    HIPC(e, hipMemcpyAsync(dst, dh, nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost, e->stream));
    HIPC(e, hipStreamSynchronize(e->stream));
    const double rho = e->N / (4.0 * e->p.Lx * e->p.Ly);
    for (int k = 0; k < nbins; ++k) {
        unsigned long long c;
        std::memcpy(&c, static_cast<const char*>(dst) + 8 * (size_t)k, 8);
        const double a = spp::kPi * ((k + 1) * dr * (k + 1) * dr - k * dr * k * dr);
        if (g) g[k] = (double)c / (e->N * rho * a);
    }
    return SPP_OK;
}

int spp_lyapunov_init(spp_engine* e, int32_t replicas, int32_t vectors, double epsilon, int32_t rescale,
                      uint64_t seed)
{
    TRY(check_ready(e));
    DeviceClock dc(e);
    if (!e->dsys) return fail(e, SPP_ERR_STATE, "Lyapunov clones need a batched engine (spp_create_batch)");
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "a split step is in progress");
    if (replicas <= 0 || vectors < 1 || vectors > lyap::kMaxK)
        return fail(e, SPP_ERR_ARG, "replicas >= 1 and 1 <= vectors <= %d required", lyap::kMaxK);
    if (!(epsilon > 0)) return fail(e, SPP_ERR_ARG, "epsilon must be positive");
    if ((int64_t)replicas * (vectors + 1) != e->nsys)
        return fail(e, SPP_ERR_ARG, "%d replicas x (1 + %d vectors) != %d systems", replicas, vectors, e->nsys);
    const int n = e->hsys[0].n;
    for (int r = 0; r < replicas; ++r) {
        const SysInfo& B = e->hsys[r * (vectors + 1)];
        for (int c = 0; c <= vectors; ++c) {
            const SysInfo& S = e->hsys[r * (vectors + 1) + c];
            bool same = S.n == n && S.G.Lx == B.G.Lx && S.G.Ly == B.G.Ly && S.G.ncx == B.G.ncx &&
                        S.G.ncy == B.G.ncy && S.v0 == B.v0 && S.g == B.g && S.noise_amp == B.noise_amp &&
                        S.seed == B.seed;
            if (!same)
                return fail(e, SPP_ERR_ARG, "system %d differs from its replica's base system %d (all systems "
                            "need n = %d; a replica's systems need equal boxes, parameters and seeds)",
                            r * (vectors + 1) + c, r * (vectors + 1), n);
        }
    }
    lyap_free(e->ly);
    LyapState& L = e->ly;
    L.R = replicas;
    L.K = vectors;
    L.n = n;
    L.eps = epsilon;
    L.rescale = rescale ? 1 : 0;
    // enough Gram workgroups to fill the chip, each with at least one pass of 256 components
    L.nchunk = std::max(1, std::min((3 * n + kBlock - 1) / kBlock, (2048 + replicas - 1) / replicas));
    const int NP = vectors * (vectors + 1) / 2;
    int rc = SPP_OK;
    auto A = [&](int r) {
        if (rc == SPP_OK && r != SPP_OK) rc = r;
    };
    A(dalloc(e, &L.inv, (size_t)e->N));
    A(dalloc(e, &L.status, (size_t)replicas));
    A(dalloc(e, &L.dev, (size_t)replicas * vectors * 3 * n));
    A(dalloc(e, &L.part, (size_t)replicas * L.nchunk * NP));
    A(dalloc(e, &L.coef, (size_t)replicas * (vectors * vectors + vectors)));
    A(dalloc(e, &L.rdiag, (size_t)replicas * vectors));
    A(dalloc(e, &L.logsum, (size_t)replicas * vectors));
    A(dalloc(e, &L.norm0, (size_t)replicas * vectors));
    if (rc != SPP_OK) {
        lyap_free(e->ly);
        return rc;
    }
    L.on = true;
    L.t_start = L.t_last = e->t;
    TRY(launch(e, "fill", k_fill<int>, nblocks(replicas), L.status, 0, replicas));
    TRY(launch(e, "fill", k_fill<double>, nblocks(replicas * vectors), L.logsum, 0.0, replicas * vectors));
    TRY(launch(e, "fill", k_fill<double>, nblocks(replicas * vectors), L.norm0, epsilon, replicas * vectors));
    lyap::Args a = lyap_args(e);
    TRY(launch(e, "lyap_inverse", lyap::k_inverse, nblocks(e->N), e->N, (const int*)e->set[e->cur].id, L.inv));
    TRY(launch(e, "lyap_random", lyap::k_random_dev, nblocks(replicas * n), a, seed));
    TRY(lyap_qr(e, true));
    e->list_valid = false;  // clones moved: the next step rebuilds
    e->persist_valid = false;
    return lyap_check_status(e);
}

int spp_lyapunov_renormalize(spp_engine* e, double* log_growth)
{
    TRY(check_ready(e));
    DeviceClock dc(e);
    LyapState& L = e->ly;
    if (!L.on) return fail(e, SPP_ERR_STATE, "call spp_lyapunov_init first");
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "a split step is in progress");
    lyap::Args a = lyap_args(e);
    TRY(launch(e, "lyap_inverse", lyap::k_inverse, nblocks(e->N), e->N, (const int*)e->set[e->cur].id, L.inv));
    TRY(launch(e, "lyap_deviation", lyap::k_deviation, nblocks(L.R * L.n), a));
    TRY(lyap_qr(e, false));
    e->list_valid = false;
    e->persist_valid = false;
    L.renorms++;
    L.t_last = e->t;
    if (log_growth) {
        TRY(lyap_check_status(e));
        HIPC(e, hipMemcpy(log_growth, L.logsum, (size_t)L.R * L.K * sizeof(double), hipMemcpyDeviceToHost));
    }
    return SPP_OK;
}

int spp_lyapunov_run(spp_engine* e, int32_t steps, int32_t interval)
{
    TRY(check_ready(e));
    DeviceClock dc(e);
    if (!e->ly.on) return fail(e, SPP_ERR_STATE, "call spp_lyapunov_init first");
    if (interval <= 0 || steps < 0 || steps % interval)
        return fail(e, SPP_ERR_ARG, "steps (%d) must be a non-negative multiple of interval (%d)", steps, interval);
    for (int s = 0; s < steps; s += interval) {
        TRY(spp_multi_step(e, interval));
        TRY(spp_lyapunov_renormalize(e, nullptr));
    }
    return SPP_OK;
}

int spp_lyapunov_get(spp_engine* e, double* sum_log, double* time)
{
    TRY(check_ready(e));
    const LyapState& L = e->ly;
    if (!L.on) return fail(e, SPP_ERR_STATE, "call spp_lyapunov_init first");
    TRY(lyap_check_status(e));
    if (sum_log)
        HIPC(e, hipMemcpy(sum_log, L.logsum, (size_t)L.R * L.K * sizeof(double), hipMemcpyDeviceToHost));
    if (time) *time = L.t_last - L.t_start;
    return SPP_OK;
}

int spp_lyapunov_vectors(spp_engine* e, int32_t replica, double* out, int64_t len)
{
    TRY(check_ready(e));
    const LyapState& L = e->ly;
    if (!L.on) return fail(e, SPP_ERR_STATE, "call spp_lyapunov_init first");
    int64_t need = (int64_t)L.K * 3 * L.n;
    if (replica < 0 || replica >= L.R || !out || len != need)
        return fail(e, SPP_ERR_ARG, "replica in [0, %d) and an array of %lld doubles required", L.R, (long long)need);
    HIPC(e, hipStreamSynchronize(e->stream));
    HIPC(e, hipMemcpy(out, L.dev + (size_t)replica * need, need * sizeof(double), hipMemcpyDeviceToHost));
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
    if (!s.s.xy) TRY(alloc_set(e, s.s));
    TRY(copy_set(e, e->set[e->cur], s.s));
    s.step = e->step;
    s.t = e->t;
    for (int k = 0; k < 3; ++k) s.T_hist[k] = e->T_hist[k];
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
    for (int k = 0; k < 3; ++k) e->T_hist[k] = it->second.T_hist[k];
    derive_schedule(e);
    e->force_m = 0;
    e->unfused_stage = 0;
    // Box::Load ends in a cell update (parallel/box.h:241)
    return rebuild_sync(e, false);
}

int spp_get_clock(spp_engine* e, spp_clock* clock)
{
    if (!e || !clock) return e ? fail(e, SPP_ERR_ARG, "null clock") : SPP_ERR_ARG;
    clock->step = e->step;
    clock->t = e->t;
    for (int k = 0; k < 3; ++k) clock->rebuild_estimate[k] = e->T_hist[k];
    return SPP_OK;
}

int spp_set_clock(spp_engine* e, const spp_clock* clock)
{
    TRY(check_ready(e));
    if (!clock) return fail(e, SPP_ERR_ARG, "null clock");
    if (clock->step < 0) return fail(e, SPP_ERR_ARG, "step %lld < 0", (long long)clock->step);
    for (int k = 0; k < 3; ++k)
        if (!(clock->rebuild_estimate[k] >= 0.0) || !std::isfinite(clock->rebuild_estimate[k]))
            return fail(e, SPP_ERR_ARG, "rebuild estimate %d is %g", k, clock->rebuild_estimate[k]);
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "a split step is in progress");
    e->step = clock->step;
    e->t = clock->t;
    for (int k = 0; k < 3; ++k) e->T_hist[k] = clock->rebuild_estimate[k];
    derive_schedule(e);
    e->force_m = 0;
    return SPP_OK;
}

int spp_synchronize(spp_engine* e)
{
    if (!e) return SPP_ERR_ARG;
    HIPC(e, hipStreamSynchronize(e->stream));
    return SPP_OK;
}

int spp_graph_stats(spp_engine* e, int64_t* captures, int64_t* replays, int32_t* cached)
{
    if (!e) return SPP_ERR_ARG;
    if (captures) *captures = e->graph_captures;
    if (replays) *replays = e->graph_replays;
    if (cached) *cached = (int32_t)e->graphs.size();
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

int spp_time_stage(spp_engine* e, int32_t stage, int32_t reps, double* avg_us)
{
    TRY(check_ready(e));
    int integ = e->p.integrator;
    int nst = integ == SPP_EULER ? 1 : integ;
    if (reps <= 0 || !avg_us || (integ == SPP_EULER ? stage != 0 : (stage < 1 || stage > nst)))
        return fail(e, SPP_ERR_ARG, "stage %d / reps %d", stage, reps);
    if (e->unfused_stage != 0) return fail(e, SPP_ERR_STATE, "a split step is in progress");
    const SetPtrs& S = e->set[e->cur];
    StageArgs a{};
    a.N = e->N;
    a.P = e->P;
    a.sys_tab = e->dsys;
    a.sys_of = e->sys_of;
    a.abp = e->abp;
    a.inv = e->inv;
    a.nbr32 = e->nbr32;
    a.ncount = e->ncount;
    a.id = S.id;
    a.build = e->build;
    a.step = e->step;
    a.thresh2 = cur_thresh(e) * cur_thresh(e);
    if (e->wide) a.P.img_m = e->img_m_wide;
    a.blkbad = e->slotb[0].blkbad;
    a.blockmax = e->slotb[0].blkmax;
    a.gxy = stage <= 1 ? S.xy : e->xy1;
    a.gth = stage <= 1 ? S.th : e->th1;
    a.th0 = S.th;
    a.U_in = S.U;
    // every output goes to scratch: the engine state is left as it was
    a.oxy = e->xy2;
    a.oth = e->th2;
    a.U_out = reinterpret_cast<double2*>(e->aux_alt[0]);
    a.dth = stage <= 1 ? e->dth_alt : e->dth;
    a.kacc = e->aux_alt[1];
    a.stream = e->N <= kStreamMaxN;
    StageFn fn = pick_stage(e->p.model, e->p.repulsion, integ, stage, e->variant, e->deep || e->N < kSmallN,
                            e->one_cut);
    a.grid = e->nblk;
    hipEvent_t t0, t1;
    HIPC(e, hipEventCreate(&t0));
    HIPC(e, hipEventCreate(&t1));
    bool prof = e->profiling;
    e->profiling = false;
    auto go = [&]() -> int {
        return launch(e, "time", fn, e->nblk, a);
    };
    int rc = go();  // warm
    if (rc == SPP_OK) rc = hipEventRecord(t0, e->stream) == hipSuccess ? SPP_OK : SPP_ERR_HIP;
    for (int r = 0; r < reps && rc == SPP_OK; ++r) rc = go();
    if (rc == SPP_OK) rc = hipEventRecord(t1, e->stream) == hipSuccess ? SPP_OK : SPP_ERR_HIP;
    // the timed launches may have marked slot 0's per-block bad steps: clear them for its next interval
    if (rc == SPP_OK) rc = launch(e, "fill", k_fill<int>, nblocks(e->nstat), e->slotb[0].blkbad, INT_MAX, e->nstat);
    float ms = 0;
    if (rc == SPP_OK) rc = hipEventSynchronize(t1) == hipSuccess ? SPP_OK : SPP_ERR_HIP;
    if (rc == SPP_OK) rc = hipEventElapsedTime(&ms, t0, t1) == hipSuccess ? SPP_OK : SPP_ERR_HIP;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->profiling = prof;
    if (rc != SPP_OK) return rc == SPP_ERR_HIP ? fail(e, rc, "event timing failed") : rc;
    *avg_us = 1e3 * ms / reps;
    return SPP_OK;
}

int spp_random_state(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y,
                     double* theta)
{
    return spp_random_formation(seed, n, Lx, Ly, 0.0, x, y, theta);
}

int spp_random_formation(uint64_t seed, int32_t n, double Lx, double Ly, double sigma, double* x, double* y,
                         double* theta)
{
    if (n < 0 || !x || !y || !theta || !(sigma >= 0) || !(sigma < Lx) || !(sigma < Ly)) return SPP_ERR_ARG;
    const double ax = Lx - sigma, ay = Ly - sigma;  // C2DVector::Rand(Lx - sigma, Ly - sigma)
    for (int32_t i = 0; i < n; ++i) {
        uint32_t c[4] = {(uint32_t)i, 0u, 0u, 0x53505069u};
        spp::philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        double u0 = (double)c[0] * (1.0 / 4294967296.0);
        double u1 = (double)c[1] * (1.0 / 4294967296.0);
        double u2 = (double)c[2] * (1.0 / 4294967296.0);
        x[i] = -ax + 2 * ax * u0;
        y[i] = -ay + 2 * ay * u1;
        theta[i] = spp::kTwoPi * u2;
    }
    return SPP_OK;
}

int spp_polar_formation(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y, double* theta)
{
    const int rc = spp_random_formation(seed, n, Lx, Ly, 0.0, x, y, theta);
    if (rc != SPP_OK) return rc;
    for (int32_t i = 0; i < n; ++i) theta[i] = 0.0;  // every heading along +x (v = (1, 0))
    return SPP_OK;
}

int spp_triangle_lattice(int32_t n, double Lx, double Ly, double sigma, double* x, double* y)
{
    if (n <= 0 || !x || !y || !(Lx > 0) || !(Ly > 0)) return SPP_ERR_ARG;
    // Triangle_Lattice_Formation, shared/set-up.h:116-145: rows of Nx sites along basis_1, rows
    // stacked along basis_2, every second pair of rows shifted back by basis_1
    const double Lx2 = 2 * Lx, Ly2 = 2 * Ly;
    const int nx = (int)std::sqrt((Lx * n * std::sqrt(3.0)) / (2 * Ly)) + 1;
    const double s = (Lx2 - sigma) / nx;
    const double b1x = 1 * s, b1y = 0 * s, b2x = 0.5 * s, b2y = std::sqrt(3.0) / 2 * s;
    for (int32_t i = 0; i < n; ++i) {
        const int a1 = i % nx, a2 = i / nx, a3 = i / (2 * nx);
        double rx = b1x * a1 + b2x * a2 - b1x * a3;
        double ry = b1y * a1 + b2y * a2 - b1y * a3;
        rx -= (Lx - sigma);
        ry -= (Ly - sigma);
        x[i] = rx - Lx2 * (double)(int)std::floor(rx / Lx2 + 0.5);
        y[i] = ry - Ly2 * (double)(int)std::floor(ry / Ly2 + 0.5);
    }
    return SPP_OK;
}

int spp_ring_membrane(int32_t nm, double bead_diameter, double radius, double* x, double* y)
{
    if (nm < 3 || !x || !y) return SPP_ERR_ARG;
    // Ring_Membrane, shared/set-up.h:292-316: radius 0.5 d / sin(pi / Nm) unless given; the angle
    // accumulates 2 pi / Nm per bead
    const double R = radius > 0 ? radius : 0.5 * bead_diameter / std::sin(spp::kPi / nm);
    if (!(R > 0)) return SPP_ERR_ARG;
    double theta = 0;
    for (int32_t i = 0; i < nm; ++i) {
        x[i] = R * std::cos(theta);
        y[i] = R * std::sin(theta);
        theta += 2 * spp::kPi / nm;
    }
    return SPP_OK;
}

int spp_disk_state(uint64_t seed, int32_t n, double radius, double* x, double* y, double* theta)
{
    if (n < 0 || !x || !y || !theta || !(radius > 0)) return SPP_ERR_ARG;
    for (int32_t i = 0; i < n; ++i) {
        uint32_t c[4] = {(uint32_t)i, 1u, 0u, 0x53505069u};
        spp::philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u0 = (double)c[0] * (1.0 / 4294967296.0), u1 = (double)c[1] * (1.0 / 4294967296.0);
        const double u2 = (double)c[2] * (1.0 / 4294967296.0);
        const double r = radius * std::sqrt(u0), phi = spp::kTwoPi * u1;
        x[i] = r * std::cos(phi);
        y[i] = r * std::sin(phi);
        theta[i] = spp::kTwoPi * u2;
    }
    return SPP_OK;
}

double spp_noise_normal(uint64_t seed, int32_t id, int64_t step)
{
    return spp::noise_normal(seed, (uint32_t)id, step);
}

}  // extern "C"
