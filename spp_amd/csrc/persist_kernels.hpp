// persist_kernels.hpp -- one workgroup runs one small system for many steps (gfx950).
//
// Included by engine.hip after the slot kernels.  For systems of up to kMaxN particles a stage
// launch costs a whole dependent chain of memory round trips for a handful of wavefronts (about
// 10 us at 1024 particles, whatever the batch size), so a step of two launches is bound by launch
// latency, not by work.  Here each workgroup of kBlock1024 threads owns one system of a batched
// engine (or the only system of a single one) and runs it to a target step inside one launch:
// the rebuild (binning, scan, scatter, id order in cells, permutation, Verlet list), the fused
// stages of every step, the displacement check and the polarisation samples, separated by
// workgroup barriers.  Everything a system touches lies in its own slot and cell-key ranges, so
// workgroups never wait on each other.
//
// The arithmetic is the slot kernels' own (stage_particle, nlist_particle) on the same layout, so
// forces and moves are those of the multi-launch engine.  Only the rebuild schedule differs: a
// system rebuilds when a stage output has moved more than the half skin from the list positions
// (the stage whose input would be unsafe is redone from the step's start after the rebuild), and
// every `period` steps in the fixed-period mode -- the reference's Multi_Step(cell_update_period)
// cadence (parallel/box.h:434-438).  The order of additions in a particle's pair sums follows
// the cell order at its list build, so trajectories agree with the multi-launch engine's to
// rounding, not bitwise.
#pragma once

namespace persist {

constexpr int kBlock1024 = 1024;
constexpr int kMaxN = 4096;  // particles per system (four per thread)

enum : int { kRunning = 0, kNeedCap = 1, kSkinJump = 2 };

// Per-system progress, kept on the device between launches.
struct Status {
    long long step;  // absolute step the system has reached
    int since;       // steps since its list was built (-1: no valid list)
    int code;        // kRunning, kNeedCap (list longer than the capacity), kSkinJump
    int need_cap;    // longest list of the failed build
    int rebuilds;
    int max_len;     // longest list of the last build
    int pad;
    long long nnz;   // entries of the last build
};

struct Args {
    StageArgs st;         // constants, list and noise buffers; per-stage pointers set in-kernel
    NlistFlatArgs nl;     // list arrays; positions set in-kernel
    const SysInfo* sys_tab;
    SetPtrs cur, spare;   // the engine's state set (in and out) and a set for the permutation
    double2* xy1;
    double2* xy2;
    double* th1;
    double* th2;
    int* key;
    int* count;           // cell counts, zero between rebuilds
    int* start;
    int* perm;
    int* perm2;
    int* inv;
    long long target;     // absolute step every system runs to
    int period;           // > 0: rebuild every `period` steps and skip the displacement check
    int nsys;
    double thresh2;
    double2* samples;     // heading sums every sample_every steps (0: none)
    long long sample_base, sample_cap;
    int sample_every;
    int pad;
    Status* status;
};

// Workgroup-wide reductions (every thread gets the result).  sh has kBlock1024/64 slots.
__device__ __forceinline__ int wg_or(int v, int* sh)
{
    v = __any(v) ? 1 : 0;
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < kBlock1024 / 64; ++k) r |= sh[k];
    __syncthreads();
    return r;
}

__device__ __forceinline__ int wg_max(int v, int* sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < kBlock1024 / 64; ++k) r = max(r, sh[k]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ long long wg_sum(long long v, long long* sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    long long r = 0;
    for (int k = 0; k < kBlock1024 / 64; ++k) r += sh[k];
    __syncthreads();
    return r;
}

// Exclusive scan of one value per thread; returns this thread's prefix.
__device__ __forceinline__ int wg_exclusive_scan(int v, int* sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; ++k) before += sh[k];
    __syncthreads();
    return before + incl - v;
}

// Box::Update_Cells + Update_Neighbor_List for one system: counting sort of its slots by cell
// (keys koff.., id order inside a cell), permutation from `from` into `to`, Verlet list on `to`.
// Returns the longest list; nnz gets the entry count.
__device__ int rebuild_system(const Args& a, const SysInfo& S, const SetPtrs& from, const SetPtrs& to,
                              int* sh, long long* shl, long long& nnz)
{
    const int t = threadIdx.x, off = S.pid_off, end = S.pid_off + S.n;
    const Grid& G = S.G;
    const int koff = S.key_off, nk = G.ncx * G.ncy;
    for (int i = off + t; i < end; i += kBlock1024) {  // k_bin
        const double2 p = from.xy[i];
        const int cx = cell_coord(p.x, G.Lx, G.Lx2, G.invLx2, G.ncx, G.quirk);
        const int cy = cell_coord(p.y, G.Ly, G.Ly2, G.invLy2, G.ncy, G.quirk);
        const int k = koff + cy * G.ncx + cx;
        a.key[i] = k;
        atomicAdd(&a.count[k], 1);
    }
    __syncthreads();
    {  // exclusive scan of the system's counts: thread t owns cells [t q, t q + q)
        const int q = (nk + kBlock1024 - 1) / kBlock1024;
        const int c0 = min(nk, t * q), c1 = min(nk, c0 + q);
        int local = 0;
        for (int c = c0; c < c1; ++c) local += a.count[koff + c];
        int run = off + wg_exclusive_scan(local, sh);
        for (int c = c0; c < c1; ++c) {
            a.start[koff + c] = run;
            run += a.count[koff + c];
        }
        // the end of the last cell is the next system's first start, which never changes
        if (t == 0) a.start[koff + nk] = end;
    }
    __syncthreads();
    for (int i = off + t; i < end; i += kBlock1024) {  // k_scatter (counts return to zero)
        const int k = a.key[i];
        a.perm[a.start[k] + atomicSub(&a.count[k], 1) - 1] = i;
    }
    __syncthreads();
    for (int j = off + t; j < end; j += kBlock1024) {  // k_rank_permute, ranking
        const int me = a.perm[j], c = a.key[me], mine = from.id[me];
        const int s0 = a.start[c], e0 = a.start[c + 1];
        int r = 0;
        for (int q = s0; q < e0; ++q) r += from.id[a.perm[q]] < mine;
        a.perm2[s0 + r] = me;
    }
    __syncthreads();
    for (int k = off + t; k < end; k += kBlock1024) {  // k_rank_permute, copying
        const int o = a.perm2[k];
        to.xy[k] = from.xy[o];
        to.th[k] = from.th[o];
        to.U[k] = from.U[o];
        const int pid = from.id[o];
        to.id[k] = pid;
        a.inv[pid] = k;
    }
    __syncthreads();
    NlistFlatArgs nl = a.nl;
    nl.pos = to.xy;
    int mlen = 0;
    long long cnt = 0;
    for (int i = off + t; i < end; i += kBlock1024) {  // k_nlist_flat
        const int c = nlist_particle(nl, i, G, koff);
        mlen = max(mlen, c);
        cnt += c;
    }
    mlen = wg_max(mlen, sh);
    nnz = wg_sum(cnt, shl);
    return mlen;
}

// One stage over the system's slots (stage pointers wired as the multi-launch step wires them,
// with the step's input and output in the same set); returns whether an output left the skin.
template <int MODEL, int REP, int INTEG, int STAGE>
__device__ int stage_system(const Args& a, const SysInfo& S, const Phys& P, const SetPtrs& cur, long long step,
                            int* sh)
{
    StageArgs sa = a.st;
    sa.step = step;
    sa.id = cur.id;
    sa.id_out = nullptr;
    sa.inj = nullptr;
    sa.th0 = cur.th;
    sa.U_in = cur.U;
    sa.U_out = cur.U;
    if (INTEG == SPP_EULER) {
        sa.gxy = cur.xy;
        sa.gth = cur.th;
        sa.oxy = a.xy1;
        sa.oth = a.th1;
    } else if (INTEG == SPP_RK2) {
        sa.gxy = STAGE == 1 ? cur.xy : a.xy1;
        sa.gth = STAGE == 1 ? cur.th : a.th1;
        sa.oxy = STAGE == 1 ? a.xy1 : cur.xy;
        sa.oth = STAGE == 1 ? a.th1 : cur.th;
    } else {
        sa.gxy = STAGE == 1 ? cur.xy : (STAGE == 3 ? a.xy2 : a.xy1);
        sa.gth = STAGE == 1 ? cur.th : (STAGE == 3 ? a.th2 : a.th1);
        sa.oxy = STAGE == 2 ? a.xy2 : (STAGE == 4 ? cur.xy : a.xy1);
        sa.oth = STAGE == 2 ? a.th2 : (STAGE == 4 ? cur.th : a.th1);
    }
    const int off = S.pid_off, end = S.pid_off + S.n;
    int bad = 0;
    for (int i = off + threadIdx.x; i < end; i += kBlock1024) {
        double x, y;
        stage_particle<MODEL, REP, INTEG, STAGE>(sa, P, i, off, x, y);
        if (a.period == 0) {
            const float2 b = a.nl.build[i];
            const double dx = spp::min_image(x - (double)b.x, P.Lx2, P.invLx2);
            const double dy = spp::min_image(y - (double)b.y, P.Ly2, P.invLy2);
            bad |= dx * dx + dy * dy > a.thresh2;
        }
    }
    bad = wg_or(bad, sh);  // also the barrier before the next stage reads these outputs
    if (INTEG == SPP_EULER) {
        for (int i = off + threadIdx.x; i < end; i += kBlock1024) {
            cur.xy[i] = a.xy1[i];
            cur.th[i] = a.th1[i];
        }
        __syncthreads();
    }
    return bad;
}

// The stages of one step from `STAGE` on: 0 done, 1 an intermediate output left the skin (redo the
// step after a rebuild), 2 the final output did (rebuild before the next step).
template <int MODEL, int REP, int INTEG, int STAGE>
__device__ int step_from(const Args& a, const SysInfo& S, const Phys& P, const SetPtrs& cur, long long step, int* sh)
{
    constexpr int last = INTEG == SPP_EULER ? 0 : INTEG;
    const int bad = stage_system<MODEL, REP, INTEG, STAGE>(a, S, P, cur, step, sh);
    if constexpr (STAGE < last) {
        if (bad) return 1;
        return step_from<MODEL, REP, INTEG, STAGE + 1>(a, S, P, cur, step, sh);
    } else {
        return bad ? 2 : 0;
    }
}

template <int MODEL, int REP, int INTEG>
__global__ __launch_bounds__(kBlock1024) void k_persist(Args a)
{
    __shared__ int sh[kBlock1024 / 64];
    __shared__ long long shl[kBlock1024 / 64];
    __shared__ double sx[4], sy[4];
    const int s = blockIdx.x;
    const SysInfo S = a.sys_tab[s];
    Phys P = a.st.P;
    apply_sys(P, S);
    Status st = a.status[s];
    long long step = st.step;
    int since = st.since;
    bool need = since < 0;
    SetPtrs cur = a.cur, spare = a.spare;
    bool moved = false;
    while (step < a.target) {
        if (need) {
            long long nnz = 0;
            const int mlen = rebuild_system(a, S, cur, spare, sh, shl, nnz);
            const SetPtrs tmp = cur;
            cur = spare;
            spare = tmp;
            moved = !moved;
            since = 0;
            need = false;
            st.rebuilds++;
            st.max_len = mlen;
            st.nnz = nnz;
            if (mlen > a.nl.cap) {
                st.code = kNeedCap;
                st.need_cap = mlen;
                since = -1;
                break;
            }
        }
        constexpr int first = INTEG == SPP_EULER ? 0 : 1;
        const int r = step_from<MODEL, REP, INTEG, first>(a, S, P, cur, step, sh);
        if (r == 1) {
            if (since == 0) {  // a single stage moved a particle farther than the half skin
                st.code = kSkinJump;
                break;
            }
            need = true;  // rebuild from this step's start (untouched) and redo the step
            continue;
        }
        ++step;
        ++since;
        if (r == 2 || (a.period > 0 && since >= a.period)) need = true;
        if (a.sample_every > 0 && (step - a.sample_base) % a.sample_every == 0) {
            // heading sums as k_polar_sys forms them (256 threads, fixed order)
            const long long k = (step - a.sample_base) / a.sample_every - 1;
            if (k >= 0 && k < a.sample_cap && threadIdx.x < kBlock) {
                double vx = 0, vy = 0;
                for (int i = S.pid_off + threadIdx.x; i < S.pid_off + S.n; i += kBlock) {
                    double c, sn;
                    spp::heading_unit(cur.th[i], c, sn);
                    vx += c;
                    vy += sn;
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
            }
            __syncthreads();
            if (k >= 0 && k < a.sample_cap && threadIdx.x == 0) {
                double ax = 0, ay = 0;
                for (int w = 0; w < kBlock / 64; ++w) {
                    ax += sx[w];
                    ay += sy[w];
                }
                a.samples[k * a.nsys + s] = make_double2(ax, ay);
            }
            __syncthreads();
        }
    }
    if (moved) {  // the state ends in the engine's set
        for (int i = S.pid_off + threadIdx.x; i < S.pid_off + S.n; i += kBlock1024) {
            a.cur.xy[i] = cur.xy[i];
            a.cur.th[i] = cur.th[i];
            a.cur.U[i] = cur.U[i];
            a.cur.id[i] = cur.id[i];
        }
    }
    if (threadIdx.x == 0) {
        st.step = step;
        st.since = need ? -1 : since;  // -1: the list does not cover the current state
        a.status[s] = st;
    }
}

}  // namespace persist
