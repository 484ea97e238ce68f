/*
 * spp_engine.h -- C ABI of the MI355X particle engine for 2-D self-propelled particles.
 *
 * The reference (hamidallaei/SPP) has no FFI: its "operator API" is the header-only C++ Box /
 * Particle class surface.  Every entry point below replaces one of those methods; the comment on
 * each names the reference member it stands in for (file:line in the reference tree).  The C++
 * facade in spp_amd/csrc/box.hpp keeps the reference's method names on top of these calls.
 *
 * Conventions
 *   - Plain C types only; all arrays are host pointers in ORIGINAL particle order (index i is the
 *     i-th particle the caller handed to spp_set_state).  The engine keeps its device copy sorted
 *     by cell and un-permutes on every getter.
 *   - Box geometry follows the reference: Lx, Ly are HALF widths, the box is [-Lx, Lx) x [-Ly, Ly)
 *     (shared/parameters.h:68-74, shared/c2dvector.h:67-73).
 *   - Every call returns SPP_OK (0) or a negative error code; spp_last_error() gives the text.
 *     Nothing exits the process (the reference calls exit(0), e.g. shared/cell.h:45-49).
 *   - Calls are ordered on the engine's HIP stream; getters synchronise.  One engine per host
 *     thread; several engines may coexist on one GPU.
 */
#ifndef SPP_ENGINE_H
#define SPP_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPP_ABI_VERSION 1

/* ---- error codes ---------------------------------------------------------------------------- */
enum {
    SPP_OK = 0,
    SPP_ERR_ARG = -1,         /* bad argument / parameter combination                              */
    SPP_ERR_HIP = -2,         /* HIP runtime error (message has the HIP error string)              */
    SPP_ERR_STATE = -3,       /* call not valid in the engine's current state                      */
    SPP_ERR_CAPACITY = -4,    /* caller buffer too small (size needed is reported)                 */
    SPP_ERR_UNSUPPORTED = -5, /* model / integrator combination not implemented                    */
    SPP_ERR_NO_DEVICE = -6    /* no HIP device visible                                             */
};

/* ---- particle models ------------------------------------------------------------------------ */
enum {
    /* RepulsiveParticle (shared/particle.h:499-786): f = v0*vhat + sum F_rep(d < rc),
       tau = sum g*sin(theta_j - theta_i)/(pi*R_align^2) over d < R_align.                       */
    SPP_MODEL_REPULSIVE = 0,
    /* ContinuousParticle torque law (shared/particle.h:296-338): tau = g*sum[(1-alpha)
       sin(theta_j - theta_i)/pi - alpha*(dr x vhat_i)/(pi d)] over d^2 < R_align^2,
       f = v0*vhat (no pair force).                                                              */
    SPP_MODEL_CONTINUOUS = 1,
    /* Single-bead ActiveBrownianChain swimmers plus an optional closed ring membrane
       (shared/particle.h:792-1057, parallel/beadbox.h:185-196).                                 */
    SPP_MODEL_ABP = 2
};

/* ---- integrators (Box::Move / Move_Runge_Kutta2_* / Move_Runge_Kutta4_*) --------------------- */
enum { SPP_EULER = 1, SPP_RK2 = 2, SPP_RK4 = 4 };

/* ---- pair repulsion laws (shared/force-fields.h) ------------------------------------------- */
enum {
    SPP_REP_NONE = 0,
    SPP_REP_R12_TRUNCATED = 1, /* R12_Repulsive_Truncated, force-fields.h:14-39                   */
    SPP_REP_HARMONIC = 2,      /* Spring(dr, d, rc, k) gated by d < rc: k (rc - d) dr/d            */
    SPP_REP_R12 = 3,           /* R12_Repulsive, force-fields.h:41-58 (gated by d < rc)            */
    SPP_REP_YUKAWA_TRUNCATED = 4, /* Yukawa_Truncated, force-fields.h:60-73                        */
    SPP_REP_YUKAWA = 5         /* Yukawa, force-fields.h:75-85 (gated by d < rc)                   */
};

/* ---- cell binning ---------------------------------------------------------------------------- */
enum {
    SPP_BIN_EXACT = 0,     /* cell = floor((x + Lx) * ncell_x / (2 Lx))                            */
    SPP_BIN_REF_QUIRK = 1  /* cell = (int)((int)(x + Lx) * ncell_x / (2 Lx)), serial/box.h:116-117 */
};

/* ---- neighbour-list rebuild policy ------------------------------------------------------------ */
enum {
    SPP_REBUILD_DISPLACEMENT = 0, /* rebuild when any particle moved more than (r_list - r_cut)/2   */
    SPP_REBUILD_PERIOD = 1        /* rebuild every rebuild_period steps (Multi_Step(cell_update_period)
                                     then Update_Cells, parallel/box.h:394-439)                   */
};

/* ---- how a rebuild sorts the particles into cells ---------------------------------------------- */
enum {
    SPP_REBUILD_SORT_AUTO = 0,   /* one-launch tile sort where it applies (single box, slot kernels,
                                    no interval graphs): a particle moves less than a cell between
                                    builds, so each tile of a cell row gathers its particles from the
                                    previous order; the general sort takes over when that fails     */
    SPP_REBUILD_SORT_GENERAL = 1 /* always the general counting sort (binning, scan, scatter,
                                    in-cell order, permutation); same slots bit for bit            */
};

/* ---- angular noise source --------------------------------------------------------------------- */
enum {
    SPP_NOISE_STREAM = 0,  /* Philox4x32-10 stream addressed by (seed, particle id, step), Box-Muller */
    SPP_NOISE_INJECTED = 1 /* caller supplies d_theta for each step with spp_inject_noise            */
};

/* ---- interaction-kernel variants (same forces, different gather strategy or launch shape) ----- */
enum {
    SPP_KERNEL_SLOT = 0,       /* one lane per particle, neighbour records gathered from L2         */
    /* 1 and 2 are retired (SPP_ERR_UNSUPPORTED): a heading fetched only inside the cutoff, and cell
       strips staged in LDS; both were slower than SLOT at every measured size                       */
    SPP_KERNEL_PERSIST = 3,    /* one workgroup per system (<= 4096 particles, Vicsek models, noise
                                  stream) runs a whole spp_multi_step in one launch: rebuilds when a
                                  system's own displacement (or period) requires; single or batched */
    SPP_KERNEL_SLOT_DEEP = 4   /* as SLOT with eight neighbour records in flight per lane at every
                                  size (SLOT keeps eight below 393,216 particles, four above); same
                                  results bit for bit; single or batched, Vicsek models            */
};

typedef struct spp_params {
    int32_t abi_version;    /* must be SPP_ABI_VERSION                                              */
    int32_t model;          /* SPP_MODEL_*                                                          */
    int32_t integrator;     /* SPP_EULER / SPP_RK2 / SPP_RK4                                        */
    int32_t repulsion;      /* SPP_REP_*                                                            */
    int32_t n;              /* number of particles (ring beads first for SPP_MODEL_ABP)             */
    int32_t n_membrane;     /* SPP_MODEL_ABP: number of ring beads (0 = no membrane)                */
    int32_t binning;        /* SPP_BIN_*                                                            */
    int32_t rebuild;        /* SPP_REBUILD_*                                                        */
    int32_t rebuild_period; /* steps between rebuilds for SPP_REBUILD_PERIOD                        */
    int32_t ncell_x;        /* cells along x (reference divisor_x); 0 = floor(2 Lx / r_list)        */
    int32_t ncell_y;        /* cells along y (reference divisor_y); 0 = floor(2 Ly / r_list)        */
    int32_t noise_mode;     /* SPP_NOISE_*                                                          */
    uint64_t seed;          /* noise stream key                                                     */
    double Lx, Ly;          /* half widths of the periodic box                                      */
    double dt;              /* time step                                                            */
    double v0;              /* self-propulsion: f starts at v0 * (cos theta, sin theta)             */
    double g;               /* alignment coupling                                                   */
    double R_align;         /* alignment radius                                                     */
    double alpha;           /* SPP_MODEL_CONTINUOUS self-torque fraction (0 = pure alignment)       */
    double A_rep;           /* repulsion amplitude (k for SPP_REP_HARMONIC)                         */
    double sigma_rep;       /* repulsion length scale sigma                                         */
    double rc_rep;          /* repulsion cutoff radius                                              */
    double D_rot;           /* rotational diffusion: d_theta ~ N(0, sqrt(2 D_rot dt))               */
    double r_list;          /* Verlet list radius (reference rv), >= every interaction cutoff       */
    /* SPP_MODEL_ABP only (v0 is the swimmers' self-propulsion force F0, speed mob_par * F0; every
       pair within rc_rep feels the repulsion law, as the reference's single-bead branch) --------- */
    double mob_par;         /* mobility along the heading (reference m_parallel, 0.5 for nb = 1)   */
    double mob_perp;        /* mobility across the heading (reference m_perpendicular)             */
    double mob_rot;         /* rotational mobility (reference k_perpendicular, 1 for nb = 1)       */
    double torque0;         /* intrinsic torque (reference torque0)                                 */
    double k_bond;          /* ring bond stiffness (reference membrane_elasticity)                  */
    double l_bond;          /* ring bond rest length                                                */
    double kappa_bend;      /* ring bending: E = kappa sum_i (1 - cos phi_i), phi_i the turn angle at
                               bead i (no reference counterpart; 0 = none)                          */
    int32_t ell_capacity;   /* neighbour slots per particle on the device (0 = default 32)          */
    int32_t kernel_variant; /* SPP_KERNEL_*: how the interaction kernels gather neighbours          */
    int32_t graph_intervals;/* 1: with SPP_REBUILD_PERIOD (RK2/RK4, no sampling), capture each full
                               interval (rebuild + rebuild_period steps) once as a HIP graph per
                               buffer layout and replay it; 0: launch every kernel                  */
    int32_t rebuild_sort;   /* SPP_REBUILD_SORT_*                                                   */
    /* SPP_MODEL_ABP multi-bead swimmers (ActiveBrownianChain with nb > 1, shared/particle.h:792-1057).
       nb_swimmer <= 1: single-bead swimmers, every pair feels the repulsion law above (the reference's
       nb = 1 branch).  nb_swimmer >= 2: swimmer i carries nb_swimmer beads at s_k = v_i bead_sep
       ((1 - nb)/2 + k); pairs of two single beads (ring beads) keep the law above, every other pair
       sums R12_Repulsive_Truncated(sigma_bead, rc_bead, A_bead) over bead pairs, with torques s_k x F
       (particle.h:983-1011); swimmers move with the mobilities of Set_nb(nb_swimmer) (:856-876), ring
       beads with mob_par / mob_perp / mob_rot.                                                    */
    int32_t nb_swimmer;
    int32_t device_timing;  /* 1: HIP events on the engine's stream around each outermost stepping or
                               measuring call, summed into spp_info.device_ms; 0 (default): none, and
                               device_ms stays 0 (an event record idles the GPU ~5 us per call)    */
    int32_t reserved_i[2];
    double bead_sep;        /* bead separation (reference ActiveBrownianChain::separation)          */
    double sigma_bead;      /* bead repulsion sigma (sigma_p)                                        */
    double rc_bead;         /* bead repulsion cutoff (repulsion_radius)                              */
    double A_bead;          /* bead repulsion amplitude (A_p)                                        */
    double reserved_d[7];
} spp_params;

typedef struct spp_engine spp_engine;

/* One independent periodic box of a batched engine (spp_create_batch).  Everything not listed
   here (model, integrator, laws, radii, dt, r_list, rebuild policy) comes from spp_params.        */
typedef struct spp_system {
    int32_t n;              /* particles of this box                                                */
    int32_t ncell_x;        /* 0 = floor(2 Lx / r_list)                                             */
    int32_t ncell_y;
    int32_t reserved;
    double Lx, Ly;          /* half widths                                                          */
    double v0, g, D_rot;    /* self-propulsion, alignment coupling, rotational diffusion            */
    uint64_t seed;          /* noise stream key; systems with equal seeds draw identical noise      */
} spp_system;

typedef struct spp_info {
    int64_t steps;          /* completed steps (also the noise-stream step counter)                 */
    int64_t rebuilds;       /* neighbour-list rebuilds so far                                       */
    int64_t rollbacks;      /* chunks re-run because a displacement check failed late              */
    int64_t host_syncs;     /* host round trips taken by spp_multi_step                            */
    double t;               /* simulated time                                                       */
    int32_t ncell_x, ncell_y;
    int32_t ell_capacity;
    int32_t max_neighbors;  /* longest full neighbour list at the last rebuild                     */
    double mean_neighbors;  /* mean full neighbour-list length at the last rebuild                 */
    int64_t tile_rebuilds;  /* rebuilds sorted by the one-launch tile path (SPP_REBUILD_SORT_AUTO)  */
    int64_t tile_fallbacks; /* tile sorts that failed and were redone by the general sort           */
    double device_ms;       /* with spp_params.device_timing: device time of the engine's stepping
                               and measuring calls so far (spp_set_state, spp_multi_step,
                               spp_multi_step_sampled, spp_batch_polarization, spp_lyapunov_*): HIP
                               events on the engine's stream around each outermost call, summed  */
    int64_t wide_rebuilds;  /* rebuilds that built the wide list (5x5-cell stencil, two cell widths):
                               after a step moved a particle beyond the half skin right after a
                               rebuild, until the particles slow down                              */
} spp_info;

/* ---- lifetime ----------------------------------------------------------------------------------- */

/* Fill *p with the defaults of a model (reference statics in shared/particle.h:774-786,
   1048-1056 and globals in shared/parameters.h:60-99).                                             */
int spp_params_default(spp_params* p, int32_t model);

/* Box::Box + Box::Init (parallel/box.h:69-151): validates parameters, allocates device state.     */
int spp_create(const spp_params* p, spp_engine** out);
/* Many independent boxes in one engine: every kernel launch advances all of them (the reference
   runs independent realizations as separate MPI processes, serial/mpi-main.cpp:157-160).  Caller
   arrays are the systems' particles concatenated in order.  p->n, p->Lx, p->Ly, p->v0, p->g,
   p->D_rot and p->seed are ignored in favour of the per-system values.                           */
int spp_create_batch(const spp_params* p, const spp_system* systems, int32_t nsys, spp_engine** out);
void spp_destroy(spp_engine* e);
/* Text of the last error on this engine (e may be NULL for errors raised by spp_create).          */
const char* spp_last_error(const spp_engine* e);
int spp_get_info(spp_engine* e, spp_info* info);

/* ---- state ------------------------------------------------------------------------------------- */

/* Particle positions and headings, host -> device.  xw/yw are wrapped positions (reference r),
   xu/yu unwrapped (reference r_original), theta the heading.  Any of xu/yu may be NULL: they are
   then taken equal to xw/yw.  Also rebuilds cells and the neighbour list (Box::Load,
   parallel/box.h:222-243, which ends in Full_Update_Cells).  SPP_ERR_ARG, with the engine's state
   unchanged, if a wrapped position lies outside its box [-L, L] or any value is NaN or infinite.  */
int spp_set_state(spp_engine* e, const double* xw, const double* yw, const double* xu,
                  const double* yu, const double* theta, int32_t n);
/* Device -> host in original order.  Any pointer may be NULL.                                     */
int spp_get_state(spp_engine* e, double* xw, double* yw, double* xu, double* yu, double* theta,
                  int32_t n);

/* ---- neighbour structures (Box::Update_Cells + Update_Neighbor_List, serial/box.h:107-179) ---- */

int spp_build_neighbors(spp_engine* e);
/* Full (symmetric) Verlet list in original ids: offsets[n+1] CSR, idx[nnz].  Order inside a row
   is unspecified.  Returns SPP_ERR_CAPACITY and sets *nnz when cap is too small.                 */
int spp_get_neighbors(spp_engine* e, int32_t* offsets, int32_t* idx, int64_t cap, int64_t* nnz);
/* Cell index of every particle as the last build binned it (original order).                     */
int spp_get_cells(spp_engine* e, int32_t* cell_x, int32_t* cell_y, int32_t n);

/* ---- per-stage operations (the reference's Box / Particle methods) ----------------------------- */

/* Box::Interact (serial/box.h:322-371): forces and torques at the current stage state.           */
int spp_interact(spp_engine* e);
/* f (including the self-propulsion term) and tau of the last spp_interact, original order.       */
int spp_get_forces(spp_engine* e, double* fx, double* fy, double* tau, int32_t n);
/* The Particle-level surface, for callers that loop over pairs themselves.
   spp_reset_forces = Particle::Reset for every particle (particle.h:559-564, 878-883): f = v0 v
   (F0 v for swimmers, 0 for ring beads), torque 0 (torque0 for SPP_MODEL_ABP).
   spp_pair_interact = particles[i].Interact(particles[j]) (particle.h:728-757, 296-338, 964-1012):
   adds the pair's force and torque to both particles' accumulated ones, whatever their distance
   (the law's own cutoffs apply).  Ring bonds and bending are Box-level (spp_interact) only.  The
   accumulated forces are what spp_get_forces reads and spp_move_stage applies, so
   reset + pair calls + move stages is the reference's hand-written Interact loop.                */
int spp_reset_forces(spp_engine* e);
int spp_pair_interact(spp_engine* e, int32_t i, int32_t j);
/* Noise for the next step (original order): replaces Noise_Gen (particle.h:566-569) when
   noise_mode == SPP_NOISE_INJECTED.                                                              */
int spp_inject_noise(spp_engine* e, const double* dtheta, int32_t n);
/* Box::Move (stage 0, Euler), Move_Runge_Kutta2_{1,2} (stage 1..2) or Move_Runge_Kutta4_{1..4}
   (stage 1..4), using the forces of the last spp_interact.  The last stage of a step advances the
   step counter.                                                                                   */
int spp_move_stage(spp_engine* e, int32_t stage);
/* Box::One_Step (serial/box.h:427-455) through the fused interact+move kernels.                   */
int spp_one_step(spp_engine* e);
/* Box::Multi_Step (serial/box.h:469-506): steps through the fused kernels, rebuilding the
   neighbour list by the configured policy.                                                        */
int spp_multi_step(spp_engine* e, int32_t steps);

/* ---- observables and snapshots ----------------------------------------------------------------- */

/* Box::Compute_Polarization (serial/box.h:531-538): P = |sum vhat| / N, plus the summed vector.   */
int spp_polarization(spp_engine* e, double* p, double* sum_x, double* sum_y);
/* spp_multi_step plus a polarisation sample of every system after each `every` steps, recorded on
   the device (no extra host round trip): polar[k * nsys + s] after step k * every of this call
   (Box::Save_Polarization, parallel/box.h:467-473, called in the data_gathering loop).  nsys = 1
   for a single-box engine; steps must be a multiple of every; len = (steps / every) * nsys.      */
int spp_multi_step_sampled(spp_engine* e, int32_t steps, int32_t every, double* polar, int64_t len);

/* Polarization of each system of a batched engine: p[s] = |sum vhat| / n_s.                     */
int spp_batch_polarization(spp_engine* e, double* p, int32_t nsys);

/* Isotropic pair distribution g(r) over [0, r_max) in nbins bins (r_max <= r_list), from the Verlet
   list: counts[k] = ordered pairs (i, j) with distance in bin k; g[k] = counts[k] / (N rho pi
   (r_{k+1}^2 - r_k^2)), rho = N / (4 Lx Ly).  Either output may be null.  Single-box engines.
   (The reference has a body-frame 2-D pair distribution, analyze/analyze.h:365-423, and a radial
   density about the origin, :332-362; the isotropic g(r) is new.)                              */
int spp_pair_distribution(spp_engine* e, double r_max, int32_t nbins, int64_t* counts, double* g);
/* spp_pair_distribution without any allocation.  No counterpart in the reference: it is for callers
   that cannot allocate device (or host) memory -- code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc.  The histogram lives in a static device array of 8192
   bins (64 KB per GPU, held until the process ends): nbins above 8192 is SPP_ERR_ARG.  The list
   must be current (SPP_ERR_STATE otherwise: a rebuild could grow it; call spp_build_neighbors
   first), and profiling must be off.  counts and g as in spp_pair_distribution, bit for bit; the
   counts pass through the caller's counts array (or, if it is NULL, through g).  All engines on one
   GPU share the array: calls that overlap on one GPU (engines on different streams or threads)
   race.                                                                                          */
int spp_pair_distribution_noalloc(spp_engine* e, double r_max, int32_t nbins, int64_t* counts, double* g);

/* ---- Lyapunov exponents by finite-difference clones ----------------------------------------------
   The reference evolves perturbed copies of the state with identical noise and re-orthonormalises
   their deviations by Gram-Schmidt (parallel/lyapunovbox.h:128-194, shared/vector-set.h:107-127,
   parallel/separate-lyapunovbox.h:224-305).  Here a batched engine holds `replicas` groups of
   (1 + vectors) systems, [base, clone_1 .. clone_K], all with identical parameters and seeds.
   Deviations live in the 3n space (x, y, theta) with the reference's inner product
   (state-hyper-vector.h:157-167) and minimum-image / wrapped-angle differences (:96-106, :223).   */
/* Sets every clone to base + epsilon * (orthonormalised random deviation).  rescale = 1: every
   renormalisation restarts the clones at distance epsilon (Benettin).  rescale = 0: the reference's
   choice -- the projected deviation keeps its grown length (lyapunovbox.h:177-183).  Both
   accumulate the same log growth in the linear regime.  Needs an engine from spp_create_batch
   with replicas * (1 + vectors) systems, all with the same n; a replica's systems need equal
   boxes, parameters and seeds.  Clone states are overwritten; the bases keep theirs.             */
int spp_lyapunov_init(spp_engine* e, int32_t replicas, int32_t vectors, double epsilon, int32_t rescale,
                      uint64_t seed);
/* Gram-Schmidt (QR) of the current deviations, adds log(R_kk / start length) to the accumulators
   and resets the clones to base + length * Q.  Optional output (synchronises): the accumulated
   sums, replicas x vectors.                                                                     */
int spp_lyapunov_renormalize(spp_engine* e, double* log_growth);
/* `steps` steps, renormalising every `interval` steps (steps must be a multiple of interval).    */
int spp_lyapunov_run(spp_engine* e, int32_t steps, int32_t interval);
/* Accumulated log growth per (replica, vector) and the time it covers: lambda = sum / time.      */
int spp_lyapunov_get(spp_engine* e, double* sum_log, double* time);
/* The orthonormal directions of one replica after the last renormalisation, laid out
   [vector][x|y|theta][particle] (len = vectors * 3 * n).                                          */
int spp_lyapunov_vectors(spp_engine* e, int32_t replica, double* out, int64_t len);

/* Box::Save / Box::Load (parallel/box.h:222-267) into device-resident slots, including the noise
   step counter (the counter-based stream makes replay exact).                                    */
int spp_snapshot(spp_engine* e, int32_t slot);
int spp_restore(spp_engine* e, int32_t slot);
/* What a host-side copy of the state needs besides positions and headings to replay a run bit for
   bit.  Box::Load(const State_Hyper_Vector&) restores the saved copy of the random generator with
   the state (Set_C2DVector_Rand_Generator, state-hyper-vector.h:169-172, parallel/box.h:236); with
   the counter-based stream that copy is the step counter.  The rebuild estimates decide at which
   steps the neighbour list is rebuilt under SPP_REBUILD_DISPLACEMENT, and with that the order in
   which pair terms are summed (forces agree to rounding for any schedule).                       */
typedef struct spp_clock {
    int64_t step;               /* noise-stream step counter (spp_info.steps)                          */
    double t;                   /* simulated time                                                      */
    double rebuild_estimate[3]; /* last three estimates of the steps until a particle crosses the half
                                   skin (0 = none yet)                                                 */
} spp_clock;
int spp_get_clock(spp_engine* e, spp_clock* clock);
/* SPP_ERR_ARG for a negative step or a negative or non-finite estimate; SPP_ERR_STATE inside a
   split step.  A host-side state restored with spp_set_state plus this call replays the run.      */
int spp_set_clock(spp_engine* e, const spp_clock* clock);

/* ---- timing -------------------------------------------------------------------------------------- */

int spp_synchronize(spp_engine* e);
/* Interval graphs (spp_params.graph_intervals): intervals captured, intervals launched from a
   graph (captures included), graphs held.  No reference counterpart.                            */
int spp_graph_stats(spp_engine* e, int64_t* captures, int64_t* replays, int32_t* cached);
/* Per-kernel HIP-event timing on the engine's stream: enable, run work, read averages.
   names: '\n'-separated kernel names written into buf (cap bytes).  ms_avg/launches have room
   for max_kernels entries; *nk receives the number of kernels.                                    */
int spp_profile_enable(spp_engine* e, int32_t on);
int spp_profile_read(spp_engine* e, char* names, int32_t cap, double* ms_avg, int64_t* launches,
                     int32_t max_kernels, int32_t* nk);

/* Average duration of one fused stage kernel (stage as in spp_move_stage) over `reps` back-to-back
   launches bracketed by two HIP events on the engine's stream.  The launches read the current
   state and write scratch buffers, so the state is unchanged.  No reference counterpart: this is
   the measurement hook for the roofline report.                                                   */
int spp_time_stage(spp_engine* e, int32_t stage, int32_t reps, double* avg_us);

/* ---- helpers ------------------------------------------------------------------------------------ */

/* Uniform positions in [-Lx, Lx) x [-Ly, Ly) and headings in [0, 2 pi) from a seed
   (Random_Formation, shared/set-up.h:152-163, with a counter-based generator).                   */
int spp_random_state(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y,
                     double* theta);
/* Random_Formation(particle, N, sigma) (set-up.h:152-163): positions uniform in
   [-(Lx - sigma), Lx - sigma) x [-(Ly - sigma), Ly - sigma), headings uniform in [0, 2 pi);
   sigma = 0 gives spp_random_state's positions.  0 <= sigma < min(Lx, Ly).                       */
int spp_random_formation(uint64_t seed, int32_t n, double Lx, double Ly, double sigma, double* x, double* y,
                         double* theta);
/* Polar_Formation (set-up.h:181-192): Random_Formation's positions, every heading 0 (v = (1, 0)). */
int spp_polar_formation(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y, double* theta);

/* Triangle_Lattice_Formation (shared/set-up.h:116-145): n sites of a triangular lattice with
   spacing (2 Lx - sigma) / Nx, Nx = (int)sqrt(Lx n sqrt(3) / (2 Ly)) + 1, offset by -(L - sigma)
   and wrapped into the box, in the reference's site order.                                      */
int spp_triangle_lattice(int32_t n, double Lx, double Ly, double sigma, double* x, double* y);
/* Ring_Membrane (shared/set-up.h:292-316): nm beads on a circle about the origin; radius
   0.5 bead_diameter / sin(pi / nm) (touching beads) unless radius > 0 is given.                 */
int spp_ring_membrane(int32_t nm, double bead_diameter, double radius, double* x, double* y);
/* n points uniform in the disk of the given radius about the origin, headings uniform in
   [0, 2 pi), from the Philox stream keyed by seed (swimmers placed inside a ring membrane; the
   reference's Confined_In_Ring_Membrane, set-up.h:319-366, reads an uninitialised swimmer length
   and is not restated).                                                                          */
int spp_disk_state(uint64_t seed, int32_t n, double radius, double* x, double* y, double* theta);
/* Standard-normal variate of the angular-noise stream for (seed, particle id, step); the engine's
   d_theta is sqrt(2 D_rot dt) times this value.                                                   */
double spp_noise_normal(uint64_t seed, int32_t id, int64_t step);

#ifdef __cplusplus
}
#endif

#endif /* SPP_ENGINE_H */
