/*
 * spp_oracle.c -- CPU oracle for the per-time-step particle path (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, in plain C, the reference's per-step algorithm.  Every function names the reference
 * file:line it follows.  The reference itself does not compile (missing GSL/Boost/MPI headers and
 * source-level errors such as shared/set-up.h:29-30 assigning a const int), so this restatement is
 * the oracle.  It is pinned by closed forms, brute force and the published Philox known-answer
 * vectors in tests/test_oracle.py.
 *
 * Deliberate, documented departures from the reference (each is a project definition, not a
 * reference behaviour):
 *   - noise comes from the project's counter-based stream (Philox4x32-10 + Box-Muller addressed by
 *     (seed, particle id, step)) instead of the global GSL mt19937 stream (c2dvector.h:19-25);
 *   - self-propulsion speed v0 scales the Reset force (particle.h:559-564 hard-codes speed 1);
 *   - SPP_REP_HARMONIC (k (rc - d) dr / d for d < rc) is Spring (force-fields.h:87-97) gated by
 *     d < rc, which the reference does not ship as a pair law;
 *   - SPP_MODEL_CONTINUOUS integrates the ContinuousParticle torque law with the RepulsiveParticle
 *     RK2/RK4/Euler updates (the reference has Euler only for that law);
 *   - cell indices are clamped to the grid (the reference indexes out of range when a wrapped
 *     coordinate rounds onto +Lx);
 *   - SPP_MODEL_ABP adds a ring bending energy kappa * sum_i (1 - cos phi_i), which the
 *     reference's bead boxes do not have;
 *   - ring beads move their unwrapped position with the wrapped one (the reference's
 *     Positioning_Particles copies r_original into r, serial/beadbox.h:255-259).
 */
#include "spp_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846 /* the reference's PI = M_PI, shared/parameters.h:62 */
#endif

static char g_err[512];

static void set_err(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char* orc_error(void) { return g_err; }

/* ------------------------------------------------------------------------------------------------
 * Vec_template::Periodic_Transform, shared/c2dvector.h:67-73 (one component):
 *     x -= Lx2*((int) floor(x / Lx2 + 0.5));
 * ---------------------------------------------------------------------------------------------- */
double orc_periodic(double x, double L2)
{
    return x - L2 * (double)((int)floor(x / L2 + 0.5));
}

/* Theta wrap of the RK stages, shared/particle.h:599:
 *     theta = theta - floor(theta/(2*M_PI) + 0.5)*2*M_PI;                                         */
static double wrap_angle(double th)
{
    return th - floor(th / (2 * M_PI) + 0.5) * 2 * M_PI;
}

/* ------------------------------------------------------------------------------------------------
 * Cell binning, serial/box.h:116-117 (also parallel/node.h:758-759, 896-897):
 *     x = (int) (particle[i].r.x + Lx)*divisor_x / Lx2;
 * The cast binds to (r.x + Lx) only, so the coordinate is truncated before scaling.  The result is
 * clamped to the grid (see the header comment).
 * ---------------------------------------------------------------------------------------------- */
int orc_cell_quirk(double x, double L, int divisor)
{
    int c = (int)((double)((int)(x + L) * divisor) / (2 * L));
    if (c < 0) c = 0;
    if (c >= divisor) c = divisor - 1;
    return c;
}

/* Exact binning: floor((x + L) * divisor / (2 L)), clamped.                                       */
int orc_cell_exact(double x, double L, int divisor)
{
    int c = (int)floor((x + L) * divisor / (2 * L));
    if (c < 0) c = 0;
    if (c >= divisor) c = divisor - 1;
    return c;
}

/* ------------------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11).  Round: (hi0, lo0) = M0*c0, (hi1, lo1) = M1*c2,
 *                 c' = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0); key += Weyl between rounds.
 * ---------------------------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Project noise stream: counter (id, step_lo, step_hi, 'SPPn'), key (seed_lo, seed_hi); two
 * 53-bit uniforms u_a, u_b in [0, 1); z = sqrt(-2 log(1 - u_a)) cos(2 pi u_b).                     */
double orc_noise_normal(uint64_t seed, int32_t id, int64_t step)
{
    uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32),
                       0x5350506Eu};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    double ua = ((double)(w[0] >> 5) * 67108864.0 + (double)(w[1] >> 6)) * (1.0 / 9007199254740992.0);
    double ub = ((double)(w[2] >> 5) * 67108864.0 + (double)(w[3] >> 6)) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(1.0 - ua)) * cos(6.283185307179586 * ub);
}

/* Initial condition: uniform positions in [-Lx, Lx) x [-Ly, Ly), headings in [0, 2 pi)
 * (the role of Random_Formation, shared/set-up.h:152-163), from the tag 'SPPi' stream.            */
void orc_random_state(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y,
                      double* theta)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int32_t i = 0; i < n; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, 0u, 0u, 0x53505069u};
        uint32_t w[4];
        orc_philox4x32_10(ctr, key, w);
        double u0 = (double)w[0] * (1.0 / 4294967296.0);
        double u1 = (double)w[1] * (1.0 / 4294967296.0);
        double u2 = (double)w[2] * (1.0 / 4294967296.0);
        x[i] = -Lx + 2 * Lx * u0;
        y[i] = -Ly + 2 * Ly * u1;
        theta[i] = 6.283185307179586 * u2;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Pair force laws, shared/force-fields.h.  dr = r_this - r_other, d = |dr|.
 * ---------------------------------------------------------------------------------------------- */
void orc_force_law(int kind, double dx, double dy, double d, double rc, double sigma, double A,
                   double* fx, double* fy)
{
    double s;
    switch (kind) {
    case SPP_REP_R12_TRUNCATED: { /* force-fields.h:14-39 */
        double sd = sigma / d, sd2 = sd * sd, sd4 = sd2 * sd2, sd8 = sd4 * sd4;
        double sd13 = sd8 * sd4 * sd;
        double sc = sigma / rc, sc2 = sc * sc, sc4 = sc2 * sc2, sc8 = sc4 * sc4;
        double sc13 = sc8 * sc4 * sc;
        s = A * (sd13 - sc13);
        *fx = s * dx / d;
        *fy = s * dy / d;
        return;
    }
    case SPP_REP_R12: { /* force-fields.h:41-58 */
        double sd = sigma / d, sd2 = sd * sd, sd4 = sd2 * sd2, sd8 = sd4 * sd4;
        double sd13 = sd8 * sd4 * sd;
        s = A * sd13;
        *fx = s * dx / d;
        *fy = s * dy / d;
        return;
    }
    case SPP_REP_YUKAWA_TRUNCATED: { /* force-fields.h:60-73 */
        double d2 = d * d, c2 = rc * rc;
        *fx = A * (exp(-d / sigma) * (1. / d2 + 1. / (sigma * d)) -
                   exp(-rc / sigma) * (1. / c2 + 1. / (sigma * rc))) * dx / d;
        *fy = A * (exp(-d / sigma) * (1. / d2 + 1. / (sigma * d)) -
                   exp(-rc / sigma) * (1. / c2 + 1. / (sigma * rc))) * dy / d;
        return;
    }
    case SPP_REP_YUKAWA: { /* force-fields.h:75-85 */
        double d2 = d * d;
        *fx = A * (exp(-d / sigma) * (1. / d2 + 1. / (sigma * d))) * dx / d;
        *fy = A * (exp(-d / sigma) * (1. / d2 + 1. / (sigma * d))) * dy / d;
        return;
    }
    case SPP_REP_HARMONIC: /* Spring(dr, d, rc, k), force-fields.h:87-97 */
        *fx = A * (rc - d) * dx / d;
        *fy = A * (rc - d) * dy / d;
        return;
    default:
        *fx = 0;
        *fy = 0;
        return;
    }
}

/* Brute-force O(N^2) pair enumeration with the reference's Verlet predicate sqrt(d2) < r
 * (shared/cell.h:84-85) and minimum-image differences.  Pairs are reported with i < j.           */
int64_t orc_brute_pairs(const double* x, const double* y, int32_t n, double Lx, double Ly,
                        double r, int32_t* pi, int32_t* pj, int64_t cap)
{
    int64_t m = 0;
    double L2x = 2 * Lx, L2y = 2 * Ly;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = i + 1; j < n; ++j) {
            double dx = orc_periodic(x[i] - x[j], L2x);
            double dy = orc_periodic(y[i] - y[j], L2y);
            double d = sqrt(dx * dx + dy * dy);
            if (d < r) {
                if (m < cap) {
                    pi[m] = i;
                    pj[m] = j;
                }
                ++m;
            }
        }
    return m;
}

/* ------------------------------------------------------------------------------------------------
 * Box state.  Particle fields follow RepulsiveParticle (shared/particle.h:499-520) and
 * BasicParticle/BasicDynamicParticle (:39-71); the cell keeps a pid list (shared/cell.h:8-41).
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    double x, y;
} vec2;

typedef struct {
    vec2 r, v, r_original, f, r_old;
    double theta, theta_old, torque, dtheta;
    vec2 k1_f, k2_f, k3_f;
    double k1_t, k2_t, k3_t;
    int32_t* nbr; /* neighbor_id */
    int32_t nn, ncap;
} particle;

typedef struct {
    int32_t* pid;
    int32_t n, cap;
} cell;

struct orc_sys {
    spp_params p;
    int32_t N;
    particle* P;
    cell* C;
    int32_t ncx, ncy;
    double Lx2, Ly2, half_dt, dt_over_6, noise_amp;
    double r_cut, half_skin;
    int64_t step;
    int64_t rebuilds;
    int64_t steps_since_build;
    double t;
    vec2* build_r; /* wrapped positions at the last build (displacement check) */
    double* injected;
    int has_injected;
    int has_state;
};

static void push(int32_t** a, int32_t* n, int32_t* cap, int32_t v)
{
    if (*n == *cap) {
        int32_t nc = *cap ? 2 * *cap : 8;
        *a = (int32_t*)realloc(*a, (size_t)nc * sizeof(int32_t));
        *cap = nc;
    }
    (*a)[(*n)++] = v;
}

static double max_cutoff(const spp_params* p)
{
    if (p->model == SPP_MODEL_ABP) {
        double c = p->repulsion != SPP_REP_NONE ? p->rc_rep : 0.0;
        /* two chains interact while bead pairs, up to bead_sep (nb - 1) / 2 from either centre, are
           closer than rc_bead */
        if (p->nb_swimmer >= 2) {
            double cb = p->rc_bead + p->bead_sep * (p->nb_swimmer - 1);
            if (cb > c) c = cb;
        }
        return c;
    }
    double c = p->R_align;
    if (p->repulsion != SPP_REP_NONE && p->rc_rep > c) c = p->rc_rep;
    return c;
}

orc_sys* orc_create(const spp_params* p)
{
    if (!p || p->n <= 0 || p->Lx <= 0 || p->Ly <= 0 || p->dt <= 0) {
        set_err("orc_create: bad parameters");
        return NULL;
    }
    if (p->model != SPP_MODEL_REPULSIVE && p->model != SPP_MODEL_CONTINUOUS && p->model != SPP_MODEL_ABP) {
        set_err("orc_create: model %d not in the oracle", p->model);
        return NULL;
    }
    if (p->model == SPP_MODEL_ABP &&
        (p->integrator == SPP_RK4 || p->n_membrane < 0 || p->n_membrane > p->n ||
         (p->n_membrane > 0 && p->n_membrane < 3))) {
        set_err("orc_create: active Brownian model needs Euler/RK2 and n_membrane 0 or in [3, n]");
        return NULL;
    }
    if (p->model == SPP_MODEL_ABP && p->nb_swimmer >= 2 &&
        (p->nb_swimmer > 16 || !(p->bead_sep > 0) || !(p->sigma_bead > 0) || !(p->rc_bead > 0))) {
        set_err("orc_create: chains need 2 <= nb_swimmer <= 16 and positive bead_sep, sigma_bead, rc_bead");
        return NULL;
    }
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->p = *p;
    s->N = p->n;
    s->Lx2 = 2 * p->Lx;
    s->Ly2 = 2 * p->Ly;
    /* shared/parameters.h:77-79 */
    s->half_dt = p->dt / 2;
    s->dt_over_6 = p->dt / 6;
    /* BasicDynamicParticle::Set_Dr, shared/particle.h:106-110 */
    s->noise_amp = sqrt(2 * p->D_rot * p->dt);
    s->r_cut = max_cutoff(p);
    s->half_skin = 0.5 * (p->r_list - s->r_cut);
    s->ncx = p->ncell_x > 0 ? p->ncell_x : (int32_t)floor(s->Lx2 / p->r_list);
    s->ncy = p->ncell_y > 0 ? p->ncell_y : (int32_t)floor(s->Ly2 / p->r_list);
    /* Cell::Cell, shared/cell.h:45-49: cells must not be narrower than rv; three or more cells per
       side keep every neighbour cell distinct (two would double-count pairs).                    */
    if (s->ncx < 3 || s->ncy < 3 || s->Lx2 / s->ncx < p->r_list || s->Ly2 / s->ncy < p->r_list) {
        set_err("orc_create: cell grid %dx%d invalid for r_list %g", s->ncx, s->ncy, p->r_list);
        free(s);
        return NULL;
    }
    if (p->r_list < s->r_cut) {
        set_err("orc_create: r_list below interaction cutoff");
        free(s);
        return NULL;
    }
    s->P = (particle*)calloc((size_t)s->N, sizeof(particle));
    s->C = (cell*)calloc((size_t)s->ncx * s->ncy, sizeof(cell));
    s->build_r = (vec2*)calloc((size_t)s->N, sizeof(vec2));
    s->injected = (double*)calloc((size_t)s->N, sizeof(double));
    return s;
}

void orc_destroy(orc_sys* s)
{
    if (!s) return;
    for (int32_t i = 0; i < s->N; ++i) free(s->P[i].nbr);
    for (int64_t c = 0; c < (int64_t)s->ncx * s->ncy; ++c) free(s->C[c].pid);
    free(s->P);
    free(s->C);
    free(s->build_r);
    free(s->injected);
    free(s);
}

/* Reset: RepulsiveParticle::Reset (particle.h:559-564), f = v (times v0), torque = 0;
 * ActiveBrownianChain::Reset (:878-883), torque = torque0, f = F0 v with F0 = 0 for ring beads
 * (membrane.cpp:144-149).                                                                         */
static void reset(orc_sys* s, int32_t i)
{
    particle* q = &s->P[i];
    if (s->p.model == SPP_MODEL_ABP) {
        double F0 = i < s->p.n_membrane ? 0.0 : s->p.v0;
        q->torque = s->p.torque0;
        q->f.x = q->v.x * F0;
        q->f.y = q->v.y * F0;
        return;
    }
    q->torque = 0;
    q->f.x = q->v.x * s->p.v0;
    q->f.y = q->v.y * s->p.v0;
}

int orc_set_state(orc_sys* s, const double* xw, const double* yw, const double* xu,
                  const double* yu, const double* theta, int32_t n)
{
    if (n != s->N) {
        set_err("orc_set_state: n mismatch");
        return SPP_ERR_ARG;
    }
    /* Box::Positioning_Particles tail, serial/box.h:251-258 */
    for (int32_t i = 0; i < n; ++i) {
        particle* q = &s->P[i];
        q->r.x = xw[i];
        q->r.y = yw[i];
        q->r_original.x = xu ? xu[i] : xw[i];
        q->r_original.y = yu ? yu[i] : yw[i];
        q->theta = theta[i];
        q->v.x = cos(q->theta);
        q->v.y = sin(q->theta);
        reset(s, i);
    }
    s->has_state = 1;
    return orc_update_cells(s);
}

int orc_get_state(orc_sys* s, double* xw, double* yw, double* xu, double* yu, double* theta,
                  int32_t n)
{
    if (n != s->N) return SPP_ERR_ARG;
    for (int32_t i = 0; i < n; ++i) {
        if (xw) xw[i] = s->P[i].r.x;
        if (yw) yw[i] = s->P[i].r.y;
        if (xu) xu[i] = s->P[i].r_original.x;
        if (yu) yu[i] = s->P[i].r_original.y;
        if (theta) theta[i] = s->P[i].theta;
    }
    return SPP_OK;
}

static int cell_of(const orc_sys* s, const vec2* r, int32_t* cx, int32_t* cy)
{
    if (s->p.binning == SPP_BIN_REF_QUIRK) {
        *cx = orc_cell_quirk(r->x, s->p.Lx, s->ncx);
        *cy = orc_cell_quirk(r->y, s->p.Ly, s->ncy);
    } else {
        *cx = orc_cell_exact(r->x, s->p.Lx, s->ncx);
        *cy = orc_cell_exact(r->y, s->p.Ly, s->ncy);
    }
    return 0;
}

/* Cell::Neighbor_List(Cell*), shared/cell.h:74-89: pairs across two cells, min-image, d < rv.   */
static void nlist_cross(orc_sys* s, const cell* a, const cell* b)
{
    double rv = s->p.r_list;
    for (int32_t i = 0; i < a->n; ++i) {
        particle* pi = &s->P[a->pid[i]];
        for (int32_t j = 0; j < b->n; ++j) {
            const particle* pj = &s->P[b->pid[j]];
            double dx = orc_periodic(pi->r.x - pj->r.x, s->Lx2);
            double dy = orc_periodic(pi->r.y - pj->r.y, s->Ly2);
            double d = sqrt(dx * dx + dy * dy);
            if (d < rv) push(&pi->nbr, &pi->nn, &pi->ncap, b->pid[j]);
        }
    }
}

/* Cell::Neighbor_List(), shared/cell.h:91-103: pairs inside one cell, no wrap, d < rv.          */
static void nlist_self(orc_sys* s, const cell* a)
{
    double rv = s->p.r_list;
    for (int32_t i = 0; i < a->n; ++i) {
        particle* pi = &s->P[a->pid[i]];
        for (int32_t j = i + 1; j < a->n; ++j) {
            const particle* pj = &s->P[a->pid[j]];
            double dx = pi->r.x - pj->r.x, dy = pi->r.y - pj->r.y;
            double d = sqrt(dx * dx + dy * dy);
            if (d < rv) push(&pi->nbr, &pi->nn, &pi->ncap, a->pid[j]);
        }
    }
}

/* Box::Update_Cells + Update_Neighbor_List, serial/box.h:107-179 (periodic branch).            */
int orc_update_cells(orc_sys* s)
{
    int32_t X = s->ncx, Y = s->ncy;
    for (int64_t c = 0; c < (int64_t)X * Y; ++c) s->C[c].n = 0; /* Cell::Delete */
    for (int32_t i = 0; i < s->N; ++i) {
        int32_t cx, cy;
        cell_of(s, &s->P[i].r, &cx, &cy);
        cell* c = &s->C[(int64_t)cx * Y + cy];
        push(&c->pid, &c->n, &c->cap, i); /* Cell::Add */
    }
    for (int32_t x = 0; x < X; ++x)
        for (int32_t y = 0; y < Y; ++y) {
            cell* c = &s->C[(int64_t)x * Y + y];
            for (int32_t i = 0; i < c->n; ++i) s->P[c->pid[i]].nn = 0; /* Clear_Neighbor_List */
            nlist_self(s, c);
        }
    for (int32_t x = 0; x < X; ++x)
        for (int32_t y = 0; y < Y; ++y) {
            cell* c = &s->C[(int64_t)x * Y + y];
            nlist_cross(s, c, &s->C[(int64_t)((x + 1) % X) * Y + y]);
            nlist_cross(s, c, &s->C[(int64_t)x * Y + (y + 1) % Y]);
            nlist_cross(s, c, &s->C[(int64_t)((x + 1) % X) * Y + (y + 1) % Y]);
            nlist_cross(s, c, &s->C[(int64_t)((x + 1) % X) * Y + (y - 1 + Y) % Y]);
        }
    for (int32_t i = 0; i < s->N; ++i) s->build_r[i] = s->P[i].r;
    s->steps_since_build = 0;
    s->rebuilds++;
    return SPP_OK;
}

int orc_get_cells(orc_sys* s, int32_t* cx, int32_t* cy, int32_t n)
{
    if (n != s->N) return SPP_ERR_ARG;
    for (int32_t i = 0; i < n; ++i) cell_of(s, &s->build_r[i], &cx[i], &cy[i]);
    return SPP_OK;
}

int64_t orc_half_list(orc_sys* s, int32_t* owner, int32_t* other, int64_t cap)
{
    int64_t m = 0;
    for (int32_t i = 0; i < s->N; ++i)
        for (int32_t k = 0; k < s->P[i].nn; ++k) {
            if (m < cap) {
                owner[m] = i;
                other[m] = s->P[i].nbr[k];
            }
            ++m;
        }
    return m;
}

/* Beads of particle i: ring beads (i < n_membrane) are single; swimmers carry nb_swimmer beads. */
static int nbeads(const orc_sys* s, int32_t i)
{
    return (s->p.nb_swimmer >= 2 && i >= s->p.n_membrane) ? s->p.nb_swimmer : 1;
}

/* ActiveBrownianChain::Set_nb (particle.h:856-876): mobilities of a chain of nb beads. */
static void chain_mobilities(int nb, double sep, double sigma, double* mpar, double* mperp, double* kperp)
{
    if (nb == 1) {
        *mpar = 0.5;
        *mperp = 0.5;
        *kperp = 1;
        return;
    }
    double aspect_ratio = 1 / ((sep * (nb - 1) / sigma) + 1);
    double e2 = 1 - aspect_ratio * aspect_ratio;
    double e = sqrt(e2);
    double L = log((1 + e) / (1 - e));
    *mpar = 1.0;
    *mperp = 0.5 * (2 * e + (3 * e2 - 1) * L) / ((1 + e2) * L - 2 * e);
    *kperp = 6 / (1 + aspect_ratio * aspect_ratio) / sigma / sigma;
}

void orc_chain_mobilities(int32_t nb, double sep, double sigma, double* mpar, double* mperp, double* kperp)
{
    chain_mobilities(nb, sep, sigma, mpar, mperp, kperp);
}

/* Pair interaction.  SPP_MODEL_REPULSIVE follows RepulsiveParticle::Interact
 * (particle.h:728-757); SPP_MODEL_CONTINUOUS follows ContinuousParticle::Interact
 * (particle.h:296-338).  Newton-3: both particles are updated.                                    */
static void pair(orc_sys* s, particle* a, particle* b)
{
    const spp_params* p = &s->p;
    vec2 dr = {a->r.x - b->r.x, a->r.y - b->r.y};
    dr.x = orc_periodic(dr.x, s->Lx2);
    dr.y = orc_periodic(dr.y, s->Ly2);
    double d2 = dr.x * dr.x + dr.y * dr.y;
    if (p->model == SPP_MODEL_ABP) {
        const int na = nbeads(s, (int32_t)(a - s->P)), nbb = nbeads(s, (int32_t)(b - s->P));
        if (na == 1 && nbb == 1) {
            /* ActiveBrownianChain::Interact, nb == 1 on both sides (particle.h:968-981) */
            double d = sqrt(d2);
            if (p->repulsion != SPP_REP_NONE && d < p->rc_rep) {
                double fx, fy;
                orc_force_law(p->repulsion, dr.x, dr.y, d, p->rc_rep, p->sigma_rep, p->A_rep, &fx, &fy);
                a->f.x += fx;
                a->f.y += fy;
                b->f.x -= fx;
                b->f.y -= fy;
            }
            return;
        }
        /* bead pairs (particle.h:983-1011): s = v separation ((1 - nb)/2 + i), dr = dr0 + s_this -
           s_that, R12_Repulsive_Truncated within repulsion_radius, torques s x F */
        const double sep = p->bead_sep;
        for (int i = 0; i < na; i++)
            for (int j = 0; j < nbb; j++) {
                vec2 st = {a->v.x * (sep * ((1 - na) / 2.0 + i)), a->v.y * (sep * ((1 - na) / 2.0 + i))};
                vec2 su = {b->v.x * (sep * ((1 - nbb) / 2.0 + j)), b->v.y * (sep * ((1 - nbb) / 2.0 + j))};
                vec2 db = {dr.x + st.x - su.x, dr.y + st.y - su.y};
                double d = sqrt(db.x * db.x + db.y * db.y);
                if (d < p->rc_bead) {
                    double fx, fy;
                    orc_force_law(SPP_REP_R12_TRUNCATED, db.x, db.y, d, p->rc_bead, p->sigma_bead, p->A_bead, &fx, &fy);
                    a->f.x += fx;
                    a->f.y += fy;
                    b->f.x -= fx;
                    b->f.y -= fy;
                    a->torque += st.x * fy - st.y * fx;
                    b->torque += -su.x * fy + su.y * fx;
                }
            }
        return;
    }
    if (p->model == SPP_MODEL_REPULSIVE) {
        double d = sqrt(d2);
        if (p->repulsion != SPP_REP_NONE && d < p->rc_rep) {
            double fx, fy;
            orc_force_law(p->repulsion, dr.x, dr.y, d, p->rc_rep, p->sigma_rep, p->A_rep, &fx, &fy);
            a->f.x += fx;
            a->f.y += fy;
            b->f.x -= fx;
            b->f.y -= fy;
        }
        if (d < p->R_align) {
            double ti = p->g * sin(b->theta - a->theta) / (M_PI * p->R_align * p->R_align);
            a->torque += ti;
            b->torque -= ti;
        }
    } else {
        if (d2 < p->R_align * p->R_align) {
            double d = sqrt(d2);
            double ti = (1 - p->alpha) * sin(b->theta - a->theta) / (M_PI);
            a->torque += ti;
            b->torque -= ti;
            if (p->alpha != 0) {
                a->torque -= p->alpha * (dr.x * a->v.y - dr.y * a->v.x) / (M_PI * d);
                b->torque += p->alpha * (dr.x * b->v.y - dr.y * b->v.x) / (M_PI * d);
            }
        }
    }
}

/* Minimum-image bond vector r_b - r_a. */
static vec2 bond(const orc_sys* s, int32_t a, int32_t b)
{
    vec2 t = {s->P[b].r.x - s->P[a].r.x, s->P[b].r.y - s->P[a].r.y};
    t.x = orc_periodic(t.x, s->Lx2);
    t.y = orc_periodic(t.y, s->Ly2);
    return t;
}

/* Ring forces.  Bonds: Box::Interact_Membrane_Beads (serial/beadbox.h:286-297), Spring(dr, d, l, k)
 * = k (l - d) dr / d with dr = r_i - r_{i+1} (force-fields.h:87-97), added to bead i and
 * subtracted from bead i+1.  Bending: E = kappa * sum_i (1 - cos phi_i) with cos phi_i = a.b/(|a||b|),
 * a = r_i - r_{i-1}, b = r_{i+1} - r_i; force = -grad E = kappa * grad cos phi, so angle i adds
 * kappa * (-dc/da) to bead i-1, kappa * (dc/da - dc/db) to bead i and kappa * dc/db to bead i+1.   */
static void ring_forces(orc_sys* s)
{
    const int32_t M = s->p.n_membrane;
    const double K = s->p.k_bond, l0 = s->p.l_bond, kappa = s->p.kappa_bend;
    for (int32_t i = 0; i < M; ++i) {
        int32_t j = (i + 1) % M;
        vec2 t = bond(s, i, j);
        vec2 dr = {-t.x, -t.y};
        double d = sqrt(dr.x * dr.x + dr.y * dr.y);
        double fx = K * (l0 - d) * dr.x / d, fy = K * (l0 - d) * dr.y / d;
        s->P[i].f.x += fx;
        s->P[i].f.y += fy;
        s->P[j].f.x -= fx;
        s->P[j].f.y -= fy;
    }
    if (kappa == 0) return;
    for (int32_t i = 0; i < M; ++i) {
        int32_t im = (i + M - 1) % M, ip = (i + 1) % M;
        vec2 a = bond(s, im, i), b = bond(s, i, ip);
        double aa = a.x * a.x + a.y * a.y, bb = b.x * b.x + b.y * b.y, ab = a.x * b.x + a.y * b.y;
        double inv = 1.0 / sqrt(aa * bb), c = ab * inv;
        vec2 ga = {b.x * inv - c * a.x / aa, b.y * inv - c * a.y / aa};
        vec2 gb = {a.x * inv - c * b.x / bb, a.y * inv - c * b.y / bb};
        s->P[im].f.x -= kappa * ga.x;
        s->P[im].f.y -= kappa * ga.y;
        s->P[i].f.x += kappa * (ga.x - gb.x);
        s->P[i].f.y += kappa * (ga.y - gb.y);
        s->P[ip].f.x += kappa * gb.x;
        s->P[ip].f.y += kappa * gb.y;
    }
}

/* Ring energy: sum over bonds of k/2 (d - l)^2 plus the bending energy above. */
double orc_ring_energy(orc_sys* s)
{
    const int32_t M = s->p.n_membrane;
    double E = 0;
    for (int32_t i = 0; i < M; ++i) {
        vec2 t = bond(s, i, (i + 1) % M);
        double d = sqrt(t.x * t.x + t.y * t.y);
        E += 0.5 * s->p.k_bond * (d - s->p.l_bond) * (d - s->p.l_bond);
    }
    for (int32_t i = 0; i < M && s->p.kappa_bend != 0; ++i) {
        vec2 a = bond(s, (i + M - 1) % M, i), b = bond(s, i, (i + 1) % M);
        double c = (a.x * b.x + a.y * b.y) / sqrt((a.x * a.x + a.y * a.y) * (b.x * b.x + b.y * b.y));
        E += s->p.kappa_bend * (1 - c);
    }
    return E;
}

/* Box::Interact with the Verlet list, serial/box.h:324-327 -> Cell::Interact, cell.h:105-110.   */
int orc_interact(orc_sys* s)
{
    if (!s->has_state) {
        set_err("orc_interact: no state");
        return SPP_ERR_STATE;
    }
    for (int32_t i = 0; i < s->N; ++i) reset(s, i);
    if (s->p.model == SPP_MODEL_ABP && s->p.n_membrane > 0) ring_forces(s);
    int32_t X = s->ncx, Y = s->ncy;
    for (int32_t x = 0; x < X; ++x)
        for (int32_t y = 0; y < Y; ++y) {
            const cell* c = &s->C[(int64_t)x * Y + y];
            for (int32_t i = 0; i < c->n; ++i) {
                particle* a = &s->P[c->pid[i]];
                for (int32_t k = 0; k < a->nn; ++k) pair(s, a, &s->P[a->nbr[k]]);
            }
        }
    return SPP_OK;
}

/* Torque as integrated: ContinuousParticle::Move applies g to the summed torque (particle.h:273). */
static double drift_torque(const orc_sys* s, const particle* q)
{
    return s->p.model == SPP_MODEL_CONTINUOUS ? s->p.g * q->torque : q->torque;
}

int orc_get_forces(orc_sys* s, double* fx, double* fy, double* tau, int32_t n)
{
    if (n != s->N) return SPP_ERR_ARG;
    for (int32_t i = 0; i < n; ++i) {
        if (fx) fx[i] = s->P[i].f.x;
        if (fy) fy[i] = s->P[i].f.y;
        if (tau) tau[i] = drift_torque(s, &s->P[i]);
    }
    return SPP_OK;
}

int orc_set_noise(orc_sys* s, const double* dtheta, int32_t n)
{
    if (n != s->N) return SPP_ERR_ARG;
    memcpy(s->injected, dtheta, (size_t)n * sizeof(double));
    s->has_injected = 1;
    return SPP_OK;
}

/* RepulsiveParticle::Noise_Gen, particle.h:566-569, with the project stream (or injected).       */
static int noise_gen(orc_sys* s)
{
    if (s->p.noise_mode == SPP_NOISE_INJECTED) {
        if (!s->has_injected) {
            set_err("noise: injected mode but no noise was injected for this step");
            return SPP_ERR_STATE;
        }
        for (int32_t i = 0; i < s->N; ++i) s->P[i].dtheta = s->injected[i];
        s->has_injected = 0;
    } else {
        for (int32_t i = 0; i < s->N; ++i)
            s->P[i].dtheta = s->noise_amp * orc_noise_normal(s->p.seed, i, s->step);
    }
    return SPP_OK;
}

static void wrap_r(const orc_sys* s, vec2* r)
{
    r->x = orc_periodic(r->x, s->Lx2);
    r->y = orc_periodic(r->y, s->Ly2);
}

static void set_v(particle* q)
{
    q->v.x = cos(q->theta);
    q->v.y = sin(q->theta);
}

/* ActiveBrownianChain::Move (particle.h:898-914, heading updated first, projection on the new
 * heading), Move_Runge_Kutta2_1 (:917-940, projection on the old heading) and _2 (:942-962,
 * projection on the midpoint heading).  k_perpendicular = mob_rot, m_parallel = mob_par,
 * m_perpendicular = mob_perp.                                                                       */
static void abp_move(orc_sys* s, int32_t stage)
{
    const double dt = s->p.dt, half_dt = s->half_dt;
    /* ring beads (and single-bead swimmers) move with the configured mobilities, chains of
       nb_swimmer beads with Set_nb's */
    double c_par, c_perp, c_rot;
    chain_mobilities(s->p.nb_swimmer >= 2 ? s->p.nb_swimmer : 1, s->p.bead_sep, s->p.sigma_bead, &c_par, &c_perp,
                     &c_rot);
    for (int32_t i = 0; i < s->N; ++i) {
        particle* q = &s->P[i];
        const int chain = nbeads(s, i) > 1;
        const double kp = chain ? c_rot : s->p.mob_rot, mpar = chain ? c_par : s->p.mob_par,
                     mperp = chain ? c_perp : s->p.mob_perp;
        double tq = q->torque;
        if (stage == 0) {
            q->theta += kp * dt * (tq);
            q->theta += q->dtheta;
            set_v(q);
            q->r_original.x += q->f.x * (dt * mperp);
            q->r_original.y += q->f.y * (dt * mperp);
            double fv = q->f.x * q->v.x + q->f.y * q->v.y;
            q->r_original.x += q->v.x * (fv * (dt * (mpar - mperp)));
            q->r_original.y += q->v.y * (fv * (dt * (mpar - mperp)));
            q->r = q->r_original;
            wrap_r(s, &q->r);
        } else if (stage == 1) {
            q->theta_old = q->theta;
            q->r_old = q->r_original;
            q->theta += kp * half_dt * (tq);
            q->theta += q->dtheta / 2;
            q->theta = wrap_angle(q->theta);
            q->r.x += q->f.x * (half_dt * mperp);
            q->r.y += q->f.y * (half_dt * mperp);
            double fv = q->f.x * q->v.x + q->f.y * q->v.y;
            q->r.x += q->v.x * (fv * (half_dt * (mpar - mperp)));
            q->r.y += q->v.y * (fv * (half_dt * (mpar - mperp)));
            set_v(q);
            wrap_r(s, &q->r);
        } else {
            q->theta = q->theta_old + kp * dt * (tq);
            q->theta += q->dtheta;
            q->theta = wrap_angle(q->theta);
            q->r_original.x = q->r_old.x + q->f.x * (dt * mperp);
            q->r_original.y = q->r_old.y + q->f.y * (dt * mperp);
            double fv = q->f.x * q->v.x + q->f.y * q->v.y;
            q->r_original.x += q->v.x * (fv * (dt * (mpar - mperp)));
            q->r_original.y += q->v.y * (fv * (dt * (mpar - mperp)));
            q->r = q->r_original;
            set_v(q);
            wrap_r(s, &q->r);
        }
        reset(s, i);
    }
}

/* Move stages.  stage 0: RepulsiveParticle::Move (particle.h:571-587); RK2 stages 1-2:
 * Move_Runge_Kutta2_1/2 (:589-631); RK4 stages 1-4: Move_Runge_Kutta4_1..4 (:634-725).           */
int orc_move_stage(orc_sys* s, int32_t stage)
{
    int integ = s->p.integrator;
    int last;
    if (integ == SPP_EULER) {
        if (stage != 0) return SPP_ERR_ARG;
        last = 1;
    } else if (integ == SPP_RK2) {
        if (stage < 1 || stage > 2) return SPP_ERR_ARG;
        last = stage == 2;
    } else if (integ == SPP_RK4) {
        if (stage < 1 || stage > 4) return SPP_ERR_ARG;
        last = stage == 4;
    } else
        return SPP_ERR_ARG;
    if (stage <= 1) {
        int rc = noise_gen(s);
        if (rc) return rc;
    }
    double dt = s->p.dt, half_dt = s->half_dt, dt6 = s->dt_over_6;
    if (s->p.model == SPP_MODEL_ABP) {
        abp_move(s, stage);
        if (last) {
            s->step++;
            s->t += dt;
            s->steps_since_build++;
        }
        return SPP_OK;
    }
    for (int32_t i = 0; i < s->N; ++i) {
        particle* q = &s->P[i];
        double tq = drift_torque(s, q);
        if (integ == SPP_EULER) {
            q->theta += dt * (tq);
            q->theta += q->dtheta;
            set_v(q);
            q->r_original.x += q->f.x * dt;
            q->r_original.y += q->f.y * dt;
            q->r = q->r_original;
            wrap_r(s, &q->r);
        } else if (stage == 1) { /* RK2_1 and RK4_1 share the update */
            q->theta_old = q->theta;
            q->r_old = q->r_original;
            q->theta += half_dt * tq;
            q->theta += q->dtheta / 2;
            q->theta = wrap_angle(q->theta);
            set_v(q);
            q->r.x += q->f.x * half_dt;
            q->r.y += q->f.y * half_dt;
            wrap_r(s, &q->r);
            if (integ == SPP_RK4) {
                q->k1_f = q->f;
                q->k1_t = tq;
            }
        } else if (integ == SPP_RK2) { /* stage 2 */
            q->theta = q->theta_old + dt * tq;
            q->theta += q->dtheta;
            q->theta = wrap_angle(q->theta);
            set_v(q);
            q->r_original.x = q->r_old.x + q->f.x * dt;
            q->r_original.y = q->r_old.y + q->f.y * dt;
            q->r = q->r_original;
            wrap_r(s, &q->r);
        } else if (stage == 2) { /* RK4_2 */
            q->theta = q->theta_old + half_dt * tq;
            q->theta += q->dtheta / 2;
            q->theta = wrap_angle(q->theta);
            set_v(q);
            q->r_original.x = q->r_old.x + q->f.x * half_dt;
            q->r_original.y = q->r_old.y + q->f.y * half_dt;
            q->r = q->r_original;
            wrap_r(s, &q->r);
            q->k2_f = q->f;
            q->k2_t = tq;
        } else if (stage == 3) { /* RK4_3 */
            q->theta = q->theta_old + dt * tq;
            q->theta += q->dtheta;
            q->theta = wrap_angle(q->theta);
            set_v(q);
            q->r_original.x = q->r_old.x + q->f.x * dt;
            q->r_original.y = q->r_old.y + q->f.y * dt;
            q->r = q->r_original;
            wrap_r(s, &q->r);
            q->k3_f = q->f;
            q->k3_t = tq;
        } else { /* RK4_4 */
            q->theta = q->theta_old + (q->k1_t + q->k2_t * 2 + q->k3_t * 2 + tq) * dt6;
            q->theta += q->dtheta;
            q->theta = wrap_angle(q->theta);
            set_v(q);
            q->r_original.x = q->r_old.x + (q->k1_f.x + q->k2_f.x * 2 + q->k3_f.x * 2 + q->f.x) * dt6;
            q->r_original.y = q->r_old.y + (q->k1_f.y + q->k2_f.y * 2 + q->k3_f.y * 2 + q->f.y) * dt6;
            q->r = q->r_original;
            wrap_r(s, &q->r);
        }
        reset(s, i);
    }
    if (last) {
        s->step++;
        s->t += dt;
        s->steps_since_build++;
    }
    return SPP_OK;
}

/* Box::One_Step, serial/box.h:427-455 */
int orc_one_step(orc_sys* s)
{
    int integ = s->p.integrator;
    int nst = integ == SPP_EULER ? 1 : integ;
    for (int st = 1; st <= nst; ++st) {
        int rc = orc_interact(s);
        if (rc) return rc;
        rc = orc_move_stage(s, integ == SPP_EULER ? 0 : st);
        if (rc) return rc;
    }
    return SPP_OK;
}

static double max_displacement(const orc_sys* s)
{
    double m = 0;
    for (int32_t i = 0; i < s->N; ++i) {
        double dx = orc_periodic(s->P[i].r.x - s->build_r[i].x, s->Lx2);
        double dy = orc_periodic(s->P[i].r.y - s->build_r[i].y, s->Ly2);
        double d = sqrt(dx * dx + dy * dy);
        if (d > m) m = d;
    }
    return m;
}

/* Box::Multi_Step, serial/box.h:469-506 (steps, then Update_Cells + Update_Neighbor_List).
 * SPP_REBUILD_PERIOD rebuilds every rebuild_period steps exactly like the reference.
 * SPP_REBUILD_DISPLACEMENT rebuilds at a step boundary once a particle has moved half of the
 * skin's half since the last build (a conservative trigger: any valid list gives the same
 * forces), and fails loudly if a stage state ever moves farther than the half skin, because the
 * list could then miss pairs.                                                                    */
int orc_multi_step(orc_sys* s, int32_t steps)
{
    for (int32_t k = 0; k < steps; ++k) {
        if (s->p.rebuild == SPP_REBUILD_DISPLACEMENT) {
            if (max_displacement(s) > 0.5 * s->half_skin) orc_update_cells(s);
            int integ = s->p.integrator;
            int nst = integ == SPP_EULER ? 1 : integ;
            for (int st = 1; st <= nst; ++st) {
                int rc = orc_interact(s);
                if (rc) return rc;
                rc = orc_move_stage(s, integ == SPP_EULER ? 0 : st);
                if (rc) return rc;
                if (max_displacement(s) > s->half_skin) {
                    set_err("orc_multi_step: a particle moved beyond the half skin within a step");
                    return SPP_ERR_STATE;
                }
            }
        } else {
            int rc = orc_one_step(s);
            if (rc) return rc;
            if (s->steps_since_build >= s->p.rebuild_period) orc_update_cells(s);
        }
    }
    return SPP_OK;
}

/* Box::Compute_Polarization, serial/box.h:531-538 */
double orc_polarization(orc_sys* s, double* sx, double* sy)
{
    vec2 v = {0, 0};
    for (int32_t i = 0; i < s->N; ++i) {
        v.x += s->P[i].v.x;
        v.y += s->P[i].v.y;
    }
    if (sx) *sx = v.x;
    if (sy) *sy = v.y;
    v.x /= s->N;
    v.y /= s->N;
    return sqrt(v.x * v.x + v.y * v.y);
}

int64_t orc_step_count(orc_sys* s) { return s->step; }
int64_t orc_rebuild_count(orc_sys* s) { return s->rebuilds; }
