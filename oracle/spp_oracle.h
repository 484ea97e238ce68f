/*
 * spp_oracle.h -- CPU oracle for the per-time-step particle path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker or the timed CPU baseline -- never as the
 * product path.  It restates the reference algorithm (hamidallaei/SPP) in plain C, keeping the
 * reference's data structures (AoS particles, per-cell id lists, Newton-3 half Verlet lists) so
 * that it is also a faithful speed proxy of the reference's CPU path.
 *
 * Parameters come in the product's spp_params struct (include/spp_engine.h) so that both sides
 * are driven by the same numbers.
 */
#ifndef SPP_ORACLE_H
#define SPP_ORACLE_H

#include <stdint.h>
#include "../include/spp_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

const char* orc_error(void);

/* --- pure functions ------------------------------------------------------------------------- */
double orc_periodic(double x, double L2);
int orc_cell_quirk(double x, double L, int divisor);
int orc_cell_exact(double x, double L, int divisor);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_noise_normal(uint64_t seed, int32_t id, int64_t step);
void orc_force_law(int kind, double dx, double dy, double d, double rc, double sigma, double A,
                   double* fx, double* fy);
int64_t orc_brute_pairs(const double* x, const double* y, int32_t n, double Lx, double Ly,
                        double r, int32_t* pi, int32_t* pj, int64_t cap);
void orc_random_state(uint64_t seed, int32_t n, double Lx, double Ly, double* x, double* y,
                      double* theta);

/* --- simulation box ----------------------------------------------------------------------------- */
orc_sys* orc_create(const spp_params* p);
void orc_destroy(orc_sys* s);
int orc_set_state(orc_sys* s, const double* xw, const double* yw, const double* xu,
                  const double* yu, const double* theta, int32_t n);
int orc_get_state(orc_sys* s, double* xw, double* yw, double* xu, double* yu, double* theta,
                  int32_t n);
int orc_update_cells(orc_sys* s);
int orc_get_cells(orc_sys* s, int32_t* cx, int32_t* cy, int32_t n);
int64_t orc_half_list(orc_sys* s, int32_t* owner, int32_t* other, int64_t cap);
int orc_interact(orc_sys* s);
int orc_get_forces(orc_sys* s, double* fx, double* fy, double* tau, int32_t n);
int orc_set_noise(orc_sys* s, const double* dtheta, int32_t n);
int orc_move_stage(orc_sys* s, int32_t stage);
int orc_one_step(orc_sys* s);
int orc_multi_step(orc_sys* s, int32_t steps);
double orc_polarization(orc_sys* s, double* sx, double* sy);
int64_t orc_step_count(orc_sys* s);
int64_t orc_rebuild_count(orc_sys* s);
double orc_ring_energy(orc_sys* s);
/* ActiveBrownianChain::Set_nb mobilities (m_parallel, m_perpendicular, k_perpendicular) */
void orc_chain_mobilities(int32_t nb, double sep, double sigma, double* mpar, double* mperp, double* kperp);

#ifdef __cplusplus
}
#endif

#endif
