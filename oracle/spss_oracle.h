/*
 * spss_oracle.h -- CPU oracle for the right-hand-side hot path of the
 * self-propelled-particle / Kuramoto-coupling simulator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference algorithm, used as the checker by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product path
 * (the HIP library) never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  State layout is
 * the reference's: array-of-structures FP64, s values per particle,
 *   s = 4: (x, y, phi, omega)          DynamicalSystems/ParticleSystemFirstOrderPtr.cpp
 *   s = 6: (x, y, vx, vy, phi, omega)  DynamicalSystems/ParticleSystemSecondOrderPtr.cpp
 *   s = 3: (x, y, phi)                 first-order (m = 0) Kuramoto, no reference counterpart
 *
 * Built with -O2 -ffp-contract=off so that every expression is evaluated as
 * the separate IEEE operations written in the reference, in its order.
 */
#ifndef SPSS_ORACLE_H
#define SPSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* neighbour strategies (same numbering as the product C ABI) */
#define ORC_METHOD_AUTO 0
#define ORC_METHOD_GLOBAL_FIELD 1
#define ORC_METHOD_ALL_PAIRS 2
#define ORC_METHOD_LINKED_LIST 3
#define ORC_METHOD_VERLET 4
#define ORC_METHOD_LINKED_LIST_VERLET 5

/* Verlet semantics */
#define ORC_VERLET_COMPAT 0 /* r_max = rho + 0.05 fixed, trigger 0.5*(r_max-r_min) */
#define ORC_VERLET_EXACT 1  /* r_max = rho + skin, candidates from cells >= r_max */

typedef struct orc_system orc_system;

typedef struct orc_params {
  int model;      /* 3, 4 or 6 (= state width s) */
  int n;          /* particles */
  double v0;
  double xi;      /* s=4: phase friction xi; s=6: xi_phi; s=3: unused */
  double sigma;
  double rho;
  double alpha;
  double D_phi;
  double xi_r;    /* s=6 translational friction */
  double D_r;     /* s=6 translational noise */
  double x_size, y_size; /* box; the reference fixes 1.0 for the cell grid */
  double verlet_skin;    /* exact mode: r_max - rho; compat mode ignores it (0.05) */
  int verlet_mode;
  int natfreq_local;     /* 1: add natural frequencies in local branches too */
} orc_params;

orc_system *orc_create(const orc_params *p);
void orc_destroy(orc_system *sys);
void orc_set_natural_frequencies(orc_system *sys, const double *w);
void orc_set_higher_order(orc_system *sys, int on);
/* worker threads of the linked-list particle loop (contiguous particle blocks, as the reference's
 * ranks own them, Parallelization/ThreadSharedMemory.cpp:20-23); results equal the serial ones */
void orc_set_threads(orc_system *sys, int threads);
int orc_uses_verlet(const orc_system *sys);
int orc_should_update_lists(const orc_system *sys);
double orc_accumulated_displacement(const orc_system *sys);
int orc_num_cells_x(const orc_system *sys);
const double *orc_complementary_force(const orc_system *sys);

/* ---- periodic boundary conditions (BoundaryConfiguration/PeriodicBoundaryConditions.cpp) */
void orc_pbc_class_a(double xs, double ys, double xi, double yi, double xj, double yj, double *dx, double *dy);
void orc_pbc_class_b_signed(double xs, double ys, double xi, double yi, double xj, double yj, double *dx, double *dy);
void orc_pbc_class_b_unsigned(double xs, double ys, double xi, double yi, double xj, double yj, double *dx, double *dy);
void orc_pbc_class_c_signed(double xs, double ys, double xi, double yi, double xj, double yj, double *dx, double *dy);
void orc_pbc_class_c_unsigned(double xs, double ys, double xi, double yi, double xj, double yj, double *dx, double *dy);
void orc_pbc_apply(double xs, double ys, double x, double y, double *xp, double *yp);
void orc_pbc_apply_state(double xs, double ys, double *state, int n, int s);

/* ---- right-hand side (DynamicalSystems directory) */
void orc_evaluate_rhs(orc_system *sys, int method, const double *state, const double *k_prev,
                      double *k_next, double k_coef, double dt);
void orc_calculate_max_displacement(orc_system *sys, const double *total_displacement);

/* ---- steppers (Steppers directory) */
void orc_rk4_step(orc_system *sys, int method, double *state, double *k1, double *k2, double *k3,
                  double *k4, double *total_disp, double dt);
/* s=4: one normal per particle; s=6: three per particle (slots 2,3,5); s=3: one. */
void orc_euler_step(orc_system *sys, int method, double *state, double *derivative, double *total_disp,
                    double dt, const double *normals);
/* s=4 only: two normals per particle (U1, U2) interleaved. */
void orc_srk_step(orc_system *sys, int method, double *state, double *derivative, double *total_disp,
                  double dt, const double *normals);

/* ---- neighbour sets, exported as (i, j) pairs sorted by (i, j) */
long orc_pairs_bruteforce(const double *stage, int n, int s, double rho, double xs, double ys,
                          int32_t *out, long cap);
long orc_pairs_linked_list(orc_system *sys, const double *stage, int32_t *out, long cap);
/* current Verlet lists (unsorted, in list order); returns number of entries */
long orc_verlet_lists(orc_system *sys, int32_t *out_i, int32_t *out_j, long cap);

/* ---- observers (Observers directory) */
void orc_order_parameter(const double *state, int n, int s, int slot, double *re, double *im);
void orc_order_parameter_ref_float(const double *state, int n, int s, int slot, float *re, float *im);
void orc_averaging_accumulate(double *acc, const double *state, long size);

/* ---- first-order Kuramoto (s=3) and global field helpers */
void orc_polar_fields(const double *stage, int n, int s, int slot_phi, int slot_omega, double *zre,
                      double *zim, double *wre, double *wim);

#ifdef __cplusplus
}
#endif
#endif
