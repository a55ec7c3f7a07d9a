/*
 * spss_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the RHS hot path.
 *
 * A plain-C FP64 restatement of the reference algorithm.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker; the HIP product never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Where the
 * reference has a serial twin and an MPI twin with slightly different code,
 * the MPI twin (the class family the reference's default program runs,
 * main.cpp:27-31) is followed and the difference is noted.
 *
 * Compile with -ffp-contract=off: the expressions below are written as the
 * reference writes them, and each must round as a separate IEEE operation.
 */
#include "spss_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

struct ivec {
  int *v;
  int n, cap;
};

static void ivec_push(struct ivec *a, int x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 16;
    a->v = (int *)realloc(a->v, sizeof(int) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}

struct orc_system {
  orc_params p;
  int s;
  double rho2;
  double cos_a, sin_a;
  double x_size, y_size;      /* particle-system box: the reference fixes 1.0 */
  double pbc_xs, pbc_ys;      /* PBC object box */
  int ncx, ncy;               /* cell grid of the linked list */
  int ncvx, ncvy;             /* cell grid used to build Verlet candidates */
  double r_min, r_max;
  int should_update;
  double acc_disp;
  double last_step_max;
  int uses_verlet;
  int higher_order;
  int threads;                /* worker threads of the linked-list particle loop (1: serial) */
  double *nat;                /* natural frequencies [n] */
  double *stage;              /* stage state [s*n] */
  double *align;              /* alignment force [n] */
  double *card;               /* neighbourhood cardinality [n] (a double, as in the reference) */
  double *compl_force;        /* complementary alignment force [n] */
  int *cell_of;               /* pre_linked_list_ */
  struct ivec *buckets;       /* linked_list_ [ncx*ncy] */
  struct ivec *vbuckets;      /* buckets on the Verlet-candidate grid (exact mode) */
  struct ivec *verlet;        /* verlet_list_ [n] */
};

/* ---------------------------------------------------------------------------
 * Periodic boundary conditions: BoundaryConfiguration/PeriodicBoundaryConditions.cpp
 * -------------------------------------------------------------------------*/

/* PeriodicBoundaryConditions.cpp:26-38 (ClassAEffectiveParticleDistance) */
void orc_pbc_class_a(double xs, double ys, double xi, double yi, double xj, double yj, double *dx,
                     double *dy) {
  double xr = 1.0 / xs, yr = 1.0 / ys;
  double ddx = xj - xi;
  ddx -= (double)((int)(ddx * 2.0 * xr)) * xs;
  double ddy = yj - yi;
  ddy -= (double)((int)(ddy * 2.0 * yr)) * ys;
  *dx = ddx;
  *dy = ddy;
}

/* PeriodicBoundaryConditions.cpp:41-65 (ClassBEffectiveParticleDistanceSigned) */
void orc_pbc_class_b_signed(double xs, double ys, double xi, double yi, double xj, double yj,
                            double *dx, double *dy) {
  double ddx = xj - xi;
  if (ddx > 0.5 * xs)
    ddx -= xs;
  else if (ddx < -0.5 * xs)
    ddx += xs;
  double ddy = yj - yi;
  if (ddy > 0.5 * ys)
    ddy -= ys;
  else if (ddy < -0.5 * ys)
    ddy += ys;
  *dx = ddx;
  *dy = ddy;
}

/* PeriodicBoundaryConditions.cpp:68-86 (ClassBEffectiveParticleDistanceUnsigned) */
void orc_pbc_class_b_unsigned(double xs, double ys, double xi, double yi, double xj, double yj,
                              double *dx, double *dy) {
  double ddx = fabs(xj - xi);
  if (ddx > 0.5 * xs) ddx -= xs;
  double ddy = fabs(yj - yi);
  if (ddy > 0.5 * ys) ddy -= ys;
  *dx = ddx;
  *dy = ddy;
}

/* PeriodicBoundaryConditions.cpp:90-102 (ClassCEffectiveParticleDistanceSigned; std::nearbyint: round-half-even in the default mode) */
void orc_pbc_class_c_signed(double xs, double ys, double xi, double yi, double xj, double yj,
                            double *dx, double *dy) {
  double xr = 1.0 / xs, yr = 1.0 / ys;
  double ddx = xj - xi;
  ddx -= xs * nearbyint(ddx * xr);
  double ddy = yj - yi;
  ddy -= ys * nearbyint(ddy * yr);
  *dx = ddx;
  *dy = ddy;
}

/* PeriodicBoundaryConditions.cpp:107-119 (ClassCEffectiveParticleDistanceUnsigned) */
void orc_pbc_class_c_unsigned(double xs, double ys, double xi, double yi, double xj, double yj,
                              double *dx, double *dy) {
  double xr = 1.0 / xs, yr = 1.0 / ys;
  double ddx = fabs(xj - xi);
  ddx -= (double)((int)(ddx * xr + 0.5)) * xs;
  double ddy = fabs(yj - yi);
  ddy -= (double)((int)(ddy * yr + 0.5)) * ys;
  *dx = ddx;
  *dy = ddy;
}

/* PeriodicBoundaryConditions.cpp:121-125 */
void orc_pbc_apply(double xs, double ys, double x, double y, double *xp, double *yp) {
  double xr = 1.0 / xs, yr = 1.0 / ys;
  *xp = x - floor(x * xr) * xs;
  *yp = y - floor(y * yr) * ys;
}

/* PeriodicBoundaryConditions.cpp:147-155 (the size argument is unused there) */
void orc_pbc_apply_state(double xs, double ys, double *state, int n, int s) {
  double xr = 1.0 / xs, yr = 1.0 / ys;
  for (int i = 0; i < n; ++i) {
    int ii = i * s;
    state[ii] -= floor(state[ii] * xr) * xs;
    state[ii + 1] -= floor(state[ii + 1] * yr) * ys;
  }
}

/* ---------------------------------------------------------------------------
 * System object
 * -------------------------------------------------------------------------*/

static int slot_phi(int s) { return s == 6 ? 4 : 2; }
static int slot_omega(int s) { return s == 6 ? 5 : 3; }

/* ParticleSystemFirstOrderPtr.cpp:10-103 (members, nc = int(1.0/rho), r_max = rho + 0.05) */
orc_system *orc_create(const orc_params *p) {
  orc_system *sys = (orc_system *)calloc(1, sizeof(orc_system));
  sys->p = *p;
  sys->s = p->model;
  int n = p->n, s = sys->s;
  sys->rho2 = p->rho * p->rho;
  sys->cos_a = cos(p->alpha);
  sys->sin_a = sin(p->alpha);
  sys->x_size = 1.0;
  sys->y_size = 1.0;
  sys->pbc_xs = p->x_size > 0 ? p->x_size : 1.0;
  sys->pbc_ys = p->y_size > 0 ? p->y_size : 1.0;
  sys->ncx = (int)(1.0 / p->rho);
  sys->ncy = (int)(1.0 / p->rho);
  if (sys->ncx < 1) sys->ncx = 1;
  if (sys->ncy < 1) sys->ncy = 1;
  sys->r_min = p->rho;
  if (p->verlet_mode == ORC_VERLET_EXACT) {
    sys->r_max = p->rho + p->verlet_skin;
  } else if (s == 6) {
    sys->r_max = 2 * p->rho; /* ParticleSystemSecondOrderPtr.cpp:96-97 (never used: Verlet is a stub) */
  } else {
    sys->r_max = p->rho + 0.05; /* ParticleSystemFirstOrderPtr.cpp:97-98 */
  }
  if (p->verlet_mode == ORC_VERLET_EXACT) {
    sys->ncvx = (int)(1.0 / sys->r_max);
    sys->ncvy = sys->ncvx;
    if (sys->ncvx < 1) sys->ncvx = sys->ncvy = 1;
  } else {
    sys->ncvx = sys->ncx;
    sys->ncvy = sys->ncy;
  }
  sys->should_update = 1;
  sys->acc_disp = 0.0;
  sys->threads = 1;
  sys->nat = (double *)calloc((size_t)n, sizeof(double));
  sys->stage = (double *)calloc((size_t)n * (size_t)s, sizeof(double));
  sys->align = (double *)calloc((size_t)n, sizeof(double));
  sys->card = (double *)calloc((size_t)n, sizeof(double));
  sys->compl_force = (double *)calloc((size_t)n, sizeof(double));
  sys->cell_of = (int *)calloc((size_t)n, sizeof(int));
  sys->buckets = (struct ivec *)calloc((size_t)sys->ncx * sys->ncy, sizeof(struct ivec));
  sys->vbuckets = (struct ivec *)calloc((size_t)sys->ncvx * sys->ncvy, sizeof(struct ivec));
  sys->verlet = (struct ivec *)calloc((size_t)n, sizeof(struct ivec));
  return sys;
}

void orc_destroy(orc_system *sys) {
  if (!sys) return;
  for (int c = 0; c < sys->ncx * sys->ncy; ++c) free(sys->buckets[c].v);
  for (int c = 0; c < sys->ncvx * sys->ncvy; ++c) free(sys->vbuckets[c].v);
  for (int i = 0; i < sys->p.n; ++i) free(sys->verlet[i].v);
  free(sys->buckets);
  free(sys->vbuckets);
  free(sys->verlet);
  free(sys->nat);
  free(sys->stage);
  free(sys->align);
  free(sys->card);
  free(sys->compl_force);
  free(sys->cell_of);
  free(sys);
}

/* ParticleSystemFirstOrderPtr.cpp:126-129 */
void orc_set_natural_frequencies(orc_system *sys, const double *w) {
  memcpy(sys->nat, w, sizeof(double) * (size_t)sys->p.n);
}
void orc_set_higher_order(orc_system *sys, int on) { sys->higher_order = on; }
int orc_uses_verlet(const orc_system *sys) { return sys->uses_verlet; }
int orc_should_update_lists(const orc_system *sys) { return sys->should_update; }
double orc_accumulated_displacement(const orc_system *sys) { return sys->acc_disp; }
int orc_num_cells_x(const orc_system *sys) { return sys->ncx; }
const double *orc_complementary_force(const orc_system *sys) { return sys->compl_force; }

/* stage state y + k_coef*dt*k_prev: ParticleSystemFirstOrderPtr.cpp:234-241 */
static void make_stage(orc_system *sys, const double *state, const double *k_prev, double k_coef,
                       double dt) {
  long m = (long)sys->p.n * sys->s;
  for (long k = 0; k < m; ++k) sys->stage[k] = state[k] + k_coef * dt * k_prev[k];
}

/* wrap + cell index, ParticleSystemFirstOrderPtr.cpp:374-381.  The reference
 * leaves ix == nc unguarded when x - floor(x) rounds to 1.0; it is clamped here. */
static int cell_key(int nc_x, int nc_y, double xw, double yw, int *cx, int *cy) {
  int ix = (int)(xw / 1.0 * nc_x);
  int iy = (int)(yw / 1.0 * nc_y);
  if (ix >= nc_x) ix = nc_x - 1;
  if (iy >= nc_y) iy = nc_y - 1;
  if (ix < 0) ix = 0;
  if (iy < 0) iy = 0;
  *cx = ix;
  *cy = iy;
  return iy * nc_x + ix;
}

/* ParticleSystemFirstOrderPtr.cpp:817-834 */
static void adjust_cell(int nx, int ny, int *cx, int *cy) {
  if (*cx < 0)
    *cx = nx - 1;
  else if (*cx >= nx)
    *cx = 0;
  if (*cy < 0)
    *cy = ny - 1;
  else if (*cy >= ny)
    *cy = 0;
}

/* stencil order, ParticleSystemFirstOrderPtr.cpp:83-88 */
static const int kStencil[9][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {0, 0},
                                   {1, 0},   {-1, 1}, {0, 1},  {1, 1}};

/* wrap used for the cell keys.  s = 4/3: unconditional (ParticleSystemFirstOrderPtr.cpp:374);
 * s = 6: only if outside [0, L) (ParticleSystemSecondOrderPtr.cpp:292-295). */
static void key_wrap(const orc_system *sys, double x, double y, double *xw, double *yw) {
  if (sys->s == 6) {
    if (x < 0.0 || x >= sys->x_size || y < 0.0 || y >= sys->y_size) {
      orc_pbc_apply(sys->pbc_xs, sys->pbc_ys, x, y, xw, yw);
    } else {
      *xw = x;
      *yw = y;
    }
  } else {
    orc_pbc_apply(sys->pbc_xs, sys->pbc_ys, x, y, xw, yw);
  }
}

/* linked-list build: ParticleSystemFirstOrderPtr.cpp:351-391 */
static void build_buckets(orc_system *sys, const double *stage, struct ivec *buckets, int nx, int ny) {
  int n = sys->p.n, s = sys->s;
  for (int c = 0; c < nx * ny; ++c) buckets[c].n = 0;
  for (int i = 0; i < n; ++i) {
    double xw, yw;
    int cx, cy;
    key_wrap(sys, stage[i * s], stage[i * s + 1], &xw, &yw);
    sys->cell_of[i] = cell_key(nx, ny, xw, yw, &cx, &cy);
  }
  for (int i = 0; i < n; ++i) ivec_push(&buckets[sys->cell_of[i]], i);
}

/* the pair kernel body shared by every local branch.
 * s = 4/3: ParticleSystemFirstOrderPtr.cpp:285-299 (ClassC unsigned, sin/cos split of alpha)
 * s = 6:   ParticleSystemSecondOrderPtr.cpp:208-215 (ClassA, sin(phi_j - phi_i - alpha)) */
static inline void pair_accumulate(orc_system *sys, int i, double x_i, double y_i, double phi_i,
                                   double omega_i, int j, const double *stage) {
  int s = sys->s;
  int jj = j * s;
  double x_j = stage[jj], y_j = stage[jj + 1];
  double dx, dy;
  if (s == 6) {
    /* the s = 6 linked list passes a stage whose out-of-box x_j are wrapped
     * (ParticleSystemSecondOrderPtr.cpp:356-359); the all-pairs branch wraps nothing */
    orc_pbc_class_a(sys->pbc_xs, sys->pbc_ys, x_i, y_i, x_j, y_j, &dx, &dy);
  } else {
    orc_pbc_class_c_unsigned(sys->pbc_xs, sys->pbc_ys, x_i, y_i, x_j, y_j, &dx, &dy);
  }
  double dr2 = dx * dx + dy * dy;
  if (dr2 <= sys->rho2) {
    double phi_j = stage[jj + slot_phi(s)];
    if (s == 6) {
      sys->align[i] += sin(phi_j - phi_i - sys->p.alpha);
      ++sys->card[i];
    } else {
      sys->align[i] += sin(phi_j - phi_i) * sys->cos_a - cos(phi_j - phi_i) * sys->sin_a;
      ++sys->card[i];
      if (sys->higher_order && s == 4) {
        double omega_j = stage[jj + 3];
        sys->compl_force[i] +=
            (omega_j - omega_i) * (cos(phi_j - phi_i) * sys->cos_a + sin(phi_j - phi_i) * sys->sin_a);
      }
    }
  }
}

/* output row: ParticleSystemFirstOrderPtr.cpp:303-313 (s=4),
 * ParticleSystemSecondOrderPtr.cpp:219-227 (s=6); s=3 is the m = 0 model. */
static void write_k(orc_system *sys, int i, const double *stage, double *k_next) {
  int s = sys->s, ii = i * s;
  const orc_params *p = &sys->p;
  if (s == 4) {
    double phi_i = stage[ii + 2], omega_i = stage[ii + 3];
    k_next[ii] = p->v0 * cos(phi_i);
    k_next[ii + 1] = p->v0 * sin(phi_i);
    k_next[ii + 2] = omega_i;
    if (p->natfreq_local) {
      if (sys->card[i] != 0.0)
        k_next[ii + 3] = -p->xi * omega_i + sys->nat[i] + p->sigma * sys->align[i] / sys->card[i];
      else
        k_next[ii + 3] = -p->xi * omega_i + sys->nat[i];
    } else if (sys->card[i] != 0.0) {
      k_next[ii + 3] = -p->xi * omega_i + p->sigma * sys->align[i] / sys->card[i];
    }
    if (sys->card[i] != 0.0 && sys->higher_order) {
      sys->compl_force[i] *= p->sigma / sys->card[i];
    }
  } else if (s == 6) {
    double vx = stage[ii + 2], vy = stage[ii + 3], phi_i = stage[ii + 4], omega_i = stage[ii + 5];
    k_next[ii] = vx;
    k_next[ii + 1] = vy;
    k_next[ii + 2] = -p->xi_r * vx + p->v0 * cos(phi_i);
    k_next[ii + 3] = -p->xi_r * vy + p->v0 * sin(phi_i);
    k_next[ii + 4] = omega_i;
    if (sys->card[i] != 0.0) k_next[ii + 5] = -p->xi * omega_i + p->sigma * sys->align[i] / sys->card[i];
  } else { /* s == 3 */
    double phi_i = stage[ii + 2];
    k_next[ii] = p->v0 * cos(phi_i);
    k_next[ii + 1] = p->v0 * sin(phi_i);
    if (sys->card[i] != 0.0)
      k_next[ii + 2] = sys->nat[i] + p->sigma * sys->align[i] / sys->card[i];
    else
      k_next[ii + 2] = sys->nat[i];
  }
}

static void clear_accumulators(orc_system *sys) {
  int n = sys->p.n;
  memset(sys->align, 0, sizeof(double) * (size_t)n);
  memset(sys->card, 0, sizeof(double) * (size_t)n);
  if (sys->higher_order) memset(sys->compl_force, 0, sizeof(double) * (size_t)n);
}

static double omega_of(const orc_system *sys, const double *stage, int i) {
  return sys->s == 3 ? 0.0 : stage[i * sys->s + slot_omega(sys->s)];
}

/* ParticleSystemFirstOrderPtr.cpp:225-316 (s=4: x_i wrapped at :270, j ascending, j != i);
 * ParticleSystemSecondOrderPtr.cpp:155-230 (s=6: nothing wrapped, ClassA). */
static void rhs_all_pairs(orc_system *sys, const double *stage, double *k_next) {
  int n = sys->p.n, s = sys->s;
  clear_accumulators(sys);
  for (int i = 0; i < n; ++i) {
    int ii = i * s;
    double x_i = stage[ii], y_i = stage[ii + 1];
    double phi_i = stage[ii + slot_phi(s)], omega_i = omega_of(sys, stage, i);
    if (s != 6) orc_pbc_apply(sys->pbc_xs, sys->pbc_ys, x_i, y_i, &x_i, &y_i);
    for (int j = 0; j < n; ++j) {
      if (j != i) pair_accumulate(sys, i, x_i, y_i, phi_i, omega_i, j, stage);
    }
    write_k(sys, i, stage, k_next);
  }
}

/* the s = 6 linked list wraps x_j before ClassA when it is outside the box
 * (ParticleSystemSecondOrderPtr.cpp:356-360); build a wrapped copy for it. */
static const double *s6_wrapped_stage(orc_system *sys, const double *stage, double **tmp) {
  int n = sys->p.n, s = sys->s;
  double *w = (double *)malloc(sizeof(double) * (size_t)n * (size_t)s);
  memcpy(w, stage, sizeof(double) * (size_t)n * (size_t)s);
  for (int j = 0; j < n; ++j) {
    double xj = w[j * s], yj = w[j * s + 1];
    if (xj < 0.0 || xj >= sys->x_size || yj < 0.0 || yj >= sys->y_size)
      orc_pbc_apply(sys->pbc_xs, sys->pbc_ys, xj, yj, &w[j * s], &w[j * s + 1]);
  }
  *tmp = w;
  return w;
}

/* the particle loop of the linked-list branch over particles [i0, i1): ParticleSystemFirstOrderPtr.cpp:398-466 */
static void linked_list_particles(orc_system *sys, const double *stage, const double *jstage, double *k_next,
                                  int i0, int i1) {
  int s = sys->s;
  for (int i = i0; i < i1; ++i) {
    int ii = i * s;
    double x_i, y_i;
    key_wrap(sys, stage[ii], stage[ii + 1], &x_i, &y_i);
    double phi_i = stage[ii + slot_phi(s)], omega_i = omega_of(sys, stage, i);
    int cx, cy;
    cell_key(sys->ncx, sys->ncy, x_i, y_i, &cx, &cy);
    for (int c = 0; c < 9; ++c) {
      int jx = cx + kStencil[c][0], jy = cy + kStencil[c][1];
      adjust_cell(sys->ncx, sys->ncy, &jx, &jy);
      const struct ivec *b = &sys->buckets[jy * sys->ncx + jx];
      for (int t = 0; t < b->n; ++t) {
        int j = b->v[t];
        if (j != i) pair_accumulate(sys, i, x_i, y_i, phi_i, omega_i, j, jstage);
      }
    }
    write_k(sys, i, stage, k_next);
  }
}

struct ll_task {
  orc_system *sys;
  const double *stage, *jstage;
  double *k_next;
  int i0, i1;
};

static void *ll_worker(void *arg) {
  struct ll_task *t = (struct ll_task *)arg;
  linked_list_particles(t->sys, t->stage, t->jstage, t->k_next, t->i0, t->i1);
  return NULL;
}

/* ParticleSystemFirstOrderPtr.cpp:318-468 (s=4), ParticleSystemSecondOrderPtr.cpp:240-383 (s=6).
 * With threads > 1 the particle loop runs over contiguous blocks of particles, one per thread, as
 * the reference's ranks own them (Parallelization/ThreadSharedMemory.cpp:20-23, with the last block
 * taking the remainder); every particle's sums are the serial ones, so results do not change. */
static void rhs_linked_list(orc_system *sys, const double *stage, double *k_next) {
  int n = sys->p.n, s = sys->s;
  clear_accumulators(sys);
  build_buckets(sys, stage, sys->buckets, sys->ncx, sys->ncy);
  double *tmp = NULL;
  const double *jstage = stage;
  if (s == 6) jstage = s6_wrapped_stage(sys, stage, &tmp);
  int T = sys->threads;
  if (T > n) T = n;
  if (T <= 1) {
    linked_list_particles(sys, stage, jstage, k_next, 0, n);
  } else {
    pthread_t th[256];
    struct ll_task task[256];
    int running[256];
    if (T > 256) T = 256;
    int per = n / T;
    for (int t = 0; t < T; ++t) {
      task[t].sys = sys;
      task[t].stage = stage;
      task[t].jstage = jstage;
      task[t].k_next = k_next;
      task[t].i0 = t * per;
      task[t].i1 = (t == T - 1) ? n : (t + 1) * per;
      running[t] = t > 0 && pthread_create(&th[t], NULL, ll_worker, &task[t]) == 0;
      if (t > 0 && !running[t]) ll_worker(&task[t]); /* no thread: the block runs here */
    }
    ll_worker(&task[0]);
    for (int t = 1; t < T; ++t)
      if (running[t]) pthread_join(th[t], NULL);
  }
  free(tmp);
}

void orc_set_threads(orc_system *sys, int threads) { sys->threads = threads < 1 ? 1 : threads; }

/* Verlet list build over all pairs: ParticleSystemFirstOrderPtr.cpp:497-543 (x_i not wrapped) */
static void build_verlet_all_pairs(orc_system *sys, const double *stage) {
  int n = sys->p.n, s = sys->s;
  double rmax2 = sys->r_max * sys->r_max;
  for (int i = 0; i < n; ++i) {
    sys->verlet[i].n = 0;
    double x_i = stage[i * s], y_i = stage[i * s + 1];
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx, dy;
      orc_pbc_class_c_unsigned(sys->pbc_xs, sys->pbc_ys, x_i, y_i, stage[j * s], stage[j * s + 1], &dx, &dy);
      double dr2 = dx * dx + dy * dy;
      if (dr2 <= rmax2) ivec_push(&sys->verlet[i], j);
    }
  }
}

/* Verlet list build from the cell stencil: ParticleSystemFirstOrderPtr.cpp:643-743
 * (x_i wrapped, candidates only from the 9 cells).  In exact mode the cells are
 * at least r_max wide, so the candidate set contains every j within r_max. */
static void build_verlet_cells(orc_system *sys, const double *stage) {
  int n = sys->p.n, s = sys->s;
  int nx = sys->ncvx, ny = sys->ncvy;
  struct ivec *bk = (sys->p.verlet_mode == ORC_VERLET_EXACT) ? sys->vbuckets : sys->buckets;
  build_buckets(sys, stage, bk, nx, ny);
  double rmax2 = sys->r_max * sys->r_max;
  for (int i = 0; i < n; ++i) {
    sys->verlet[i].n = 0;
    double x_i, y_i;
    key_wrap(sys, stage[i * s], stage[i * s + 1], &x_i, &y_i);
    int cx, cy;
    cell_key(nx, ny, x_i, y_i, &cx, &cy);
    for (int c = 0; c < 9; ++c) {
      int jx = cx + kStencil[c][0], jy = cy + kStencil[c][1];
      adjust_cell(nx, ny, &jx, &jy);
      const struct ivec *b = &bk[jy * nx + jx];
      for (int t = 0; t < b->n; ++t) {
        int j = b->v[t];
        if (j == i) continue;
        double dx, dy;
        orc_pbc_class_c_unsigned(sys->pbc_xs, sys->pbc_ys, x_i, y_i, stage[j * s], stage[j * s + 1], &dx, &dy);
        double dr2 = dx * dx + dy * dy;
        if (dr2 <= rmax2) ivec_push(&sys->verlet[i], j);
      }
    }
  }
}

/* Verlet pair loop: ParticleSystemFirstOrderPtr.cpp:545-613 / :745-814 (x_i not wrapped) */
static void rhs_verlet(orc_system *sys, const double *stage, double *k_next, int cells) {
  int n = sys->p.n, s = sys->s;
  if (sys->should_update) {
    if (cells)
      build_verlet_cells(sys, stage);
    else
      build_verlet_all_pairs(sys, stage);
    sys->should_update = 0;
  }
  clear_accumulators(sys);
  for (int i = 0; i < n; ++i) {
    int ii = i * s;
    double x_i = stage[ii], y_i = stage[ii + 1];
    double phi_i = stage[ii + slot_phi(s)], omega_i = omega_of(sys, stage, i);
    const struct ivec *l = &sys->verlet[i];
    for (int t = 0; t < l->n; ++t) {
      int j = l->v[t];
      if (j != i) pair_accumulate(sys, i, x_i, y_i, phi_i, omega_i, j, stage);
    }
    write_k(sys, i, stage, k_next);
  }
}

/* complex helpers with the libstdc++ operations the reference uses */
static double cabs_d(double re, double im) { return hypot(re, im); }
static double carg_d(double re, double im) { return atan2(im, re); }

void orc_polar_fields(const double *stage, int n, int s, int slot_phi_, int slot_omega_, double *zre,
                      double *zim, double *wre, double *wim) {
  double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
  for (int i = 0; i < n; ++i) {
    double phi = stage[s * i + slot_phi_];
    double c = cos(phi), sn = sin(phi);
    ar += c;
    ai += sn;
    if (slot_omega_ >= 0) {
      double om = stage[s * i + slot_omega_];
      br += om * c;
      bi += om * sn;
    }
  }
  *zre = ar / n;
  *zim = ai / n;
  *wre = br / n;
  *wim = bi / n;
}

/* ParticleSystemFirstOrderPtr.cpp:155-223 (self included, normalised by N) */
static void rhs_global(orc_system *sys, const double *stage, double *k_next) {
  int n = sys->p.n, s = sys->s;
  const orc_params *p = &sys->p;
  double zr, zi, wr, wi;
  orc_polar_fields(stage, n, s, slot_phi(s), s == 4 ? 3 : -1, &zr, &zi, &wr, &wi);
  double absZ = cabs_d(zr, zi), argZ = carg_d(zr, zi);
  double absW = cabs_d(wr, wi), argW = carg_d(wr, wi);
  for (int i = 0; i < n; ++i) {
    int ii = s * i;
    double phi_i = stage[ii + slot_phi(s)];
    if (s == 4) {
      double omega_i = stage[ii + 3];
      k_next[ii] = p->v0 * cos(phi_i);
      k_next[ii + 1] = p->v0 * sin(phi_i);
      k_next[ii + 2] = omega_i;
      k_next[ii + 3] = -p->xi * omega_i + sys->nat[i] +
                       p->sigma * absZ * (sin(argZ - phi_i) * sys->cos_a - cos(argZ - phi_i) * sys->sin_a);
      if (sys->higher_order) {
        /* :215-219, including the arg(W) in the last sine as the reference has it */
        sys->compl_force[i] =
            p->sigma * (absW * (cos(argW - phi_i) * sys->cos_a + sin(argW - phi_i) * sys->sin_a) -
                        omega_i * absZ * (cos(argZ - phi_i) * sys->cos_a + sin(argW - phi_i) * sys->sin_a));
      }
    } else if (s == 3) {
      k_next[ii] = p->v0 * cos(phi_i);
      k_next[ii + 1] = p->v0 * sin(phi_i);
      k_next[ii + 2] = sys->nat[i] +
                       p->sigma * absZ * (sin(argZ - phi_i) * sys->cos_a - cos(argZ - phi_i) * sys->sin_a);
    }
  }
}

/* ParticleSystemFirstOrderPtr.cpp:131-153 (dispatch on rho) and the s = 6 dispatch
 * ParticleSystemSecondOrderPtr.cpp:125-138 */
static int resolve_method(const orc_system *sys, int method) {
  if (method != ORC_METHOD_AUTO) return method;
  double rho = sys->p.rho;
  if (sys->s == 6) return rho <= 0.25 ? ORC_METHOD_LINKED_LIST : ORC_METHOD_ALL_PAIRS;
  if (rho >= M_SQRT1_2) return ORC_METHOD_GLOBAL_FIELD;
  if (rho <= 0.25) return ORC_METHOD_LINKED_LIST_VERLET;
  if (rho <= 0.5) return ORC_METHOD_VERLET;
  return ORC_METHOD_ALL_PAIRS;
}

void orc_evaluate_rhs(orc_system *sys, int method, const double *state, const double *k_prev,
                      double *k_next, double k_coef, double dt) {
  int m = resolve_method(sys, method);
  make_stage(sys, state, k_prev, k_coef, dt);
  const double *st = sys->stage;
  switch (m) {
    case ORC_METHOD_GLOBAL_FIELD: rhs_global(sys, st, k_next); break;
    case ORC_METHOD_ALL_PAIRS: rhs_all_pairs(sys, st, k_next); break;
    case ORC_METHOD_LINKED_LIST: rhs_linked_list(sys, st, k_next); break;
    case ORC_METHOD_VERLET:
      rhs_verlet(sys, st, k_next, 0);
      sys->uses_verlet = 1;
      break;
    case ORC_METHOD_LINKED_LIST_VERLET:
      rhs_verlet(sys, st, k_next, 1);
      sys->uses_verlet = 1;
      break;
    default: break;
  }
}

/* ParticleSystemFirstOrderPtr.cpp:836-861.  Exact mode adds the step's own
 * displacement as a margin, since RK stages of the next step move that far
 * beyond the positions the accumulated sum covers. */
void orc_calculate_max_displacement(orc_system *sys, const double *total_displacement) {
  int n = sys->p.n, s = sys->s;
  double max2 = 0.0;
  for (int i = 0; i < n; ++i) {
    double dx = total_displacement[s * i + 0];
    double dy = total_displacement[s * i + 1];
    double dist2 = dx * dx + dy * dy;
    if (max2 < dist2) max2 = dist2;
  }
  double step = sqrt(max2);
  sys->acc_disp += step;
  sys->last_step_max = step;
  int flag = 0;
  if (sys->p.verlet_mode == ORC_VERLET_EXACT) {
    if (sys->acc_disp + step > 0.5 * (sys->r_max - sys->r_min)) {
      sys->acc_disp = 0.0;
      flag = 1;
    }
  } else if (sys->acc_disp > 0.5 * (sys->r_max - sys->r_min)) {
    sys->acc_disp = 0.0;
    flag = 1;
  }
  sys->should_update = flag;
}

/* RungeKutta4StepperPtr.cpp:68-92 (s=4) and :94-115 (s=6, no displacement check) */
void orc_rk4_step(orc_system *sys, int method, double *state, double *k1, double *k2, double *k3,
                  double *k4, double *total_disp, double dt) {
  long m = (long)sys->p.n * sys->s;
  orc_evaluate_rhs(sys, method, state, k1, k1, 0.0, dt);
  orc_evaluate_rhs(sys, method, state, k1, k2, 0.5, dt);
  orc_evaluate_rhs(sys, method, state, k2, k3, 0.5, dt);
  orc_evaluate_rhs(sys, method, state, k3, k4, 1.0, dt);
  for (long k = 0; k < m; ++k) {
    total_disp[k] = (k1[k] + 2.0 * k2[k] + 2.0 * k3[k] + k4[k]) * dt / 6.0;
    state[k] += total_disp[k];
  }
  if (sys->uses_verlet && sys->s != 6) orc_calculate_max_displacement(sys, total_disp);
}

/* s=4: StochasticEulerStepperPtr.hpp:40-73; s=6: StochasticEulerStepper.hpp:33-51 with
 * ParticleSystemSecondOrder.cpp:439-451 (AddNoise draws slots 2, 3, 5 per particle);
 * s=3: the same scheme on the phase. */
void orc_euler_step(orc_system *sys, int method, double *state, double *derivative, double *total_disp,
                    double dt, const double *normals) {
  int n = sys->p.n, s = sys->s;
  long m = (long)n * s;
  memset(derivative, 0, sizeof(double) * (size_t)m);
  orc_evaluate_rhs(sys, method, state, derivative, derivative, 0.0, dt);
  if (s == 4) {
    double intensity_phi = sqrt(2.0 * sys->p.D_phi);
    for (int i = 0; i < n; ++i) {
      double dW = sqrt(dt) * normals[i];
      int ii = s * i;
      total_disp[ii] = derivative[ii] * dt;
      total_disp[ii + 1] = derivative[ii + 1] * dt;
      total_disp[ii + 2] = derivative[ii + 2] * dt;
      total_disp[ii + 3] = derivative[ii + 3] * dt + intensity_phi * dW;
      state[ii] += total_disp[ii];
      state[ii + 1] += total_disp[ii + 1];
      state[ii + 2] += total_disp[ii + 2];
      state[ii + 3] += total_disp[ii + 3];
    }
    if (sys->uses_verlet) orc_calculate_max_displacement(sys, total_disp);
  } else if (s == 6) {
    double ir = sqrt(2.0 * sys->p.D_r), ip = sqrt(2.0 * sys->p.D_phi);
    double sdt = sqrt(dt);
    for (int i = 0; i < n; ++i) {
      int ii = s * i;
      double g[6] = {0.0, 0.0, ir * normals[3 * i], ir * normals[3 * i + 1], 0.0, ip * normals[3 * i + 2]};
      for (int k = 0; k < 6; ++k) state[ii + k] += dt * derivative[ii + k] + sdt * g[k];
    }
  } else {
    double intensity_phi = sqrt(2.0 * sys->p.D_phi);
    for (int i = 0; i < n; ++i) {
      double dW = sqrt(dt) * normals[i];
      int ii = s * i;
      total_disp[ii] = derivative[ii] * dt;
      total_disp[ii + 1] = derivative[ii + 1] * dt;
      total_disp[ii + 2] = derivative[ii + 2] * dt + intensity_phi * dW;
      state[ii] += total_disp[ii];
      state[ii + 1] += total_disp[ii + 1];
      state[ii + 2] += total_disp[ii + 2];
    }
    if (sys->uses_verlet) orc_calculate_max_displacement(sys, total_disp);
  }
}

/* StochasticRungeKuttaStepperPtr.hpp:40-76 (strong order 1.5 Taylor scheme), s = 4 */
void orc_srk_step(orc_system *sys, int method, double *state, double *derivative, double *total_disp,
                  double dt, const double *normals) {
  int n = sys->p.n, s = sys->s;
  long m = (long)n * s;
  memset(derivative, 0, sizeof(double) * (size_t)m);
  sys->higher_order = 1;
  orc_evaluate_rhs(sys, method, state, derivative, derivative, 0.0, dt);
  const double *C = sys->compl_force;
  double noise = sys->p.D_phi, friction = sys->p.xi;
  for (int i = 0; i < n; ++i) {
    double U_1 = normals[2 * i], U_2 = normals[2 * i + 1];
    double delta_W_i = U_1 * sqrt(dt);
    double delta_Z_i = 0.5 * pow(dt, 1.5) * (U_1 + 1.0 / sqrt(3.0) * U_2);
    int ii = s * i;
    const double *f = derivative + ii;
    total_disp[ii] = f[0] * dt - 0.5 * f[2] * f[1] * dt * dt;
    total_disp[ii + 1] = f[1] * dt + 0.5 * f[2] * f[0] * dt * dt;
    total_disp[ii + 2] = f[2] * dt + 0.5 * f[3] * dt * dt + sqrt(2.0 * noise) * delta_Z_i;
    total_disp[ii + 3] = f[3] * dt + 0.5 * (C[i] - friction * f[3]) * dt * dt +
                         sqrt(2.0 * noise) * (delta_W_i - friction * delta_Z_i);
    state[ii] += total_disp[ii];
    state[ii + 1] += total_disp[ii + 1];
    state[ii + 2] += total_disp[ii + 2];
    state[ii + 3] += total_disp[ii + 3];
  }
  if (sys->uses_verlet) orc_calculate_max_displacement(sys, total_disp);
}

/* ---------------------------------------------------------------------------
 * Neighbour-set export
 * -------------------------------------------------------------------------*/

static int cmp_pair(const void *a, const void *b) {
  const int32_t *x = (const int32_t *)a, *y = (const int32_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

/* the ClassC in-range test of the pair loops over all j (the set a correct cell or Verlet
 * search must reproduce): PeriodicBoundaryConditions.cpp:107-119 + ParticleSystemFirstOrderPtr.cpp:285-288 */
long orc_pairs_bruteforce(const double *stage, int n, int s, double rho, double xs, double ys,
                          int32_t *out, long cap) {
  double rho2 = rho * rho;
  long cnt = 0;
  for (int i = 0; i < n; ++i) {
    double x_i, y_i;
    orc_pbc_apply(xs, ys, stage[i * s], stage[i * s + 1], &x_i, &y_i);
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx, dy;
      orc_pbc_class_c_unsigned(xs, ys, x_i, y_i, stage[j * s], stage[j * s + 1], &dx, &dy);
      double dr2 = dx * dx + dy * dy;
      if (dr2 <= rho2) {
        if (cnt < cap) {
          out[2 * cnt] = i;
          out[2 * cnt + 1] = j;
        }
        ++cnt;
      }
    }
  }
  if (cnt <= cap) qsort(out, (size_t)cnt, 2 * sizeof(int32_t), cmp_pair);
  return cnt;
}

/* pairs found by the linked-list branch (ParticleSystemFirstOrderPtr.cpp:398-453), sorted */
long orc_pairs_linked_list(orc_system *sys, const double *stage, int32_t *out, long cap) {
  int n = sys->p.n, s = sys->s;
  build_buckets(sys, stage, sys->buckets, sys->ncx, sys->ncy);
  double *tmp = NULL;
  const double *jst = stage;
  if (s == 6) jst = s6_wrapped_stage(sys, stage, &tmp);
  long cnt = 0;
  for (int i = 0; i < n; ++i) {
    double x_i, y_i;
    key_wrap(sys, stage[i * s], stage[i * s + 1], &x_i, &y_i);
    int cx, cy;
    cell_key(sys->ncx, sys->ncy, x_i, y_i, &cx, &cy);
    for (int c = 0; c < 9; ++c) {
      int jx = cx + kStencil[c][0], jy = cy + kStencil[c][1];
      adjust_cell(sys->ncx, sys->ncy, &jx, &jy);
      const struct ivec *b = &sys->buckets[jy * sys->ncx + jx];
      for (int t = 0; t < b->n; ++t) {
        int j = b->v[t];
        if (j == i) continue;
        double dx, dy;
        if (s == 6)
          orc_pbc_class_a(sys->pbc_xs, sys->pbc_ys, x_i, y_i, jst[j * s], jst[j * s + 1], &dx, &dy);
        else
          orc_pbc_class_c_unsigned(sys->pbc_xs, sys->pbc_ys, x_i, y_i, jst[j * s], jst[j * s + 1], &dx, &dy);
        double dr2 = dx * dx + dy * dy;
        if (dr2 <= sys->rho2) {
          if (cnt < cap) {
            out[2 * cnt] = i;
            out[2 * cnt + 1] = j;
          }
          ++cnt;
        }
      }
    }
  }
  free(tmp);
  if (cnt <= cap) { /* pairs come out by ascending i: sort each particle's j in place */
    long a = 0;
    while (a < cnt) {
      long b = a + 1;
      while (b < cnt && out[2 * b] == out[2 * a]) ++b;
      for (long u = a + 1; u < b; ++u) { /* insertion sort of a few dozen entries */
        int32_t j = out[2 * u + 1];
        long v = u;
        while (v > a && out[2 * (v - 1) + 1] > j) {
          out[2 * v + 1] = out[2 * (v - 1) + 1];
          --v;
        }
        out[2 * v + 1] = j;
      }
      a = b;
    }
  }
  return cnt;
}

long orc_verlet_lists(orc_system *sys, int32_t *out_i, int32_t *out_j, long cap) {
  long cnt = 0;
  for (int i = 0; i < sys->p.n; ++i) {
    for (int t = 0; t < sys->verlet[i].n; ++t) {
      if (cnt < cap) {
        out_i[cnt] = i;
        out_j[cnt] = sys->verlet[i].v[t];
      }
      ++cnt;
    }
  }
  return cnt;
}

/* ---------------------------------------------------------------------------
 * Observers
 * -------------------------------------------------------------------------*/

/* FP64 order parameter (1/N) sum e^{i phi}; BinaryObserverFirstOrderPtr.cpp:181-200 restated
 * in FP64 with a compensated (Neumaier) sum, the accuracy the reference's float sum lacks. */
void orc_order_parameter(const double *state, int n, int s, int slot, double *re, double *im) {
  double sr = 0.0, cr = 0.0, si = 0.0, ci = 0.0;
  for (int i = 0; i < n; ++i) {
    double phi = state[s * i + slot];
    double a = cos(phi), b = sin(phi);
    double t = sr + a;
    cr += fabs(sr) >= fabs(a) ? (sr - t) + a : (a - t) + sr;
    sr = t;
    t = si + b;
    ci += fabs(si) >= fabs(b) ? (si - t) + b : (b - t) + si;
    si = t;
  }
  *re = (sr + cr) / n;
  *im = (si + ci) / n;
}

/* the reference's own summary statistic: complex<float> accumulation of float-cast
 * cos/sin (BinaryObserverFirstOrderPtr.cpp:188-194) */
void orc_order_parameter_ref_float(const double *state, int n, int s, int slot, float *re, float *im) {
  float ar = 0.0f, ai = 0.0f;
  for (int i = 0; i < n; ++i) {
    double phi = state[s * i + slot];
    ar += (float)cos(phi);
    ai += (float)sin(phi);
  }
  *re = ar / (float)n;
  *im = ai / (float)n;
}

/* AveragingObserver.cpp:66-74 */
void orc_averaging_accumulate(double *acc, const double *state, long size) {
  for (long k = 0; k < size; ++k) acc[k] = state[k] + acc[k];
}
