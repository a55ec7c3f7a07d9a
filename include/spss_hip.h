/*
 * spss_hip.h -- C ABI of the MI355X (gfx950) right-hand-side hot path for
 * self-propelled particles with Kuramoto-type phase coupling.
 *
 * The library (libspss_hip.so) keeps one particle system per handle with its
 * state resident in HBM and runs hand-written HIP kernels.  Signatures use
 * plain C types only: pointers, sizes and POD structs.  Every entry point
 * returns SPSS_OK or an error code; spss_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Host pointers are borrowed for the duration of a call.  Device pointers
 * (the *_device variants) are read/written on the handle's stream.  One
 * handle is bound to one HIP stream: calls on a handle are stream-ordered and
 * not thread-safe; different handles may be used from different threads.
 *
 * State layout everywhere is the reference's: array-of-structures FP64 with
 * s values per particle
 *   model 4: (x, y, phi, omega)          DynamicalSystems/ParticleSystemFirstOrderPtr.hpp:12-114
 *   model 6: (x, y, vx, vy, phi, omega)  DynamicalSystems/ParticleSystemSecondOrderPtr.hpp:12-116
 *   model 3: (x, y, phi)                 first-order (m = 0) Kuramoto; no reference counterpart
 */
#ifndef SPSS_HIP_H
#define SPSS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPSS_ABI_VERSION 1

/* status codes */
#define SPSS_OK 0
#define SPSS_ERR_INVALID 1     /* bad argument */
#define SPSS_ERR_DEVICE 2      /* a HIP runtime call failed */
#define SPSS_ERR_UNSUPPORTED 3 /* combination the reference does not define */
#define SPSS_ERR_CAPACITY 4    /* a fixed-size buffer would overflow */

/* neighbour strategies; AUTO is the reference dispatch on rho
 * (ParticleSystemFirstOrderPtr.cpp:131-153, ParticleSystemSecondOrderPtr.cpp:125-138) */
#define SPSS_METHOD_AUTO 0
#define SPSS_METHOD_GLOBAL_FIELD 1      /* EvaluateInteractionsWithGlobalPolarOrderField, :155-223 */
#define SPSS_METHOD_ALL_PAIRS 2         /* EvaluateInteractionsWithAllPairs, :225-316 */
#define SPSS_METHOD_LINKED_LIST 3       /* EvaluateInteractionsWithLinkedList, :318-468 */
#define SPSS_METHOD_VERLET 4            /* EvaluateInteractionsWithVerletNeighborList, :470-614 */
#define SPSS_METHOD_LINKED_LIST_VERLET 5 /* EvaluateInteractionsWithLinkedListAndVerletNeighborList, :616-815 */

/* arithmetic of the pair sums */
#define SPSS_PRECISION_FP64 0  /* FP64, the reference's operation order */
#define SPSS_PRECISION_FP32C 1 /* models 3 and 4, the linked-list method (RHS and fused stochastic
                                  steps) and exact-mode Verlet lists (fused stochastic steps): FP32
                                  distance tests (guard band decided in FP64, so the neighbour sets
                                  equal the FP64 ones), coupling sums through the neighbours' unit
                                  phasors in FP32 partial sums combined in FP64, FP64 state and
                                  stage combination */

/* Verlet semantics */
#define SPSS_VERLET_COMPAT 0 /* reference: r_max = rho + 0.05, candidates from the 9 cells */
#define SPSS_VERLET_EXACT 1  /* r_max = rho + verlet_skin, candidates from cells >= r_max */

/* steppers */
#define SPSS_STEPPER_RK4 0   /* RungeKutta4Stepper[Ptr]::DoStep, Steppers/RungeKutta4StepperPtr.cpp:31-82 */
#define SPSS_STEPPER_EULER 1 /* StochasticEulerStepperPtr.hpp:40-73 (s=4), StochasticEulerStepper.hpp:33-51 (s=6) */
#define SPSS_STEPPER_SRK 2   /* StochasticRungeKuttaStepperPtr.hpp:40-76 (s=4 only) */

/* where the normal draws of a stochastic step come from */
#define SPSS_NOISE_NONE 0   /* all draws are 0 */
#define SPSS_NOISE_HOST 1   /* host array of the draws for this step (noise injection) */
#define SPSS_NOISE_DEVICE 2 /* device array of the draws for this step */
#define SPSS_NOISE_PHILOX 3 /* counter-based Philox4x32-10 + Box-Muller on the device,
                               key = seed, counter = (particle, draw block, step) */

typedef struct spss_system spss_system;

/* Constructor arguments of ParticleSystemFirstOrderPtr (ParticleSystemFirstOrderPtr.hpp:16-24),
 * ParticleSystemSecondOrderPtr (ParticleSystemSecondOrderPtr.hpp:16-26) and
 * PeriodicBoundaryConditions (PeriodicBoundaryConditions.hpp:15-19), plus build options. */
typedef struct spss_params {
  int model;       /* 4, 6 or 3 (= number of state variables s) */
  int n;           /* number of particles */
  double v0;       /* self-propulsion speed v_0 */
  double xi;       /* model 4: phase friction xi; model 6: xi_phi */
  double sigma;    /* coupling strength */
  double rho;      /* interaction radius */
  double alpha;    /* phase lag */
  double D_phi;    /* phase noise intensity */
  double xi_r;     /* model 6: translational friction */
  double D_r;      /* model 6: translational noise */
  double x_size;   /* periodic box (the reference's cell grid assumes 1.0) */
  double y_size;
  double verlet_skin;  /* exact Verlet mode: r_max - rho */
  int verlet_mode;     /* SPSS_VERLET_* */
  int natfreq_local;   /* 1: natural frequencies enter the local branches too (model 4) */
  int precision;       /* SPSS_PRECISION_* */
  int cell_subdivision; /* FP32C cell list: cells of side 1/(k*nc) with a (2k+1)^2 stencil; 0 -> 1 */
  uint64_t seed;       /* Philox key */
} spss_params;

/* fill with the reference defaults of ParticleSystemFirstOrderPtr (model 4, L = 1) */
int spss_params_default(spss_params *p);

/* ParticleSystemFirstOrderPtr::ParticleSystemFirstOrderPtr / ~ParticleSystemFirstOrderPtr
 * (ParticleSystemFirstOrderPtr.cpp:10-124); the device owns the state and scratch. */
int spss_create(const spss_params *p, spss_system **out);
int spss_destroy(spss_system *sys);
const char *spss_last_error(void);
int spss_abi_version(void);

/* stream binding; the stream is a hipStream_t passed as void* (NULL = a stream of the library's own) */
int spss_set_stream(spss_system *sys, void *hip_stream);
int spss_synchronize(spss_system *sys);

/* device-resident master state (SimulationEngineFirstOrderPtr.cpp:29 `new Real[system_state_size_]`) */
int spss_set_state(spss_system *sys, const double *host_state);
int spss_get_state(spss_system *sys, double *host_state);
int spss_set_state_device(spss_system *sys, const double *device_state);
/* the device buffer of the master state (s*n doubles, original particle order), for the caller
 * to read or write between calls.  FP32C stochastic steps advance a copy kept in cell order;
 * this call writes that copy back first and drops it (the next such step re-sorts). */
double *spss_state_device(spss_system *sys);

/* ParticleSystemFirstOrderPtr::SetNaturalFrequencies (ParticleSystemFirstOrderPtr.cpp:126-129) */
int spss_set_natural_frequencies(spss_system *sys, const double *host_w, int n);
/* the public flag is_higher_order_sde_integration_ (ParticleSystemFirstOrderPtr.hpp:72) */
int spss_set_higher_order(spss_system *sys, int on);
/* GetComplimentaryAlignmentForce (ParticleSystemFirstOrderPtr.cpp:863-866) */
int spss_get_complementary_force(spss_system *sys, double *host_out);

/* EvaluateRhs (ParticleSystemFirstOrderPtr.cpp:131-153) and the EvaluateInteractionsWith*
 * methods it dispatches to, on host buffers of s*n doubles.  k_next may alias k_prev. */
int spss_evaluate_rhs(spss_system *sys, int method, const double *state, const double *k_prev,
                      double *k_next, double k_coef, double dt);
/* the same on device buffers of s*n doubles */
int spss_evaluate_rhs_device(spss_system *sys, int method, const double *d_state, const double *d_k_prev,
                             double *d_k_next, double k_coef, double dt);

/* DoStep of the three steppers on the device-resident state.  `noise` holds the draws
 * of this step for SPSS_NOISE_HOST/DEVICE: model 4 Euler n, model 6 Euler 3n (slots 2,3,5),
 * model 3 Euler n, SRK 2n (U1, U2 interleaved). */
int spss_rk4_step(spss_system *sys, int method, double dt);
int spss_euler_step(spss_system *sys, int method, double dt, int noise_mode, const double *noise);
int spss_srk_step(spss_system *sys, int method, double dt, int noise_mode, const double *noise);
/* `nsteps` steps of one stepper (the time loop of SimulationEngineFirstOrderPtr::RunSimulation,
 * SimulationEngineFirstOrderPtr.cpp:128-137, without observers).  Noise: NONE or PHILOX. */
int spss_run(spss_system *sys, int stepper, int method, long nsteps, double dt, int noise_mode);
/* No reference counterpart (the reference has no graphs): readies spss_run of the FP32C linked-list
 * Euler/SRK steps with Philox noise -- enters the cell-ordered resident copy of the state and
 * captures the HIP graphs every later run of this stepper and dt replays, from both buffer
 * parities, without taking a step.  A no-op for runs that step eagerly (other precisions, methods
 * or noise).  spss_run then replays every step from graphs (the first step of a run enters the
 * resident order eagerly when it is not there yet). */
int spss_prepare_run(spss_system *sys, int stepper, int method, double dt, int noise_mode);
/* No reference counterpart: steps launched one by one and steps replayed from graphs since the
 * last reset, and graphs captured since then; reset != 0 clears the counts after reading */
int spss_run_stats(spss_system *sys, long *eager_steps, long *replayed_steps, long *graphs_captured, int reset);
/* number of steps taken so far (the Philox step counter) */
int spss_get_step(spss_system *sys, long *step);
int spss_set_step(spss_system *sys, long step);

/* cell-size autotuner of the FP32C linked-list path (no reference counterpart): times
 * `probe_steps` stochastic Euler (stepper = SPSS_STEPPER_EULER) or Runge-Kutta steps for each
 * cell layout (k, xsplit) of SPSS_CELL_LAYOUTS (see spss_set_cell_layout) and keeps the fastest;
 * ms per step into table_ms[i] for layout i (-1 for a grid that is too fine; at most table_cap
 * entries), the chosen k into *chosen (spss_get_cell_layout gives both).  State, step counter and
 * complementary force are restored.  Neighbour sets do not depend on the layout; the coupling
 * sums change within the FP32C tolerance (candidate order). */
int spss_autotune_cells(spss_system *sys, int stepper, double dt, int probe_steps, double *table_ms, int table_cap,
                        int *chosen);
int spss_get_cell_subdivision(spss_system *sys, int *k);
/* The FP32C cell grid (no reference counterpart; the reference's grid is k = 1, xsplit = 1):
 * cells 1/(k*xsplit*int(1/rho)) wide and 1/(k*int(1/rho)) high, k in 1..3, xsplit 1, 2 or 4.  A
 * particle's candidates are the 2k+1 rows of 2k*xsplit+1 cells around its own: every pair within
 * rho still, fewer candidates beyond it at a narrower xsplit.  The layouts the autotuners time, in
 * table order: */
#define SPSS_CELL_LAYOUTS {{1, 1}, {2, 1}, {3, 1}, {1, 2}, {1, 4}}
#define SPSS_CELL_LAYOUT_COUNT 5
/* rebuilds the FP32C cell path with that grid (the state is kept; the next fused step re-enters
 * the cell order); neighbour sets do not depend on it, the coupling sums change within the FP32C
 * tolerance (candidate order) */
int spss_set_cell_layout(spss_system *sys, int k, int xsplit);
int spss_get_cell_layout(spss_system *sys, int *k, int *xsplit);
/* all-pairs autotuner (no reference counterpart): times `probe_evals` RHS evaluations of the current
 * state by EvaluateInteractionsWithAllPairs (ParticleSystemFirstOrderPtr.cpp:225-316) for each of
 * the kernel's workgroup / j-tile shapes {64,64}, {64,256}, {128,128}, {128,512}, {256,256},
 * {256,512} (particles per workgroup, j per LDS tile) and keeps the fastest for the handle's
 * all-pairs evaluations; ms per evaluation into table_ms (at most table_cap entries), the chosen
 * index into *chosen.  Every shape visits j in ascending order, so results do not depend on it. */
int spss_autotune_all_pairs(spss_system *sys, int probe_evals, double *table_ms, int table_cap, int *chosen);
int spss_get_all_pairs_config(spss_system *sys, int *block, int *jtile);
/* selects shape `config` (the index into the list above) by hand */
int spss_set_all_pairs_config(spss_system *sys, int config);

/* CalculateMaxDisplacement (ParticleSystemFirstOrderPtr.cpp:836-861) on a host total-displacement
 * vector, and the Verlet bookkeeping it updates */
int spss_calculate_max_displacement(spss_system *sys, const double *host_total_displacement);
int spss_verlet_status(spss_system *sys, int *should_update, double *accumulated_displacement,
                       int *uses_verlet_list);
/* the current Verlet lists: counts[n], then up to `cap` entries in list order */
int spss_get_verlet_lists(spss_system *sys, int *host_counts, int *host_flat, long cap, long *total);

/* neighbour set of one stage state (host, s*n doubles) found by `method`, as (i, j) pairs sorted
 * by (i, j): at most `cap` pairs into host_pairs[2*cap]; *count = number found.  With
 * host_pairs == NULL only the count is computed. */
int spss_neighbour_pairs(spss_system *sys, int method, const double *host_stage, int32_t *host_pairs,
                         long cap, long *count);

/* SaveSummaryStatistics' complex order parameter (BinaryObserverFirstOrderPtr.cpp:126-146),
 * accumulated in FP64 on the device: (1/N) sum exp(i * state[s*k + slot]) */
int spss_order_parameter(spss_system *sys, int slot, double *re, double *im);
/* ApplyPeriodicBoundaryConditions(Real*, long) on the device state (PeriodicBoundaryConditions.cpp:147-155) */
int spss_apply_periodic_boundary_conditions(spss_system *sys);

/* Instrumentation with no reference counterpart (the reference only prints wall-clock
 * seconds per output frame, BinaryObserverFirstOrderPtr.cpp:103-105): HIP events on the
 * handle's stream around every launch of the RHS pair kernel between start and stop. */
int spss_kernel_timer_start(spss_system *sys);
int spss_kernel_timer_stop(spss_system *sys, double *total_ms, long *launches);
/* attainable HBM bandwidth of a plain 16-byte-vector copy of `bytes` bytes, repeated, in GB/s
 * (read + write bytes / time); no reference counterpart */
int spss_device_copy_bandwidth(long bytes, int repeats, double *gbps);

/* The counter-based noise of the kernels, on the host (no reference counterpart: the reference
 * draws from std::mt19937, StochasticEulerStepperPtr.hpp:53-62).  One Philox4x32-10 block
 * (Salmon et al., SC'11) of counter ctr[4] under key key[2]; and the two standard normals the
 * kernels draw for particles [first, first + count) at `step` from counter block `block`
 * (Euler: z0 of block 0; SRK: U1 = z0, U2 = z1 of block 0; model 6: z0, z1 of block 0 and
 * z0 of block 1).  The uniforms are bit-identical to the device's; the Box-Muller transform
 * uses the host libm and agrees with the device to a few ulp. */
int spss_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int spss_philox_normals(uint64_t seed, uint64_t step, uint32_t block, long first, long count, double *z0,
                        double *z1);

/* ---- batches of independent replicas (no reference counterpart: the reference runs one
 * system per job and sweeps parameters in a host loop, SimulationEngineFirstOrder.cpp:105-113).
 * The replicas of a batch share the model (3 or 4), v0, rho, alpha and the cell grid; each has
 * its own particle count, phase friction, coupling, noise intensity and Philox seed.  FP32 cell-
 * list path (linked-list neighbours, stochastic Euler, Philox noise).  The state and the natural
 * frequencies are the replicas' arrays one after another (AoS FP64, reference layout).  Replica
 * r of a batch steps exactly like a single system (spss_create with precision FP32C, the same
 * parameters and seed) run through spss_run with SPSS_NOISE_PHILOX. */
typedef struct spss_batch spss_batch;
typedef struct spss_replica {
  int n;          /* particles */
  double xi;      /* phase friction (model 4) */
  double sigma;   /* coupling */
  double D_phi;   /* noise intensity */
  uint64_t seed;  /* Philox key */
} spss_replica;
typedef struct spss_batch_params {
  int model;            /* 3 or 4 */
  double v0;
  double rho;
  double alpha;
  int cell_subdivision; /* cells of side 1/(k*int(1/rho)); 0 -> 1 */
  int natfreq_local;    /* 1: natural frequencies in the local branches (model 4) */
} spss_batch_params;
int spss_batch_create(const spss_batch_params *bp, int nrep, const spss_replica *reps, spss_batch **out);
int spss_batch_destroy(spss_batch *b);
int spss_batch_set_state(spss_batch *b, const double *host_state);
int spss_batch_get_state(spss_batch *b, double *host_state);
int spss_batch_set_natural_frequencies(spss_batch *b, const double *host_w);
/* clear the time averages and set the step counter (the Philox step of the next step) */
int spss_batch_reset(spss_batch *b, long step);
/* `nsteps` stochastic Euler steps of every replica.  Every `observe_every` steps (0: never)
 * each replica's order parameter Z = <e^{i phi}> is read on the device; readings taken once the
 * step counter has reached `observe_from` enter the time averages of |Z| and |Z|^2.  use_graph:
 * the step sequence is captured in a HIP graph once and replayed. */
int spss_batch_run(spss_batch *b, long nsteps, double dt, long observe_every, long observe_from, int use_graph);
/* per replica: time-mean |Z|, its variance, the number of readings, the last Z; the step counter */
int spss_batch_order_stats(spss_batch *b, double *mean_R, double *var_R, long *count, double *last_re,
                           double *last_im, long *step);
/* tile width (cells per pair workgroup) of every replica */
int spss_batch_get_tile_widths(spss_batch *b, int *W);
int spss_batch_set_tile_widths(spss_batch *b, const int *W);
/* autotuner: times `probe_steps` steps for 6 tile-width scales (0.5x .. 2x the default
 * particles per workgroup), keeps the fastest; ms per step into table_ms[<= table_cap], the
 * chosen index into *chosen.  State and step counter are restored, the averages cleared. */
int spss_batch_autotune(spss_batch *b, double dt, int probe_steps, double *table_ms, int table_cap, int *chosen);
/* cell-size autotuner: times `probe_steps` steps for each cell layout of SPSS_CELL_LAYOUTS (default
 * tile widths), keeps the fastest; ms per step into table_ms[i] for layout i (at most table_cap
 * entries; -1 where a layout does not build), the chosen subdivision into *chosen
 * (spss_batch_get_cell_layout gives both).  State and step counter are restored, the averages
 * cleared.  Later results depend on the cell layout to the FP32C accuracy; a replica stays
 * bit-identical to a single FP32C system of the same layout. */
int spss_batch_autotune_cells(spss_batch *b, double dt, int probe_steps, double *table_ms, int table_cap,
                              int *chosen);
int spss_batch_get_cell_subdivision(spss_batch *b, int *k);
/* the batch's cell grid as spss_set_cell_layout (state and step counter kept, averages cleared) */
int spss_batch_set_cell_layout(spss_batch *b, int k, int xsplit);
int spss_batch_get_cell_layout(spss_batch *b, int *k, int *xsplit);

/* ---- the network of the learned sweep screen (no reference counterpart).  A fully connected
 * FP32 network of widths dims[0] -> ... -> dims[ndims-1] (at most 8 layers, widths <= 4096):
 * GELU (erf form) after every hidden layer, linear output, mean-squared-error loss, Adam.
 * Parameters are one flat float array, layer after layer: W_l [dims[l+1] x dims[l]] row-major,
 * then b_l.  Rows of inputs and targets are row-major float arrays.  All kernels sum in a fixed
 * order without atomics: results are reproducible bit for bit. */
typedef struct spss_mlp spss_mlp;
/* W and b ~ U(-1/sqrt(in), 1/sqrt(in)) from Philox4x32-10 keyed by `seed`; max_batch bounds the
 * rows of one forward/backward pass */
int spss_mlp_create(spss_mlp **out, const int *dims, int ndims, int max_batch, uint64_t seed);
int spss_mlp_destroy(spss_mlp *m);
int spss_mlp_param_count(const spss_mlp *m, long *count);
int spss_mlp_set_params(spss_mlp *m, const float *params, long count);
int spss_mlp_get_params(spss_mlp *m, float *params, long count);
/* outputs of n rows (any n; processed max_batch rows at a time) */
int spss_mlp_predict(spss_mlp *m, const float *x, long n, float *y);
/* gradient of the loss mean((y - t)^2) over batch x dims[last] outputs w.r.t. every parameter
 * (flat layout), and the loss */
int spss_mlp_gradients(spss_mlp *m, const float *x, const float *t, int batch, float *grad, float *loss);
/* training rows kept on the device; the trainer walks them in order, `batch` at a time, wrapping */
int spss_mlp_set_data(spss_mlp *m, const float *x, const float *t, long n);
/* nsteps Adam steps (torch.optim.Adam's update, no weight decay); the loss of each step's batch
 * (before its update) into losses[nsteps] (may be NULL); nsteps <= 65536 */
int spss_mlp_train(spss_mlp *m, int nsteps, int batch, float lr, float beta1, float beta2, float eps,
                   float *losses);
/* zero the Adam moments and step count */
int spss_mlp_reset_optimizer(spss_mlp *m);
int spss_mlp_adam_steps(const spss_mlp *m, long *steps);

/* ---- the reference's output files and restart.
 * A frame is a float time and the s*n state values as floats, as written by
 * BinaryObserverFirstOrderPtr::SaveSystemState (BinaryObserverFirstOrderPtr.cpp:108-121): for
 * models 3 and 4 x and y wrapped into the box by the float overload of
 * ApplyPeriodicBoundaryConditions (PeriodicBoundaryConditions.cpp:137-145) and phi wrapped by
 * phi -= floor(phi / 2 pi) * 2 pi, both in double on the float value; model 6 the raw float copy
 * (BinaryObserverSecondOrderPtr.cpp:108-113).  spss_frame fills host_frame[0] with t and
 * host_frame[1 .. s*n] with the state; spss_frame_append appends the frame to a binary file. */
int spss_frame(spss_system *sys, double t, float *host_frame);
int spss_frame_append(spss_system *sys, const char *path, double t);
/* SaveSummaryStatistics (BinaryObserverFirstOrderPtr.cpp:126-146; model 6 phi in slot 4,
 * BinaryObserverSecondOrderPtr.cpp:175-193): Z = sum_i ((float) cos phi_i, (float) sin phi_i)
 * accumulated in complex<float> in particle order, divided by n; |Z| and arg Z as floats.
 * spss_summary_append appends the line "t \t |Z| \t arg Z" (default stream formatting, %g). */
int spss_summary(spss_system *sys, float *abs_z, float *arg_z);
int spss_summary_append(spss_system *sys, const char *path, double t);
/* SimulationEngineFirstOrderPtr::InitializeSystemStateFromPreviousSolution
 * (SimulationEngineFirstOrderPtr.cpp:70-94): the last complete frame of a file of frames of
 * s*n values; its floats widened to double.  spss_load_last_frame sets the system's state
 * from it and returns its time. */
int spss_read_last_frame(const char *path, int s, int n, float *t, double *state);
int spss_load_last_frame(spss_system *sys, const char *path, double *t);

/* AveragingObserver::operator() and the sum behind SaveAveragedSystemState (Observers/AveragingObserver.cpp:29-37,
 * 39-68; AveragingObserverPtr.cpp:46-79): running FP64 sum of the state and the step count */
int spss_averaging_reset(spss_system *sys);
int spss_averaging_accumulate(spss_system *sys);
int spss_averaging_get(spss_system *sys, double *host_sum, long *count);

/* Adaptive Dormand-Prince 5(4) with dense output, the reference's Boost.Odeint calls
 * (SimulationEngineFirstOrder.cpp:158-165: integrate_const(make_dense_output(1e-6, 1e-6,
 * runge_kutta_dopri5<std::vector<Real>>()), particle_system, state, t_0, t_1, dt, observer);
 * :273-281: integrate_adaptive(make_dense_output(1e-10, 1e-10, ...), particle_system, state, t_0,
 * t_1, dt)).  The system is the reference's functor (ParticleSystemFirstOrder.cpp:114-120):
 * EvaluateRhs(state, zeros, derivative, 0, t) with `method` (SPSS_METHOD_AUTO: the reference's
 * dispatch).  SPSS_DOPRI5_CONST observes t_0 + j dt (j = 0, 1, ... while <= t_1) from the dense
 * output and leaves the last observed state as the system's state; SPSS_DOPRI5_ADAPTIVE observes
 * t_0 and every accepted step and leaves the state at t_1.  The step control is Boost's
 * default_error_checker (a_x = a_dxdt = 1) and default_step_adjuster; 500 rejected trials in a
 * row are an error, as Boost's failed_step_checker throws.  The observer gets the state
 * (s*n doubles, AoS) as a host pointer (observe_on_host = 1) or a device pointer, valid for the
 * call; during the call the system's own state is the observed state (so spss_frame_append and
 * spss_summary_append write it); a nonzero return stops the integration with SPSS_ERR_INVALID.
 * steps / rejected (may be NULL) count accepted steps and rejected trials. */
#define SPSS_DOPRI5_CONST 0
#define SPSS_DOPRI5_ADAPTIVE 1
typedef int (*spss_observer_fn)(void *user, double t, const double *state, int n, int s);
int spss_dopri5_integrate(spss_system *sys, int method, int mode, double t0, double t1, double dt, double eps_abs,
                          double eps_rel, spss_observer_fn observer, void *user, int observe_on_host, long *steps,
                          long *rejected);

#ifdef __cplusplus
}
#endif
#endif /* SPSS_HIP_H */
