// spss_reference_api.hpp -- the reference's C++ class interface on top of the C ABI of
// libspss_hip.so (include/spss_hip.h).  Header-only; link with -lspss_hip.
//
// A host program written against the reference classes
//   Thread / ThreadSharedMemory            (Parallelization/Thread.hpp, ThreadSharedMemory.hpp)
//   PeriodicBoundaryConditions             (BoundaryConfiguration/PeriodicBoundaryConditions.hpp:12-48)
//   ParticleSystemFirstOrderPtr            (DynamicalSystems/ParticleSystemFirstOrderPtr.hpp:12-98)
//   ParticleSystemSecondOrderPtr           (DynamicalSystems/ParticleSystemSecondOrderPtr.hpp:16-37)
//   RungeKutta4StepperPtr                  (Steppers/RungeKutta4StepperPtr.hpp:15-35)
//   StochasticEulerStepperPtr              (Steppers/StochasticEulerStepperPtr.hpp:22-73)
//   StochasticRungeKuttaStepperPtr         (Steppers/StochasticRungeKuttaStepperPtr.hpp:22-76)
//   AveragingObserverPtr                   (Observers/AveragingObserverPtr.hpp:13-40)
//   SimulationEngineFirstOrderPtr          (ControlEngines/SimulationEngineFirstOrderPtr.hpp:15-40)
// compiles against this header unchanged: same class names, constructor arguments and
// method signatures.  The particle system owns one device handle; EvaluateRhs and the
// steppers run on the GPU.
//
// Host memory semantics.  The reference keeps the state in host memory and a stepper updates
// it in place.  Here every DoStep uploads the caller's state, steps on the device and writes
// the result back (exact drop-in semantics, PCIe-bound).  A caller that does not read the
// host state every step calls SetHostSync(false) on the system: DoStep then keeps the state
// on the device and DownloadState() refreshes the host copy on demand (the engine below does
// this, reading the order parameter on the device).
//
// Errors: the reference has none (void returns, asserts); here a failing call throws
// std::runtime_error with spss_last_error().
//
// Noise: the stochastic steppers draw from the library's Philox4x32-10 stream (seed, step,
// particle), not from the reference's std::mt19937; DoStep overloads taking the normals
// reproduce the reference's draws when they are supplied.
#pragma once

#include <cmath>
#include <complex>
#include <cstdio>
#include <iomanip>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "spss_hip.h"

typedef double Real;
typedef float RealOutput;

namespace spss_shim {

inline void check(int rc) {
  if (rc != SPSS_OK) throw std::runtime_error(std::string("spss: ") + spss_last_error());
}

// options that the reference fixes in code (method dispatch, precision, seed)
struct Options {
  int method = SPSS_METHOD_AUTO;  // EvaluateRhs's dispatch on rho unless set
  int precision = SPSS_PRECISION_FP64;
  int verlet_mode = SPSS_VERLET_COMPAT;
  double verlet_skin = 0.05;
  int natfreq_local = 0;  // the reference applies natural frequencies in the global branch only
  int cell_subdivision = 1;
  uint64_t seed = 0x5EED0000ull;
  // where the binary observers write; the reference's folder is fixed per machine
  // (BinaryObserverFirstOrderPtr.cpp:42-52)
  std::string output_folder = "./";
};

inline Options &default_options() {
  static Options o;
  return o;
}

}  // namespace spss_shim

// ---------------------------------------------------------------------------
// one device: rank 0 of one, every particle owned (Thread.cpp:22-23)
class Thread {
 public:
  Thread(int /*argc*/, char ** /*argv*/, int number_of_particles, int number_of_state_variables)
      : n_(number_of_particles), s_(number_of_state_variables), loop_indices_(number_of_particles) {
    for (int i = 0; i < n_; ++i) loop_indices_[i] = i;
  }
  virtual ~Thread() = default;
  int GetRank() { return 0; }
  int GetNumberOfMpichThreads() { return 1; }
  int GetNumberOfParticlesPerMpichThread() { return n_; }
  const std::vector<int> &GetLoopIndices() const { return loop_indices_; }
  virtual bool IsRoot() const { return true; }
  virtual void SynchronizeVector(std::vector<Real> &) {}
  virtual void SynchronizeVectorThoughBuffer(std::vector<Real> &, std::vector<Real> &) {}
  virtual void BroadcastVector(std::vector<Real> &) {}
  virtual void SynchronizePrelinkedList(std::vector<int> &) const {}

 protected:
  int n_, s_;
  std::vector<int> loop_indices_;
};

class ThreadSharedMemory : public Thread {
 public:
  ThreadSharedMemory(int argc, char **argv, int number_of_particles, int number_of_state_variables)
      : Thread(argc, argv, number_of_particles, number_of_state_variables) {}
  int GetRootRank() { return 0; }
  int GetSharedRootRank() { return 0; }
  bool IsRoot() const override { return true; }
  bool IsSharedRoot() { return true; }
  void SynchronizeVector(std::vector<Real> &) override {}
  void SynchronizeVector(Real *const, long) {}
  void BroadcastVector(std::vector<Real> &) override {}
  void BroadcastVector(Real *const, long) {}
};

// ---------------------------------------------------------------------------
// PeriodicBoundaryConditions.cpp:7-165: scalar helpers for host code (observers, output);
// the kernels carry their own copies of these formulas
class PeriodicBoundaryConditions {
 public:
  PeriodicBoundaryConditions(Thread *thread, int number_of_particles, int number_of_state_variables, Real x_size,
                             Real y_size)
      : thread_(thread), n_(number_of_particles), s_(number_of_state_variables), x_size_(x_size), y_size_(y_size),
        x_rsize_(1.0 / x_size), y_rsize_(1.0 / y_size) {}

  void ClassAEffectiveParticleDistance(Real x_i, Real y_i, Real x_j, Real y_j, Real &dx, Real &dy) const {
    dx = x_j - x_i;
    dx -= static_cast<int>(dx * 2.0 * x_rsize_) * x_size_;
    dy = y_j - y_i;
    dy -= static_cast<int>(dy * 2.0 * y_rsize_) * y_size_;
  }
  void ClassCEffectiveParticleDistanceUnsigned(Real x_i, Real y_i, Real x_j, Real y_j, Real &dx, Real &dy) const {
    dx = std::fabs(x_j - x_i);
    dx -= static_cast<int>(dx * x_rsize_ + 0.5) * x_size_;
    dy = std::fabs(y_j - y_i);
    dy -= static_cast<int>(dy * y_rsize_ + 0.5) * y_size_;
  }
  void ApplyPeriodicBoundaryConditions(Real x, Real y, Real &x_pbc, Real &y_pbc) const {
    x_pbc = x - std::floor(x * x_rsize_) * x_size_;
    y_pbc = y - std::floor(y * y_rsize_) * y_size_;
  }
  // wraps positions of a host state vector (output copies, PeriodicBoundaryConditions.cpp:127-155)
  template <class T>
  void ApplyPeriodicBoundaryConditions(std::vector<T> &system_state) {
    for (int i = 0; i < n_; ++i) {
      const long ii = (long)s_ * i;
      system_state[ii] -= std::floor(system_state[ii] * x_rsize_) * x_size_;
      system_state[ii + 1] -= std::floor(system_state[ii + 1] * y_rsize_) * y_size_;
    }
  }
  Real GetXSize() const { return x_size_; }
  Real GetYSize() const { return y_size_; }

 private:
  Thread *thread_;
  int n_, s_;
  Real x_size_, y_size_, x_rsize_, y_rsize_;
};

// ---------------------------------------------------------------------------
namespace spss_shim {

// one device handle and the host-side flags the reference exposes as public members
class DeviceSystem {
 public:
  DeviceSystem(const spss_params &p) : p_(p) { check(spss_create(&p_, &h_)); }
  ~DeviceSystem() { spss_destroy(h_); }
  DeviceSystem(const DeviceSystem &) = delete;
  DeviceSystem &operator=(const DeviceSystem &) = delete;

  spss_system *handle() const { return h_; }
  int n() const { return p_.n; }
  int s() const { return p_.model; }
  int method() const { return method_; }
  void SetMethod(int m) { method_ = m; }
  void SetHostSync(bool on) { host_sync_ = on; }
  bool host_sync() const { return host_sync_; }

  // k_next entries left untouched by the RHS (k[3] of isolated particles) keep their values
  void Rhs(int method, const Real *state, const std::vector<Real> &k_prev, std::vector<Real> &k_next, Real k_coef,
           Real dt, bool higher_order, bool &uses_verlet) {
    check(spss_set_higher_order(h_, higher_order ? 1 : 0));
    check(spss_evaluate_rhs(h_, method, state, k_prev.data(), k_next.data(), k_coef, dt));
    int flag = 0;
    check(spss_verlet_status(h_, nullptr, nullptr, &flag));
    uses_verlet = flag != 0;
  }
  void Upload(const Real *state) {
    if (host_sync_ || !resident_) check(spss_set_state(h_, state));
    resident_ = true;
  }
  void Download(Real *state) {
    if (host_sync_) check(spss_get_state(h_, state));
  }
  void DownloadState(Real *state) { check(spss_get_state(h_, state)); }

 private:
  spss_params p_;
  spss_system *h_ = nullptr;
  int method_ = SPSS_METHOD_AUTO;
  bool host_sync_ = true;
  bool resident_ = false;
};

inline spss_params make_params(int model, int n, Real v0, Real xi, Real sigma, Real rho, Real alpha, Real D_phi,
                               Real xi_r, Real D_r, const PeriodicBoundaryConditions &pbc) {
  spss_params p;
  check(spss_params_default(&p));
  const Options &o = default_options();
  p.model = model;
  p.n = n;
  p.v0 = v0;
  p.xi = xi;
  p.sigma = sigma;
  p.rho = rho;
  p.alpha = alpha;
  p.D_phi = D_phi;
  p.xi_r = xi_r;
  p.D_r = D_r;
  p.x_size = pbc.GetXSize();
  p.y_size = pbc.GetYSize();
  p.verlet_skin = o.verlet_skin;
  p.verlet_mode = o.verlet_mode;
  p.natfreq_local = o.natfreq_local;
  p.precision = o.precision;
  p.cell_subdivision = o.cell_subdivision;
  p.seed = o.seed;
  return p;
}

}  // namespace spss_shim

// ---------------------------------------------------------------------------
// ParticleSystemFirstOrderPtr.hpp:12-98 (s = 4; s = 3 is the first-order phase model)
class ParticleSystemFirstOrderPtr {
 public:
  ParticleSystemFirstOrderPtr(ThreadSharedMemory *thread, Real v_0, Real xi, Real sigma, Real rho, Real alpha,
                              Real D_phi, PeriodicBoundaryConditions &pbc_config, int number_of_particles,
                              int number_of_state_variables)
      : uses_verlet_list_(false), is_higher_order_sde_integration_(false), thread_(thread),
        pbc_config_(pbc_config), xi_(xi), rho_(rho), alpha_(alpha), D_phi_(D_phi),
        dev_(spss_shim::make_params(number_of_state_variables, number_of_particles, v_0, xi, sigma, rho, alpha,
                                    D_phi, 1.0, 0.0, pbc_config)) {
    dev_.SetMethod(spss_shim::default_options().method);
  }

  void SetNaturalFrequencies(const std::vector<Real> &natural_frequencies) {
    spss_shim::check(spss_set_natural_frequencies(dev_.handle(), natural_frequencies.data(),
                                                  (int)natural_frequencies.size()));
  }

  void EvaluateRhs(const Real *const system_state, const std::vector<Real> &k_prev, std::vector<Real> &k_next,
                   Real k_coef, Real dt) {
    Eval(dev_.method(), system_state, k_prev, k_next, k_coef, dt);
  }
  void EvaluateInteractionsWithGlobalPolarOrderField(const Real *const s, const std::vector<Real> &kp,
                                                     std::vector<Real> &kn, Real c, Real dt) {
    Eval(SPSS_METHOD_GLOBAL_FIELD, s, kp, kn, c, dt);
  }
  void EvaluateInteractionsWithAllPairs(const Real *const s, const std::vector<Real> &kp, std::vector<Real> &kn,
                                        Real c, Real dt) {
    Eval(SPSS_METHOD_ALL_PAIRS, s, kp, kn, c, dt);
  }
  void EvaluateInteractionsWithLinkedList(const Real *const s, const std::vector<Real> &kp, std::vector<Real> &kn,
                                          Real c, Real dt) {
    Eval(SPSS_METHOD_LINKED_LIST, s, kp, kn, c, dt);
  }
  void EvaluateInteractionsWithVerletNeighborList(const Real *const s, const std::vector<Real> &kp,
                                                  std::vector<Real> &kn, Real c, Real dt) {
    Eval(SPSS_METHOD_VERLET, s, kp, kn, c, dt);
  }
  void EvaluateInteractionsWithLinkedListAndVerletNeighborList(const Real *const s, const std::vector<Real> &kp,
                                                               std::vector<Real> &kn, Real c, Real dt) {
    Eval(SPSS_METHOD_LINKED_LIST_VERLET, s, kp, kn, c, dt);
  }

  Real xi() const { return xi_; }
  Real rho() const { return rho_; }
  Real D_phi() const { return D_phi_; }
  Real alpha() const { return alpha_; }

  bool uses_verlet_list_;
  void CalculateMaxDisplacement(const std::vector<Real> &total_displacement) {
    spss_shim::check(spss_calculate_max_displacement(dev_.handle(), total_displacement.data()));
  }
  bool is_higher_order_sde_integration_;
  const std::vector<Real> &GetComplimentaryAlignmentForce() const {
    compl_.resize(dev_.n());
    spss_shim::check(spss_get_complementary_force(dev_.handle(), compl_.data()));
    return compl_;
  }

  // beyond the reference: device residency and method choice
  spss_shim::DeviceSystem &device() { return dev_; }
  void SetHostSync(bool on) { dev_.SetHostSync(on); }
  void SetMethod(int method) { dev_.SetMethod(method); }
  void DownloadState(Real *system_state) { dev_.DownloadState(system_state); }
  std::complex<Real> OrderParameter() {
    double re = 0.0, im = 0.0;
    spss_shim::check(spss_order_parameter(dev_.handle(), dev_.s() == 6 ? 4 : 2, &re, &im));
    return {re, im};
  }

 private:
  void Eval(int method, const Real *s, const std::vector<Real> &kp, std::vector<Real> &kn, Real c, Real dt) {
    dev_.Rhs(method, s, kp, kn, c, dt, is_higher_order_sde_integration_, uses_verlet_list_);
  }

  ThreadSharedMemory *thread_;
  PeriodicBoundaryConditions &pbc_config_;
  Real xi_, rho_, alpha_, D_phi_;
  spss_shim::DeviceSystem dev_;
  mutable std::vector<Real> compl_;
};

// ParticleSystemSecondOrderPtr.hpp:16-37 (s = 6)
class ParticleSystemSecondOrderPtr {
 public:
  ParticleSystemSecondOrderPtr(ThreadSharedMemory *thread, Real v_0, Real xi_r, Real D_r, Real xi_phi, Real D_phi,
                               Real sigma, Real rho, Real alpha, PeriodicBoundaryConditions &pbc_config,
                               int number_of_particles, int number_of_state_variables)
      : thread_(thread), pbc_config_(pbc_config), rho_(rho), alpha_(alpha), D_phi_(D_phi), D_r_(D_r),
        dev_(spss_shim::make_params(number_of_state_variables, number_of_particles, v_0, xi_phi, sigma, rho, alpha,
                                    D_phi, xi_r, D_r, pbc_config)) {
    dev_.SetMethod(spss_shim::default_options().method);
  }
  void EvaluateRhs(const Real *const s, const std::vector<Real> &kp, std::vector<Real> &kn, Real c, Real dt) {
    bool v = false;
    dev_.Rhs(dev_.method(), s, kp, kn, c, dt, false, v);
  }
  void EvaluateInteractionsWithAllPairs(const Real *const s, const std::vector<Real> &kp, std::vector<Real> &kn,
                                        Real c, Real dt) {
    bool v = false;
    dev_.Rhs(SPSS_METHOD_ALL_PAIRS, s, kp, kn, c, dt, false, v);
  }
  void EvaluateInteractionsWithLinkedList(const Real *const s, const std::vector<Real> &kp, std::vector<Real> &kn,
                                          Real c, Real dt) {
    bool v = false;
    dev_.Rhs(SPSS_METHOD_LINKED_LIST, s, kp, kn, c, dt, false, v);
  }
  Real rho() const { return rho_; }
  Real alpha() const { return alpha_; }
  Real D_phi() const { return D_phi_; }
  Real D_r() const { return D_r_; }

  spss_shim::DeviceSystem &device() { return dev_; }
  void SetHostSync(bool on) { dev_.SetHostSync(on); }
  void SetMethod(int method) { dev_.SetMethod(method); }
  void DownloadState(Real *system_state) { dev_.DownloadState(system_state); }

 private:
  ThreadSharedMemory *thread_;
  PeriodicBoundaryConditions &pbc_config_;
  Real rho_, alpha_, D_phi_, D_r_;
  spss_shim::DeviceSystem dev_;
};

// ---------------------------------------------------------------------------
// steppers: the step runs on the device (the k-buffers live in the handle)
namespace spss_shim {
template <class System>
inline void stochastic_step(System &sys, Real *const state, Real dt, bool srk, const std::vector<Real> *normals) {
  DeviceSystem &d = sys.device();
  d.Upload(state);
  const int mode = normals ? SPSS_NOISE_HOST : SPSS_NOISE_PHILOX;
  const double *z = normals ? normals->data() : nullptr;
  if (srk)
    check(spss_srk_step(d.handle(), d.method(), dt, mode, z));
  else
    check(spss_euler_step(d.handle(), d.method(), dt, mode, z));
  d.Download(state);
}
}  // namespace spss_shim

// RungeKutta4StepperPtr.hpp:15-35
class RungeKutta4StepperPtr {
 public:
  RungeKutta4StepperPtr(ThreadSharedMemory *thread, int number_of_particles, int number_of_state_variables)
      : thread_(thread), n_(number_of_particles), s_(number_of_state_variables) {}
  template <class System>
  void DoStep(System &particle_system, Real *const system_state, Real /*t*/, Real dt) {
    spss_shim::DeviceSystem &d = particle_system.device();
    d.Upload(system_state);
    spss_shim::check(spss_rk4_step(d.handle(), d.method(), dt));
    d.Download(system_state);
  }

 private:
  ThreadSharedMemory *thread_;
  int n_, s_;
};

// StochasticEulerStepperPtr.hpp:22-73; normals: one per particle (s = 3, 4), three for s = 6
class StochasticEulerStepperPtr {
 public:
  StochasticEulerStepperPtr(ThreadSharedMemory *thread, int number_of_particles, int number_of_state_variables)
      : thread_(thread), n_(number_of_particles), s_(number_of_state_variables) {}
  template <class System>
  void DoStep(System &particle_system, Real *const system_state, Real /*t*/, Real dt) {
    spss_shim::stochastic_step(particle_system, system_state, dt, false, nullptr);
  }
  template <class System>
  void DoStep(System &particle_system, Real *const system_state, Real /*t*/, Real dt,
              const std::vector<Real> &normals) {
    spss_shim::stochastic_step(particle_system, system_state, dt, false, &normals);
  }

 private:
  ThreadSharedMemory *thread_;
  int n_, s_;
};

// StochasticRungeKuttaStepperPtr.hpp:22-76; normals: (U1, U2) per particle
class StochasticRungeKuttaStepperPtr {
 public:
  StochasticRungeKuttaStepperPtr(ThreadSharedMemory *thread, int number_of_particles, int number_of_state_variables)
      : thread_(thread), n_(number_of_particles), s_(number_of_state_variables) {}
  void DoStep(ParticleSystemFirstOrderPtr &particle_system, Real *const system_state, Real /*t*/, Real dt) {
    particle_system.is_higher_order_sde_integration_ = true;  // StochasticRungeKuttaStepperPtr.hpp:44
    spss_shim::stochastic_step(particle_system, system_state, dt, true, nullptr);
  }
  void DoStep(ParticleSystemFirstOrderPtr &particle_system, Real *const system_state, Real /*t*/, Real dt,
              const std::vector<Real> &normals) {
    particle_system.is_higher_order_sde_integration_ = true;
    spss_shim::stochastic_step(particle_system, system_state, dt, true, &normals);
  }

 private:
  ThreadSharedMemory *thread_;
  int n_, s_;
};

// ---------------------------------------------------------------------------
// AveragingObserverPtr.hpp:13-40: running sum of the state on the device of the system it
// observes (Bind), or of the host state it is given (uploaded through the bound system)
class AveragingObserverPtr {
 public:
  AveragingObserverPtr(ThreadSharedMemory *thread, int number_of_particles, int number_of_state_variables)
      : thread_(thread), n_(number_of_particles), s_(number_of_state_variables) {}
  void Bind(ParticleSystemFirstOrderPtr &system) {
    dev_ = &system.device();
    spss_shim::check(spss_averaging_reset(dev_->handle()));
  }
  // with host sync the device already holds system_state; without it the device state is the
  // current one by construction
  void operator()(Real *const /*system_state*/, long /*size*/, Real /*t*/) {
    if (!dev_) throw std::runtime_error("spss: AveragingObserverPtr needs Bind(particle_system)");
    spss_shim::check(spss_averaging_accumulate(dev_->handle()));
  }
  // time-averaged state (sum / count)
  std::vector<Real> AveragedSystemState() const {
    std::vector<Real> sum((size_t)n_ * s_);
    long count = 0;
    spss_shim::check(spss_averaging_get(dev_->handle(), sum.data(), &count));
    for (Real &v : sum) v /= (count > 0 ? (Real)count : 1.0);
    return sum;
  }

 private:
  ThreadSharedMemory *thread_;
  int n_, s_;
  spss_shim::DeviceSystem *dev_ = nullptr;
};

// ---------------------------------------------------------------------------
// BinaryObserverFirstOrderPtr.cpp:19-146 and BinaryObserverSecondOrderPtr.cpp:16-193: the
// trajectory file (every `threshold`-th SaveSystemState call: a float time and the float state,
// models 3/4 wrapped into the box and phi into [0, 2 pi)) and the summary file (every
// threshold-th SaveSummaryStatistics call: "t \t |Z| \t arg Z"), named from the parameters as
// the reference names them, in default_options().output_folder.  The frames and sums come from
// the device state of the bound particle system (Bind); the host pointer arguments are kept for
// the reference's signatures.
namespace spss_shim {
class BinaryFiles {
 public:
  BinaryFiles(const std::string &sim, const std::string &summary, int threshold)
      : sim_(sim), summary_(summary), threshold_(threshold) {
    std::remove(sim_.c_str());
    std::remove(summary_.c_str());
  }
  void Bind(DeviceSystem &dev) { dev_ = &dev; }
  void SaveSystemState(Real t) {
    if (!(counter_[0]++ % threshold_)) check(spss_frame_append(device().handle(), sim_.c_str(), t));
  }
  void SaveSummaryStatistics(Real t) {
    if (!(counter_[1]++ % threshold_)) check(spss_summary_append(device().handle(), summary_.c_str(), t));
  }
  const std::string &simulation_file_name() const { return sim_; }
  const std::string &summary_statistics_file_name() const { return summary_; }

 private:
  DeviceSystem &device() const {
    if (!dev_) throw std::runtime_error("spss: the binary observer needs Bind(particle_system)");
    return *dev_;
  }
  std::string sim_, summary_;
  int threshold_;
  long counter_[2] = {0, 0};
  DeviceSystem *dev_ = nullptr;
};
}  // namespace spss_shim

class BinaryObserverFirstOrderPtr {
 public:
  BinaryObserverFirstOrderPtr(ThreadSharedMemory *thread, Real v_0, Real xi, Real sigma, Real rho, Real alpha,
                              Real D_phi, PeriodicBoundaryConditions &pbc_config, int number_of_particles,
                              int number_of_state_variables, Real dt, int trial)
      : files_(name(v_0, xi, sigma, rho, alpha, D_phi, number_of_particles, trial, "", ".bin"),
               name(v_0, xi, sigma, rho, alpha, D_phi, number_of_particles, trial, "", ".txt"), 100) {
    (void)thread, (void)pbc_config, (void)number_of_state_variables, (void)dt;
  }
  void Bind(ParticleSystemFirstOrderPtr &system) { files_.Bind(system.device()); }
  void SaveSystemState(Real *const /*system_state*/, long /*size*/, Real t) { files_.SaveSystemState(t); }
  void SaveSummaryStatistics(const Real *const /*system_state*/, long /*size*/, Real t) {
    files_.SaveSummaryStatistics(t);
  }
  const std::string &simulation_file_name() const { return files_.simulation_file_name(); }
  const std::string &summary_statistics_file_name() const { return files_.summary_statistics_file_name(); }

 private:
  static std::string name(Real v_0, Real xi, Real sigma, Real rho, Real alpha, Real D_phi, int n, int trial,
                          const char *prefix, const char *ext) {
    std::ostringstream b;  // BinaryObserverFirstOrderPtr.cpp:54-60, 70-76
    b << spss_shim::default_options().output_folder << prefix << "v0_" << v_0 << "_xi_" << xi << "_sigma_" << sigma
      << "_rho_" << rho << "_alpha_" << alpha << "_Dphi_" << D_phi << "_N_" << n << "_" << 0 << "_" << trial << ext;
    return b.str();
  }
  spss_shim::BinaryFiles files_;
};

class BinaryObserverSecondOrderPtr {
 public:
  BinaryObserverSecondOrderPtr(ThreadSharedMemory *thread, Real v_0, Real xi_r, Real D_r, Real xi_phi, Real D_phi,
                               Real sigma, Real rho, Real alpha, PeriodicBoundaryConditions &pbc_config,
                               int number_of_particles, int number_of_state_variables, Real dt, int trial)
      : files_(name(v_0, xi_r, D_r, xi_phi, D_phi, sigma, rho, alpha, number_of_particles, trial, "", ".bin"),
               name(v_0, xi_r, D_r, xi_phi, D_phi, sigma, rho, alpha, number_of_particles, trial,
                    "summary_statistics_", ".txt"),
               10) {
    (void)thread, (void)pbc_config, (void)number_of_state_variables, (void)dt;
  }
  void Bind(ParticleSystemSecondOrderPtr &system) { files_.Bind(system.device()); }
  void SaveSystemState(Real *const /*system_state*/, long /*size*/, Real t) { files_.SaveSystemState(t); }
  void SaveSummaryStatistics(const Real *const /*system_state*/, long /*size*/, Real t) {
    files_.SaveSummaryStatistics(t);
  }
  const std::string &simulation_file_name() const { return files_.simulation_file_name(); }
  const std::string &summary_statistics_file_name() const { return files_.summary_statistics_file_name(); }

 private:
  static std::string name(Real v_0, Real xi_r, Real D_r, Real xi_phi, Real D_phi, Real sigma, Real rho, Real alpha,
                          int n, int trial, const char *prefix, const char *ext) {
    std::ostringstream b;  // BinaryObserverSecondOrderPtr.cpp:51-57, 66-72
    b << spss_shim::default_options().output_folder << prefix << "v0_" << v_0 << "_xir_" << xi_r << "_Dr_" << D_r
      << "_xiphi_" << xi_phi << "_Dphi_" << D_phi << "_sigma_" << sigma << "_rho_" << rho << "_alpha_" << alpha
      << "_N_" << n << "_" << 0 << "_" << trial << ext;
    return b.str();
  }
  spss_shim::BinaryFiles files_;
};

// ---------------------------------------------------------------------------
// BinaryObserverFirstOrder.cpp:18-82,101-148 (the serial observer that Boost.Odeint calls): the
// same frames and summary lines as BinaryObserverFirstOrderPtr, every 100th call each, under the
// serial class's file names (alpha and what follows it in the trajectory name printed with
// setprecision(10)); operator()(state, t) saves both.  Frames and sums come from the device state
// of the bound system, which during an integration's observer call is the observed state.
class BinaryObserverFirstOrder {
 public:
  BinaryObserverFirstOrder(Thread *thread, Real v_0, Real xi, Real sigma, Real rho, Real alpha, Real D_phi,
                           PeriodicBoundaryConditions &pbc_config, int number_of_particles,
                           int number_of_state_variables, Real dt, int trial)
      : files_(name(v_0, xi, sigma, rho, alpha, D_phi, number_of_particles, trial, true),
               name(v_0, xi, sigma, rho, alpha, D_phi, number_of_particles, trial, false), 100) {
    (void)thread, (void)pbc_config, (void)number_of_state_variables, (void)dt;
  }
  void Bind(spss_shim::DeviceSystem &dev) { files_.Bind(dev); }
  void SaveSystemState(const std::vector<Real> & /*system_state*/, Real t) { files_.SaveSystemState(t); }
  void SaveSummaryStatistics(const std::vector<Real> & /*system_state*/, Real t) { files_.SaveSummaryStatistics(t); }
  void operator()(Real t) {
    files_.SaveSystemState(t);
    files_.SaveSummaryStatistics(t);
  }
  const std::string &simulation_file_name() const { return files_.simulation_file_name(); }
  const std::string &summary_statistics_file_name() const { return files_.summary_statistics_file_name(); }

 private:
  static std::string name(Real v_0, Real xi, Real sigma, Real rho, Real alpha, Real D_phi, int n, int trial,
                          bool trajectory) {
    std::ostringstream b;  // BinaryObserverFirstOrder.cpp:50-60, 66-76
    b << spss_shim::default_options().output_folder << "v0_" << v_0 << "_xi_" << xi << "_sigma_" << sigma
      << "_rho_" << rho << "_alpha_";
    if (trajectory) b << std::setprecision(10);
    b << alpha << "_Dphi_" << D_phi << "_N_" << n << "_" << 0 << "_" << trial << (trajectory ? ".bin" : ".txt");
    return b.str();
  }
  spss_shim::BinaryFiles files_;
};

// SimulationEngineFirstOrder.cpp:96-165 (the serial engine): random initial state and natural
// frequencies ~ N(0, 0.5) (:35-64), then Boost.Odeint's integrate_const(make_dense_output(1e-6,
// 1e-6, runge_kutta_dopri5<vector<Real>>()), particle_system, state, t_0, t_1, dt,
// boost::ref(binary_observer)) as spss_dopri5_integrate(SPSS_DOPRI5_CONST) on the device, the
// observer writing the reference's files at every 100th observation.
class SimulationEngineFirstOrder {
 public:
  SimulationEngineFirstOrder(Thread *thread, int number_of_particles, int number_of_state_variables)
      : thread_(thread), pbc_config_(thread, number_of_particles, number_of_state_variables, 1.0, 1.0),
        n_(number_of_particles), s_(number_of_state_variables),
        system_state_((size_t)number_of_particles * number_of_state_variables, 0.0),
        natural_frequencies_(number_of_particles, 0.0) {}

  struct Run {
    Real v_0 = 1.0, xi = 0.1, sigma = 5.0, rho = 0.8, alpha = 0.0, D_phi = 0.0;  // :115-120
    Real t_0 = 0.0, t_1 = 10000.0, dt = 0.01;                                   // :123-125
    Real eps_abs = 1.0e-6, eps_rel = 1.0e-6;                                    // :159
    unsigned seed = 0;
    int trial = 0;
    int method = SPSS_METHOD_AUTO;  // the reference's EvaluateRhs dispatch
  };

  void RunSimulation() { RunSimulation(Run{}); }
  void RunSimulation(const Run &r) {
    InitializeRandomSystemState(r.sigma, r.xi, r.seed);
    initial_state_ = system_state_;
    spss_shim::DeviceSystem dev(
        spss_shim::make_params(s_, n_, r.v_0, r.xi, r.sigma, r.rho, r.alpha, r.D_phi, 1.0, 0.0, pbc_config_));
    spss_shim::check(spss_set_natural_frequencies(dev.handle(), natural_frequencies_.data(), n_));
    spss_shim::check(spss_set_state(dev.handle(), system_state_.data()));
    BinaryObserverFirstOrder observer(thread_, r.v_0, r.xi, r.sigma, r.rho, r.alpha, r.D_phi, pbc_config_, n_, s_,
                                      r.dt, r.trial);
    observer.Bind(dev);
    files_ = {observer.simulation_file_name(), observer.summary_statistics_file_name()};
    spss_shim::check(spss_dopri5_integrate(dev.handle(), r.method, SPSS_DOPRI5_CONST, r.t_0, r.t_1, r.dt, r.eps_abs,
                                           r.eps_rel, &observe, &observer, 0, &steps_, &rejected_));
    dev.DownloadState(system_state_.data());
  }
  const std::vector<Real> &system_state() const { return system_state_; }
  const std::vector<Real> &initial_state() const { return initial_state_; }
  const std::vector<Real> &natural_frequencies() const { return natural_frequencies_; }
  const std::vector<std::string> &files() const { return files_; }
  long steps() const { return steps_; }
  long rejected() const { return rejected_; }

 private:
  static int observe(void *user, double t, const double * /*state*/, int /*n*/, int /*s*/) {
    try {
      (*static_cast<BinaryObserverFirstOrder *>(user))(t);
      return 0;
    } catch (...) {
      return 1;
    }
  }
  void InitializeRandomSystemState(Real sigma, Real xi, unsigned seed) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<Real> unif(0.0, 1.0);
    std::uniform_real_distribution<Real> ang(0.0, 2.0 * M_PI);
    std::uniform_real_distribution<Real> om(-sigma / xi, sigma / xi);
    for (int i = 0; i < n_; ++i) {
      system_state_[(size_t)s_ * i] = unif(gen);
      system_state_[(size_t)s_ * i + 1] = unif(gen);
      system_state_[(size_t)s_ * i + 2] = ang(gen);
      if (s_ == 4) system_state_[(size_t)s_ * i + 3] = om(gen);
    }
    std::normal_distribution<Real> nat(0.0, 0.5);
    for (int i = 0; i < n_; ++i) natural_frequencies_[i] = nat(gen);
  }

  Thread *thread_;
  PeriodicBoundaryConditions pbc_config_;
  int n_, s_;
  std::vector<Real> system_state_, initial_state_, natural_frequencies_;
  std::vector<std::string> files_;
  long steps_ = 0, rejected_ = 0;
};

// ---------------------------------------------------------------------------
// SimulationEngineFirstOrderPtr.cpp:96-165: random initial state (x, y ~ U[0,1), phi ~
// U[0,2pi), omega ~ U[-sigma/xi, sigma/xi]), then `while (t <= t_1) { t += dt; DoStep; observe }`
// with the stochastic Runge-Kutta stepper.  The run keeps the state on the device and reads
// the order parameter there every `summary_every` steps (the binary observer's cadence).
class SimulationEngineFirstOrderPtr {
 public:
  SimulationEngineFirstOrderPtr(ThreadSharedMemory *thread, int number_of_particles, int number_of_state_variables)
      : thread_(thread), pbc_config_(thread, number_of_particles, number_of_state_variables, 1.0, 1.0),
        n_(number_of_particles), s_(number_of_state_variables),
        system_state_((size_t)number_of_particles * number_of_state_variables, 0.0) {}

  struct Run {
    Real v_0 = 1.0, xi = 0.1, sigma = 1.0, rho = 0.8, alpha = 0.3, D_phi = 0.01;  // :103-108
    Real t_0 = 0.0, t_1 = 1000.0, dt = 0.01;                                        // :110-112
    int summary_every = 100;
    unsigned seed = 0;
    // write the reference's trajectory and summary files (BinaryObserverFirstOrderPtr, :125-138)
    bool write_files = false;
    int trial = 0;
  };

  void RunSimulation() { RunSimulation(Run{}); }
  void RunSimulation(const Run &r) {
    InitializeRandomSystemState(r.sigma, r.xi, r.seed);
    StochasticRungeKuttaStepperPtr stepper(thread_, n_, s_);
    ParticleSystemFirstOrderPtr particle_system(thread_, r.v_0, r.xi, r.sigma, r.rho, r.alpha, r.D_phi, pbc_config_,
                                                n_, s_);
    particle_system.SetHostSync(false);
    summary_.clear();
    std::unique_ptr<BinaryObserverFirstOrderPtr> observer;
    if (r.write_files) {
      observer.reset(new BinaryObserverFirstOrderPtr(thread_, r.v_0, r.xi, r.sigma, r.rho, r.alpha, r.D_phi,
                                                     pbc_config_, n_, s_, r.dt, r.trial));
      observer->Bind(particle_system);
      files_ = {observer->simulation_file_name(), observer->summary_statistics_file_name()};
    }
    Real t = r.t_0;
    long k = 0;
    if (observer) {
      particle_system.device().Upload(system_state_.data());  // the device holds the initial state
      observer->SaveSystemState(system_state_.data(), (long)system_state_.size(), t);
    }
    while (t <= r.t_1) {
      t += r.dt;
      stepper.DoStep(particle_system, system_state_.data(), t, r.dt);
      if (++k % r.summary_every == 0) summary_.push_back({t, particle_system.OrderParameter()});
      if (observer) {
        observer->SaveSystemState(system_state_.data(), (long)system_state_.size(), t);
        observer->SaveSummaryStatistics(system_state_.data(), (long)system_state_.size(), t);
      }
    }
    particle_system.DownloadState(system_state_.data());
  }
  // the trajectory and summary files of the last run with write_files
  const std::vector<std::string> &files() const { return files_; }
  // SimulationEngineFirstOrderPtr.cpp:167-217: from the last frame of a trajectory file, RK4 runs
  // for alpha = alpha_0 - alpha_step * ia, ia = 1 .. n_alpha, each continuing from the previous
  // one's state, with a summary line per step and the final frame in its own observer files
  struct Continuation {
    Real v_0 = 1.0, xi = 0.1, sigma = 1.0, rho = 0.8, D_phi = 0.0;  // :174-179
    Real t_0 = 0.0, t_1 = 200.0, dt = 0.01;                         // :181-183
    Real alpha_0 = 0.6, alpha_step = 0.05;                          // :191
    int n_alpha = 5, trial = 0;
    std::string initial_file;
  };
  void RunContinuationMethod(const Continuation &c) {
    InitializeSystemStateFromPreviousSolution(c.initial_file);
    for (int ia = 1; ia <= c.n_alpha; ++ia) {
      const Real alpha = c.alpha_0 - c.alpha_step * ia;
      RungeKutta4StepperPtr stepper(thread_, n_, s_);
      ParticleSystemFirstOrderPtr particle_system(thread_, c.v_0, c.xi, c.sigma, c.rho, alpha, c.D_phi, pbc_config_,
                                                  n_, s_);
      particle_system.SetHostSync(false);
      BinaryObserverFirstOrderPtr observer(thread_, c.v_0, c.xi, c.sigma, c.rho, alpha, c.D_phi, pbc_config_, n_, s_,
                                           c.dt, c.trial);
      observer.Bind(particle_system);
      particle_system.device().Upload(system_state_.data());
      Real t = c.t_0;
      while (t <= c.t_1) {
        t += c.dt;
        stepper.DoStep(particle_system, system_state_.data(), t, c.dt);
        observer.SaveSummaryStatistics(system_state_.data(), (long)system_state_.size(), t);
      }
      observer.SaveSystemState(system_state_.data(), (long)system_state_.size(), t);
      particle_system.DownloadState(system_state_.data());
      continuation_files_.push_back(observer.simulation_file_name());
    }
  }
  const std::vector<std::string> &continuation_files() const { return continuation_files_; }
  // SimulationEngineFirstOrderPtr.cpp:70-94: the state from the last frame of a trajectory file
  void InitializeSystemStateFromPreviousSolution(const std::string &file_name) {
    float t = 0.f;
    spss_shim::check(spss_read_last_frame(file_name.c_str(), s_, n_, &t, system_state_.data()));
    started_from_ = t;
  }
  Real started_from() const { return started_from_; }

  const std::vector<Real> &system_state() const { return system_state_; }
  struct Summary {
    Real t;
    std::complex<Real> order_parameter;
  };
  const std::vector<Summary> &summary() const { return summary_; }

 private:
  void InitializeRandomSystemState(Real sigma, Real xi, unsigned seed) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<Real> unif(0.0, 1.0);
    std::uniform_real_distribution<Real> ang(0.0, 2.0 * M_PI);
    std::uniform_real_distribution<Real> om(-sigma / xi, sigma / xi);
    for (int i = 0; i < n_; ++i) {
      system_state_[(size_t)s_ * i] = unif(gen);
      system_state_[(size_t)s_ * i + 1] = unif(gen);
      system_state_[(size_t)s_ * i + 2] = ang(gen);
      if (s_ == 4) system_state_[(size_t)s_ * i + 3] = om(gen);
    }
  }

  ThreadSharedMemory *thread_;
  PeriodicBoundaryConditions pbc_config_;
  int n_, s_;
  std::vector<Real> system_state_;
  std::vector<Summary> summary_;
  std::vector<std::string> files_, continuation_files_;
  Real started_from_ = 0.0;
};
