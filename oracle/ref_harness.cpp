// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A driver of our own that links the reference's unmodified source files
// (compiled in place from the reference checkout by oracle/Makefile) so that
// the reference code itself produces golden vectors and a CPU timing.  Built
// twice:
//   REF_PTR=0 -> ref_serial: the MPI-free class family (ParticleSystemFirstOrder,
//                ParticleSystemSecondOrder, RungeKutta4Stepper, StochasticEulerStepper)
//   REF_PTR=1 -> ref_ptr:    the MPI-3 class family the reference's default program runs
//                (ParticleSystemFirstOrderPtr, RungeKutta4StepperPtr,
//                StochasticEulerStepperPtr, StochasticRungeKuttaStepperPtr), as a
//                one-rank MPI singleton.
//
// Usage: ref_xxx <config.txt> <state_in.bin> <out.bin>
// config.txt holds key=value lines (see parse_config).  The input file holds
// the initial AoS FP64 state, then (optional) N natural frequencies.  The
// output holds, per recorded step, [int64 step][s*N doubles state], then for
// stochastic steppers the normal draws the reference consumed, and, when
// asked, the Verlet lists / linked-list pair sets read from the reference
// objects.  "bench" mode times the serial linked-list RHS + Euler update.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#if REF_PTR
#include <mpi.h>
#endif

// read the reference objects' lists for the neighbour-set fixtures
#define private public
#define protected public
#include "Definitions.hpp"
#include "BoundaryConfiguration/PeriodicBoundaryConditions.hpp"
#include "Parallelization/Thread.hpp"
#if REF_PTR
#include "Parallelization/ThreadSharedMemory.hpp"
#include "Parallelization/Parallelization.hpp"
#include "DynamicalSystems/ParticleSystemFirstOrderPtr.hpp"
#include "DynamicalSystems/ParticleSystemSecondOrderPtr.hpp"
#include "Steppers/RungeKutta4StepperPtr.hpp"
#include "Steppers/StochasticEulerStepperPtr.hpp"
#include "Steppers/StochasticRungeKuttaStepperPtr.hpp"
#else
#include "DynamicalSystems/ParticleSystemFirstOrder.hpp"
#include "DynamicalSystems/ParticleSystemSecondOrder.hpp"
#include "Steppers/RungeKutta4Stepper.hpp"
#include "Steppers/StochasticEulerStepper.hpp"
#endif
#undef private
#undef protected

std::mt19937 mersenne_twister_generator(12345u);

namespace {

struct Config {
  std::map<std::string, std::string> kv;
  double d(const std::string &k, double def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::atof(it->second.c_str());
  }
  long l(const std::string &k, long def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::atol(it->second.c_str());
  }
  std::string s(const std::string &k, const std::string &def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
  }
};

Config parse_config(const char *path) {
  Config c;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    auto p = line.find('=');
    if (p == std::string::npos) continue;
    c.kv[line.substr(0, p)] = line.substr(p + 1);
  }
  return c;
}

std::vector<long> parse_list(const std::string &s) {
  std::vector<long> v;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) {
    if (!tok.empty()) v.push_back(std::atol(tok.c_str()));
  }
  return v;
}

struct Out {
  std::ofstream f;
  explicit Out(const char *p) : f(p, std::ios::binary | std::ios::trunc) {}
  void tag(const char *t) {
    char b[8] = {0};
    std::strncpy(b, t, 8);
    f.write(b, 8);
  }
  void i64(long v) {
    int64_t x = v;
    f.write((const char *)&x, 8);
  }
  void dv(const double *p, size_t n) {
    i64((long)n);
    f.write((const char *)p, (std::streamsize)(n * sizeof(double)));
  }
  void iv(const std::vector<int> &v) {
    i64((long)v.size());
    f.write((const char *)v.data(), (std::streamsize)(v.size() * sizeof(int)));
  }
};

// record `count` normal draws that a fresh std::normal_distribution would
// produce from a copy of the global generator (the *Ptr steppers construct a
// new distribution in every DoStep call).
std::vector<double> peek_normals(size_t count) {
  std::mt19937 g = mersenne_twister_generator;
  std::normal_distribution<Real> d(0.0, 1.0);
  std::vector<double> r(count);
  for (size_t k = 0; k < count; ++k) r[k] = d(g);
  return r;
}

template <class PS>
void dump_verlet(Out &out, PS &ps, int n) {
  std::vector<int> flat;
  for (int i = 0; i < n; ++i) {
    flat.push_back((int)ps.verlet_list_[i].size());
    for (int j : ps.verlet_list_[i]) flat.push_back(j);
  }
  out.tag("VERLET");
  out.iv(flat);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s config state_in out\n", argv[0]);
    return 2;
  }
  Config c = parse_config(argv[1]);
  const int model = (int)c.l("model", 4);
  const int n = (int)c.l("n", 0);
  const int s = model;
  const double v0 = c.d("v0", 0.01), xi = c.d("xi", 1.0), sigma = c.d("sigma", 1.0);
  const double rho = c.d("rho", 0.1), alpha = c.d("alpha", 0.0), D_phi = c.d("D_phi", 0.0);
  const double xi_r = c.d("xi_r", 1.0), D_r = c.d("D_r", 0.0);
  const double dt = c.d("dt", 0.01);
  const long steps = c.l("steps", 1);
  const std::string stepper = c.s("stepper", "rk4");
  const std::string method = c.s("method", "dispatch");
  std::vector<long> snaps = parse_list(c.s("snapshots", ""));
  const int dump_lists = (int)c.l("dump_lists", 0);
  mersenne_twister_generator.seed((unsigned)c.l("seed", 12345));

  std::vector<double> state((size_t)s * n, 0.0), natf((size_t)n, 0.0);
  {
    std::ifstream f(argv[2], std::ios::binary);
    f.read((char *)state.data(), (std::streamsize)(state.size() * sizeof(double)));
    if (c.l("natfreq", 0)) f.read((char *)natf.data(), (std::streamsize)(natf.size() * sizeof(double)));
  }
  Out out(argv[3]);
  auto snap = [&](long step) {
    if (std::find(snaps.begin(), snaps.end(), step) != snaps.end()) {
      out.tag("STATE");
      out.i64(step);
      out.dv(state.data(), state.size());
    }
  };

#if REF_PTR
  LaunchParallelSession(argc, argv);
  {
    ThreadSharedMemory thread(argc, argv, n, s);
    PeriodicBoundaryConditions pbc(&thread, n, s, 1.0, 1.0);
    if (model == 4) {
      ParticleSystemFirstOrderPtr ps(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
      if (c.l("natfreq", 0)) ps.SetNaturalFrequencies(natf);
      snap(0);
      if (stepper == "rk4") {
        RungeKutta4StepperPtr st(&thread, n, s);
        for (long k = 1; k <= steps; ++k) {
          st.DoStep(ps, state.data(), 0.0, dt);
          snap(k);
          if (dump_lists && k == 1) dump_verlet(out, ps, n);
        }
      } else if (stepper == "rk4_direct") {
        // RungeKutta4StepperPtr::DoStep (Steppers/RungeKutta4StepperPtr.cpp:68-92) with the RHS
        // method named explicitly instead of EvaluateRhs's dispatch on rho.
        std::vector<double> k1(state.size(), 0.0), k2 = k1, k3 = k1, k4 = k1;
        auto rhs = [&](const std::vector<double> &kp, std::vector<double> &kn, double kc) {
          if (method == "all_pairs") ps.EvaluateInteractionsWithAllPairs(state.data(), kp, kn, kc, dt);
          else if (method == "linked_list") ps.EvaluateInteractionsWithLinkedList(state.data(), kp, kn, kc, dt);
          else if (method == "global") ps.EvaluateInteractionsWithGlobalPolarOrderField(state.data(), kp, kn, kc, dt);
          else { std::fprintf(stderr, "bad method\n"); std::exit(3); }
        };
        for (long k = 1; k <= steps; ++k) {
          rhs(k1, k1, 0.0);
          rhs(k1, k2, 0.5);
          rhs(k2, k3, 0.5);
          rhs(k3, k4, 1.0);
          for (size_t q = 0; q < state.size(); ++q)
            state[q] += (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]) * dt / 6.0;
          snap(k);
        }
      } else if (stepper == "euler") {
        StochasticEulerStepperPtr st(&thread, n, s);
        for (long k = 1; k <= steps; ++k) {
          std::vector<double> z = peek_normals((size_t)n);
          out.tag("NOISE");
          out.i64(k);
          out.dv(z.data(), z.size());
          st.DoStep(ps, state.data(), 0.0, dt);
          snap(k);
          if (dump_lists && k == 1) dump_verlet(out, ps, n);
        }
      } else if (stepper == "srk") {
        StochasticRungeKuttaStepperPtr st(&thread, n, s);
        for (long k = 1; k <= steps; ++k) {
          std::vector<double> z = peek_normals((size_t)2 * n);
          out.tag("NOISE");
          out.i64(k);
          out.dv(z.data(), z.size());
          st.DoStep(ps, state.data(), 0.0, dt);
          snap(k);
          if (k == 1) {
            out.tag("COMPL");
            out.i64(k);
            out.dv(ps.GetComplimentaryAlignmentForce().data(), (size_t)n);
          }
        }
      } else if (stepper == "rhs") {
        // one EvaluateRhs through the dispatch, k_prev = given derivative file (zeros), k_coef = kc
        std::vector<double> kp(state.size(), 0.0), kn(state.size(), 0.0);
        for (size_t q = 0; q < kp.size(); ++q) kp[q] = std::sin(0.37 * (double)q) * 0.5;
        ps.is_higher_order_sde_integration_ = (bool)c.l("higher_order", 0);
        ps.EvaluateRhs(state.data(), kp, kn, c.d("k_coef", 0.5), dt);
        out.tag("KPREV");
        out.i64(0);
        out.dv(kp.data(), kp.size());
        out.tag("KNEXT");
        out.i64(0);
        out.dv(kn.data(), kn.size());
        out.tag("COMPL");
        out.i64(0);
        out.dv(ps.GetComplimentaryAlignmentForce().data(), (size_t)n);
        if (dump_lists) dump_verlet(out, ps, n);
      }
    } else if (model == 6) {
      ParticleSystemSecondOrderPtr ps(&thread, v0, xi_r, D_r, xi, D_phi, sigma, rho, alpha, pbc, n, s);
      snap(0);
      RungeKutta4StepperPtr st(&thread, n, s);
      for (long k = 1; k <= steps; ++k) {
        st.DoStep(ps, state.data(), 0.0, dt);
        if (c.l("wrap", 0)) pbc.ApplyPeriodicBoundaryConditions(state.data(), (long)state.size());
        snap(k);
      }
    }
  }
  FinalizeParallelSession();
#else
  Thread thread(argc, argv, n, s);
  PeriodicBoundaryConditions pbc(&thread, n, s, 1.0, 1.0);
  if (stepper == "bench") {
    // the reference serial linked-list RHS (DynamicalSystems/ParticleSystemFirstOrder.cpp:231-353)
    // + the s = 4 Euler update; timed with steady_clock after `warmup` steps
    ParticleSystemFirstOrder ps(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
    std::vector<double> der(state.size(), 0.0);
    const long warm = c.l("warmup", 1);
    std::normal_distribution<Real> nd(0.0, 1.0);
    const double ip = std::sqrt(2.0 * D_phi), sdt = std::sqrt(dt);
    std::chrono::steady_clock::time_point t0;
    for (long k = 0; k < warm + steps; ++k) {
      if (k == warm) t0 = std::chrono::steady_clock::now();
      std::fill(der.begin(), der.end(), 0.0);
      ps.EvaluateInteractionsWithLinkedList(state, der, der, 0.0, dt);
      for (int i = 0; i < n; ++i) {
        double dW = sdt * nd(mersenne_twister_generator);
        int ii = 4 * i;
        state[ii] += der[ii] * dt;
        state[ii + 1] += der[ii + 1] * dt;
        state[ii + 2] += der[ii + 2] * dt;
        state[ii + 3] += der[ii + 3] * dt + ip * dW;
      }
    }
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"seconds\": %.6f, \"steps\": %ld, \"n\": %d, \"particle_steps_per_s\": %.6e}\n", sec,
                steps, n, (double)n * (double)steps / sec);
    return 0;
  }
  if (model == 4) {
    ParticleSystemFirstOrder ps(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
    if (c.l("natfreq", 0)) ps.SetNaturalFrequencies(natf);
    snap(0);
    if (stepper == "rk4") {
      RungeKutta4Stepper st(&thread, n, s);
      for (long k = 1; k <= steps; ++k) {
        st.DoStep(ps, state, 0.0, dt);
        snap(k);
        if (dump_lists && k == 1) dump_verlet(out, ps, n);
      }
    } else if (stepper == "pairs_ll") {
      // the linked-list buckets the reference builds, for the cell-list fixtures
      std::vector<double> kn(state.size(), 0.0), kp(state.size(), 0.0);
      ps.EvaluateInteractionsWithLinkedList(state, kp, kn, 0.0, dt);
      std::vector<int> flat;
      for (auto &b : ps.linked_list_) {
        flat.push_back((int)b.size());
        for (int j : b) flat.push_back(j);
      }
      out.tag("BUCKETS");
      out.iv(flat);
      out.tag("KNEXT");
      out.i64(0);
      out.dv(kn.data(), kn.size());
    }
  } else if (model == 6) {
    ParticleSystemSecondOrder ps(&thread, v0, xi_r, D_r, xi, D_phi, sigma, rho, alpha, pbc, n, s);
    snap(0);
    if (stepper == "rk4") {
      RungeKutta4Stepper st(&thread, n, s);
      for (long k = 1; k <= steps; ++k) {
        st.DoStep(ps, state, 0.0, dt);
        if (c.l("wrap", 0)) pbc.ApplyPeriodicBoundaryConditions(state);
        snap(k);
      }
    } else if (stepper == "euler") {
      // AddNoise keeps a static distribution (ParticleSystemSecondOrder.cpp:443) for the whole
      // run; mirror it with our own distribution on a copy of the generator, drawn in lockstep.
      std::mt19937 g2 = mersenne_twister_generator;
      std::normal_distribution<Real> d2(0.0, 1.0);
      StochasticEulerStepper st(&thread, n, s);
      for (long k = 1; k <= steps; ++k) {
        std::vector<double> z((size_t)3 * n);
        for (auto &v : z) v = d2(g2);
        out.tag("NOISE");
        out.i64(k);
        out.dv(z.data(), z.size());
        st.DoStep(ps, state, 0.0, dt);
        if (c.l("wrap", 0)) pbc.ApplyPeriodicBoundaryConditions(state);
        snap(k);
      }
    }
  }
#endif
  return 0;
}
