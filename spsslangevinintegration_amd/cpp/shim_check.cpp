// shim_check.cpp -- drives the reference-named C++ classes of spss_reference_api.hpp the way
// the reference's engine does, for the drop-in tests: reads an input state, runs one scenario,
// writes the result as raw FP64.
//
//   shim_check rk4     n rho K dt in.bin out.bin               all-pairs RK4 (method AUTO: rho > 0.5)
//   shim_check euler   n rho K dt in.bin normals.bin out.bin   linked list, injected normals
//   shim_check srk     n rho K dt in.bin out.bin               linked list, Philox noise (seed 0x5EED0000)
//   shim_check average n rho K dt in.bin out.bin               Euler + AveragingObserverPtr, time mean
//   shim_check engine  n                 out.bin               SimulationEngineFirstOrderPtr, t_1 = 1
//   shim_check files   n folder          out.bin               the same run writing the reference's trajectory and
//                                                              summary files into folder (BinaryObserverFirstOrderPtr),
//                                                              then InitializeSystemStateFromPreviousSolution on them
//   shim_check continue n folder in.traj out.bin               RunContinuationMethod from in.traj: 2 alphas, t_1 = 0.05
//   shim_check dopri5  n folder          out.bin               SimulationEngineFirstOrder (Boost.Odeint dopri5 with dense
//                                                              output, t_1 = 2) writing the reference's files into folder;
//                                                              out.bin: initial state, natural frequencies, final state
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "spss_reference_api.hpp"

static std::vector<Real> read_file(const char *path, size_t count) {
  std::vector<Real> v(count);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(count * sizeof(Real)));
  if (!f) throw std::runtime_error(std::string("cannot read ") + path);
  return v;
}

static void write_file(const char *path, const std::vector<Real> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(Real)));
}

int main(int argc, char **argv) {
  try {
    if (argc < 3) throw std::runtime_error("usage: shim_check mode n ...");
    const std::string mode = argv[1];
    const int n = std::atoi(argv[2]), s = 4;
    ThreadSharedMemory thread(argc, argv, n, s);
    PeriodicBoundaryConditions pbc(&thread, n, s, 1.0, 1.0);
    if (mode == "engine") {
      SimulationEngineFirstOrderPtr engine(&thread, n, s);
      SimulationEngineFirstOrderPtr::Run r;
      r.t_1 = 1.0;
      r.summary_every = 10;
      engine.RunSimulation(r);
      std::vector<Real> out;
      for (const auto &e : engine.summary()) {
        out.push_back(e.t);
        out.push_back(e.order_parameter.real());
        out.push_back(e.order_parameter.imag());
      }
      write_file(argv[3], engine.system_state());
      write_file((std::string(argv[3]) + ".summary").c_str(), out);
      return 0;
    }
    if (mode == "dopri5") {
      spss_shim::default_options().output_folder = std::string(argv[3]) + "/";
      SimulationEngineFirstOrder engine(&thread, n, s);
      SimulationEngineFirstOrder::Run r;
      r.t_1 = 2.0;
      r.seed = 7;
      engine.RunSimulation(r);
      std::vector<Real> out(engine.initial_state());
      out.insert(out.end(), engine.natural_frequencies().begin(), engine.natural_frequencies().end());
      out.insert(out.end(), engine.system_state().begin(), engine.system_state().end());
      write_file(argv[4], out);
      std::ofstream f(std::string(argv[4]) + ".files");
      f << engine.files()[0] << "\n" << engine.files()[1] << "\n" << engine.steps() << "\n" << engine.rejected() << "\n";
      return 0;
    }
    if (mode == "continue") {
      spss_shim::default_options().output_folder = std::string(argv[3]) + "/";
      SimulationEngineFirstOrderPtr engine(&thread, n, s);
      SimulationEngineFirstOrderPtr::Continuation c;
      c.t_1 = 0.05;
      c.n_alpha = 2;
      c.initial_file = argv[4];
      engine.RunContinuationMethod(c);
      write_file(argv[5], engine.system_state());
      for (const std::string &f : engine.continuation_files()) std::printf("%s\n", f.c_str());
      return 0;
    }
    if (mode == "files") {
      spss_shim::default_options().output_folder = std::string(argv[3]) + "/";
      SimulationEngineFirstOrderPtr engine(&thread, n, s);
      SimulationEngineFirstOrderPtr::Run r;
      r.t_1 = 1.0;
      r.write_files = true;
      engine.RunSimulation(r);
      write_file(argv[4], engine.system_state());
      SimulationEngineFirstOrderPtr restarted(&thread, n, s);
      restarted.InitializeSystemStateFromPreviousSolution(engine.files()[0]);
      write_file((std::string(argv[4]) + ".restart").c_str(), restarted.system_state());
      std::printf("%s\n%s\n%.9g\n", engine.files()[0].c_str(), engine.files()[1].c_str(),
                  (double)restarted.started_from());
      return 0;
    }
    const double rho = std::atof(argv[3]);
    const int K = std::atoi(argv[4]);
    const double dt = std::atof(argv[5]);
    std::vector<Real> state = read_file(argv[6], (size_t)n * s);
    const Real v0 = 0.02, xi = 1.3, sigma = 2.0, alpha = 0.1, D_phi = 0.07;
    if (mode == "rk4") {
      ParticleSystemFirstOrderPtr system(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
      RungeKutta4StepperPtr stepper(&thread, n, s);
      Real t = 0.0;
      for (int k = 0; k < K; ++k) {
        t += dt;
        stepper.DoStep(system, state.data(), t, dt);
      }
      write_file(argv[7], state);
    } else if (mode == "euler") {
      spss_shim::default_options().method = SPSS_METHOD_LINKED_LIST;
      std::vector<Real> z = read_file(argv[7], (size_t)n * K);
      ParticleSystemFirstOrderPtr system(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
      StochasticEulerStepperPtr stepper(&thread, n, s);
      Real t = 0.0;
      for (int k = 0; k < K; ++k) {
        t += dt;
        std::vector<Real> zk(z.begin() + (size_t)k * n, z.begin() + (size_t)(k + 1) * n);
        stepper.DoStep(system, state.data(), t, dt, zk);
      }
      write_file(argv[8], state);
    } else if (mode == "srk") {
      spss_shim::default_options().method = SPSS_METHOD_LINKED_LIST;
      ParticleSystemFirstOrderPtr system(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
      StochasticRungeKuttaStepperPtr stepper(&thread, n, s);
      Real t = 0.0;
      for (int k = 0; k < K; ++k) {
        t += dt;
        stepper.DoStep(system, state.data(), t, dt);
      }
      std::vector<Real> out = state;
      const std::vector<Real> &c = system.GetComplimentaryAlignmentForce();
      out.insert(out.end(), c.begin(), c.end());
      write_file(argv[7], out);
    } else if (mode == "average") {
      spss_shim::default_options().method = SPSS_METHOD_LINKED_LIST;
      ParticleSystemFirstOrderPtr system(&thread, v0, xi, sigma, rho, alpha, D_phi, pbc, n, s);
      system.SetHostSync(false);
      StochasticEulerStepperPtr stepper(&thread, n, s);
      AveragingObserverPtr observer(&thread, n, s);
      observer.Bind(system);
      Real t = 0.0;
      for (int k = 0; k < K; ++k) {
        t += dt;
        stepper.DoStep(system, state.data(), t, dt);
        observer(state.data(), (long)state.size(), t);
      }
      std::vector<Real> out = observer.AveragedSystemState();
      system.DownloadState(state.data());
      out.insert(out.end(), state.begin(), state.end());
      write_file(argv[7], out);
    } else {
      throw std::runtime_error("unknown mode " + mode);
    }
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "shim_check: %s\n", e.what());
    return 1;
  }
}
