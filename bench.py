#!/usr/bin/env python3
"""Headline benchmark: particle-steps/s of the N = 1,048,576 cell-list stochastic step.

Workload (the large second-order configuration, one GPU): N = 1,048,576 particles
uniform in the unit box, rho = 0.004 (250 x 250 cells, ~52.7 neighbours), inertial
phase model with m = 0.5, sigma = 1.5, D = 0.02 mapped onto the reference's
variables (xi = 1/m = 2, sigma/m = 3, D/m^2 = 0.08, natural frequencies
omega_i/m with omega_i ~ Lorentzian(0, 0.05) clipped at +-10), phi_dot(0) = 0,
v0 = 0.005, dt = 0.002, cell membership refreshed every step, stochastic Euler with
Philox noise (seed 0x5EED0003), FP32 pair arithmetic (blocked sums: FP32 partials, FP64 totals), FP64 state.

Prints one JSON line.  `value` is whole-job particle-steps/s with the state resident in HBM:
after the warm-up steps the step graphs are captured (spss_prepare_run) and the timed steps are
replayed from them; the line reports how many timed steps were replayed and how many ran eagerly.
`roofline` is the whole step against the algorithmic bytes per particle-step of the cell-list step
(104.47 B: FP64 state read + write, FP32 snapshot write + read, cell key and sorted index write +
read, cell table), with the pair kernel (fk_pairs, timed with HIP events on its stream in a
separate run of eager steps) as a sub-field.  `cpu_baseline` is the reference algorithm on the
host: all cores through the oracle's threaded restatement (contiguous particle blocks per thread,
as the reference's ranks own them), and one core through the reference's own serial linked-list
RHS + Euler update compiled from its sources with its own -O3 -flto (oracle/_ref).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s, N=1M cell list; batched vs plain at ≤1e-4 order-param err"
HBM_PEAK = 8.0e12


def pair_kernel_bytes(n, cells):
    """algorithmic HBM bytes per particle of the pair kernel of one Euler step: FP32 snapshot
    (x, y, cos phi, sin phi) read once (16), phasor sum (2 x FP64) and count written (20); per
    cell its start read (4 B x 2)"""
    return 16.0 + 20.0 + 8.0 * cells / n


# HBM bytes per launch of that kernel measured with rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled
# by the copy-kernel calibration, by cell layout (subdivision, x split)
# (profiles/fr_pairs_euler_N1M_hbm_traffic.txt)
PAIR_KERNEL_TRAFFIC = {(1, 1): 4.4389e7, (1, 2): 4.7781e7, (1, 4): 5.4909e7}
# HBM bytes of the three kernels of one resident step (fr_pairs, fr_far, fr_scatter), the same way
STEP_TRAFFIC = {(1, 1): 1.9577e8, (1, 2): 2.0656e8, (1, 4): 2.1588e8}
# VALU instructions issued per CU-cycle by that kernel (SQ_INSTS_VALU / (256 CUs x cycles),
# profiles/fr_pairs_euler_N1M_valu.txt), by cell layout: the nominal rate is one per CU-cycle
PAIR_KERNEL_VALU_PER_CU_CYCLE = {(1, 1): 0.99, (1, 2): 0.91, (1, 4): 0.86}
# algorithmic HBM bytes per particle-step of the cell-list step (s = 4 state, one RHS): FP64 state
# read + write 16*s, FP32 snapshot x, y, phi written and read 8*3, cell key and sorted index written
# and read 16, cell start/count table 8 per cell -- the compulsory traffic of the reference
# algorithm, whatever the implementation moves
def step_bytes(n, cells, s=4):
    return 16.0 * s + 8.0 * 3 + 16.0 + 8.0 * cells / n


# bytes the implementation's resident step moves per particle (its own design, for comparison):
# pair kernel 36 (snapshot read, sums and count written), scatter 132 (move word, id, FP64 state,
# natural frequency, sums read; state, id, natural frequency, snapshot, next move word written),
# and per cell 68 (the arrivals scan: arrival counts read and cleared, starts written, the 9 move
# counters of the next step cleared; the scatter: starts and move counts read)
def impl_step_bytes(n, cells):
    return 36.0 + 132.0 + 68.0 * cells / n


def headline_system(spss, n=1 << 20, precision=None, subdiv=1, seed=0x5EED0003, verlet=False):
    """the large configuration's system; verlet=True: its Verlet lists (exact mode, skin 0.2 rho)"""
    m, sigma, D = 0.5, 1.5, 0.02
    kw = dict(model=4, n=n, v0=0.005, xi=1.0 / m, sigma=sigma / m, rho=0.004, alpha=0.0, D_phi=D / m ** 2,
              natfreq_local=1, seed=seed, cell_subdivision=subdiv,
              precision=spss.PRECISION_FP32C if precision is None else precision)
    if verlet:
        kw.update(verlet_mode=spss.VERLET_EXACT, verlet_skin=0.2 * 0.004)
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 4))
    st[:, 0] = rng.random(n)
    st[:, 1] = rng.random(n)
    st[:, 2] = rng.random(n) * 2.0 * np.pi
    w = np.clip(0.05 * np.tan(np.pi * (rng.random(n) - 0.5)), -10.0, 10.0)  # Lorentzian(0, 0.05)
    g = spss.ParticleSystem(**kw)
    g.set_natural_frequencies(w / m)
    g.set_state(st.reshape(-1))
    return g, kw, st.reshape(-1), w / m


def time_steps(g, spss, stepper, steps, warmup, dt=0.002):
    """wall time of `steps` steps after `warmup` warm-up steps and the capture of the step graphs
    (outside the timed region), and how the timed steps ran: (seconds, eager, replayed, captured)"""
    g.run(stepper, spss.METHOD_LINKED_LIST, warmup, dt, spss.NOISE_PHILOX)
    g.prepare_run(stepper, spss.METHOD_LINKED_LIST, dt, spss.NOISE_PHILOX)
    g.synchronize()
    g.run_stats(reset=True)
    t0 = time.perf_counter()
    g.run(stepper, spss.METHOD_LINKED_LIST, steps, dt, spss.NOISE_PHILOX)
    g.synchronize()
    sec = time.perf_counter() - t0
    eager, replayed, captured = g.run_stats(reset=True)
    return sec, eager, replayed, captured


def time_pair_kernel(g, spss, stepper, steps, dt=0.002):
    """mean duration of the pair kernel from HIP events recorded around it on its stream (the
    events need eager launches: with the timer on, spss_run does not replay a graph)"""
    g.kernel_timer_start()
    g.run(stepper, spss.METHOD_LINKED_LIST, steps, dt, spss.NOISE_PHILOX)
    g.synchronize()
    kms, launches = g.kernel_timer_stop()
    return kms / max(launches, 1)


def ensemble_config(R=512, seed=0x5EED0004):
    """The batched ensemble: R replicas with N_r = round(2048 * 32^u), u ~ U[0,1), sigma ~ U[0,4],
    D ~ U[0,0.2], m in {0, 0.5, 1}; rho = 0.03, v0 = 0.01, Philox seed 0x5EED0004 + r; natural
    frequencies omega ~ N(0, 0.2^2) (the inertial configuration's distribution).  m = 0 replicas
    are the first-order model (s = 3); m > 0 map to model 4 with xi = 1/m, sigma/m, D/m^2, omega/m."""
    rng = np.random.default_rng(seed)
    ns = np.round(2048 * 32.0 ** rng.random(R)).astype(np.int64)
    sigma = rng.uniform(0, 4, R)
    D = rng.uniform(0, 0.2, R)
    m = rng.choice([0.0, 0.5, 1.0], R)
    seeds = seed + np.arange(R)
    groups = {}
    for model, sel in ((3, m == 0), (4, m > 0)):
        idx = np.nonzero(sel)[0]
        mm = np.where(m[idx] > 0, m[idx], 1.0)
        xi = np.where(model == 4, 1.0 / mm, 0.0)
        sig = sigma[idx] / (mm if model == 4 else 1.0)
        Dp = D[idx] / (mm ** 2 if model == 4 else 1.0)
        states, nats = [], []
        for k, r in enumerate(idx):
            st = np.zeros((ns[r], model))
            st[:, 0], st[:, 1] = rng.random(ns[r]), rng.random(ns[r])
            st[:, 2] = rng.random(ns[r]) * 2 * np.pi
            states.append(st.reshape(-1))
            w = rng.normal(0, 0.2, ns[r])
            nats.append(w / mm[k] if model == 4 else w)
        groups[model] = dict(idx=idx, ns=ns[idx], xi=xi, sigma=sig, D=Dp, seeds=seeds[idx], states=states, nats=nats)
    return groups


def split_by_size(groups, n_split):
    """Each model's replicas as two batches, N_r < n_split and N_r >= n_split (n_split 0: one batch
    per model), so that each batch's cell size is tuned for the densities it holds"""
    out = {}
    for model, g in groups.items():
        for dense in ((False, True) if n_split > 0 else (None,)):
            k = np.arange(len(g["ns"])) if dense is None else np.nonzero((g["ns"] >= n_split) == dense)[0]
            if k.size == 0:
                continue
            key = f"model{model}" + ("" if dense is None else ("_large" if dense else "_small"))
            out[key] = dict(model=model, idx=g["idx"][k], ns=g["ns"][k], xi=g["xi"][k], sigma=g["sigma"][k],
                            D=g["D"][k], seeds=g["seeds"][k], states=[g["states"][i] for i in k],
                            nats=[g["nats"][i] for i in k])
    return out


def run_batches(batches, nsteps, dt, every, observe_from, graph, concurrent):
    """every batch advanced nsteps; concurrent: each from a host thread of its own, so the batches'
    streams run side by side on the GPU (the library's error state is per thread, each batch has its
    own stream)"""
    if not concurrent or len(batches) < 2:
        for b in batches.values():
            b.run(nsteps, dt, observe_every=every, observe_from=observe_from, graph=graph)
        return
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=len(batches)) as ex:
        futs = [ex.submit(b.run, nsteps, dt, every, observe_from, graph) for b in batches.values()]
        for f in futs:
            f.result()


def batched_ensemble(spss, steps=3000, plain_steps=100, dt=0.005, v0=0.01, rho=0.03, every=10, graph=True,
                     tune_it=True, n_split=24576, concurrent=True):
    """512-replica ensemble: batched run (autotuned cell size and tiles, graph replays) against the
    plain loop (the same kernels at the same cell size through one single-system handle per
    replica, one after another).  graph=False, tune_it=False: eager launches, default tiles (a
    kernel trace of the batch's kernels).  n_split > 0: each model's replicas in two batches by
    size (split_by_size)."""
    groups = split_by_size(ensemble_config(), n_split)
    total_n = sum(int(g["ns"].sum()) for g in groups.values())
    batches, tune = {}, {}
    for key, g in groups.items():
        model = g["model"]
        b = spss.ReplicaBatch(model, g["ns"], g["xi"], g["sigma"], g["D"], g["seeds"], v0=v0, rho=rho)
        b.set_natural_frequencies(np.concatenate(g["nats"]))
        b.set_state(np.concatenate(g["states"]))
        if tune_it:
            ctable, k = b.autotune_cells(dt, probe_steps=10)
            table, chosen = b.autotune(dt, probe_steps=10)
            tune[key] = {"cell_layout_ms_per_step": [round(float(t), 4) for t in ctable],
                         "cell_layout": list(b.cell_layout), "ms_per_step": [round(float(t), 4) for t in table],
                         "chosen": int(chosen)}
        batches[key] = b
    # sample for the plain-loop comparison: plain_steps steps of every replica, both ways
    t0 = time.perf_counter()
    run_batches(batches, plain_steps, dt, every, 0, graph, concurrent)
    t_batch_sample = time.perf_counter() - t0
    zb = {key: b.order_stats() for key, b in batches.items()}
    t_plain, max_err = 0.0, 0.0
    for key, g in groups.items():
        if plain_steps <= 0:
            break
        model = g["model"]
        for k in range(len(g["ns"])):
            h = spss.ParticleSystem(model=model, n=int(g["ns"][k]), v0=v0, xi=g["xi"][k], sigma=g["sigma"][k],
                                    rho=rho, D_phi=g["D"][k], natfreq_local=1, precision=spss.PRECISION_FP32C,
                                    seed=int(g["seeds"][k]), cell_subdivision=batches[key].cell_subdivision)
            if batches[key].cell_layout[1] != 1:
                h.set_cell_layout(*batches[key].cell_layout)
            h.set_natural_frequencies(g["nats"][k])
            h.set_state(g["states"][k])
            h.synchronize()
            rs = []
            t0 = time.perf_counter()
            for _ in range(plain_steps // every):
                h.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, every, dt, spss.NOISE_PHILOX)
                rs.append(np.hypot(*h.order_parameter(2)))
            h.synchronize()
            t_plain += time.perf_counter() - t0
            max_err = max(max_err, abs(float(np.mean(rs)) - float(zb[key][0][k])))
            h.close()
    # the full batched run: time averages over the second half
    for key, b in batches.items():
        b.reset(0)
        b.set_state(np.concatenate(groups[key]["states"]))
    t0 = time.perf_counter()
    run_batches(batches, steps, dt, every, steps // 2, graph, concurrent)
    t_batch = time.perf_counter() - t0
    R_mean = np.zeros(sum(len(g["ns"]) for g in groups.values()))
    R_var = np.zeros_like(R_mean)
    for key, b in batches.items():
        mean, var, cnt, z, step = b.order_stats()
        R_mean[groups[key]["idx"]] = mean
        R_var[groups[key]["idx"]] = var
        b.close()
    return {
        "replicas": int(R_mean.size), "particles": total_n, "steps": steps, "dt": dt,
        "particle_steps_per_s": total_n * steps / t_batch,
        "sample_steps": plain_steps,
        "sample_batched_particle_steps_per_s": total_n * plain_steps / t_batch_sample,
        "sample_plain_particle_steps_per_s": total_n * plain_steps / t_plain if t_plain else None,
        "speedup_vs_plain_loop": t_plain / t_batch_sample if t_plain else None,
        "graph_replays": graph,
        "batches_concurrent": concurrent,
        "batches": {key: {"replicas": int(len(g["ns"])), "particles": int(g["ns"].sum())} for key, g in groups.items()},
        "order_param_max_abs_err_vs_plain": max_err,
        "mean_R_over_replicas": float(R_mean.mean()),
        "autotune": tune,
    }


def host_cpu():
    """CPU model name and the cores this process may use (the GPU host's share)"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return {"model": model, "nproc": cores, "os_cpu_count": os.cpu_count()}


def _cpu_state(n):
    rng = np.random.default_rng(1)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1], st[:, 2] = rng.random(n), rng.random(n), rng.random(n) * 2 * np.pi
    return st.reshape(-1)


def cpu_single_core(n=1 << 20, steps=3, warmup=1):
    """the reference's serial linked-list RHS + Euler update, compiled from its sources with its own
    -O3 -flto (oracle/_ref/ref_serial_o3), on one core; None without that build"""
    m, sigma, D = 0.5, 1.5, 0.02
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_serial_o3")
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as td:
        cfg, inp = os.path.join(td, "c.txt"), os.path.join(td, "in.bin")
        with open(cfg, "w") as f:
            f.write(f"model=4\nn={n}\nv0=0.005\nxi={1 / m!r}\nsigma={sigma / m!r}\nrho=0.004\nalpha=0.0\n"
                    f"D_phi={D / m ** 2!r}\ndt=0.002\nsteps={steps}\nwarmup={warmup}\nstepper=bench\n")
        _cpu_state(n).tofile(inp)
        cpu = sorted(os.sched_getaffinity(0))[0]
        out = subprocess.run([ref, cfg, inp, os.path.join(td, "o.bin")], capture_output=True, text=True,
                             check=True, timeout=900, preexec_fn=lambda: os.sched_setaffinity(0, {cpu})).stdout
    res = json.loads(out.strip().splitlines()[-1])
    return {"value": res["particle_steps_per_s"], "unit": "particle-steps/s", "cores": 1, "kind": "reference",
            "sample": f"N={n} rho=0.004 cell list + Euler, {steps} timed steps after {warmup} (the reference's "
                      "serial class family from its sources, g++ -O3 -flto -march=x86-64-v3)"}


def cpu_all_cores(n=1 << 20, steps=20, warmup=2, threads=None):
    """the oracle's restatement of the reference's linked-list RHS + Euler update with its particle
    loop over contiguous blocks, one per thread (ThreadSharedMemory.cpp:20-23), on the host's cores"""
    import oracle
    m, sigma, D = 0.5, 1.5, 0.02
    threads = threads or min(16, host_cpu()["nproc"])
    o = oracle.Oracle(model=4, n=n, v0=0.005, xi=1 / m, sigma=sigma / m, rho=0.004, D_phi=D / m ** 2)
    o.set_threads(threads)
    s = _cpu_state(n)
    z = np.zeros(n)
    for _ in range(warmup):
        o.euler_step(3, s, 0.002, z)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.euler_step(3, s, 0.002, z)
    sec = time.perf_counter() - t0
    return {"value": n * steps / sec, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"N={n} rho=0.004 cell list + Euler, {steps} timed steps after {warmup}, {threads} threads "
                      "(the oracle's C restatement, gcc -O2, particle blocks per thread)"}


def cpu_baseline(single_steps=3, all_steps=20):
    line = cpu_all_cores(steps=all_steps)
    line["single_core"] = cpu_single_core(steps=single_steps)
    line["host"] = host_cpu()
    return line


def verlet_srk_rate(spss, steps=200, warmup=20):
    """the large configuration as specified: N = 1,048,576, Verlet lists with skin 0.2 rho rebuilt on
    the largest displacement (decided on the device), stochastic Runge-Kutta dt = 0.002, FP32 pair
    work and FP64 stage combination; steps replayed from graphs after a warm-up"""
    g, kw, _, _ = headline_system(spss, verlet=True)
    method = spss.METHOD_LINKED_LIST_VERLET
    g.run(spss.STEPPER_SRK, method, warmup, 0.002, spss.NOISE_PHILOX)
    g.prepare_run(spss.STEPPER_SRK, method, 0.002, spss.NOISE_PHILOX)
    g.synchronize()
    g.run_stats(reset=True)
    t0 = time.perf_counter()
    g.run(spss.STEPPER_SRK, method, steps, 0.002, spss.NOISE_PHILOX)
    g.synchronize()
    sec = time.perf_counter() - t0
    eager, replayed, captured = g.run_stats(reset=True)
    g.close()
    return {"particle_steps_per_s": kw["n"] * steps / sec, "ms_per_step": sec / steps * 1e3, "steps": steps,
            "graph_replayed": replayed, "eager": eager, "skin": 0.2 * 0.004, "stepper": "SRK"}


def screen_at_scale(spss, n_train=20000, n_screen=100000, n_lines=16, steps=2000, seed=0x5C4EE):
    """the learned sweep screen at its stated scale: a network trained on the GPU (the library's
    forward/backward/Adam kernels) on n_train batched-run results of random points (sigma, D, m, v0,
    rho, N; N log-uniform in [2048, 4096]), then n_screen points screened (simulated only where the
    predicted |d mean|Z| / d sigma| > 0.5 or the predicted spread of |Z| > 0.05, scored against
    those simulations), and n_lines sigma sweeps of 40 points located with and without the screen"""
    from spsslangevinintegration_amd import screen as S
    sim = dict(steps=steps, dt=0.005, observe_every=10, observe_from=steps // 2)
    rng = np.random.default_rng(seed)
    sc = spss.SweepScreen(seed=0x5EED0005, sim_kwargs=sim)
    train = S.random_points(n_train, rng, N=(2048, 4096), seed0=0x5EED6000)
    t0 = time.perf_counter()
    y = sc.simulate(train)
    t_sim = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = sc.fit(train, y, steps=8000, batch=256, lr=2e-3)
    t_fit = time.perf_counter() - t0
    pts = S.random_points(n_screen, np.random.default_rng(seed + 1), N=(2048, 4096), seed0=0x5EED7000)
    t0 = time.perf_counter()
    out = sc.screen(pts, slope_min=0.5, spread_max=0.05, simulate=True)
    t_screen = time.perf_counter() - t0
    sel = out["selected"]
    err = out.get("error", np.zeros((0, 2)))
    G = 40
    sig = np.linspace(0.0, 3.0, G)
    plain_sims = screened_sims = 0
    worst, found, missed = 0.0, 0, 0
    lr = np.random.default_rng(seed + 2)
    for k in range(n_lines):
        D, m = lr.uniform(0.0, 0.15), lr.uniform(0.3, 1.0)
        v0, rho = lr.choice([0.005, 0.01, 0.02]), lr.choice([0.02, 0.03, 0.04])
        N = int(2 ** lr.uniform(11.0, 12.0))
        line = S.points(sig, D, m, v0, rho, N, 0x5EED8000 + 100 * k + np.arange(G, dtype=np.uint64))
        before = sc.simulations
        x_plain, _ = sc.plain_sweep(line)
        plain_sims += sc.simulations - before
        before = sc.simulations
        x_scr, _, _, _ = sc.screened_sweep(line, window=6)
        screened_sims += sc.simulations - before
        if (x_plain is None) != (x_scr is None):
            missed += 1
        elif x_plain is not None:
            found += 1
            worst = max(worst, abs(x_scr - x_plain))
    return {"training_points": int(n_train), "steps_per_simulation": steps,
            "training_simulation_seconds": t_sim, "training_particle_steps_per_s": float(train["N"].sum()) * steps / t_sim,
            "fit_seconds": t_fit, "adam_steps": 8000, "loss_first_last": [float(losses[:50].mean()), float(losses[-50:].mean())],
            "screened_points": int(n_screen), "simulated": int(sel.size), "simulations_saved": int(n_screen - sel.size),
            "screen_seconds": t_screen,
            "mean_abs_error_on_simulated": [float(err[:, 0].mean()), float(err[:, 1].mean())] if sel.size else None,
            "sweeps": {"lines": n_lines, "points_per_line": G, "plain_simulations": plain_sims,
                       "screened_simulations": screened_sims, "crossings_found": found,
                       "crossing_disagreements": missed, "max_abs_crossing_error": worst,
                       "grid_step": float(sig[1] - sig[0])}}


def fp64_step_rate(spss, steps=10, warmup=2):
    """the reference-precision (FP64, the reference's operation order) cell-list Euler step of the
    headline system: device counting sort and 9-cell pair loop every step"""
    g, kw, _, _ = headline_system(spss, precision=spss.PRECISION_FP64)
    g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, warmup, 0.002, spss.NOISE_PHILOX)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, steps, 0.002, spss.NOISE_PHILOX)
    g.synchronize()
    sec = time.perf_counter() - t0
    g.close()
    return {"particle_steps_per_s": kw["n"] * steps / sec, "ms_per_step": sec / steps * 1e3, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--subdiv", type=int, default=0,
                    help="cell subdivision 1-3; 0: the autotuner picks the cell layout")
    ap.add_argument("--xsplit", type=int, default=1, help="with --subdiv: columns per square cell (1, 2, 4)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20, help="timed steps of the all-core CPU baseline")
    ap.add_argument("--cpu-single-steps", type=int, default=3, help="timed steps of the one-core CPU baseline")
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-verlet", action="store_true")
    ap.add_argument("--no-screen", action="store_true")
    ap.add_argument("--only-batch", action="store_true",
                    help="run the batched ensemble alone (with --batch-eager: eager launches, default tiles)")
    ap.add_argument("--batch-eager", action="store_true")
    ap.add_argument("--batch-sequential", action="store_true",
                    help="run the ensemble's batches one after another (default: side by side, a host thread and "
                         "stream each)")
    ap.add_argument("--batch-split", type=int, default=24576,
                    help="replicas of at least this many particles in batches of their own, whose autotuned cell "
                         "size can differ (0: one batch per model)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the sharding by slabs of cell rows is "
              "designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import spsslangevinintegration_amd as spss
    from spsslangevinintegration_amd import _capi

    if args.only_batch:
        res = batched_ensemble(spss, steps=args.batch_steps, plain_steps=args.plain_steps,
                               graph=not args.batch_eager, tune_it=not args.batch_eager, n_split=args.batch_split,
                               concurrent=not args.batch_sequential)
        print(json.dumps({"batched": res}), flush=True)
        return

    # attainable HBM bandwidth of the build's own streaming copy kernel (1 GiB, read + write)
    copy_gbps = _capi.copy_bandwidth(1 << 30, 10)
    g, kw, st0, natf = headline_system(spss, subdiv=max(args.subdiv, 1))
    n = kw["n"]
    tune = None
    if args.subdiv == 0:  # cell layout picked by the autotuner (outside the timed region)
        table, k = g.autotune_cells(spss.STEPPER_EULER, 0.002, probe_steps=20)
        tune = {"layouts": [list(c) for c in spss.CELL_LAYOUTS],
                "ms_per_step_by_layout": [round(float(t), 4) for t in table], "chosen": list(g.cell_layout)}
        print(f"[bench] cell autotuner: {tune}", file=sys.stderr, flush=True)
    elif args.xsplit != 1:
        g.set_cell_layout(args.subdiv, args.xsplit)
    subdiv, xsplit = g.cell_layout
    cells = (250 * subdiv) ** 2 * xsplit
    print(f"[bench] headline n={n} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    wall, eager, replayed, captured = time_steps(g, spss, spss.STEPPER_EULER, args.steps, args.warmup)
    pair_ms = time_pair_kernel(g, spss, spss.STEPPER_EULER, max(args.steps // 5, 20))
    value = n * args.steps / wall
    ms_step = wall / args.steps * 1e3
    # neighbour count of the final state (pair-interactions/s = sum_i n_i * steps / s)
    pairs = C_pairs(g, spss)
    pair_achieved = pair_kernel_bytes(n, cells) * n / (pair_ms * 1e-3)
    step_achieved = step_bytes(n, 250 * 250) * value
    # secondary: the SRK stepper (complementary force, 2 draws per particle)
    g2, _, _, _ = headline_system(spss, subdiv=subdiv)
    g2.set_cell_layout(subdiv, xsplit)
    srk_steps = max(args.steps // 5, 10)
    wall2, _, _, _ = time_steps(g2, spss, spss.STEPPER_SRK, srk_steps, 10)
    g2.close()
    print(f"[bench] {ms_step:.4f} ms/step ({replayed} replayed, {eager} eager, {captured} graphs captured in the "
          f"timed region), pair kernel {pair_ms * 1e3:.1f} us", file=sys.stderr, flush=True)
    verlet = None
    if not args.no_verlet:
        print("[bench] verlet SRK (large configuration)", file=sys.stderr, flush=True)
        verlet = verlet_srk_rate(spss)
        print(f"[bench] verlet SRK {verlet['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    fp64 = None
    if not args.no_fp64:
        print("[bench] fp64 cell-list step", file=sys.stderr, flush=True)
        fp64 = fp64_step_rate(spss)
    batched = None
    if not args.no_batch:
        print("[bench] batched ensemble", file=sys.stderr, flush=True)
        batched = batched_ensemble(spss, steps=args.batch_steps, plain_steps=args.plain_steps,
                                   n_split=args.batch_split, concurrent=not args.batch_sequential)
        print(f"[bench] batched {batched['speedup_vs_plain_loop']:.1f}x plain loop", file=sys.stderr, flush=True)
    screen = None
    if not args.no_screen:
        print("[bench] learned sweep screen at scale", file=sys.stderr, flush=True)
        screen = screen_at_scale(spss)
        print(f"[bench] screen: {screen['simulated']} of {screen['screened_points']} simulated", file=sys.stderr,
              flush=True)
    # secondary: the reference's deterministic production path, Boost.Odeint's dense-output
    # dopri5 (eps 1e-6) over the global-field RHS of its default run (v0 = 1, xi = 0.1, sigma = 5,
    # rho = 0.8, natural frequencies ~ N(0, 0.5)), observing every dt = 0.01 on the device
    print("[bench] dopri5", file=sys.stderr, flush=True)
    dopri5 = dopri5_run(spss)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(single_steps=args.cpu_single_steps, all_steps=args.cpu_steps)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "N=1,048,576 uniform in the unit box, rho=0.004 cell list refreshed every step, "
                               "stochastic Euler dt=0.002, m=0.5 sigma=1.5 D=0.02, Lorentzian(0,0.05) omega, "
                               "FP32 distance tests, phasor sums in FP32 partials combined in FP64, FP64 state",
                   "n": n, "rho": 0.004, "cell_subdivision": subdiv, "cell_xsplit": xsplit, "cell_autotune": tune,
                   "parallelism": "single GPU"},
        "timed_steps": {"graph_replayed": replayed, "eager": eager, "graphs_captured": captured},
        "roofline": {"bound": "hbm", "achieved": step_achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": step_achieved / HBM_PEAK, "traffic": STEP_TRAFFIC.get((subdiv, xsplit)),
                     "kernel": "resident cell-list step (fr_pairs, fr_far, fr_scatter)",
                     "algorithmic_bytes_per_particle_step": step_bytes(n, 250 * 250),
                     "pair_kernel": {"achieved": pair_achieved / 1e9, "frac": pair_achieved / HBM_PEAK,
                                     "us": pair_ms * 1e3, "bytes_per_particle": pair_kernel_bytes(n, cells),
                                     "traffic": PAIR_KERNEL_TRAFFIC.get((subdiv, xsplit))}},
        # the pair kernel is bound by VALU issue, not HBM (no MFMA work on this path)
        "valu": {"kernel": "fr_pairs", "valu_instr_per_cu_cycle": PAIR_KERNEL_VALU_PER_CU_CYCLE.get((subdiv, xsplit))},
        "cpu_baseline": cpu,
        "pair_kernel_us": pair_ms * 1e3,
        "hbm_copy_gbps": copy_gbps,
        "impl_step_bytes_frac": impl_step_bytes(n, cells) * value / HBM_PEAK,
        "pair_interactions_per_s": pairs / n * value if pairs else None,
        "srk_particle_steps_per_s": n * srk_steps / wall2,
        "verlet_srk": verlet,
        "fp64_cell_list": fp64,
        "batched": batched,
        "screen": screen,
        "dopri5": dopri5,
    }
    g.close()
    print(json.dumps(line), flush=True)


def dopri5_run(spss, n=1 << 20, t1=0.5, dt=0.01):
    """accepted steps, rejected trials and particle-RHS evaluations per second of spss_dopri5_integrate
    (integrate_const, eps_abs = eps_rel = 1e-6) at N = 1,048,576 over [0, t1], device observer"""
    rng = np.random.default_rng(0x5EED0005)
    st = np.empty((n, 4))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.uniform(0, 2 * np.pi, n)
    st[:, 3] = rng.uniform(-50.0, 50.0, n)  # U[-sigma/xi, sigma/xi]
    g = spss.ParticleSystem(model=4, n=n, v0=1.0, xi=0.1, sigma=5.0, rho=0.8, alpha=0.0, D_phi=0.0)
    g.set_natural_frequencies(rng.normal(0.0, 0.5, n))
    g.set_state(st.reshape(-1))
    g.integrate_dopri5(0.0, 2 * dt, dt, 1e-6, 1e-6)  # warm-up
    g.set_state(st.reshape(-1))
    g.synchronize()
    t = time.perf_counter()
    steps, rejected = g.integrate_dopri5(0.0, t1, dt, 1e-6, 1e-6, method=spss.METHOD_GLOBAL_FIELD)
    g.synchronize()
    sec = time.perf_counter() - t
    g.close()
    evals = 6 * (steps + rejected) + 1  # six per trial step (first same as last) and the initial one
    return {"n": n, "t1": t1, "dt": dt, "eps": 1e-6, "method": "global field", "accepted_steps": steps,
            "rejected_trials": rejected, "seconds": sec, "particle_rhs_evals_per_s": evals * n / sec}


def C_pairs(g, spss):
    """sum_i n_i of the current state (count only, no list copy)"""
    import ctypes
    from spsslangevinintegration_amd import _capi
    st = g.get_state()
    cnt = ctypes.c_long()
    _capi.check(g.lib.spss_neighbour_pairs(g.h, spss.METHOD_LINKED_LIST, _capi.ptr(st), None, 0,
                                           ctypes.byref(cnt)))
    return cnt.value


if __name__ == "__main__":
    main()
