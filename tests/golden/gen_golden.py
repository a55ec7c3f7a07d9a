#!/usr/bin/env python3
"""Generate golden vectors by running the reference's own code.

Runs oracle/_ref/ref_ptr and oracle/_ref/ref_serial (the reference sources
compiled in place by oracle/Makefile, driven by oracle/ref_harness.cpp) on
seeded inputs and stores inputs and outputs as small .npz fixtures in this
directory, with a manifest.json holding parameters and SHA-256 sums.

Each configuration runs in its own process: the reference keeps
function-local statics (stage buffers, noise intensities) that would leak
between configurations.  Run from the repository root:

    make -C oracle all && python tests/golden/gen_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

BASE = dict(model=4, v0=0.01, xi=1.0, sigma=1.0, rho=0.1, alpha=0.0, D_phi=0.0, xi_r=1.0, D_r=0.0,
            dt=0.01, steps=1, seed=20240611)

# name -> (binary, overrides, snapshots, init seed)
CONFIGS = {
    # config-1 sizes on the inertial (s = 4) model: all-pairs RK4, 1000 steps, FP64
    "rk4_allpairs_s4": ("ref_ptr", dict(n=1024, rho=0.1, stepper="rk4_direct", method="all_pairs",
                                        steps=1000), [0, 1, 10, 100, 1000], 11),
    "rk4_linkedlist_s4": ("ref_ptr", dict(n=1024, rho=0.1, alpha=0.3, stepper="rk4_direct",
                                          method="linked_list", steps=200), [0, 1, 10, 200], 12),
    # EvaluateRhs dispatch: rho <= 0.25 -> linked list + Verlet (compat), fast movers force rebuilds
    "rk4_llverlet_s4": ("ref_ptr", dict(n=1024, rho=0.1, v0=0.5, stepper="rk4", steps=60, dump_lists=1),
                        [0, 1, 5, 6, 7, 20, 60], 13),
    "rk4_verlet_s4": ("ref_ptr", dict(n=256, rho=0.3, v0=0.3, stepper="rk4", steps=100, dump_lists=1),
                      [0, 1, 9, 10, 100], 14),
    "rk4_global_s4": ("ref_ptr", dict(n=1000, rho=0.8, alpha=0.3, xi=0.1, stepper="rk4", steps=200,
                                      natfreq=1), [0, 1, 50, 200], 15),
    "rk4_allpairs_dispatch_s4": ("ref_ptr", dict(n=256, rho=0.6, alpha=0.2, stepper="rk4", steps=100),
                                 [0, 1, 100], 16),
    "euler_s4": ("ref_ptr", dict(n=256, rho=0.1, D_phi=0.05, stepper="euler", steps=40), [0, 1, 10, 40], 17),
    "srk_s4": ("ref_ptr", dict(n=256, rho=0.1, D_phi=0.05, stepper="srk", steps=40), [0, 1, 10, 40], 18),
    "srk_global_s4": ("ref_ptr", dict(n=256, rho=0.8, alpha=0.3, xi=0.1, D_phi=0.01, stepper="srk",
                                      steps=30), [0, 1, 30], 19),
    "rhs_llverlet_s4": ("ref_ptr", dict(n=512, rho=0.1, alpha=0.4, stepper="rhs", k_coef=0.5,
                                        higher_order=1, dump_lists=1), [], 20),
    "rhs_verlet_s4": ("ref_ptr", dict(n=256, rho=0.3, alpha=0.4, stepper="rhs", k_coef=0.5,
                                      higher_order=1, dump_lists=1), [], 21),
    "rhs_global_s4": ("ref_ptr", dict(n=256, rho=0.75, alpha=0.4, stepper="rhs", k_coef=0.5,
                                      higher_order=1, natfreq=1), [], 22),
    "rhs_allpairs_s4": ("ref_ptr", dict(n=256, rho=0.6, alpha=0.4, stepper="rhs", k_coef=0.5,
                                        higher_order=1), [], 23),
    "rk4_ll_s6": ("ref_serial", dict(model=6, n=1024, rho=0.1, alpha=0.5, v0=1.0, xi_r=1.0, stepper="rk4",
                                     steps=100, wrap=1), [0, 1, 10, 100], 24),
    "rk4_allpairs_s6": ("ref_serial", dict(model=6, n=256, rho=0.3, alpha=1.54, v0=1.0, stepper="rk4",
                                           steps=100, wrap=1), [0, 1, 100], 25),
    "rk4_ll_s6_ptr": ("ref_ptr", dict(model=6, n=512, rho=0.1, alpha=1.5, v0=1.0, stepper="rk4",
                                      steps=50, wrap=1), [0, 1, 50], 26),
    "euler_s6": ("ref_serial", dict(model=6, n=256, rho=0.1, alpha=0.5, v0=1.0, D_r=0.01, D_phi=0.05,
                                    stepper="euler", steps=40, wrap=1), [0, 1, 40], 27),
    "pairs_ll_s4": ("ref_serial", dict(n=2048, rho=0.03, stepper="pairs_ll"), [], 28),
}


def init_state(model, n, sigma, xi, seed):
    """x, y ~ U[0,1), phi ~ U[0,2pi), omega ~ U[-sigma/xi, sigma/xi] (SimulationEngineFirstOrderPtr.cpp:79-99);
    s = 6: theta ~ U[0,2pi), v = (cos, sin) theta, phi = theta, omega = 0 (SimulationEngineSecondOrder.cpp:33-56)."""
    rng = np.random.default_rng(seed)
    if model == 4:
        st = np.empty((n, 4))
        st[:, 0] = rng.random(n)
        st[:, 1] = rng.random(n)
        st[:, 2] = rng.random(n) * 2.0 * np.pi
        st[:, 3] = rng.uniform(-sigma / xi, sigma / xi, n)
    else:
        st = np.empty((n, 6))
        st[:, 0] = rng.random(n)
        st[:, 1] = rng.random(n)
        th = rng.random(n) * 2.0 * np.pi
        st[:, 2] = np.cos(th)
        st[:, 3] = np.sin(th)
        st[:, 4] = th
        st[:, 5] = 0.0
    natf = rng.normal(0.0, 0.1, n)
    return st.reshape(-1), natf


def read_output(path):
    rec = {"STATE": {}, "NOISE": {}, "COMPL": {}, "KPREV": {}, "KNEXT": {}}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        tag = data[pos:pos + 8].rstrip(b"\0").decode()
        pos += 8
        if tag in ("VERLET", "BUCKETS"):
            cnt = int(np.frombuffer(data, np.int64, 1, pos)[0])
            pos += 8
            rec[tag] = np.frombuffer(data, np.int32, cnt, pos).copy()
            pos += 4 * cnt
            continue
        step = int(np.frombuffer(data, np.int64, 1, pos)[0])
        pos += 8
        cnt = int(np.frombuffer(data, np.int64, 1, pos)[0])
        pos += 8
        rec[tag][step] = np.frombuffer(data, np.float64, cnt, pos).copy()
        pos += 8 * cnt
    return rec


def run_one(name, spec):
    binary, over, snaps, seed = spec
    cfg = dict(BASE)
    cfg.update(over)
    n, model = cfg["n"], cfg["model"]
    state, natf = init_state(model, n, cfg["sigma"], cfg["xi"], seed)
    with tempfile.TemporaryDirectory() as td:
        cpath, ipath, opath = (os.path.join(td, x) for x in ("cfg.txt", "in.bin", "out.bin"))
        with open(cpath, "w") as f:
            for k, v in cfg.items():
                f.write(f"{k}={v!r}\n" if isinstance(v, float) else f"{k}={v}\n")
            f.write("snapshots=" + ",".join(str(s) for s in snaps) + "\n")
        with open(ipath, "wb") as f:
            f.write(state.tobytes())
            f.write(natf.tobytes())
        subprocess.run([os.path.join(REF_DIR, binary), cpath, ipath, opath], check=True,
                       stdout=subprocess.DEVNULL)
        rec = read_output(opath)
    arrays = {"state_init": state, "natf": natf}
    if rec["STATE"]:
        steps = sorted(rec["STATE"])
        arrays["snap_steps"] = np.array(steps, dtype=np.int64)
        arrays["snap_states"] = np.stack([rec["STATE"][k] for k in steps])
    if rec["NOISE"]:
        steps = sorted(rec["NOISE"])
        arrays["noise"] = np.stack([rec["NOISE"][k] for k in steps])
    for tag, key in (("COMPL", "compl"), ("KPREV", "kprev"), ("KNEXT", "knext")):
        if rec[tag]:
            arrays[key] = rec[tag][sorted(rec[tag])[0]]
    for tag, key in (("VERLET", "verlet_flat"), ("BUCKETS", "buckets_flat")):
        if tag in rec:
            arrays[key] = rec[tag]
    out = os.path.join(HERE, name + ".npz")
    np.savez_compressed(out, **arrays)
    with open(out, "rb") as f:
        sha = hashlib.sha256(f.read()).hexdigest()
    return dict(binary=binary, params=cfg, snapshots=snaps, init_seed=seed, sha256=sha)


def main(names=None):
    if not os.path.exists(os.path.join(REF_DIR, "ref_ptr")):
        sys.exit("oracle/_ref not built: run `make -C oracle ref` where the reference checkout exists")
    manifest = {
        "generator": "tests/golden/gen_golden.py",
        "reference_build": "oracle/Makefile: g++ -std=c++17 -O2 -ffp-contract=off (no -march=native)",
        "fixtures": {},
    }
    mpath = os.path.join(HERE, "manifest.json")
    if os.path.exists(mpath):
        with open(mpath) as f:
            manifest["fixtures"] = json.load(f).get("fixtures", {})
    for name, spec in CONFIGS.items():
        if names and name not in names:
            continue
        print("generating", name, flush=True)
        manifest["fixtures"][name] = run_one(name, spec)
    with open(mpath, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
