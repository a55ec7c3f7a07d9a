"""Replay a golden fixture (written by tests/golden/gen_golden.py from the
reference's own code) through an engine with the oracle's method names.

An engine factory is called as factory(**params) and must return an object
with rk4_step, euler_step, srk_step, evaluate_rhs, set_natural_frequencies,
set_higher_order, complementary_force and verlet_lists (as oracle.Oracle).
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

METHODS = {"all_pairs": 2, "linked_list": 3, "global": 1, "dispatch": 0}


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["fixtures"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def engine_kwargs(p):
    return dict(model=p["model"], n=p["n"], v0=p["v0"], xi=p["xi"], sigma=p["sigma"], rho=p["rho"],
                alpha=p["alpha"], D_phi=p["D_phi"], xi_r=p["xi_r"], D_r=p["D_r"])


def verlet_from_flat(flat, n):
    lists, pos = [], 0
    for _ in range(n):
        c = int(flat[pos])
        lists.append(flat[pos + 1:pos + 1 + c].copy())
        pos += 1 + c
    return lists


def relerr(a, b):
    """max |a-b| / max |b| per state slot (slots have different scales)."""
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def replay(name, factory, wrap_fn=None):
    """Run fixture `name` through factory(...); return dict of comparisons:
    {"states": [(step, got, want)], "compl": (got, want), "knext": (got, want), "verlet": (got, want)}"""
    meta = manifest()[name]
    p = meta["params"]
    fx = load(name)
    n, s = p["n"], p["model"]
    method = METHODS.get(p.get("method", "dispatch"), 0)
    eng = factory(**engine_kwargs(p))
    if p.get("natfreq"):
        eng.set_natural_frequencies(fx["natf"])
    state = fx["state_init"].copy()
    res = {"states": []}
    snaps = set(meta["snapshots"])
    stepper = p["stepper"]
    if stepper == "rhs":
        eng.set_higher_order(p.get("higher_order", 0))
        kn = np.zeros(n * s)
        eng.evaluate_rhs(0, state, fx["kprev"], kn, p["k_coef"], p["dt"])
        res["knext"] = (kn, fx["knext"])
        if p.get("higher_order"):
            res["compl"] = (eng.complementary_force(), fx["compl"])
        if "verlet_flat" in fx:
            res["verlet"] = (eng.verlet_lists(), verlet_from_flat(fx["verlet_flat"], n))
        return res
    for k in range(1, p["steps"] + 1):
        if stepper in ("rk4", "rk4_direct"):
            eng.rk4_step(method, state, p["dt"])
        elif stepper == "euler":
            eng.euler_step(method, state, p["dt"], fx["noise"][k - 1])
        elif stepper == "srk":
            eng.srk_step(method, state, p["dt"], fx["noise"][k - 1])
        else:
            raise ValueError(stepper)
        if p.get("wrap"):
            wrap_fn(state, n, s)
        if k == 1 and "verlet_flat" in fx and stepper != "srk":
            res["verlet"] = (eng.verlet_lists(), verlet_from_flat(fx["verlet_flat"], n))
        if k == 1 and stepper == "srk" and "compl" in fx:
            res["compl"] = (eng.complementary_force(), fx["compl"])
        if k in snaps:
            idx = list(fx["snap_steps"]).index(k)
            res["states"].append((k, state.copy(), fx["snap_states"][idx]))
    return res
