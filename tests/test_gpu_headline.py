"""Parity at the benchmark's own workload: N = 1,048,576 particles uniform in the unit box,
rho = 0.004 (250 x 250 cells, 16.8 particles per cell, ~52.7 neighbours), the inertial model with
m = 0.5, sigma = 1.5, D = 0.02 (reference variables xi = 2, sigma = 3, D_phi = 0.08), natural
frequencies omega_i / m with omega_i ~ Lorentzian(0, 0.05) clipped at +-10, v0 = 0.005, dt = 0.002.

* the FP32C neighbour set (the row tiles the resident step stages, wider than the fixed tiles at
  this density) equals the oracle's linked-list pairs (ParticleSystemFirstOrderPtr.cpp:318-468)
  bit for bit, on the initial state and on a state 200 steps later;
* one FP32C RHS without and with the complementary force (the SRK form) against the oracle:
  slots 0-2 within 1e-12, the coupling slot within 1e-6 absolute, the complementary force within
  1e-6 of max(1, |C|);
* the FP64 HIP path against the oracle for the first two Philox-driven Euler and SRK steps
  (pointwise, 1e-12), then FP32C and FP64 fed the same Philox draws for 5000 Euler steps (the
  large configuration's horizon) and 1000 SRK steps: |R_FP32C - R_FP64| <= 1e-4 at every
  checkpoint (the accuracy target of the FP32-compensated path).

The FP32C checks run on the reference's square cells and on the cells split 4 ways in x that the
cell autotuner picks for this workload (spss_set_cell_layout(1, 4)).
"""
import time

import numpy as np
import pytest

import oracle
from oracle.pyoracle import METHOD_LINKED_LIST
from spsslangevinintegration_amd import _capi

pytestmark = pytest.mark.gpu

N = 1 << 20
RHO = 0.004
DT = 0.002
SEED = 0x5EED0003
M, SIGMA, D = 0.5, 1.5, 0.02
KW = dict(model=4, n=N, v0=0.005, xi=1.0 / M, sigma=SIGMA / M, rho=RHO, alpha=0.0, D_phi=D / M ** 2,
          natfreq_local=1)
TOL_R = 1e-4
TOL_COUPLING = 1e-6
HOST_THREADS = 16
LAYOUTS = [(1, 1), (1, 4)]  # (subdivision, x split) of the FP32C cell grid


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


@pytest.fixture(scope="module")
def init():
    rng = np.random.default_rng(SEED)
    st = np.zeros((N, 4))
    st[:, 0] = rng.random(N)
    st[:, 1] = rng.random(N)
    st[:, 2] = rng.random(N) * 2.0 * np.pi
    w = np.clip(0.05 * np.tan(np.pi * (rng.random(N) - 0.5)), -10.0, 10.0) / M
    return st.reshape(-1), w


def _system(spss, init, precision, layout=(1, 1)):
    st, w = init
    g = spss.ParticleSystem(seed=SEED, precision=precision, **KW)
    if layout != (1, 1):
        g.set_cell_layout(*layout)
    g.set_natural_frequencies(w)
    g.set_state(st)
    return g


def _oracle(init, higher=False):
    o = oracle.Oracle(**KW)
    o.set_natural_frequencies(init[1])
    o.set_threads(HOST_THREADS)
    if higher:
        o.set_higher_order(1)
    return o


@pytest.mark.parametrize("layout", LAYOUTS)
def test_headline_pairs_bitexact(spss, init, layout):
    g = _system(spss, init, spss.PRECISION_FP32C, layout)
    o = oracle.Oracle(**KW)
    states = [("initial", init[0])]
    g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 200, DT, spss.NOISE_PHILOX)
    states.append(("after 200 steps", g.get_state()))
    for name, st in states:
        print(f"[gpu] headline pairs N={N} rho={RHO} layout={layout} ({name})", flush=True)
        t0 = time.time()
        got = g.neighbour_pairs(spss.METHOD_LINKED_LIST, st)
        t1 = time.time()
        want = o.pairs_linked_list(st, cap=N * 64)
        print(f"  {len(got)} pairs on the GPU ({t1 - t0:.1f} s), {len(want)} in the oracle "
              f"({time.time() - t1:.1f} s)", flush=True)
        assert got.shape == want.shape
        assert np.array_equal(got, want)
        del got, want
    g.close()


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("higher", [False, True])
def test_headline_rhs_vs_oracle(spss, init, higher, layout):
    print(f"[gpu] headline RHS N={N} higher_order={higher} layout={layout}", flush=True)
    st = init[0].copy().reshape(-1, 4)
    st[:, 3] = np.random.default_rng(5).normal(0.0, 0.5, N)  # omega != 0 so the SRK force is not trivial
    st = st.reshape(-1)
    g = _system(spss, init, spss.PRECISION_FP32C, layout)
    g.set_higher_order(higher)
    o = _oracle(init, higher)
    kg = np.zeros(4 * N)
    ko = np.zeros(4 * N)
    g.evaluate_rhs(spss.METHOD_LINKED_LIST, st, kg.copy(), kg, 0.0, DT)
    o.evaluate_rhs(METHOD_LINKED_LIST, st, ko.copy(), ko, 0.0, DT)
    for slot in range(3):
        assert np.max(np.abs(kg[slot::4] - ko[slot::4])) <= 1e-12, slot
    err = np.max(np.abs(kg[3::4] - ko[3::4]))
    print(f"  coupling slot max |gpu - oracle| = {err:.3e}", flush=True)
    assert err <= TOL_COUPLING
    if higher:
        cg, co = g.complementary_force(), o.complementary_force()
        errc = np.max(np.abs(cg - co))
        print(f"  complementary force max |gpu - oracle| = {errc:.3e} (max |C| {np.max(np.abs(co)):.2f})",
              flush=True)
        assert errc <= TOL_COUPLING * max(1.0, np.max(np.abs(co)))
    g.close()


@pytest.mark.parametrize("stepper", ["euler", "srk"])
def test_headline_fp64_path_vs_oracle_two_steps(spss, init, stepper):
    print(f"[gpu] headline FP64 {stepper} vs oracle, 2 steps", flush=True)
    g = _system(spss, init, spss.PRECISION_FP64)
    o = _oracle(init)
    so = init[0].copy()
    sid = spss.STEPPER_SRK if stepper == "srk" else spss.STEPPER_EULER
    g.run(sid, spss.METHOD_LINKED_LIST, 2, DT, spss.NOISE_PHILOX)
    for k in range(2):
        z0, z1 = _capi.philox_normals(SEED, k, N)
        if stepper == "srk":
            o.srk_step(METHOD_LINKED_LIST, so, DT, np.stack([z0, z1], axis=1).reshape(-1))
        else:
            o.euler_step(METHOD_LINKED_LIST, so, DT, z0)
    d = np.abs(g.get_state() - so)
    print(f"  max |gpu - oracle| = {d.max():.2e}", flush=True)
    assert d.max() <= 1e-12
    g.close()


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("stepper,steps", [("euler", 5000), ("srk", 1000)])
def test_headline_fp32c_order_parameter_vs_fp64(spss, init, stepper, steps, layout):
    print(f"[gpu] headline FP32C vs FP64 {stepper}, {steps} steps, layout {layout}", flush=True)
    sid = spss.STEPPER_SRK if stepper == "srk" else spss.STEPPER_EULER
    g32 = _system(spss, init, spss.PRECISION_FP32C, layout)
    g64 = _system(spss, init, spss.PRECISION_FP64)
    chunk = 500
    worst = 0.0
    t0 = time.time()
    for done in range(chunk, steps + 1, chunk):
        g32.run(sid, spss.METHOD_LINKED_LIST, chunk, DT, spss.NOISE_PHILOX)
        g64.run(sid, spss.METHOD_LINKED_LIST, chunk, DT, spss.NOISE_PHILOX)
        r32 = np.hypot(*g32.order_parameter(2))
        r64 = np.hypot(*g64.order_parameter(2))
        worst = max(worst, abs(r32 - r64))
        print(f"  step {done}: R_fp32c {r32:.8f} R_fp64 {r64:.8f} |diff| {abs(r32 - r64):.2e} "
              f"({time.time() - t0:.1f} s)", flush=True)
        assert abs(r32 - r64) <= TOL_R
    s32, s64 = g32.get_state().reshape(-1, 4), g64.get_state().reshape(-1, 4)
    dphi = np.abs(np.angle(np.exp(1j * (s32[:, 2] - s64[:, 2]))))
    print(f"  after {steps} steps: max |d phi| {dphi.max():.2e}, median {np.median(dphi):.2e}; "
          f"max |d x| {np.abs(s32[:, :2] - s64[:, :2]).max():.2e}; worst |dR| {worst:.2e}", flush=True)
    g32.close()
    g64.close()
