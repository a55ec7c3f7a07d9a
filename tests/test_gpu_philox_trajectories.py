"""Stochastic steps driven by the device's own Philox noise, against the oracle fed the same
draws (regenerated on the host through spss_philox_normals).

This pins the NOISE_PHILOX path end to end: counter layout (particle, block, step), key
(seed), the draw each stepper consumes, and the step counter of spss_run.  FP64 mode:
pointwise within 1e-12 over 10 steps.  FP32C mode (resident cell order): within the FP32
pair-sum tolerance.
"""
import numpy as np
import pytest

import oracle
from spsslangevinintegration_amd import _capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _state(n, seed):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.random(n) * 2 * np.pi
    st[:, 3] = rng.normal(0, 0.3, n)
    return st.reshape(-1)


@pytest.mark.parametrize("stepper,precision", [("euler", 0), ("srk", 0), ("euler", 1), ("srk", 1)])
def test_philox_trajectory_vs_oracle(spss, stepper, precision):
    n, rho, dt, seed, steps = 6000, 0.04, 0.005, 0x5EED0042, 10
    print(f"[gpu] philox trajectory {stepper} precision={precision}", flush=True)
    kw = dict(model=4, n=n, v0=0.02, xi=1.3, sigma=2.0, rho=rho, alpha=0.1, D_phi=0.07, natfreq_local=1)
    natf = np.random.default_rng(1).normal(0, 0.2, n)
    st = _state(n, 3)
    g = spss.ParticleSystem(precision=precision, seed=seed, **kw)
    g.set_natural_frequencies(natf)
    g.set_state(st)
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(natf)
    so = st.copy()
    sid = spss.STEPPER_SRK if stepper == "srk" else spss.STEPPER_EULER
    g.run(sid, spss.METHOD_LINKED_LIST, steps, dt, spss.NOISE_PHILOX)
    for k in range(steps):
        z0, z1 = _capi.philox_normals(seed, k, n)
        if stepper == "srk":
            o.srk_step(spss.METHOD_LINKED_LIST, so, dt, np.stack([z0, z1], axis=1).reshape(-1))
        else:
            o.euler_step(spss.METHOD_LINKED_LIST, so, dt, z0)
    d = np.abs(g.get_state() - so).reshape(-1, 4)
    print(f"  max |gpu - oracle|: pos {d[:, :2].max():.2e} phase {d[:, 2].max():.2e} omega {d[:, 3].max():.2e}",
          flush=True)
    if precision == 0:
        assert d.max() <= 1e-12
    else:
        assert d[:, :2].max() <= 1e-12 and d[:, 2].max() <= 1e-9 and d[:, 3].max() <= 1e-7
