"""spss_order_parameter_noalloc (the entry point that allocates nothing) against spss_order_parameter
(the handle's entry point with its own scratch): the same kernels and reduction order, so the
results must agree bit for bit; two calls in a row, the second on a smaller input, then the size
check.  The device output buffer is the state array of a one-particle handle (the library's own
allocation; torch is not loaded next to the library)."""
import ctypes as C

import numpy as np
import pytest

from spsslangevinintegration_amd import _capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _noalloc(lib):
    f = lib.spss_order_parameter_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    return f


def _device_sync():
    # the HIP runtime the library is linked against (already loaded in this process)
    hip = C.CDLL("libamdhip64.so.7")
    assert hip.hipDeviceSynchronize() == 0


def _state(n, s, seed):
    rng = np.random.default_rng(seed)
    st = rng.random((n, s))
    st[:, 2 if s != 6 else 4] = rng.normal(1.0, 0.7, n)  # partly coherent phases
    return st.reshape(-1)


@pytest.mark.parametrize("s,slot", [(4, 2), (6, 4), (3, 2)])
def test_noalloc_order_parameter_equals_handle_entry_point(spss, s, slot):
    lib = _capi.load()
    f = _noalloc(lib)
    out = spss.ParticleSystem(model=3, n=1, rho=0.05)  # its 3-double state array holds the result
    out.set_state(np.zeros(3))
    for n in (300_001, 4_099):  # two calls in a row, the second smaller
        print(f"[gpu] order parameter noalloc s={s} n={n}", flush=True)
        g = spss.ParticleSystem(model=s, n=n, rho=0.05)
        g.set_state(_state(n, s, n + s))
        want = g.order_parameter(slot)
        g.synchronize()
        rc = f(g.state_device_ptr(), n, s, slot, out.state_device_ptr(), None)
        assert rc == 0, lib.spss_last_error()
        _device_sync()  # the call went to the null stream
        got = out.get_state()[:2]
        print(f"  handle {want}, noalloc {tuple(got)}", flush=True)
        assert got[0] == want[0] and got[1] == want[1]
        g.close()
    out.close()


def test_noalloc_rejects_sizes_beyond_its_scratch(spss):
    lib = _capi.load()
    f = _noalloc(lib)
    h = spss.ParticleSystem(model=3, n=1, rho=0.05)
    p = h.state_device_ptr()
    rc = f(p, 16777216 + 1, 4, 2, p, None)
    assert rc == 4  # SPSS_ERR_CAPACITY, before any launch
    assert b"at most" in lib.spss_last_error()
    h.close()
