"""Python view of a batch of independent replicas (spss_batch_* of the C ABI).

A batch runs many systems of one model (3 or 4) on the same cell grid in single kernel
launches: per replica its own particle count, phase friction xi, coupling sigma, noise D_phi
and Philox seed.  States and natural frequencies are the replicas' arrays one after another.
"""
import ctypes as C

import numpy as np

from . import _capi as A


class ReplicaBatch:
    def __init__(self, model, ns, xi, sigma, D_phi, seeds, v0=0.01, rho=0.03, alpha=0.0, natfreq_local=1,
                 cell_subdivision=1):
        self.lib = A.load()
        R = len(ns)
        reps = (A.SpssReplica * R)()
        for r in range(R):
            reps[r].n = int(ns[r])
            reps[r].xi = float(xi[r])
            reps[r].sigma = float(sigma[r])
            reps[r].D_phi = float(D_phi[r])
            reps[r].seed = int(seeds[r])
        bp = A.SpssBatchParams(model, v0, rho, alpha, cell_subdivision, natfreq_local)
        h = C.c_void_p()
        A.check(self.lib.spss_batch_create(C.byref(bp), R, C.cast(reps, C.c_void_p), C.byref(h)))
        self.h = h
        self.R, self.s = R, model
        self.ns = np.asarray(ns, dtype=np.int64)
        self.first = np.concatenate([[0], np.cumsum(self.ns)])
        self.n = int(self.first[-1])

    def close(self):
        if getattr(self, "h", None):
            self.lib.spss_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
        assert state.size == self.s * self.n
        A.check(self.lib.spss_batch_set_state(self.h, A.ptr(state)))

    def get_state(self):
        out = np.empty(self.s * self.n)
        A.check(self.lib.spss_batch_get_state(self.h, A.ptr(out)))
        return out

    def replica_state(self, state, r):
        return state[self.s * self.first[r]:self.s * self.first[r + 1]]

    def set_natural_frequencies(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.n
        A.check(self.lib.spss_batch_set_natural_frequencies(self.h, A.ptr(w)))

    def reset(self, step=0):
        A.check(self.lib.spss_batch_reset(self.h, int(step)))

    def run(self, nsteps, dt, observe_every=0, observe_from=0, graph=True):
        A.check(self.lib.spss_batch_run(self.h, int(nsteps), float(dt), int(observe_every), int(observe_from),
                                        1 if graph else 0))

    def order_stats(self):
        """(mean |Z|, variance of |Z|, readings, last Z as complex, step counter) per replica"""
        m, v, z_re, z_im = (np.zeros(self.R) for _ in range(4))
        cnt = np.zeros(self.R, dtype=np.int64)
        step = C.c_long()
        A.check(self.lib.spss_batch_order_stats(self.h, A.ptr(m), A.ptr(v), cnt.ctypes.data, A.ptr(z_re),
                                                A.ptr(z_im), C.byref(step)))
        return m, v, cnt, z_re + 1j * z_im, step.value

    def tile_widths(self):
        w = np.zeros(self.R, dtype=np.int32)
        A.check(self.lib.spss_batch_get_tile_widths(self.h, w.ctypes.data))
        return w

    def set_tile_widths(self, w):
        w = np.ascontiguousarray(w, dtype=np.int32)
        A.check(self.lib.spss_batch_set_tile_widths(self.h, w.ctypes.data))

    def autotune_cells(self, dt, probe_steps=20):
        """times every cell layout of CELL_LAYOUTS and keeps the fastest; returns (ms per step for
        each, -1 where one does not build; the chosen subdivision; cell_layout gives the chosen
        (subdivision, x split))"""
        table = np.zeros(len(A.CELL_LAYOUTS))
        chosen = C.c_int()
        A.check(self.lib.spss_batch_autotune_cells(self.h, float(dt), int(probe_steps), A.ptr(table), table.size,
                                                   C.byref(chosen)))
        return table, chosen.value

    def set_cell_layout(self, k, xsplit):
        """cells 1/(k*xsplit*int(1/rho)) wide and 1/(k*int(1/rho)) high (state and step kept)"""
        A.check(self.lib.spss_batch_set_cell_layout(self.h, int(k), int(xsplit)))

    @property
    def cell_layout(self):
        """(subdivision, x split) of the cell grid"""
        k, x = C.c_int(), C.c_int()
        A.check(self.lib.spss_batch_get_cell_layout(self.h, C.byref(k), C.byref(x)))
        return k.value, x.value

    @property
    def cell_subdivision(self):
        k = C.c_int()
        A.check(self.lib.spss_batch_get_cell_subdivision(self.h, C.byref(k)))
        return k.value

    def autotune(self, dt, probe_steps=20):
        """times the candidate tile widths; returns (ms per step for each, chosen index)"""
        table = np.zeros(6)
        chosen = C.c_int()
        A.check(self.lib.spss_batch_autotune(self.h, float(dt), int(probe_steps), A.ptr(table), 6,
                                             C.byref(chosen)))
        return table, chosen.value
