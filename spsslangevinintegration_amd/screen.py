"""Learned sweep screen: a small network trained on steady-state order parameters of batched
runs picks where along a parameter sweep full simulations are spent (no counterpart in the
reference, which sweeps parameters in a host loop, SimulationEngineFirstOrder.cpp:105-113).

* `MLP` wraps the spss_mlp_* entry points: forward, backward and Adam run as the library's own
  HIP kernels on the device.
* `SweepScreen` maps a replica's parameters (sigma, D, m, v0, rho and the density N) to the
  time-mean and the standard deviation of |Z| in the steady state, simulates points with
  `ReplicaBatch`, and trains the network on them.  Two uses: `screen` simulates, out of many
  points, only those where the predicted |d mean|Z| / d sigma| is large or the predicted
  spread of |Z| exceeds a bound, and scores the network against those simulations;
  `screened_sweep` locates the crossing of a level of mean |Z| along a sigma sweep with few
  simulations.  The simulation is both the fallback and the arbiter: every point the screen
  reports on is simulated, and a sweep whose simulated points do not bracket the crossing (or
  disagree with the prediction by more than `max_error`) is widened until they do, up to the
  plain sweep.
"""
import ctypes as C

import numpy as np

from . import _capi as A
from .batch import ReplicaBatch

FEATURES = ("sigma", "D", "m", "v0", "rho", "log2_N")
OUTPUTS = ("mean_R", "std_R")
DIMS = (6, 128, 128, 128, 2)


class MLP:
    """FP32 network on the device: GELU hidden layers, linear output, MSE loss, Adam"""

    def __init__(self, dims=DIMS, max_batch=1024, seed=0x5EED0005):
        self.lib = A.load()
        self.dims = tuple(int(d) for d in dims)
        d = (C.c_int * len(self.dims))(*self.dims)
        h = C.c_void_p()
        A.check(self.lib.spss_mlp_create(C.byref(h), C.cast(d, C.c_void_p), len(self.dims), int(max_batch),
                                         int(seed)))
        self.h = h
        self.max_batch = int(max_batch)
        n = C.c_long()
        A.check(self.lib.spss_mlp_param_count(self.h, C.byref(n)))
        self.nparam = n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.spss_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        out = np.empty(self.nparam, dtype=np.float32)
        A.check(self.lib.spss_mlp_get_params(self.h, A.ptr(out), self.nparam))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        A.check(self.lib.spss_mlp_set_params(self.h, A.ptr(p), p.size))

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        y = np.empty((x.shape[0], self.dims[-1]), dtype=np.float32)
        A.check(self.lib.spss_mlp_predict(self.h, A.ptr(x), x.shape[0], A.ptr(y)))
        return y

    def gradients(self, x, t):
        """(loss, flat gradient) of mean((y - t)^2) over one batch"""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.dims[-1])
        assert x.shape[0] == t.shape[0]
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_float()
        A.check(self.lib.spss_mlp_gradients(self.h, A.ptr(x), A.ptr(t), x.shape[0], A.ptr(g), C.byref(loss)))
        return loss.value, g

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.dims[-1])
        assert x.shape[0] == t.shape[0]
        A.check(self.lib.spss_mlp_set_data(self.h, A.ptr(x), A.ptr(t), x.shape[0]))

    def train(self, nsteps, batch, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        """nsteps Adam steps over the resident data; the loss of every step"""
        losses = np.empty(int(nsteps), dtype=np.float32)
        A.check(self.lib.spss_mlp_train(self.h, int(nsteps), int(batch), float(lr), float(betas[0]),
                                        float(betas[1]), float(eps), A.ptr(losses) if nsteps else None))
        return losses

    def reset_optimizer(self):
        A.check(self.lib.spss_mlp_reset_optimizer(self.h))

    @property
    def adam_steps(self):
        n = C.c_long()
        A.check(self.lib.spss_mlp_adam_steps(self.h, C.byref(n)))
        return n.value


def points(sigma, D, m, v0, rho, N, seed, omega_std=0.2, coherence0=1.0):
    """structured array of sweep points (broadcast over the arguments).  By default runs start
    from identical phases (coherence0 = 1): the steady state is then that of the synchronised
    branch, which survives where coupling beats noise and frequency spread and decays
    elsewhere -- a transition that shows within a few thousand steps even with local coupling,
    where synchrony grown from random phases would need far longer runs."""
    vals = (sigma, D, m, v0, rho, N, seed, omega_std, coherence0)
    cols = np.broadcast_arrays(*(np.asarray(v) for v in vals))
    p = np.zeros(cols[0].size, dtype=[("sigma", "f8"), ("D", "f8"), ("m", "f8"), ("v0", "f8"), ("rho", "f8"),
                                      ("N", "i8"), ("seed", "u8"), ("omega_std", "f8"), ("coherence0", "f8")])
    for name, c in zip(p.dtype.names, cols):
        p[name] = c.reshape(-1)
    return p


def random_points(n, rng, sigma=(0.0, 3.0), D=(0.0, 0.2), m=(0.2, 1.0), v0=(0.005, 0.01, 0.02),
                  rho=(0.02, 0.03, 0.04), N=(2048, 8192), seed0=0x5EED1000, **fixed):
    """n points: sigma, D, m uniform in their ranges, N log-uniform, v0 and rho drawn from the
    given values (points of one (v0, rho) share a batch)"""
    u = rng.random((n, 4))
    s = sigma[0] + u[:, 0] * (sigma[1] - sigma[0])
    d = D[0] + u[:, 1] * (D[1] - D[0])
    mm = m[0] + u[:, 2] * (m[1] - m[0])
    nn = np.round(2.0 ** (np.log2(N[0]) + u[:, 3] * (np.log2(N[1]) - np.log2(N[0])))).astype(np.int64)
    vv = rng.choice(np.asarray(v0, dtype=np.float64), n)
    rr = rng.choice(np.asarray(rho, dtype=np.float64), n)
    return points(s, d, mm, vv, rr, nn, seed0 + np.arange(n, dtype=np.uint64), **fixed)


def initial_state(p, model):
    """positions U[0,1)^2, phases 2 pi (1 - c) U[0,1) (mean |Z| from 0 at c = 0 to 1 at c = 1),
    phase velocity 0; natural frequencies N(0, omega_std^2) (divided by m for model 4); all
    drawn from the point's own seed, so a point simulates the same alone or in any batch"""
    rng = np.random.default_rng(int(p["seed"]))
    n = int(p["N"])
    st = np.zeros((n, model))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = 2.0 * np.pi * (1.0 - p["coherence0"]) * rng.random(n)
    w = rng.normal(0.0, p["omega_std"], n)
    return st.reshape(-1), (w / p["m"] if model == 4 else w)


def simulate(pts, steps=3000, dt=0.005, observe_every=10, observe_from=1500):
    """(mean |Z|, std |Z|) per point from full simulations: one batch per (model, v0, rho).
    m > 0 runs model 4 with xi = 1/m, sigma/m, D/m^2, omega/m; m = 0 the first-order model 3."""
    out = np.zeros((pts.size, 2))
    model_of = np.where(pts["m"] > 0, 4, 3)
    keys = sorted(set(zip(model_of.tolist(), pts["v0"].tolist(), pts["rho"].tolist())))
    for model, v0, rho in keys:
        idx = np.nonzero((model_of == model) & (pts["v0"] == v0) & (pts["rho"] == rho))[0]
        q = pts[idx]
        mm = np.where(q["m"] > 0, q["m"], 1.0)
        xi = 1.0 / mm if model == 4 else np.zeros(idx.size)
        b = ReplicaBatch(model, q["N"], xi, q["sigma"] / mm, q["D"] / mm ** 2, q["seed"], v0=v0, rho=rho)
        try:
            init = [initial_state(p, model) for p in q]
            b.set_state(np.concatenate([s for s, _ in init]))
            b.set_natural_frequencies(np.concatenate([w for _, w in init]))
            b.reset(0)
            b.run(steps, dt, observe_every=observe_every, observe_from=observe_from)
            mean_R, var_R, _, _, _ = b.order_stats()
        finally:
            b.close()
        out[idx, 0] = mean_R
        out[idx, 1] = np.sqrt(np.maximum(var_R, 0.0))
    return out


class SweepScreen:
    """network screen over sweep points; see the module docstring"""

    # feature ranges mapped onto [-1, 1]: sigma, D, m, v0, rho, log2 N
    RANGES = np.array([(0.0, 4.0), (0.0, 0.2), (0.0, 1.0), (0.0, 0.02), (0.0, 0.05), (11.0, 16.0)])

    def __init__(self, seed=0x5EED0005, sim_kwargs=None, net=None):
        self.net = net or MLP(DIMS, max_batch=1024, seed=seed)
        self.sim_kwargs = dict(sim_kwargs or {})
        self.simulations = 0

    def features(self, pts):
        raw = np.stack([pts["sigma"], pts["D"], pts["m"], pts["v0"], pts["rho"],
                        np.log2(pts["N"].astype(np.float64))], axis=1)
        lo, hi = self.RANGES[:, 0], self.RANGES[:, 1]
        return (2.0 * (raw - lo) / (hi - lo) - 1.0).astype(np.float32)

    def simulate(self, pts):
        self.simulations += pts.size
        return simulate(pts, **self.sim_kwargs)

    def fit(self, pts, targets, steps=4000, batch=128, lr=2e-3, seed=0):
        """train on (points, simulated targets); returns the per-step losses"""
        order = np.random.default_rng(seed).permutation(pts.size)
        self.net.set_data(self.features(pts[order]), np.asarray(targets, dtype=np.float32)[order])
        self.net.reset_optimizer()
        l1 = self.net.train(int(steps * 3 // 4), batch, lr=lr)
        l2 = self.net.train(steps - int(steps * 3 // 4), batch, lr=lr * 0.1)
        return np.concatenate([l1, l2])

    def predict(self, pts):
        return self.net.predict(self.features(pts)).astype(np.float64)

    def slope(self, pts, h=0.02):
        """d mean|Z| / d sigma of the network (central difference)"""
        up, dn = pts.copy(), pts.copy()
        up["sigma"] += h
        dn["sigma"] -= h
        return (self.predict(up)[:, 0] - self.predict(dn)[:, 0]) / (2.0 * h)

    def screen(self, pts, slope_min=0.5, spread_max=0.05, simulate=True):
        """predict every point; simulate only those with |d mean|Z| / d sigma| > slope_min or a
        predicted spread of |Z| > spread_max.  Returns a dict: the prediction for all points, the
        indices selected, their simulated values and the network's error on them."""
        pred = self.predict(pts)
        sel = np.nonzero((np.abs(self.slope(pts)) > slope_min) | (pred[:, 1] > spread_max))[0]
        out = {"predicted": pred, "selected": sel}
        if simulate and sel.size:
            sim = self.simulate(pts[sel])
            out["simulated"] = sim
            out["error"] = np.abs(pred[sel] - sim)
        return out

    @staticmethod
    def crossing(sigma, R, level):
        """sigma where R first reaches `level` (linear interpolation), None if it never does"""
        hit = np.nonzero(R >= level)[0]
        if hit.size == 0:
            return None
        i = int(hit[0])
        if i == 0:
            return float(sigma[0])
        a, b = R[i - 1], R[i]
        return float(sigma[i - 1] + (level - a) / (b - a) * (sigma[i] - sigma[i - 1]))

    def plain_sweep(self, line, level=0.5):
        """the method without the network: simulate every point of the line"""
        R = self.simulate(line)[:, 0]
        return self.crossing(line["sigma"], R, level), R

    def screened_sweep(self, line, level=0.5, window=6, max_error=None):
        """locate the crossing of `level` along a line sorted by sigma with few simulations:
        simulate `window` points around the predicted crossing and widen the window (towards
        the side the simulations point to) until it brackets the first crossing or covers the
        line.  With `max_error` set, a line whose simulated points differ from the prediction by
        more falls back to the plain sweep.  Returns (sigma at the crossing or None, indices
        simulated, their simulated mean |Z|, largest |prediction - simulation| on them)."""
        G = line.size
        pred = self.predict(line)[:, 0]
        hit = np.nonzero(pred >= level)[0]
        c = int(hit[0]) if hit.size else G - 1
        b = min(G, max(0, c - window // 2) + window)
        a = max(0, b - window)
        sim = np.full(G, np.nan)

        def run(lo, hi):
            need = [k for k in range(lo, hi) if np.isnan(sim[k])]
            if need:
                sim[need] = self.simulate(line[need])[:, 0]

        run(a, b)
        while True:
            if max_error is not None and np.nanmax(np.abs(pred - sim)) > max_error:
                a, b = 0, G  # the network is off here: plain sweep of the line
                run(a, b)
                break
            first = np.nonzero(sim[a:b] >= level)[0]
            if first.size == 0:
                if b == G:
                    break  # no crossing on the line
                b = min(G, b + window)
            elif first[0] == 0 and a > 0:
                a = max(0, a - window)
            else:
                break
            run(a, b)
        idx = np.arange(a, b)
        x = self.crossing(line["sigma"][idx], sim[idx], level)
        return x, idx, sim[idx], float(np.max(np.abs(pred[idx] - sim[idx])))
