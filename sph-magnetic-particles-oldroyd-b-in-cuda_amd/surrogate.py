"""Learned surrogate of the sweep: BF16-MFMA MLP ensemble over the C-ABI.

`Ensemble` binds include/sph_surrogate.h (training and inference run in the
library's HIP kernels; there is no CPU path).  `Surrogate` adds what a sweep
needs around it: standardisation of the inputs (log10 Wi, log10 Mn, phi, V) and
of the targets (wall stress, N1, phase lag; NaN targets are masked out of the
loss), members that differ by initialisation and minibatch order (optionally
bootstrap resamples), and the ensemble spread that decides
when a point must be re-run with SPH.  `sweep_with_fallback` answers a sweep
from the surrogate and re-runs, as one batch, every point whose ensemble
relative spread of the wall stress exceeds the threshold (5 % by default).
"""
import ctypes
import math
import time

import numpy as np

import sph_system
import sweep

N_IN, N_OUT = 4, 3


def _lib():
    lib = sph_system.library()
    if getattr(lib, "_mlp_bound", False):
        return lib
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float
    lib.sph_mlp_create.restype = vp
    lib.sph_mlp_create.argtypes = [i32, ctypes.c_uint64]
    lib.sph_mlp_destroy.restype = None
    lib.sph_mlp_destroy.argtypes = [vp]
    lib.sph_mlp_last_error.restype = ctypes.c_char_p
    lib.sph_mlp_last_error.argtypes = [vp]
    lib.sph_mlp_param_count.restype = i64
    lib.sph_mlp_param_count.argtypes = [vp]
    lib.sph_mlp_layout.restype = i32
    lib.sph_mlp_layout.argtypes = [vp, vp, vp]
    for fn in ("sph_mlp_set_params", "sph_mlp_get_params"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp, vp, i64]
    lib.sph_mlp_forward.restype = i32
    lib.sph_mlp_forward.argtypes = [vp, i32, vp, vp]
    lib.sph_mlp_gradients.restype = i32
    lib.sph_mlp_gradients.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    lib.sph_mlp_train_step.restype = i32
    lib.sph_mlp_train_step.argtypes = [vp, i32, vp, vp, vp, f32, vp]
    lib.sph_mlp_synchronize.restype = i32
    lib.sph_mlp_synchronize.argtypes = [vp]
    lib._mlp_bound = True
    return lib


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Ensemble:
    """E members of 4 -> 256 -> 256 -> 256 -> 3 (SiLU) on the GPU."""

    def __init__(self, members=5, seed=0):
        self.lib = _lib()
        self.h = self.lib.sph_mlp_create(int(members), int(seed))
        if not self.h:
            raise RuntimeError("sph_mlp_create failed (bad member count or no GPU)")
        self.E = members
        self.P = self.lib.sph_mlp_param_count(self.h) // members

    def close(self):
        if self.h:
            self.lib.sph_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, call):
        if rc != 0:
            raise sph_system.SphError(rc, self.lib.sph_mlp_last_error(self.h).decode(), call)

    def layout(self):
        off = np.zeros(8, np.int64)
        shp = np.zeros(8, np.int32)
        self._check(self.lib.sph_mlp_layout(self.h, _p(off), _p(shp)), "mlp_layout")
        return off, shp

    def get_params(self):
        out = np.zeros(self.E * self.P, np.float32)
        self._check(self.lib.sph_mlp_get_params(self.h, _p(out), out.size), "mlp_get_params")
        return out

    def set_params(self, params):
        p = _f32(params)
        self._check(self.lib.sph_mlp_set_params(self.h, _p(p), p.size), "mlp_set_params")

    @staticmethod
    def _pad(a, B):
        n = a.shape[1]
        if n == B:
            return a
        out = np.zeros((a.shape[0], B) + a.shape[2:], a.dtype)
        out[:, :n] = a
        return out

    def forward(self, x):
        """x [E, n, 4] or [n, 4] (shared by all members) -> y [E, n, 3]."""
        x = _f32(x)
        if x.ndim == 2:
            x = np.broadcast_to(x, (self.E,) + x.shape)
        n = x.shape[1]
        B = max(64, (n + 63) // 64 * 64)
        xp = _f32(self._pad(x, B))
        y = np.zeros((self.E, B, N_OUT), np.float32)
        self._check(self.lib.sph_mlp_forward(self.h, B, _p(xp), _p(y)), "mlp_forward")
        return y[:, :n]

    def gradients(self, x, t, mask):
        x, t, mask = _f32(x), _f32(t), _f32(mask)
        B = x.shape[1]
        g = np.zeros(self.E * self.P, np.float32)
        loss = np.zeros(self.E, np.float32)
        self._check(self.lib.sph_mlp_gradients(self.h, B, _p(x), _p(t), _p(mask), _p(g),
                                               _p(loss)), "mlp_gradients")
        return g, loss

    def train_step(self, x, t, mask, lr, want_loss=True):
        """One Adam step; returns the members' losses before it (None when
        want_loss is False: the step is then queued without waiting)."""
        x, t, mask = _f32(x), _f32(t), _f32(mask)
        B = x.shape[1]
        loss = np.zeros(self.E, np.float32) if want_loss else None
        self._check(self.lib.sph_mlp_train_step(self.h, B, _p(x), _p(t), _p(mask),
                                                ctypes.c_float(lr),
                                                _p(loss) if want_loss else None),
                    "mlp_train_step")
        return loss


def feature_matrix(points):
    """Raw (log10 Wi, log10 Mn, phi, V) per point."""
    f = np.array([sweep.features(p) for p in points], np.float64)
    out = np.empty_like(f)
    out[:, 0] = np.log10(np.maximum(f[:, 0], 1e-6))
    out[:, 1] = np.log10(np.maximum(f[:, 1], 1e-12))
    out[:, 2] = f[:, 2]
    out[:, 3] = f[:, 3]
    return out


class Surrogate:
    """Ensemble + standardisation + training + spread-based fallback.

    The network learns the wall stress as the apparent viscosity
    sigma / (eta gamma_dot) = sigma / (2 V) (eta = 1, L1 = 1), which stays O(1)
    for every wall speed, and multiplies back by 2 V; its spread relative to its
    mean is then the relative spread of the wall stress itself.  Rows with
    V < 1e-3 leave that target out (the ratio is noise there)."""

    def __init__(self, members=5, seed=0):
        self.net = Ensemble(members, seed)
        self.seed = seed
        self.xm = self.xs = self.ym = self.ys = None

    @staticmethod
    def _to_net(X, Y):
        Y = np.array(Y, np.float64)
        V = np.asarray(X, np.float64)[:, 3]
        Y[:, 0] = np.where(V >= 1e-3, Y[:, 0] / (2.0 * np.maximum(V, 1e-3)), np.nan)
        return Y

    def fit(self, X, Y, steps=3000, batch=256, lr=1e-3, lr_end=None, bootstrap=False,
            log_every=0, log=None):
        """X [n, 4] raw features, Y [n, 3] targets (NaN = missing).  Every
        member draws its own minibatches from all rows (members differ by their
        initialisation and minibatch order: a deep ensemble); with `bootstrap`
        each draws from its own bootstrap resample of the rows instead.  (On the
        2D wall + colloid sweep the resamples roughly doubled the points whose
        spread exceeded 5 %, at the same error of the accepted points.)  With
        `lr_end` the learning rate decays from `lr` to `lr_end` on a half
        cosine."""
        X = np.asarray(X, np.float64)
        Y = self._to_net(X, Y)
        self.xm, self.xs = X.mean(0), X.std(0) + 1e-12
        self.ym = np.nanmean(Y, 0)
        self.ys = np.nanstd(Y, 0) + 1e-12
        self.ym = np.where(np.isfinite(self.ym), self.ym, 0.0)
        self.ys = np.where(np.isfinite(self.ys), self.ys, 1.0)
        Xs = (X - self.xm) / self.xs
        M = np.isfinite(Y).astype(np.float64)
        Ys = np.where(M > 0, (Y - self.ym) / self.ys, 0.0)
        rng = np.random.default_rng(self.seed)
        E, n = self.net.E, X.shape[0]
        boot = [rng.integers(0, n, n) if bootstrap else np.arange(n) for _ in range(E)]
        hist = []
        for s in range(steps):
            rate = lr if lr_end is None else (
                lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * s / steps)))
            idx = np.stack([b[rng.integers(0, n, batch)] for b in boot])
            probe = s % 10 == 0 or s == steps - 1     # read the loss back every 10 steps
            loss = self.net.train_step(Xs[idx], Ys[idx], M[idx], rate, want_loss=probe)
            if probe:
                hist.append(float(loss.mean()))
                if log and log_every and (s % log_every == 0 or s == steps - 1):
                    log(f"step {s}: loss {hist[-1]:.4e}")
        return np.array(hist)

    def predict(self, X):
        """(mean [n, 3], std [n, 3]) over the members, in target units."""
        X = np.asarray(X, np.float64)
        Xs = (X - self.xm) / self.xs
        y = self.net.forward(Xs).astype(np.float64) * self.ys + self.ym   # [E, n, 3]
        y[:, :, 0] *= 2.0 * X[:, 3]
        return y.mean(0), y.std(0)


def sweep_with_fallback(points, model, steps, threshold=0.05, precision=sph_system.FP32,
                        max_particles=4_000_000, log=None):
    """Answers every point from the surrogate; re-runs with batched SPH the
    points whose ensemble relative spread of the wall stress exceeds
    `threshold`.  Returns (targets [n, 3], info)."""
    t0 = time.perf_counter()
    mean, std = model.predict(feature_matrix(points))
    t_inf = time.perf_counter() - t0
    rel = std[:, 0] / np.maximum(np.abs(mean[:, 0]), 1e-12)
    redo = np.flatnonzero(~(rel <= threshold))
    out = mean.copy()
    t_sph = 0.0
    if redo.size:
        sub = [dict(points[i]) for i in redo]
        t1 = time.perf_counter()
        y, _ = sweep.run_points(sub, steps, precision, max_particles, log=log)
        t_sph = time.perf_counter() - t1
        out[redo] = y
    info = dict(inference_s=t_inf, fallback_s=t_sph, fallback=int(redo.size),
                seconds=time.perf_counter() - t0, rel_spread=rel)
    return out, info
