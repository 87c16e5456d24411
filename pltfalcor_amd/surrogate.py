"""The learned BSDF surrogate (no counterpart in the reference): a 22-128-128-128-16 ReLU MLP with
matrix-core forward/backward/Adam kernels (libplt_amd.so, plt_surrogate_*), its exact training
samples, target transform and training loop.

Targets: y0 = ln M00, yk = asinh(Mk / M00) (k = 1..15) of the exact Mueller matrix M (row-major
4x4), standardised per output with the training set's mean and standard deviation. (The ratios
are compressed because the coated dielectric's interference term yields matrices with
|Mk| up to hundreds of times M00.)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Material

N_IN, N_OUT = 22, 16
FAMILIES = ("coated opaque dielectric", "coated conductor", "grated conductor")


def _p(t):
    return C.c_void_p(t.data_ptr())


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mueller_to_targets(M):
    """[n, 16] Mueller matrices -> [n, 16] raw targets (ln M00, asinh(Mk / M00))."""
    M = np.asarray(M, np.float64)
    m00 = M[:, 0]
    y = np.arcsinh(M / m00[:, None])
    y[:, 0] = np.log(m00)
    return y


def targets_to_mueller(y):
    y = np.asarray(y, np.float64)
    m00 = np.exp(y[:, 0])
    M = np.sinh(y) * m00[:, None]
    M[:, 0] = m00
    return M


class ExactSamples:
    """Exact-BSDF samples on the device (torch tensors): materials, geometry planes, inputs X [n, 22]
    and Mueller matrices M [n, 16]."""

    def __init__(self, ctx, n: int, seed: int):
        import torch
        dev = torch.device("cuda")
        self.n = n
        self.materials = torch.empty((n, C.sizeof(Material)), dtype=torch.uint8, device=dev)
        self.geometry = torch.empty((9, n), dtype=torch.float32, device=dev)
        self.X = torch.empty((n, N_IN), dtype=torch.float32, device=dev)
        self.M = torch.empty((n, N_OUT), dtype=torch.float32, device=dev)
        _lib.check(_lib.load().plt_surrogate_generate(ctx.handle, n, seed, _p(self.materials), _p(self.geometry), _p(self.X),
                                                     _p(self.M), None))


class Surrogate:
    """Owns a plt_surrogate (device parameters, Adam state and work buffers)."""

    def __init__(self, ctx, max_batch: int = 16384, seed: int = 1):
        self.ctx = ctx
        self.lib = _lib.load()
        self.handle = C.c_void_p()
        _lib.check(self.lib.plt_surrogate_create(ctx.handle, max_batch, seed, C.byref(self.handle)))
        self.max_batch = max_batch
        self.stats = None

    def close(self):
        if self.handle:
            self.lib.plt_surrogate_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def param_count(self) -> int:
        return int(self.lib.plt_surrogate_param_count())

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count, np.float32)
        _lib.check(self.lib.plt_surrogate_params(self.handle, _fp(p), 0))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.param_count
        _lib.check(self.lib.plt_surrogate_params(self.handle, _fp(p), 1))

    def forward(self, X):
        import torch
        Y = torch.empty((X.shape[0], N_OUT), dtype=torch.float32, device=X.device)
        _lib.check(self.lib.plt_surrogate_forward(self.handle, X.shape[0], _p(X), _p(Y), None))
        return Y

    def gradients(self, X, T):
        g = np.zeros(self.param_count, np.float32)
        loss = C.c_float(0.0)
        _lib.check(self.lib.plt_surrogate_gradients(self.handle, X.shape[0], _p(X), _p(T), _fp(g), C.byref(loss), None))
        return g, float(loss.value)

    def train_step(self, X, T, lr: float, want_loss: bool = False):
        loss = C.c_float(0.0)
        _lib.check(self.lib.plt_surrogate_train_step(self.handle, X.shape[0], _p(X), _p(T), lr,
                                                     C.byref(loss) if want_loss else None, None))
        return float(loss.value) if want_loss else None

    def get_stats(self):
        """The statistics set on the device: (y_mean [16], y_std [16], x_lo [22], x_hi [22])."""
        arrs = [np.zeros(N_OUT, np.float32), np.zeros(N_OUT, np.float32), np.zeros(N_IN, np.float32), np.zeros(N_IN, np.float32)]
        _lib.check(self.lib.plt_surrogate_get_stats(self.handle, *[_fp(a) for a in arrs]))
        return arrs

    def set_stats(self, y_mean, y_std, x_lo, x_hi):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (y_mean, y_std, x_lo, x_hi)]
        _lib.check(self.lib.plt_surrogate_set_stats(self.handle, *[_fp(a) for a in arrs]))
        self.stats = arrs

    def eval(self, samples: ExactSamples, stokes_in, mode: int = 0):
        """Outgoing Stokes [n, 4]: mode 0 surrogate (exact for out-of-distribution inputs), 1 exact.
        Returns (stokes_out, fallback count)."""
        import torch
        out = torch.empty_like(stokes_in)
        fb = C.c_uint32(0)
        _lib.check(self.lib.plt_surrogate_eval(self.handle, samples.n, _p(samples.materials), _p(samples.geometry),
                                               _p(samples.X), _p(stokes_in), _p(out), mode, C.byref(fb), None))
        return out, int(fb.value)


# Loss weights per output, applied through the standardisation (an output scaled by 1/s weighs s^2
# in the mean-squared error): the radiance row (ln M00, then M01..M03) first.
OUTPUT_SCALE = np.r_[4.0, 2.0, 2.0, 2.0, np.ones(12)]


def fit_stats(X, M):
    """Target mean / scale and the inputs' box from a training set (host arrays)."""
    y = mueller_to_targets(M)
    mean, std = y.mean(0), y.std(0)
    std = np.where(std > 1e-6, std, 1.0) / OUTPUT_SCALE
    return mean, std, X.min(0), X.max(0)


def prepare(sur: Surrogate, samples: ExactSamples):
    """The training set on the device (plt_surrogate_prepare): standardised targets T [n, 16], the
    valid rows' indices (the first n_valid of idx), and the statistics set on `sur` (also kept as
    sur.stats). Returns (T, idx, n_valid)."""
    import torch
    n = samples.n
    T = torch.empty((n, N_OUT), dtype=torch.float32, device=samples.X.device)
    idx = torch.empty(n, dtype=torch.int32, device=samples.X.device)
    nv = C.c_uint32(0)
    _lib.check(sur.lib.plt_surrogate_prepare(sur.handle, n, _p(samples.X), _p(samples.M), _p(T), _p(idx), C.byref(nv), None))
    sur.stats = sur.get_stats()
    return T, idx, int(nv.value)


def train(sur: Surrogate, samples: ExactSamples, epochs: int, batch: int = 16384, lr: float = 2e-3, seed: int = 0,
          log=None, graph: bool = True):
    """Standardises the targets of `samples` (valid M00 > 0 only) and sets the statistics
    (`prepare`), then runs `epochs` passes of Adam with a cosine-decayed learning rate in the
    library (`plt_surrogate_train`: shuffle, gather and every step on the device, one step captured
    as a HIP graph and replayed when `graph`). Returns the per-epoch mean batch loss."""
    T, idx, nv = prepare(sur, samples)
    losses = np.zeros(epochs, np.float32)
    _lib.check(sur.lib.plt_surrogate_train(sur.handle, nv, _p(idx), _p(samples.X), _p(T), batch, epochs, lr, seed,
                                           1 if graph else 0, _fp(losses), None))
    if log:
        for ep, l in enumerate(losses):
            log(f"epoch {ep}: loss {l:.5f}")
    return [float(x) for x in losses]


__all__ = ["Surrogate", "ExactSamples", "train", "prepare", "fit_stats", "mueller_to_targets", "targets_to_mueller", "FAMILIES"]
