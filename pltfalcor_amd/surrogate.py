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


def train(sur: Surrogate, samples: ExactSamples, epochs: int, batch: int = 16384, lr: float = 2e-3, seed: int = 0,
          log=None):
    """Standardises the targets of `samples` (valid M00 > 0 only), sets the statistics, and runs
    `epochs` passes of Adam with a cosine-decayed learning rate. Returns the per-epoch mean loss
    (sampled every 16 batches)."""
    import torch
    M, X = samples.M.double(), samples.X
    ok = torch.isfinite(M).all(1) & (M[:, 0] > 1e-12)
    m00 = torch.where(ok, M[:, 0], torch.ones_like(M[:, 0]))
    y = torch.asinh(M / m00[:, None])
    y[:, 0] = torch.log(m00)
    yv = y[ok]
    mean, std = yv.mean(0), yv.std(0, unbiased=False)
    std = torch.where(std > 1e-6, std, torch.ones_like(std)) / torch.tensor(OUTPUT_SCALE, device=std.device)
    Xv = X[ok]
    sur.set_stats(mean.cpu().numpy(), std.cpu().numpy(), Xv.min(0).values.cpu().numpy(), Xv.max(0).values.cpu().numpy())
    T = ((y - mean) / std).float()
    del y, yv, Xv, M
    idx = torch.nonzero(ok).flatten()
    gen = torch.Generator(device=idx.device).manual_seed(seed)
    n = (idx.numel() // batch) * batch
    steps_total = epochs * (n // batch)
    step = 0
    losses = []
    for ep in range(epochs):
        perm = idx[torch.randperm(idx.numel(), generator=gen, device=idx.device)[:n]]
        ep_loss = []
        for b in range(0, n, batch):
            sel = perm[b:b + batch]
            xb = X.index_select(0, sel).contiguous()
            tb = T.index_select(0, sel).contiguous()
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * step / max(1, steps_total)))
            want = (b // batch) % 16 == 0
            l = sur.train_step(xb, tb, max(lr_t, 1e-6), want_loss=want)
            if want:
                ep_loss.append(l)
            step += 1
        losses.append(float(np.mean(ep_loss)))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.5f}")
    return losses


__all__ = ["Surrogate", "ExactSamples", "train", "fit_stats", "mueller_to_targets", "targets_to_mueller", "FAMILIES"]
