"""Batched PLT BSDF queries on the GPU.

Mirrors, at batch granularity, the eval + pdf pair of the reference's material interface
(Source/Falcor/Rendering/Materials/IMaterialInstance.slang:122-184) as next-event estimation
calls it (Source/RenderPasses/PLTPT/pltpt_solve.rt.slang:219-243, :289-292). Device buffers are
torch tensors (device memory and streams only); all arithmetic runs in libplt_amd.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .spectra import drude_profiles


def _stream_ptr(stream) -> C.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class Context:
    """Owns a plt_context (one GPU) and its uploaded spectral profiles."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.device = device
        h = C.c_void_p()
        _lib.check(self.lib.plt_context_create(device, C.byref(h)))
        self.handle = h
        self.profiles = []

    def set_spectral_profiles(self, profiles):
        arr = (_lib.SpectralProfile * len(profiles))(*profiles)
        _lib.check(self.lib.plt_context_set_spectral_profiles(self.handle, arr, len(profiles)))
        self.profiles = list(profiles)

    def close(self):
        if self.handle:
            self.lib.plt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BSDFBatch:
    """A batch of n BSDF queries (16 SoA planes) and its results (5 planes) in HBM."""

    def __init__(self, ctx: Context, n: int, device: str = "cuda"):
        if n <= 0 or n >= 2 ** 31:
            raise ValueError("batch size must be in [1, 2^31)")
        self.ctx = ctx
        self.n = n
        self.queries = torch.empty((_lib.Q_PLANES, n), dtype=torch.int32, device=device)
        self.results = torch.empty((_lib.R_PLANES, n), dtype=torch.float32, device=device)

    def generate(self, seed: int = 0, stream=None):
        """Synthetic batch (see plt_bsdf_generate_queries in include/plt_abi.h)."""
        _lib.check(self.ctx.lib.plt_bsdf_generate_queries(self.ctx.handle, self.n, seed,
                                                          C.c_void_p(self.queries.data_ptr()), _stream_ptr(stream)))
        return self

    def load(self, queries: torch.Tensor):
        """Copies caller-provided query planes ([16, n] int32/uint32 words)."""
        if tuple(queries.shape) != (_lib.Q_PLANES, self.n):
            raise ValueError(f"expected queries of shape (16, {self.n})")
        self.queries.copy_(queries.view(torch.int32))
        return self

    def eval(self, stream=None) -> torch.Tensor:
        """Runs eval (NonDelta lobe) + pdf for every query; returns the [5, n] result planes."""
        _lib.check(self.ctx.lib.plt_bsdf_eval_queries(self.ctx.handle, self.n, C.c_void_p(self.queries.data_ptr()),
                                                      C.c_void_p(self.results.data_ptr()), _stream_ptr(stream)))
        return self.results


def default_context(device: int = 0) -> Context:
    """Context with the synthetic Drude conductor's n,k profiles at ids 0 and 1."""
    ctx = Context(device)
    ctx.set_spectral_profiles(list(drude_profiles()))
    return ctx


def sample_generator_stream(n_streams: int, width: int, sample_number: int, draws: int, stream=None) -> torch.Tensor:
    """Raw xoshiro128** draws of UniformSampleGenerator(pixel=(s % width, s // width), sample_number)."""
    lib = _lib.load()
    out = torch.empty((draws, n_streams), dtype=torch.int32, device="cuda")
    _lib.check(lib.plt_sample_generator_stream(n_streams, width, sample_number, draws, C.c_void_p(out.data_ptr()),
                                               _stream_ptr(stream)))
    return out


class MaterialTable:
    """A device array of plt_material blocks for the batched material interface
    (plt_material_eval_batch / plt_material_sample_batch)."""

    def __init__(self, ctx: Context, materials: bytes, device: str = "cuda"):
        self.ctx = ctx
        self.n = len(materials) // C.sizeof(_lib.Material)
        self.data = torch.frombuffer(bytearray(materials), dtype=torch.uint8).to(device)

    def eval(self, queries: torch.Tensor, stream=None) -> torch.Tensor:
        """queries: float32 [PLT_MQ_PLANES, n] on the device; returns [PLT_MR_PLANES, n]."""
        if queries.dtype != torch.float32 or queries.dim() != 2 or queries.shape[0] != _lib.MQ_PLANES or not queries.is_cuda:
            raise ValueError(f"queries must be a float32 CUDA tensor of shape [{_lib.MQ_PLANES}, n]")
        q = queries.contiguous()
        n = q.shape[1]
        out = torch.empty((_lib.MR_PLANES, n), dtype=torch.float32, device=q.device)
        _lib.check(_lib.load().plt_material_eval_batch(self.ctx.handle, C.c_void_p(self.data.data_ptr()), self.n, n,
                                                       C.c_void_p(q.data_ptr()), C.c_void_p(out.data_ptr()),
                                                       _stream_ptr(stream)))
        return out

    def sample(self, queries: torch.Tensor, frame: int = 0, stream=None) -> torch.Tensor:
        """queries: float32 [PLT_MSQ_PLANES, n] (wi xyz, material index bits, wavelength); returns
        [PLT_MSR_PLANES, n] (wo xyz, pdf, weight, lobe bits, valid bits)."""
        if queries.dtype != torch.float32 or queries.dim() != 2 or queries.shape[0] != _lib.MSQ_PLANES or not queries.is_cuda:
            raise ValueError(f"queries must be a float32 CUDA tensor of shape [{_lib.MSQ_PLANES}, n]")
        q = queries.contiguous()
        n = q.shape[1]
        out = torch.empty((_lib.MSR_PLANES, n), dtype=torch.float32, device=q.device)
        _lib.check(_lib.load().plt_material_sample_batch(self.ctx.handle, C.c_void_p(self.data.data_ptr()), self.n, n,
                                                         frame, C.c_void_p(q.data_ptr()), C.c_void_p(out.data_ptr()),
                                                         _stream_ptr(stream)))
        return out
