"""Host mirror of the reference's PLTPT render pass (Source/RenderPasses/PLTPT/PLTPT.{h,cpp}).

`PLTPT(ctx, dict)` accepts the reference's dictionary keys (PLTPT.cpp:102-134); unknown keys
warn and are ignored as the reference's parseDictionary does (PLTPT.cpp:173). `execute` renders
one frame into caller-owned RGBA32F device buffers and advances the frame counter.
"""
from __future__ import annotations

import ctypes as C
import warnings

import torch

from . import _lib


class RenderData(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("color", C.c_void_p), ("albedo", C.c_void_p),
                ("normal", C.c_void_p), ("vbuffer", C.c_void_p)]


class PLTPT:
    """The PLT path tracer pass on one GPU (ctx None: options only, execute raises)."""

    def __init__(self, ctx, options: dict | None = None):
        self.lib = _lib.load()
        self.ctx = ctx
        h = C.c_void_p()
        _lib.check(self.lib.plt_pass_create(ctx.handle if ctx is not None else None, b"PLTPT", C.byref(h)))
        self.handle = h
        self.scene = None
        for k, v in (options or {}).items():
            self.set_option(k, v)

    def set_option(self, key: str, value) -> bool:
        rc = self.lib.plt_pass_set_option(self.handle, key.encode(), float(value))
        if rc == _lib.PLT_WARN_UNKNOWN_KEY:
            warnings.warn(f"Unknown field '{key}' in PLTPathTracer dictionary.")
            return False
        _lib.check(rc)
        return True

    def get_option(self, key: str) -> float:
        v = C.c_double()
        _lib.check(self.lib.plt_pass_get_option(self.handle, key.encode(), C.byref(v)))
        return v.value

    def get_scripting_dictionary(self) -> dict:
        out, i = {}, 0
        while True:
            name = self.lib.plt_pass_option_name(i)
            if not name:
                return out
            out[name.decode()] = self.get_option(name.decode())
            i += 1

    def set_scene(self, scene):
        _lib.check(self.lib.plt_pass_set_scene(self.handle, scene.handle if scene is not None else None))
        self.scene = scene

    def set_frame(self, frame: int):
        _lib.check(self.lib.plt_pass_set_frame(self.handle, frame))

    def execute(self, width: int, height: int, color: torch.Tensor, albedo: torch.Tensor, normal: torch.Tensor,
                stream=None):
        for t in (color, albedo, normal):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != 4 * width * height:
                raise ValueError("outputs must be contiguous float32 CUDA tensors of width*height*4 elements")
        rd = RenderData(width, height, color.data_ptr(), albedo.data_ptr(), normal.data_ptr(), None)
        if stream is None:
            stream = torch.cuda.current_stream()
        _lib.check(self.lib.plt_pass_execute(self.handle, C.byref(rd), C.c_void_p(stream.cuda_stream)))

    def render(self, width: int, height: int, frames: int = 1, first_frame: int = 0):
        """Renders `frames` frames and returns their mean colour (an FP64 accumulation of the outputs) and the
        last frame's albedo/normal, all as CPU tensors [height, width, 4]."""
        dev = torch.device("cuda")
        color = torch.empty((height, width, 4), dtype=torch.float32, device=dev)
        albedo = torch.empty_like(color)
        normal = torch.empty_like(color)
        acc = torch.zeros((height, width, 4), dtype=torch.float64, device=dev)
        self.set_frame(first_frame)
        for _ in range(frames):
            self.execute(width, height, color, albedo, normal)
            acc += color.double()
        torch.cuda.synchronize()
        return (acc / frames).cpu(), albedo.cpu(), normal.cpu()

    def bounce_records(self, width: int, height: int):
        """The last frame's bounce records as uint32 [maxBounces+1, 9, width*height]: per bounce the hit
        triangle, barycentrics, octahedral wi, lobe, flags, path contribution and last non-delta pdf."""
        import numpy as np
        nb = int(self.get_option("maxBounces")) + 1
        out = np.empty((nb, 9, width * height), np.uint32)
        torch.cuda.synchronize()
        _lib.check(self.lib.plt_pass_copy_bounce_records(self.handle, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    MEGAKERNEL, WAVEFRONT = 0, 1
    CENTER, DIRECTX, HALTON, STRATIFIED = 0, 1, 2, 3
    SOLVE_MEGAKERNEL, SOLVE_SPLIT = 0, 1

    def set_sample_pattern(self, pattern: int, sample_count: int = 16):
        """Camera jitter of the G-buffer pass before PLTPT (GBufferRT's samplePattern / sampleCount;
        the reference graph uses Halton, 16): each execute jitters the camera by the next offset."""
        _lib.check(self.lib.plt_pass_set_sample_pattern(self.handle, int(pattern), int(sample_count)))

    def jitter(self):
        """(jitterX, jitterY) of the last execute, as fractions of the frame size."""
        xy = (C.c_float * 2)()
        _lib.check(self.lib.plt_pass_get_jitter(self.handle, xy))
        return xy[0], xy[1]

    def set_schedule(self, schedule: int):
        """Sample-pass schedule: PLTPT.MEGAKERNEL (default; one kernel, a lane per path) or
        PLTPT.WAVEFRONT (material-sorted per-bounce queues); both write identical records."""
        _lib.check(self.lib.plt_pass_set_schedule(self.handle, int(schedule)))

    def set_launch_config(self, schedule: int | None = None, sample_block: int | None = None,
                          wf_blocks_per_cu: int | None = None, graph: bool | None = None,
                          solve_schedule: int | None = None, solve_events_per_path: int | None = None,
                          solve_hits_per_path: int | None = None):
        """Launch parameters (plt_launch_config); None keeps the current value."""
        c = self.launch_config()
        for k, v in (("schedule", schedule), ("sample_block", sample_block), ("wf_blocks_per_cu", wf_blocks_per_cu),
                     ("graph", graph), ("solve_schedule", solve_schedule),
                     ("solve_events_per_path", solve_events_per_path), ("solve_hits_per_path", solve_hits_per_path)):
            if v is not None:
                c[k] = int(v)
        lc = _lib.LaunchConfig(*(c[k] for k, _ in _lib.LaunchConfig._fields_))
        _lib.check(self.lib.plt_pass_set_launch_config(self.handle, C.byref(lc)))

    def launch_config(self) -> dict:
        lc = _lib.LaunchConfig()
        _lib.check(self.lib.plt_pass_get_launch_config(self.handle, C.byref(lc)))
        return {k: getattr(lc, k) for k, _ in lc._fields_}

    def set_graph(self, enable: bool):
        """Replay the sample and solve passes from captured HIP graphs (per-frame arguments updated)."""
        _lib.check(self.lib.plt_pass_set_graph(self.handle, int(bool(enable))))

    def autotune(self, width: int, height: int, color: torch.Tensor, albedo: torch.Tensor, normal: torch.Tensor,
                 frames: int = 2, stream=None):
        """Times the candidate launch configurations on `frames` frames, keeps the fastest (frame
        counter and jitter pattern restored). Returns (chosen config, default ms/frame, best ms/frame)."""
        rd = RenderData(width, height, color.data_ptr(), albedo.data_ptr(), normal.data_ptr(), None)
        if stream is None:
            stream = torch.cuda.current_stream()
        best = _lib.LaunchConfig()
        d, b = C.c_double(), C.c_double()
        _lib.check(self.lib.plt_pass_autotune(self.handle, C.byref(rd), int(frames), C.c_void_p(stream.cuda_stream),
                                              C.byref(best), C.byref(d), C.byref(b)))
        return {k: getattr(best, k) for k, _ in best._fields_}, d.value, b.value

    def set_profiling(self, enable: bool):
        _lib.check(self.lib.plt_pass_set_profiling(self.handle, int(bool(enable))))

    def kernel_times(self):
        """(sample_ms, solve_ms, executes) summed since the previous call; waits for the recorded events."""
        a, b, n = C.c_double(), C.c_double(), C.c_uint32()
        _lib.check(self.lib.plt_pass_get_kernel_times(self.handle, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def kernel_times_split(self):
        """(sample_ms, light_ms, walk_ms, unwind_ms, megakernel_ms, executes): the solve pass resolved
        into the split pass's light-sampling, manifold-walk and unwind kernels and the solve megakernel
        (see the C ABI)."""
        v = [C.c_double() for _ in range(5)]
        n = C.c_uint32()
        _lib.check(self.lib.plt_pass_get_kernel_times_split(self.handle, *(C.byref(x) for x in v), C.byref(n)))
        return (*(x.value for x in v), n.value)

    SURROGATE_NETWORK, SURROGATE_TEACHER = 0, 1

    def set_surrogate(self, surrogate, mode: int = 0):
        """Renders with the learned BSDF surrogate (a trained pltfalcor_amd.surrogate.Surrogate) standing
        for the coated and grated scattering lobes, or exactly again with None. mode SURROGATE_TEACHER
        puts the exact lobe's matrix in the network's place (a check of the in-render plumbing)."""
        self._surrogate = surrogate  # kept alive while the pass may use it
        _lib.check(self.lib.plt_pass_set_surrogate(self.handle, None if surrogate is None else surrogate.handle, int(mode)))

    def surrogate_stats(self) -> dict:
        """Rows the network evaluated and eligible steps evaluated exactly, in the last execute."""
        r, x = C.c_uint64(), C.c_uint64()
        _lib.check(self.lib.plt_pass_get_surrogate_stats(self.handle, C.byref(r), C.byref(x)))
        return {"rows": r.value, "exact_steps": x.value}

    def surrogate_queries(self, max_rows: int = 1 << 22):
        """The last execute's surrogate query rows [rows, 22] (numpy float32)."""
        import numpy as np
        out = np.zeros((max_rows, 22), np.float32)
        n = C.c_uint64()
        _lib.check(self.lib.plt_pass_copy_surrogate_queries(self.handle, C.c_void_p(out.ctypes.data), max_rows, C.byref(n)))
        return out[:n.value]

    def solve_stats(self) -> dict:
        """Counts of the last execute's split solve pass (waits for the device): light events, stored
        walk vertices, replayed vertices (sum of vertex index + 1 over events), overflowed paths,
        manifold walks run by the walk kernel, those that failed, and the rays the walk kernel cast."""
        v = [C.c_uint64() for _ in range(7)]
        _lib.check(self.lib.plt_pass_get_solve_stats(self.handle, *(C.byref(x) for x in v)))
        keys = ("events", "walk_vertices", "event_vertices", "overflowed_paths", "walks", "failed_walks", "walk_rays")
        return dict(zip(keys, (x.value for x in v)))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.plt_pass_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
