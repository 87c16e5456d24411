"""Host mirrors of the two passes that consume PLTPT's output in the reference's render graph
(scripts/PLTPathTracer.py:19-27,36-53): AccumulatePass (Source/RenderPasses/AccumulatePass) and
ToneMapper (Source/RenderPasses/ToneMapper), over the C ABI of libplt_amd.so.

Both take the reference's dictionary keys (unknown keys warn and are ignored, as their
parseDictionary does) and expose the reference's scripting properties. Images are contiguous CUDA
tensors of [height, width, 4] texels.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
import warnings

import torch

from . import _lib


class AccumulatePrecision(enum.IntEnum):
    """AccumulatePass::Precision (AccumulatePass.h:72-77)."""
    Double = 0
    Single = 1
    SingleCompensated = 2


class AccumulateOverflowMode(enum.IntEnum):
    """AccumulatePass::OverflowMode (AccumulatePass.h:79-84)."""
    Stop = 0
    Reset = 1
    EMA = 2


class ToneMapOp(enum.IntEnum):
    """ToneMapperOperator (ToneMapperParams.slang:33-41)."""
    Linear = 0
    Reinhard = 1
    ReinhardModified = 2
    HejiHableAlu = 3
    HableUc2 = 4
    Aces = 5


class ExposureMode(enum.IntEnum):
    """ToneMapper::ExposureMode (ToneMapper.h:47-51)."""
    AperturePriority = 0
    ShutterPriority = 1


_INPUT_FORMATS = {torch.float32: 0, torch.int32: 2}


def _stream_handle(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _check_image(t: torch.Tensor, what: str, dtypes=(torch.float32,)):
    if t.dtype not in dtypes or not t.is_cuda or not t.is_contiguous() or t.dim() != 3 or t.shape[2] != 4:
        raise ValueError(f"{what} must be a contiguous CUDA tensor [height, width, 4] of {dtypes}")


class _Pass:
    _prefix = ""
    _dict_name = ""

    def __init__(self, ctx, options: dict | None = None):
        self.lib = _lib.load()
        self.ctx = ctx
        h = C.c_void_p()
        _lib.check(getattr(self.lib, f"plt_{self._prefix}_create")(ctx.handle if ctx is not None else None, C.byref(h)))
        self.handle = h
        for k, v in (options or {}).items():
            self.set_option(k, v)

    def set_option(self, key: str, value) -> bool:
        rc = getattr(self.lib, f"plt_{self._prefix}_set_option")(self.handle, key.encode(), float(value))
        if rc == _lib.PLT_WARN_UNKNOWN_KEY:
            warnings.warn(f"Unknown field '{key}' in {self._dict_name} dictionary.")
            return False
        _lib.check(rc)
        return True

    def get_option(self, key: str) -> float:
        v = C.c_double()
        _lib.check(getattr(self.lib, f"plt_{self._prefix}_get_option")(self.handle, key.encode(), C.byref(v)))
        return v.value

    def get_scripting_dictionary(self) -> dict:
        out, i = {}, 0
        name_of = getattr(self.lib, f"plt_{self._prefix}_option_name")
        while True:
            name = name_of(i)
            if not name:
                return out
            out[name.decode()] = self.get_option(name.decode())
            i += 1

    def close(self):
        if getattr(self, "handle", None):
            getattr(self.lib, f"plt_{self._prefix}_destroy")(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AccumulatePass(_Pass):
    """Temporal accumulation (AccumulatePass.{h,cpp}): running average of the input frames in Double,
    Single or compensated Single precision, with an optional frame limit and overflow mode."""

    _prefix = "accumulate"
    _dict_name = "AccumulatePass"

    @property
    def enabled(self) -> bool:
        return bool(self.get_option("enabled"))

    @enabled.setter
    def enabled(self, value: bool):
        _lib.check(self.lib.plt_accumulate_set_enabled(self.handle, int(bool(value))))

    @property
    def maxAccmulatedFrames(self) -> int:  # the reference's property name (AccumulatePass.cpp:37)
        return int(self.get_option("maxFrameCount"))

    @maxAccmulatedFrames.setter
    def maxAccmulatedFrames(self, value: int):
        self.set_option("maxFrameCount", int(value))

    def reset(self):
        _lib.check(self.lib.plt_accumulate_reset(self.handle))

    def set_scene(self, scene):
        """setScene resets the accumulation (AccumulatePass.cpp:356-362)."""
        self.reset()

    @property
    def frame_count(self) -> int:
        n = C.c_uint32()
        _lib.check(self.lib.plt_accumulate_get_frame_count(self.handle, C.byref(n)))
        return n.value

    def execute(self, src: torch.Tensor, dst: torch.Tensor | None = None, refresh: bool = False, unsigned: bool = False,
                stream=None) -> torch.Tensor:
        """Accumulates `src` (float32, or int32 texels read as RGBA32Sint -- RGBA32Uint with
        `unsigned`) and writes the running average to `dst` (allocated if None). `refresh` reports
        refresh flags or scene changes (a reset when autoReset is on)."""
        _check_image(src, "src", (torch.float32, torch.int32))
        h, w = src.shape[0], src.shape[1]
        if dst is None:
            dst = torch.empty((h, w, 4), dtype=torch.float32, device=src.device)
        _check_image(dst, "dst")
        if dst.shape != src.shape:
            raise ValueError("dst must have the shape of src (output size Default)")
        fmt = _INPUT_FORMATS[src.dtype]
        if unsigned:
            if src.dtype != torch.int32:
                raise ValueError("unsigned input is passed as int32 texels")
            fmt = 1
        _lib.check(self.lib.plt_accumulate_execute(self.handle, C.c_void_p(src.data_ptr()), fmt, w, h,
                                                   C.c_void_p(dst.data_ptr()), int(bool(refresh)),
                                                   _stream_handle(stream)))
        return dst


class ToneMapper(_Pass):
    """Tone mapping (ToneMapper.{h,cpp}): optional auto exposure from the mean log luminance, white
    balance and exposure as one colour transform, one of six operators, optional clamp."""

    _prefix = "tonemapper"
    _dict_name = "a ToneMapping"

    def set_scene_metadata(self, film_iso: float | None = None, f_number: float | None = None,
                           shutter_speed: float | None = None):
        """setScene with the scene's metadata (ToneMapper.cpp:395-405); None = not in the metadata."""
        nan = math.nan
        _lib.check(self.lib.plt_tonemapper_set_scene_metadata(
            self.handle, nan if film_iso is None else film_iso, nan if f_number is None else f_number,
            nan if shutter_speed is None else shutter_speed))

    def color_transform(self):
        import numpy as np
        m = (C.c_float * 9)()
        _lib.check(self.lib.plt_tonemapper_get_color_transform(self.handle, m))
        return np.array(list(m), np.float32).reshape(3, 3)

    def execute(self, src: torch.Tensor, dst: torch.Tensor | None = None, mean_log_luminance: torch.Tensor | None = None,
                stream=None, noalloc: bool = False) -> torch.Tensor:
        """Tone-maps src into dst. noalloc=True calls the entry point that allocates no device memory
        (its auto-exposure reduction uses the library's fixed per-GPU workspace; images up to a
        bit_floor of 8192 per side)."""
        _check_image(src, "src")
        h, w = src.shape[0], src.shape[1]
        if dst is None:
            dst = torch.empty_like(src)
        _check_image(dst, "dst")
        if dst.shape != src.shape:
            raise ValueError("dst must have the shape of src (output size Default)")
        ml = None
        if mean_log_luminance is not None:
            if mean_log_luminance.dtype != torch.float32 or not mean_log_luminance.is_cuda:
                raise ValueError("mean_log_luminance must be a float32 CUDA tensor")
            ml = C.c_void_p(mean_log_luminance.data_ptr())
        fn = self.lib.plt_tonemapper_execute_noalloc if noalloc else self.lib.plt_tonemapper_execute
        _lib.check(fn(self.handle, C.c_void_p(src.data_ptr()), w, h, C.c_void_p(dst.data_ptr()), ml,
                      _stream_handle(stream)))
        return dst


def _make_property(key, cast):
    def fget(self):
        return cast(self.get_option(key))

    def fset(self, value):
        self.set_option(key, value)
    return property(fget, fset)


# The reference's scripting properties (ToneMapper.cpp:99-116).
for _key, _cast in [("exposureCompensation", float), ("autoExposure", bool), ("exposureValue", float),
                    ("filmSpeed", float), ("whiteBalance", bool), ("whitePoint", float), ("operator", ToneMapOp),
                    ("clamp", bool), ("whiteMaxLuminance", float), ("whiteScale", float), ("fNumber", float),
                    ("shutter", float), ("exposureMode", ExposureMode)]:
    setattr(ToneMapper, _key, _make_property(_key, lambda v, c=_cast: c(int(v)) if c in (ToneMapOp, ExposureMode) else c(v)))
