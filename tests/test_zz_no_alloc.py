"""The tone mapper's entry point that allocates no device memory (plt_tonemapper_execute_noalloc),
called through the C ABI, against plt_tonemapper_execute: the same kernels, so the outputs and the
mean log luminance are bit-identical. Two calls in a row, the second with a smaller image, so the
fixed workspace is reused with stale contents from the first call; and a size the workspace cannot
hold is refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest


def _run(lib, handle, img_t, noalloc):
    import torch
    out = torch.empty_like(img_t)
    ml = torch.full((1,), float("nan"), dtype=torch.float32, device="cuda")
    fn = lib.plt_tonemapper_execute_noalloc if noalloc else lib.plt_tonemapper_execute
    rc = fn(handle, C.c_void_p(img_t.data_ptr()), img_t.shape[1], img_t.shape[0], C.c_void_p(out.data_ptr()),
            C.c_void_p(ml.data_ptr()), None)
    assert rc == 0, lib.plt_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy(), ml.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("op", [1, 5])
def test_gpu_tonemapper_noalloc_matches_allocating_entry_point(gpu, op):
    import torch
    from pltfalcor_amd import _lib
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.postfx import ToneMapper
    lib = _lib.load()
    t = ToneMapper(default_context(), {"autoExposure": True, "operator": op, "clamp": True, "exposureCompensation": 0.5})
    rng = np.random.default_rng(11 + op)
    for shape in [(1080, 1920), (37, 91)]:  # the second call is smaller: stale workspace contents
        img = torch.from_numpy(rng.lognormal(-1, 1.5, shape + (4,)).astype(np.float32)).cuda()
        want, want_ml = _run(lib, t.handle, img, noalloc=False)
        got, got_ml = _run(lib, t.handle, img, noalloc=True)
        assert np.isfinite(got_ml).all() and np.array_equal(got_ml.view(np.uint32), want_ml.view(np.uint32)), (shape, got_ml, want_ml)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), shape
    # bit_floor(16384) / 32 squared = 262,144 tiles > 65,536: refused before any launch (no buffers touched)
    rc = lib.plt_tonemapper_execute_noalloc(t.handle, C.c_void_p(16), 16384, 16384, C.c_void_p(16), None, None)
    assert rc == _lib.PLT_EINVAL and b"65536" in lib.plt_last_error()
    t.close()
