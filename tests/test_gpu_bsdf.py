"""GPU parity of the batched BSDF path (HIP, through the C ABI) against the CPU oracle.

Bars (stated per test):
- sample generator: bit-exact uint32 streams;
- query generator: integer/binary16 planes bit-exact, linear float planes bit-exact, planes that
  pass through cos/sin within 4 ulp;
- eval + pdf: per query, max_c |gpu_c - ref_c| <= 1e-4 * |ref_S0| for the Stokes output and
  |gpu_pdf - ref_pdf| <= 1e-4 * |ref_pdf| against the FP64 oracle, except where the reference's
  own FP32 arithmetic is worse than that (the FP32 oracle's deviation from FP64 sets the
  envelope: the GPU may not exceed 4x the FP32 oracle's error there).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _ctx():
    from pltfalcor_amd.bsdf import default_context
    return default_context(0)


def _profiles_bytes():
    from pltfalcor_amd.spectra import drude_profiles
    n, k = drude_profiles()
    return bytes(n) + bytes(k)


def _cmp_stats(gpu, o32, o64):
    """Per-query relative errors (Stokes normalised by |S0|, pdf by |pdf|) vs FP64."""
    s0 = np.maximum(np.abs(o64[0]), 1e-30)
    e_gpu_s = np.max(np.abs(gpu[:4] - o64[:4]), axis=0) / s0
    e_o32_s = np.max(np.abs(o32[:4] - o64[:4]), axis=0) / s0
    p = np.maximum(np.abs(o64[4]), 1e-30)
    e_gpu_p = np.abs(gpu[4] - o64[4]) / p
    e_o32_p = np.abs(o32[4] - o64[4]) / p
    return e_gpu_s, e_o32_s, e_gpu_p, e_o32_p


def test_sample_generator_bit_exact(gpu):
    from pltfalcor_amd.bsdf import sample_generator_stream
    w, h, draws = 64, 16, 24
    for sn in (0, 1, 77):
        out = sample_generator_stream(w * h, w, sn, draws).cpu().numpy().view(np.uint32)
        for s in range(0, w * h, 37):
            ref = O.usg_seq(s % w, s // w, sn, draws)
            assert np.array_equal(out[:, s], ref), (sn, s)


def test_tiny_sample_generator_bit_exact(gpu):
    """TinyUniformSampleGenerator streams (sampleGenerator 0) on the GPU equal the oracle's."""
    import ctypes as C
    from pltfalcor_amd import _lib
    lib = _lib.load()
    w, h, draws = 64, 16, 24
    for sn in (0, 1, 77, 1 << 30):
        out = torch.empty((draws, w * h), dtype=torch.int32, device="cuda")
        _lib.check(lib.plt_sample_generator_stream_typed(0, w * h, w, sn, draws, C.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        g = out.cpu().numpy().view(np.uint32)
        ref = O.sg_streams(0, 0, 0, w, w * h, sn, draws)
        assert np.array_equal(g.T, ref), sn
    with pytest.raises(_lib.PLTError):
        _lib.check(lib.plt_sample_generator_stream_typed(2, 4, 4, 0, 1, C.c_void_p(out.data_ptr()), None))


def test_query_generator_matches_oracle(gpu):
    from pltfalcor_amd.bsdf import BSDFBatch
    ctx = _ctx()
    n = 1 << 16
    b = BSDFBatch(ctx, n).generate(seed=0)
    torch.cuda.synchronize()
    g = b.queries.cpu().numpy().view(np.uint32)
    ref = O.c0_generate(n, 0).view(np.uint32)
    exact = [6, 7, 8, 9, 10, 11, 12, 15]  # lambda, mat, param, q, l, A_c, S0, S3 (no trig)
    for p in exact:
        assert np.array_equal(g[p], ref[p]), f"plane {p}"
    gf, rf = g.view(np.float32), ref.view(np.float32)
    for p in (0, 1, 3, 4, 13, 14):  # cos/sin of the same argument: a few ulp
        ulp = np.spacing(np.maximum(np.abs(rf[p]), 1e-30).astype(np.float32))
        assert np.all(np.abs(gf[p] - rf[p]) <= 4 * np.maximum(ulp, 1e-7)), f"plane {p}"
    for p in (2, 5):  # z = sqrt(1 - x^2 - y^2): error of x,y amplified by |x|/|z| near grazing
        tol = 4e-7 * (1.0 + 1.0 / np.maximum(np.abs(rf[p]), 1e-4))
        assert np.all(np.abs(gf[p] - rf[p]) <= tol), f"plane {p}"


@pytest.mark.parametrize("seed", [0, 3])
def test_eval_parity_vs_oracle(gpu, seed):
    from pltfalcor_amd.bsdf import BSDFBatch
    ctx = _ctx()
    n = 1 << 16
    b = BSDFBatch(ctx, n).generate(seed=seed)
    res = b.eval().cpu().numpy().astype(np.float64)
    q = b.queries.cpu().numpy()
    prof = _profiles_bytes()
    o32 = O.c0_eval(q, prof, 2).astype(np.float64)
    o64 = O.c0_eval(q, prof, 2, fp64=True)
    assert np.all(np.isfinite(res))
    e_gs, e_os, e_gp, e_op = _cmp_stats(res, o32, o64)
    lam = q.view(np.float32)[6]
    live = lam < 720.0
    # Exact zeros beyond the reference's spectral range (SpectrumUtils.slang:108-118).
    assert np.all(res[:4, ~live] == 0.0)
    ok_s = (e_gs <= TOL) | (e_gs <= 4 * e_os)
    ok_p = (e_gp <= TOL) | (e_gp <= 4 * e_op)
    bad = np.where(~(ok_s & ok_p))[0]
    print(f"seed {seed}: gpu-vs-fp64 Stokes max {e_gs[live].max():.3e} p99.9 {np.quantile(e_gs[live], .999):.3e}; "
          f"fp32 oracle max {e_os[live].max():.3e}; pdf max {e_gp.max():.3e} (fp32 oracle {e_op.max():.3e})")
    assert bad.size == 0, f"{bad.size} queries outside tolerance, first {bad[:8]}"
    # The envelope clause is a fallback only: in the bulk the GPU meets 1e-4 against FP64 outright
    # and is no less accurate than the reference's own FP32 arithmetic.
    assert np.mean(e_gs[live] <= TOL) >= 0.9999 and np.mean(e_gp <= TOL) >= 0.9999
    assert np.quantile(e_gs[live], 0.99) <= max(np.quantile(e_os[live], 0.99), 1e-6)


def test_eval_full_size_properties(gpu):
    """2^20 queries (the benchmark batch): finite, exact zeros past 720 nm, Stokes physicality."""
    from pltfalcor_amd.bsdf import BSDFBatch
    ctx = _ctx()
    n = 1 << 20
    b = BSDFBatch(ctx, n).generate(seed=0)
    r = b.eval()
    torch.cuda.synchronize()
    res = r.cpu().numpy()
    lam = b.queries.cpu().numpy().view(np.float32)[6]
    assert np.all(np.isfinite(res))
    assert np.all(res[:4, lam >= 720.0] == 0.0)
    s0 = res[0].astype(np.float64)
    assert np.all(s0 >= 0.0)
    # degree of polarisation <= 1 (ratios in float64: squaring tiny float32 values would underflow)
    nz = s0 > 1e-30
    dop = np.sqrt(sum((res[c, nz].astype(np.float64) / s0[nz]) ** 2 for c in (1, 2, 3)))
    assert np.all(dop <= 1 + 1e-3), f"max DoP {dop.max()}"
    assert np.all(res[4] >= 0.0)
    # Deterministic: a second evaluation is bit-identical.
    res2 = b.eval().cpu().numpy()
    assert np.array_equal(res, res2)


def test_eval_edge_cases(gpu):
    """Grazing / normal incidence, tiny batch, and n not a multiple of the block size."""
    from pltfalcor_amd.bsdf import BSDFBatch
    ctx = _ctx()
    n = 300
    q = O.c0_generate(n, 9)
    f = q.view(np.float32)
    # normal incidence for the first 10, near-grazing for the next 10
    f[0, :10] = 0; f[1, :10] = 0; f[2, :10] = 1
    f[0, 10:20] = 0.999; f[1, 10:20] = 0; f[2, 10:20] = np.sqrt(1 - 0.999 ** 2)
    b = BSDFBatch(ctx, n).load(torch.from_numpy(q.view(np.int32).copy()).cuda())
    res = b.eval().cpu().numpy().astype(np.float64)
    prof = _profiles_bytes()
    o32 = O.c0_eval(q, prof, 2).astype(np.float64)
    o64 = O.c0_eval(q, prof, 2, fp64=True)
    e_gs, e_os, e_gp, e_op = _cmp_stats(res, o32, o64)
    assert np.all(((e_gs <= TOL) | (e_gs <= 4 * e_os)) & ((e_gp <= TOL) | (e_gp <= 4 * e_op)))
