"""C-ABI library checks that need no GPU: it loads, exports every declared symbol, host-side
entry points validate their arguments, and the host spectral pipeline matches the oracle."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(plt_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from pltfalcor_amd import _lib
    lib = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 9
    for name in decl:
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(decl) <= bound, f"unbound in _lib.SIGNATURES: {set(decl) - bound}"


def test_plane_constants_match_header():
    from pltfalcor_amd import _lib
    hdr = (ROOT / "include" / "plt_abi.h").read_text()
    assert "PLT_Q_S3, \n    PLT_Q_PLANES" not in hdr
    assert _lib.Q_PLANES == 16 and _lib.R_PLANES == 5
    assert C.sizeof(_lib.SpectralProfile) == O.lib().oracle_spectral_profile_size() == 4 * (3 + 1 + 1 + 2 + 256)


def test_profile_builder_matches_oracle_bit_exact():
    from pltfalcor_amd.spectra import blackbody_table, build_profile, drude_table
    wl, n, k = drude_table()
    for vals, icdf in ((n, False), (k, False), (n, True)):
        assert bytes(build_profile(wl, vals, icdf)) == O.build_profile(wl, vals, icdf)
    bw, bv = blackbody_table()
    assert bytes(build_profile(bw, bv, True)) == O.build_profile(bw, bv, True)
    # ragged, narrow table inside the range
    wl2 = np.array([400, 401, 450, 600, 610], np.float32)
    v2 = np.array([1, 2, 0.5, 3, 0.1], np.float32)
    assert bytes(build_profile(wl2, v2, True)) == O.build_profile(wl2, v2, True)


def test_profile_builder_known_values():
    """Constant spectrum -> every bin equals the constant; range clamped to 380-720 nm."""
    from pltfalcor_amd.spectra import build_profile
    wl = np.linspace(300, 800, 101, dtype=np.float32)  # 5 nm table: bin size max(5, 5/2) = 5 nm
    p = build_profile(wl, np.full(101, 2.5, np.float32), False)
    assert p.min_wavelength == 380 and p.max_wavelength == 720
    assert p.bins == 68  # 100 samples over 500 nm, round(100 * 340/500) kept
    # two points: bin size = 250 nm -> 2 samples -> round(2 * 340/500) = 1 bin (reference semantics)
    p2 = build_profile(np.array([300, 800], np.float32), np.array([2.5, 2.5], np.float32), False)
    assert p2.bins == 1
    assert np.allclose(np.array(p.pdf[: p.bins]), 2.5)


def test_host_entry_points_validate_arguments():
    from pltfalcor_amd import _lib
    lib = _lib.load()
    out = _lib.SpectralProfile()
    wl = np.array([500, 400], np.float32)
    v = np.array([1, 1], np.float32)
    rc = lib.plt_spectral_profile_build(2, wl.ctypes.data_as(C.POINTER(C.c_float)),
                                        v.ctypes.data_as(C.POINTER(C.c_float)), 0, C.byref(out))
    assert rc == -1 and b"non-decreasing" in lib.plt_last_error()
    rc = lib.plt_spectral_profile_build(1, wl.ctypes.data_as(C.POINTER(C.c_float)),
                                        v.ctypes.data_as(C.POINTER(C.c_float)), 0, C.byref(out))
    assert rc == -1
    with pytest.raises(_lib.PLTError):
        _lib.check(rc)
    assert lib.plt_context_create(0, None) == -1
    assert lib.plt_bsdf_eval_queries(None, 4, None, None, None) == -1
    assert lib.plt_abi_version() >= 100
