"""Pins the oracle's sample generator (CPU only).

- interleave_32bit: the known answers of Tests/Utils/BitTricksTests.cpp:57-58 and the test's own
  bit-by-bit reference on std::mt19937 (default seed) inputs, as BitTricksTests.cpp does.
- SplitMix64: the published sequence for seed 1234567.
- xoshiro128** and SplitMix64 on 256 std::mt19937-seeded instances x 64 draws, against the CPU
  reference implementations of Tests/Sampling/PseudorandomTests.cpp (the reference's method).
- std::mt19937 itself is pinned by the C++ standard's 10000th-output value.
- UniformSampleGenerator = SplitMix64 seeding of xoshiro128** (UniformSampleGenerator.slang:51-69).
"""
import numpy as np

import oracle_lib as O

M32 = 0xFFFFFFFF
M64 = 0xFFFFFFFFFFFFFFFF


def mt19937(seed=5489):
    """std::mt19937 (default-constructed: seed 5489)."""
    mt = [0] * 624
    mt[0] = seed
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & M32
    idx = 624
    while True:
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        idx += 1
        yield y & M32


def test_mt19937_standard_value():
    g = mt19937()
    for _ in range(9999):
        next(g)
    assert next(g) == 4123659995  # [rand.predef]: 10000th invocation of default mt19937


def ref_interleave(x, y, m):
    r = 0
    for i in range(m):
        r |= ((x >> i) & 1) << (2 * i)
        r |= ((y >> i) & 1) << (2 * i + 1)
    return r


def test_interleave_known_answers():
    L = O.lib()
    assert L.oracle_interleave_32bit(0xE38E, 0xBE8B) == 0xDEADC0DE
    assert ref_interleave(0xE38E, 0xBE8B, 12) == 0x00ADC0DE
    assert L.oracle_interleave_32bit(0xE38E & 0xFFF, 0xBE8B & 0xFFF) == 0x00ADC0DE


def test_interleave_vs_bitloop():
    L = O.lib()
    g = mt19937()
    for _ in range(4096):
        a, b = next(g), next(g)
        assert L.oracle_interleave_32bit(a, b) == ref_interleave(a & 0xFFFF, b & 0xFFFF, 16)


def splitmix_ref(x, n):
    out = []
    for _ in range(n):
        x = (x + 0x9E3779B97F4A7C15) & M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_published_vector():
    out = np.zeros(5, dtype=np.uint64)
    O.lib().oracle_splitmix64_seq(1234567, 5, O.ptr(out, O.C.c_uint64))
    assert [int(v) for v in out] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                     4593380528125082431, 16408922859458223821]


def test_splitmix64_vs_reference_test_impl():
    g = mt19937()
    for _ in range(256):
        lo, hi = next(g), next(g)
        seed = (hi << 32) | lo
        out = np.zeros(64, dtype=np.uint64)
        O.lib().oracle_splitmix64_seq(seed, 64, O.ptr(out, O.C.c_uint64))
        assert [int(v) for v in out] == splitmix_ref(seed, 64)


def xoshiro_ref(s, n):
    s = list(s)
    rotl = lambda x, k: ((x << k) | (x >> (32 - k))) & M32  # noqa: E731
    out = []
    for _ in range(n):
        r = (rotl((s[0] * 5) & M32, 7) * 9) & M32
        t = (s[1] << 9) & M32
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 11)
        out.append(r)
    return out


def test_xoshiro_vs_reference_test_impl():
    g = mt19937()
    for _ in range(256):
        st = np.array([next(g) for _ in range(4)], dtype=np.uint32)
        out = np.zeros(64, dtype=np.uint32)
        O.lib().oracle_xoshiro128ss_seq(O.ptr(st, O.C.c_uint32), 64, O.ptr(out, O.C.c_uint32))
        assert [int(v) for v in out] == xoshiro_ref([int(v) for v in st], 64)


def test_uniform_sample_generator_composition():
    for (px, py, sn) in [(0, 0, 0), (17, 3, 0), (511, 511, 5), (0xFFFF, 0x7FFF, 123456)]:
        seed = (sn << 32) | ref_interleave(px, py, 16)
        a, b = splitmix_ref(seed, 2)
        st = [a & M32, a >> 32, b & M32, b >> 32]
        assert [int(v) for v in O.usg_seq(px, py, sn, 32)] == xoshiro_ref(st, 32)


def test_sample_statistics():
    """SampleGeneratorTests.cpp: per-dimension mean within 0.01 of 0.5 and |corr| <= 0.002 (scaled)."""
    n_px, dims = 64 * 64, 16
    u = np.stack([(O.usg_seq(i % 64, i // 64, 0, dims) >> 8).astype(np.float64) * 2.0 ** -24
                  for i in range(n_px)])
    assert np.all(np.abs(u.mean(axis=0) - 0.5) < 0.01)
    c = np.corrcoef(u.T)
    off = c[~np.eye(dims, dtype=bool)]
    assert np.max(np.abs(off)) < 0.05  # 4096 samples: sampling noise ~1/sqrt(4096)


# ---------------------------------------------------------------- TinyUniformSampleGenerator

def test_tea_published_vector():
    """TEA (Wheeler & Needham 1994) with the all-zero key and plaintext and the full 32 cycles gives
    the published ciphertext 41ea3a0a 94baa940; a slip in either half-round (shift, key word, sum)
    changes it. blockCipherTEA (HashUtils.slang:59-71) is this cipher with Falcor's key and 16 cycles."""
    assert O.tea(0, 0, [0, 0, 0, 0], 32) == (0x41EA3A0A, 0x94BAA940)


def lcg_ref(state, n):
    """The Numerical Recipes LCG of PseudorandomTests.cpp:85-97."""
    out = []
    for _ in range(n):
        state = (1664525 * state + 1013904223) & M32
        out.append(state)
    return out


def test_tiny_generator_is_tea_seeded_lcg():
    """TinyUniformSampleGenerator(pixel, sampleNumber) (TinyUniformSampleGenerator.slang:55-66):
    seed = blockCipherTEA(interleave_32bit(pixel), sampleNumber).x with Falcor's key
    (a341316c c8013ea4 ad90777d 7e95761e) and 16 cycles, then the LCG of the reference's own CPU
    test (PseudorandomTests.cpp:85-97)."""
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]
    for (px, py, sn) in [(0, 0, 0), (17, 3, 0), (511, 511, 5), (0xFFFF, 0x7FFF, 123456), (3, 9, 1 << 31)]:
        seed, _ = O.tea(ref_interleave(px, py, 16), sn, key, 16)
        got = O.sg_streams(0, px, py, 1, 1, sn, 48)[0]
        assert [int(v) for v in got] == lcg_ref(seed, 48)


def _reference_generator_statistics(gen_type, corr_threshold):
    """SampleGeneratorTests.cpp:73-141: a 64 x 64 x 16 dispatch, 32 dimensions per instance drawn as
    1D, 1D, 1D, 2D, 3D; every sample in [0, 1), mean within 0.01 of 0.5, and |Pearson correlation|
    at strides of 1-8 dimensions, the 15 nearest pixels and 1-4 instances below the threshold."""
    dims, nx, ny, nz = 32, 64, 64, 16
    per = []
    for z in range(nz):
        per.append(O.sg_streams(gen_type, 0, 0, nx, nx * ny, z, dims))
    raw = np.stack(per)  # [z, pixel, dim], pixel = y * 64 + x
    u = (raw >> 8).astype(np.float64) * 2.0 ** -24
    flat = u.reshape(-1).astype(np.float32)
    assert np.all((flat >= 0) & (flat < 1))
    assert abs(flat.astype(np.float64).mean() - 0.5) <= 0.01

    def corr(stride):
        x = flat[:-stride].astype(np.float64)
        y = flat[stride:].astype(np.float64)
        n = x.size
        sx, sy = x.sum(), y.sum()
        return (n * (x * y).sum() - sx * sy) / (np.sqrt(n * (x * x).sum() - sx * sx) * np.sqrt(n * (y * y).sum() - sy * sy))
    for i in range(1, 9):
        assert abs(corr(i)) <= corr_threshold, ("dimension", i)
    for y in range(4):
        for x in range(4):
            if x or y:
                assert abs(corr(x * dims + y * nx * dims)) <= corr_threshold, ("pixel", x, y)
    for i in range(1, 5):
        assert abs(corr(i * nx * ny * dims)) <= corr_threshold, ("instance", i)


def test_tiny_generator_statistics_as_reference_test():
    _reference_generator_statistics(0, 0.0025)  # SampleGenerator_TinyUniform


def test_uniform_generator_statistics_as_reference_test():
    _reference_generator_statistics(1, 0.002)  # SampleGenerator_Uniform
