"""Pins the oracle's sin / cos / pow (the functions the shading kernels use in
place of libm's cosf / sinf / powf, main_gpu.cu:30-31, :174, :283) against
float64 numpy, which is an independent implementation: a wrong constant, a
dropped term or a wrong quadrant fails the ulp bounds below."""
import math

import numpy as np
import pytest


def _ulp_err(got, ref):
    ref = np.asarray(ref, np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_sincos_exact_points(orc):
    assert orc.sincos(0.0) == (0.0, 1.0)
    for x in (1e-30, 1e-10, 1e-5):
        s, c = orc.sincos(x)
        assert s == np.float32(x) and c == 1.0
    # quadrant boundaries: k = rint(x 2/pi) switches the polynomial
    for k in range(1, 9):
        x = np.float32(k * math.pi / 2)
        for xx in (np.nextafter(x, np.float32(0)), x, np.nextafter(x, np.float32(100))):
            s, c = orc.sincos(float(xx))
            assert abs(s - math.sin(float(xx))) < 1.2e-7 and abs(c - math.cos(float(xx))) < 1.2e-7


def test_sincos_accuracy_on_the_sampler_domain(orc):
    """phi = (xi * 2) * pi for every 24-bit xi in [0, 1), the sampler's whole
    input set, and a dense sweep of the float bit patterns in [2^-30, 2 pi)."""
    xi = (np.arange(0, 1 << 24, dtype=np.float64) / 2**24).astype(np.float32)
    phi = (xi * np.float32(2)) * np.float32(math.pi)
    bits = np.arange(np.float32(2.0**-30).view(np.uint32), np.float32(2 * math.pi).view(np.uint32), 97,
                     dtype=np.uint32)
    x = np.concatenate([phi, bits.view(np.float32)])
    s, c = orc.sincos_batch(x)
    x64 = x.astype(np.float64)
    es, ec = _ulp_err(s, np.sin(x64)), _ulp_err(c, np.cos(x64))
    assert es.max() <= 2.5 and ec.max() <= 2.5, (es.max(), ec.max())
    assert np.abs(s.astype(np.float64) - np.sin(x64)).max() < 1.2e-7
    assert np.abs(c.astype(np.float64) - np.cos(x64)).max() < 1.2e-7
    # the bulk is correctly rounded or one ulp off
    assert (es <= 1.0).mean() > 0.99 and (ec <= 1.0).mean() > 0.99


def test_sincos_wide_arguments(orc):
    x = np.random.default_rng(3).uniform(-3000, 3000, 2_000_000).astype(np.float32)
    s, c = orc.sincos_batch(x)
    x64 = x.astype(np.float64)
    assert np.abs(s - np.sin(x64)).max() < 2e-7 and np.abs(c - np.cos(x64)).max() < 2e-7
    # odd / even symmetry is exact: the reduction and polynomials are odd / even in x
    s2, c2 = orc.sincos_batch(-x)
    assert np.array_equal(s2, -s) and np.array_equal(c2, c)


def test_log2_exp2_exact_points(orc):
    L = orc.lib()
    for k in range(-149, 128):
        x = float(np.float32(2.0**k))
        assert L.orc_log2_pos(x) == k
        assert L.orc_exp2_fin(float(k)) == x
    assert L.orc_exp2_fin(0.5) == np.float32(math.sqrt(2))
    assert L.orc_exp2_fin(200.0) == math.inf and L.orc_exp2_fin(-200.0) == 0.0


def test_pow_special_values(orc):
    L = orc.lib()
    y = float(np.float32(1 / 2.2))
    assert L.orc_pow_pos(0.0, y) == 0.0 and L.orc_pow_pos(-0.0, y) == 0.0
    assert L.orc_pow_pos(-0.5, y) == 0.0
    assert L.orc_pow_pos(1.0, y) == 1.0
    assert L.orc_pow_pos(math.inf, y) == math.inf
    assert math.isnan(L.orc_pow_pos(math.nan, y))
    for x in (0.25, 3.0, 7.5):
        assert L.orc_pow_pos(x, 1.0) == pytest.approx(x, rel=2e-7)
        assert L.orc_pow_pos(x, 2.0) == pytest.approx(x * x, rel=4e-7)


@pytest.mark.parametrize("lo,hi,bound", [(2.0**-8, 1.0, 2.0), (2.0**-24, 2.0**-8, 2.0), (1.0, 128.0, 2.0), (2.0**-126, 2.0**-24, 2.0)])
def test_pow_gamma_accuracy(orc, lo, hi, bound):
    """x^(1/2.2) (the gamma of the AO and pack kernels) on every float of a
    dense stride through [lo, hi)."""
    y = np.float32(1 / 2.2)
    bits = np.arange(np.float32(lo).view(np.uint32), np.float32(hi).view(np.uint32), 5, dtype=np.uint32)
    x = bits.view(np.float32)
    got = orc.pow_batch(x, float(y))
    e = _ulp_err(got, np.power(x.astype(np.float64), np.float64(y)))
    assert e.max() <= bound, e.max()
    assert np.percentile(e, 99) <= 1.0


def test_pow_packs_like_the_exact_power(orc):
    """What reaches the image: the 8-bit level of (1 - ao)^(1/2.2) for every
    AO fraction k/N the kernels produce (N <= 256) equals the level of the
    float64 power."""
    L = orc.lib()
    y = np.float32(1 / 2.2)
    for n in (4, 16, 64, 255, 256):
        for k in range(n + 1):
            ao = np.float32(0)
            for _ in range(k):
                ao = np.float32(ao + np.float32(1.0) / np.float32(n))
            a = np.float32(np.float32(1) - ao)
            got = L.orc_pack_rgba(L.orc_pow_pos(float(a), float(y)), 0, 0, 1) & 0xFF
            ref = float(np.power(np.float64(max(a, 0)), np.float64(y)))
            assert got == min(max(int(np.float32(np.float32(ref) * np.float32(255) + np.float32(0.5))), 0), 255)
