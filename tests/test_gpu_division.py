"""The shortened division of the fast-division stage variants (div_inrange)
against the IEEE division '/' compiles to, bit for bit, over the operand
ranges the stage kernels rely on (mgad_check_division)."""
import numpy as np
import pytest

from multigpu_advectiondiffusion_amd import _capi

pytestmark = pytest.mark.gpu

MANT = np.uint64((1 << 52) - 1)


def doubles(rng, exps, mant=None):
    """Doubles 2^e * 1.m with e from exps (ints) and random mantissas."""
    n = len(exps)
    m = rng.integers(0, 1 << 52, size=n, dtype=np.uint64) if mant is None else mant
    bits = ((np.asarray(exps, dtype=np.int64) + 1023).astype(np.uint64) << np.uint64(52)) | (m & MANT)
    return bits.view(np.float64)


def run(hip, mode, a, b):
    n = a.size
    da, db = hip.DeviceArray.from_host(a), hip.DeviceArray.from_host(b)
    qf, qi = hip.DeviceArray((n,)), hip.DeviceArray((n,))
    rc = _capi.lib().mgad_check_division(None, mode, da.ptr, db.ptr, qf.ptr, qi.ptr, n)
    assert rc == 0
    return qf.download(), qi.download()


def same_bits(x, y):
    return x.view(np.uint64) == y.view(np.uint64)


def test_division_in_range(hipdev):
    """0 < b in [2^-1000, 2^1000], a = +0 or |a| in [2^-960, 2^1000], exponent gap <= 760."""
    rng = np.random.default_rng(11)
    n = 1 << 22
    eb = rng.integers(-1000, 1000, size=n)
    ea = np.clip(eb + rng.integers(-760, 761, size=n), -960, 999)
    # mantissas: random, all ones, zero (powers of two), and a*b near-exact quotients
    mb = rng.integers(0, 1 << 52, size=n, dtype=np.uint64)
    mb[: n // 16] = MANT
    mb[n // 16: n // 8] = 0
    b = doubles(rng, eb, mb)
    a = doubles(rng, ea) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    k = n // 4
    eq = np.clip(rng.integers(-380, 380, size=k), -958 - eb[-k:], 997 - eb[-k:])
    prod = doubles(rng, eq) * b[-k:]                                       # quotient close to a double
    a[-k:] = np.nextafter(prod, np.where(rng.random(k) < 0.5, np.inf, -np.inf))
    a[:1024] = 0.0
    qf, qi = run(hipdev, 0, a, b)
    bad = ~same_bits(qf, qi)
    assert not bad.any(), (a[bad][:4], b[bad][:4], qf[bad][:4], qi[bad][:4])
    assert np.array_equal(qi[1024:2048], a[1024:2048] / b[1024:2048])  # IEEE on the device = numpy


@pytest.mark.parametrize("p", [1, 2])
def test_zweight_ranges(hipdev, p):
    """WENO-Z weights 0.3*(1 + (tau5/b)^p) for b in [2^-300, 2^167] and tau5 = 0,
    tiny (down to subnormal) or up to 2^167: identical either way."""
    rng = np.random.default_rng(12 + p)
    n = 1 << 21
    b = doubles(rng, rng.integers(-300, 167, size=n))
    t = doubles(rng, rng.integers(-380, 167, size=n))
    tiny = doubles(rng, rng.integers(-1022, -380, size=n // 4))
    t[: n // 4] = tiny
    t[n // 4: n // 4 + 4096] = rng.integers(1, 1 << 52, size=4096, dtype=np.uint64).view(np.float64)  # subnormal
    t[-1024:] = 0.0
    qf, qi = run(hipdev, p, t, b)
    bad = ~same_bits(qf, qi)
    assert not bad.any(), (t[bad][:4], b[bad][:4], qf[bad][:4], qi[bad][:4])
