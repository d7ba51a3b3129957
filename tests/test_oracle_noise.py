"""Pins the oracle's noise generator: Philox4x32-10 known answers, accuracy of
the in-house log and sin/cos against libm, and the N(0,1) distribution."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "java_potential_vectors.json")))


@pytest.mark.parametrize("case", GOLD["philox4x32_10_kat"]["cases"])
def test_philox_known_answers(case):
    ctr = [int(x, 16) for x in case["ctr"]]
    key = [int(x, 16) for x in case["key"]]
    out = orc.philox(ctr, key)
    assert [int(x) for x in out] == [int(x, 16) for x in case["out"]]


def test_log_matches_libm():
    L = orc.lib()
    xs = (np.arange(1, 1 << 24, 97, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    got = np.array([L.orc_log_f32(float(x)) for x in xs[::50]], dtype=np.float64)
    ref = np.log(xs[::50].astype(np.float64))
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)
    assert err.max() < 4e-7


def test_sincos_matches_libm():
    import ctypes as C
    L = orc.lib()
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for t in np.linspace(0.0, 1.0, 20001, endpoint=False, dtype=np.float32):
        L.orc_sincos_turns_f32(float(t), C.byref(s), C.byref(c))
        a = 2.0 * math.pi * float(t)
        worst = max(worst, abs(s.value - math.sin(a)), abs(c.value - math.cos(a)))
    assert worst < 2.5e-7


def _ncdf(x):
    return 0.5 * (1.0 + np.vectorize(math.erf)(x / math.sqrt(2.0)))


def test_normal_distribution():
    z = orc.noise(seed=12345, stream=0, step0=3, nsteps=4, n=50_000).ravel().astype(np.float64)
    n = z.size
    assert abs(z.mean()) < 5.0 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 6.0 * math.sqrt(2.0 / n)
    kurt = np.mean(z ** 4)
    assert abs(kurt - 3.0) < 0.05
    zs = np.sort(z[:100_000])
    F = _ncdf(zs)
    D = np.max(np.abs(F - np.arange(1, zs.size + 1) / zs.size))
    assert D < 1.63 / math.sqrt(zs.size)          # KS at 1%


def test_noise_layout_and_streams():
    a = orc.noise(7, 0, 10, 2, 1000)
    b = orc.noise(7, 0, 11, 1, 1000)
    assert np.array_equal(a[1], b[0])               # step counter, not call count
    c = orc.noise(7, 1, 10, 1, 1000)
    assert not np.array_equal(a[0], c[0])           # stream separates replicas
    d = orc.noise(8, 0, 10, 1, 1000)
    assert not np.array_equal(a[0], d[0])
    # dimension rows are not copies of each other
    assert abs(np.corrcoef(a[0, 0], a[0, 1])[0, 1]) < 0.1
