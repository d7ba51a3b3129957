"""Pins the oracle's potentials: the reference's Java known-answer vectors,
finite differences of the energies, continuity and closed-form cases."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd.model import (ATTRACTIVE, BOX, CYLINDER, DISK, HARMONIC, PIECEWISE_WELL, REPULSIVE, SPHERE,
                                 SPRING, WEAK_HARMONIC)

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "java_potential_vectors.json")))
SUB = {"attractive": ATTRACTIVE, "repulsive": REPULSIVE, "spring": SPRING}


def vec(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return a, a.ctypes.data_as(C.c_void_p)


def pair_eval(pot, r0, d):
    dv, dp = vec(d)
    g, gp = vec(np.zeros(3))
    e = orc.lib().orc_pair_eval(C.byref(pot), r0, dp, gp)
    return g.copy(), e


def ext_eval(pot, prad, x):
    xv, xp = vec(x)
    g, gp = vec(np.zeros(3))
    e = orc.lib().orc_ext_eval(C.byref(pot), prad, xp, gp)
    return g.copy(), e


@pytest.mark.parametrize("case", GOLD["harmonic"]["cases"])
def test_harmonic_java_vectors(case):
    H = GOLD["harmonic"]
    p1, p2 = np.array(H["p1"]), np.array(H["p2"])
    assert abs(np.linalg.norm(p2 - p1) - H["dist"]) < 1e-5
    pot = orc.make_pot2(HARMONIC, SUB[case["subtype"]], H["k"])
    g1, e = pair_eval(pot, case["r0"], p1 - p2)          # grad U wrt particle 1
    # the Java gradient is the force on particle 1 = -grad_1 U
    assert np.allclose(-g1, case["grad"], atol=1e-4)
    assert abs(e - case["energy"]) < 1e-4


@pytest.mark.parametrize("case", GOLD["disk"]["cases"])
def test_disk_java_vectors(case):
    Dk = GOLD["disk"]
    pot = orc.make_pot1(DISK, SUB[case["subtype"]], Dk["k"], Dk["center"], Dk["normal"], radius=Dk["radius"])
    xv, xp = vec(Dk["points"][case["point"]])
    f, fp = vec(np.zeros(3))
    e = orc.lib().orc_disk_java(C.byref(pot), float(case["prad"]), xp, fp)
    assert np.allclose(f, case["grad"], atol=1e-4)
    assert abs(e - case["energy"]) < 1e-4


def test_disk_gpu_form_equals_java_form_without_radius():
    """The GPU-form disk (no particle radius) is the Java form at prad = 0."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = rng.normal(size=3)
        pot = orc.make_pot1(DISK, ATTRACTIVE, 2.5, rng.normal(size=3), n, radius=rng.uniform(0.5, 3.0))
        x = rng.normal(size=3) * 3
        g, e = ext_eval(pot, 0.0, x)
        xv, xp = vec(x)
        f, fp = vec(np.zeros(3))
        ej = orc.lib().orc_disk_java(C.byref(pot), 0.0, xp, fp)
        assert np.allclose(g, -f, atol=1e-10)
        assert abs(e - ej) < 1e-10


@pytest.mark.parametrize("case", GOLD["disk_displacement"]["cases"])
def test_disk_displacement_java_vectors(case):
    """EM drift a * (-grad U) with a = dt D / (kB T) (CudaSimulation.cu:1852)."""
    Dd = GOLD["disk_displacement"]
    pot = orc.make_pot1(DISK, ATTRACTIVE, Dd["k"], Dd["center"], Dd["normal"], radius=Dd["radius"])
    g, _ = ext_eval(pot, 0.0, case["point"])
    assert np.allclose(-g, case["force"], atol=1e-4)
    a = Dd["dt"] * Dd["D"] / (Dd["kB"] * Dd["T"])
    assert np.allclose(-a * g, case["displacement"], atol=1e-5)


def _fd_check(fun, x, h=1e-6, tol=2e-5):
    g, _ = fun(x)
    num = np.zeros(3)
    for d in range(3):
        xp = np.array(x, float); xm = np.array(x, float)
        xp[d] += h; xm[d] -= h
        num[d] = (fun(xp)[1] - fun(xm)[1]) / (2 * h)
    assert np.allclose(g, num, atol=tol * max(1.0, np.abs(g).max())), (g, num)


PAIR_CASES = [
    (HARMONIC, ATTRACTIVE, 10.0, 0.0, 0.0),
    (HARMONIC, REPULSIVE, 10.0, 0.0, 0.0),
    (HARMONIC, SPRING, 3.0, 0.0, 0.0),
    (WEAK_HARMONIC, ATTRACTIVE, 4.0, 1.5, 0.0),
    (PIECEWISE_WELL, ATTRACTIVE, 10.0, 1.0, 1.0),
    (PIECEWISE_WELL, ATTRACTIVE, 0.0, 2.0, 0.5),
]


@pytest.mark.parametrize("kind,sub,k,length,depth", PAIR_CASES)
def test_pair_gradient_is_energy_derivative(kind, sub, k, length, depth):
    pot = orc.make_pot2(kind, sub, k, length, depth, cutoff=100.0)
    r0 = 2.5
    rng = np.random.default_rng(kind * 10 + sub)
    for r in np.concatenate([rng.uniform(0.3, r0 + length + 1.0, 60), [r0 * 0.5, r0 * 1.3]]):
        if min(abs(r - r0), abs(r - r0 - length), abs(r - r0 - 0.5 * length)) < 1e-3:
            continue
        u = rng.normal(size=3)
        d = u / np.linalg.norm(u) * r
        _fd_check(lambda x: pair_eval(pot, r0, x), d)


def test_piecewise_well_continuity_and_depth():
    pot = orc.make_pot2(PIECEWISE_WELL, ATTRACTIVE, 10.0, 1.0, 1.0, cutoff=100.0)
    r0, l = 3.0, 1.0
    e = lambda r: pair_eval(pot, r0, np.array([r, 0.0, 0.0]))[1]
    assert abs(e(r0) + 1.0) < 1e-12                       # minimum = -depth at r0
    for rb in (r0, r0 + 0.5 * l, r0 + l):
        assert abs(e(rb - 1e-9) - e(rb + 1e-9)) < 1e-6
    assert e(r0 + l + 0.01) == 0.0


def test_weak_harmonic_closed_form():
    """CudaSimulation.cu:1021-1031: |grad U| = k (r-r0)(ri-r) r inside (r0, ri)."""
    pot = orc.make_pot2(WEAK_HARMONIC, ATTRACTIVE, 2.0, 1.5, 0.0, cutoff=100.0)
    r0 = 2.0
    ri = r0 + 1.5
    for r in (2.2, 2.75, 3.4):
        g, _ = pair_eval(pot, r0, np.array([0.0, r, 0.0]))
        assert abs(g[1] - 2.0 * (r - r0) * (ri - r) * r) < 1e-12
    g, e = pair_eval(pot, r0, np.array([0.0, 0.0, 4.0]))
    assert np.all(g == 0) and e == 0.0


def test_pair_cutoff_and_zero_distance():
    pot = orc.make_pot2(HARMONIC, ATTRACTIVE, 10.0, cutoff=10.0)
    g, e = pair_eval(pot, 2.0, np.array([10.5, 0.0, 0.0]))     # r > maxCutoff is skipped
    assert np.all(g == 0) and e == 0
    g, e = pair_eval(pot, 2.0, np.zeros(3))                     # r == 0 is skipped
    assert np.all(g == 0) and e == 0


EXT_CASES = [
    (DISK, ATTRACTIVE, dict(normal=(0, 0, 1), radius=2.0)),
    (DISK, ATTRACTIVE, dict(normal=(1, -2, 0.5), radius=1.0)),
    (CYLINDER, ATTRACTIVE, dict(normal=(0, 1, 0), radius=1.5, height=2.0)),
    (CYLINDER, REPULSIVE, dict(normal=(0, 0, 1), radius=3.0, height=4.0)),
    (SPHERE, ATTRACTIVE, dict(radius=2.0)),
    (SPHERE, REPULSIVE, dict(radius=2.0)),
    (BOX, ATTRACTIVE, dict(extension=(4.0, 5.0, 6.0))),
    (BOX, REPULSIVE, dict(extension=(4.0, 5.0, 6.0))),
]


@pytest.mark.parametrize("kind,sub,kw", EXT_CASES)
def test_external_gradient_is_energy_derivative(kind, sub, kw):
    pot = orc.make_pot1(kind, sub, 3.0, origin=(0.5, -0.25, 1.0), **kw)
    rng = np.random.default_rng(kind * 7 + sub)
    for _ in range(80):
        x = rng.uniform(-6, 6, size=3)
        _fd_check(lambda y: ext_eval(pot, 0.4, y), x, h=1e-7)


def test_sphere_closed_form():
    pot = orc.make_pot1(SPHERE, ATTRACTIVE, 2.0, origin=(1.0, 1.0, 1.0), radius=3.0)
    g, e = ext_eval(pot, 0.0, [1.0, 1.0, 6.0])          # 5 from the centre, 2 outside
    assert np.allclose(g, [0, 0, 2.0 * 2.0]) and abs(e - 0.5 * 2.0 * 4.0) < 1e-12
    g, e = ext_eval(pot, 0.0, [1.0, 2.0, 1.0])          # inside: free
    assert np.all(g == 0) and e == 0
