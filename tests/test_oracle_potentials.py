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


# ---------------------------------------------------------------------------
# Closed forms of the faces of the BOX potential (CudaSimulation.cu:1191-1228)
# and of the repulsive sphere and cylinder; the attractive box, the repulsive
# sphere and the repulsive cylinder are also compared with the Java forms of
# P1_Cube.java:187-273, P1_Sphere.java:222-271 and P1_Cylinder.java:212-284,
# which are written differently (min/max corners, pointOnDiskPlane, force
# convention) and so catch a mistake in the oracle's geometry.
# ---------------------------------------------------------------------------
BOX_O = np.array([1.0, -2.0, 0.5])
BOX_E = np.array([4.0, 5.0, 6.0])
PRAD = 0.7


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", ["upper", "lower"])
@pytest.mark.parametrize("delta", [0.05, 1.3])
def test_box_attractive_faces_closed_form(axis, side, delta):
    """Attractive box: a particle of radius a is confined to [o + a, o + e - a]
    per axis; a point delta beyond one face has E = k delta^2 / 2 and a
    gradient of +-k delta on that axis only (pointing out of the box)."""
    k = 3.0
    pot = orc.make_pot1(BOX, ATTRACTIVE, k, origin=BOX_O, extension=BOX_E)
    x = BOX_O + 0.5 * BOX_E                             # centre: free
    g, e = ext_eval(pot, PRAD, x)
    assert np.all(g == 0) and e == 0
    if side == "upper":
        x[axis] = BOX_O[axis] + BOX_E[axis] - PRAD + delta
        sign = 1.0
    else:
        x[axis] = BOX_O[axis] + PRAD - delta
        sign = -1.0
    g, e = ext_eval(pot, PRAD, x)
    want = np.zeros(3)
    want[axis] = sign * k * delta
    assert np.allclose(g, want, atol=1e-12), (g, want)
    assert abs(e - 0.5 * k * delta * delta) < 1e-12
    # inside the face by the same margin: free (the particle radius is what moves the face)
    x2 = BOX_O + 0.5 * BOX_E
    x2[axis] = (BOX_O[axis] + BOX_E[axis] - PRAD - delta) if side == "upper" else (BOX_O[axis] + PRAD + delta)
    g, e = ext_eval(pot, PRAD, x2)
    assert np.all(g == 0) and e == 0


def test_box_attractive_corner_adds_faces():
    """Beyond two faces at once the energies add and each axis keeps its own term."""
    k = 2.0
    pot = orc.make_pot1(BOX, ATTRACTIVE, k, origin=BOX_O, extension=BOX_E)
    x = BOX_O + 0.5 * BOX_E
    x[0] = BOX_O[0] + BOX_E[0] - PRAD + 0.4
    x[2] = BOX_O[2] + PRAD - 0.9
    g, e = ext_eval(pot, PRAD, x)
    assert np.allclose(g, [k * 0.4, 0.0, -k * 0.9], atol=1e-12)
    assert abs(e - 0.5 * k * (0.16 + 0.81)) < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", ["upper", "lower"])
def test_box_repulsive_faces_closed_form(axis, side):
    """Repulsive box of the GPU reference (CudaSimulation.cu:1213-1226): a
    point delta beyond the face moved out by the particle radius,
    o + e + a (upper) or o - a (lower), is pushed further out: the gradient is
    -+k delta on that axis (the step moves it away) and the energy, the
    integral of that gradient, is -k delta^2 / 2.  Points between those faces
    feel nothing."""
    k, delta = 3.0, 0.6
    pot = orc.make_pot1(BOX, REPULSIVE, k, origin=BOX_O, extension=BOX_E)
    x = BOX_O + 0.5 * BOX_E
    for inside in (BOX_O[axis] - PRAD + 0.01, BOX_O[axis] + BOX_E[axis] + PRAD - 0.01):
        x[axis] = inside
        g, e = ext_eval(pot, PRAD, x)
        assert np.all(g == 0) and e == 0
    if side == "upper":
        x[axis] = BOX_O[axis] + BOX_E[axis] + PRAD + delta
        sign = -1.0
    else:
        x[axis] = BOX_O[axis] - PRAD - delta
        sign = 1.0
    g, e = ext_eval(pot, PRAD, x)
    want = np.zeros(3)
    want[axis] = sign * k * delta
    assert np.allclose(g, want, atol=1e-12), (g, want)
    assert abs(e + 0.5 * k * delta * delta) < 1e-12


def _java_cube_force(k, origin, extension, prad, x, attractive):
    """P1_Cube.java:187-273 with considerParticleRadius: the Java 'gradient'
    (the force) and energy; min/max are the sorted corners (:148-156)."""
    o, e = np.asarray(origin, float), np.asarray(extension, float)
    lo, hi = np.minimum(o, o + e), np.maximum(o, o + e)
    f = np.zeros(3)
    en = 0.0
    for i in range(3):
        if x[i] - prad >= lo[i] and x[i] + prad <= hi[i]:
            continue
        if x[i] - prad < lo[i]:
            r, r0 = x[i] - prad, lo[i]
        else:
            r, r0 = x[i] + prad, hi[i]
        f[i] += k * (r - r0)
        en += 0.5 * k * (r - r0) ** 2
    if attractive:
        f = -f
    return f, en


def test_box_attractive_equals_java_cube():
    rng = np.random.default_rng(41)
    for _ in range(400):
        k = rng.uniform(0.5, 5.0)
        a = rng.uniform(0.0, 1.0)
        pot = orc.make_pot1(BOX, ATTRACTIVE, k, origin=BOX_O, extension=BOX_E)
        x = BOX_O + rng.uniform(-3.0, 1.0 + 3.0 / 4.0, 3) * BOX_E
        g, e = ext_eval(pot, a, x)
        f, ej = _java_cube_force(k, BOX_O, BOX_E, a, x, True)
        assert np.allclose(g, -f, atol=1e-10), (x, g, f)
        assert abs(e - ej) < 1e-9


def test_sphere_repulsive_closed_form_and_java():
    """Repulsive sphere: inside at distance d < R, E = k (R - d)^2 / 2 and the
    gradient points to the centre with magnitude k (R - d) (the step pushes
    the particle out); P1_Sphere.java:250-262 at particle radius 0 (the GPU
    sphere has none, CudaSimulation.cu:1186-1190)."""
    k, R = 4.0, 3.0
    o = np.array([1.0, 2.0, -1.0])
    pot = orc.make_pot1(SPHERE, REPULSIVE, k, origin=o, radius=R)
    for d, u in ((1.0, np.array([0.0, 0.0, 1.0])), (2.5, np.array([0.6, -0.8, 0.0]))):
        g, e = ext_eval(pot, 0.0, o + d * u)
        assert np.allclose(g, -k * (R - d) * u, atol=1e-12)
        assert abs(e - 0.5 * k * (R - d) ** 2) < 1e-12
    g, e = ext_eval(pot, 0.0, o + 3.5 * np.array([1.0, 0.0, 0.0]))   # outside: free
    assert np.all(g == 0) and e == 0
    rng = np.random.default_rng(5)
    for _ in range(200):
        x = o + rng.uniform(-4, 4, 3)
        d = np.linalg.norm(x - o)
        fj = np.zeros(3)
        ej = 0.0
        if d < R:                                       # Java: force = k (r - r0)/r (center - x)
            fj = k * (d - R) / d * (o - x)
            ej = 0.5 * k * (d - R) ** 2
        g, e = ext_eval(pot, 0.0, x)
        assert np.allclose(g, -fj, atol=1e-10) and abs(e - ej) < 1e-10


def _java_cylinder_repulsive_force(k, center, normal, R, height, x):
    """P1_Cylinder.java:250-274 at particle radius 0 (set_coordinates:
    pointOnDiskPlane = x - n (x.n - c.n), distToDiskPlane = |x.n - c.n|,
    distToCenterWithinDiskPlane = |pointOnDiskPlane - c|)."""
    c = np.asarray(center, float)
    n = np.asarray(normal, float) / np.linalg.norm(normal)
    fac = x @ n - c @ n
    pp = x - n * fac
    dpl = abs(fac)
    din = np.linalg.norm(pp - c)
    f = np.zeros(3)
    if din < R:
        f += k * (din - R) / din * (c - pp)
    if dpl < 0.5 * height:
        f += k * (dpl - 0.5 * height) / dpl * (pp - x)
    return f


def test_cylinder_repulsive_closed_form_and_java():
    """Repulsive cylinder (P1_Cylinder.java sign, as CudaSimulation.cu:1121-1145
    would be with the sign that pushes particles out): inside, at radial
    distance r < R and axial offset s with |s| < h/2, the energy is
    k (R - r)^2 / 2 + k (h/2 - |s|)^2 / 2 and the gradient's radial and axial
    parts point inwards with magnitudes k (R - r) and k (h/2 - |s|)."""
    k, R, h = 2.0, 3.0, 4.0
    c = np.array([0.5, -1.0, 2.0])
    n = np.array([0.0, 0.0, 1.0])
    pot = orc.make_pot1(CYLINDER, REPULSIVE, k, origin=c, normal=n, radius=R, height=h)
    x = c + np.array([1.2, 0.0, 0.5])                  # r = 1.2, s = 0.5
    g, e = ext_eval(pot, 0.0, x)
    assert np.allclose(g, [-k * (R - 1.2), 0.0, -k * (0.5 * h - 0.5)], atol=1e-12)
    assert abs(e - 0.5 * k * ((R - 1.2) ** 2 + (0.5 * h - 0.5) ** 2)) < 1e-12
    x = c + np.array([0.0, -2.0, -2.5])                # r = 2 inside, |s| = 2.5 beyond h/2: radial only
    g, e = ext_eval(pot, 0.0, x)
    assert np.allclose(g, [0.0, k * (R - 2.0), 0.0], atol=1e-12)
    rng = np.random.default_rng(9)
    for nrm in (n, np.array([1.0, 1.0, 0.0]), np.array([0.3, -0.4, 0.8])):
        pot = orc.make_pot1(CYLINDER, REPULSIVE, k, origin=c, normal=nrm, radius=R, height=h)
        for _ in range(200):
            x = c + rng.uniform(-4, 4, 3)
            g, _ = ext_eval(pot, 0.0, x)
            f = _java_cylinder_repulsive_force(k, c, nrm, R, h, x)
            assert np.allclose(g, -f, atol=1e-10), (x, g, f)


# ---------------------------------------------------------------------------
# Pair displacements of PotentialDisplacementComputerTest.java:245-305
# ---------------------------------------------------------------------------
PDC = GOLD["harmonic_pair_displacement"]


def _pdc_tol(case):
    # the attractive vectors are given to 4-6 digits in the Java test
    return 1e-6 if case["subtype"] == "attractive" else 1e-12


@pytest.mark.parametrize("case", PDC["cases"])
def test_harmonic_pair_displacement_java_vectors(case):
    """Drift of each particle of a pair: -dt D_i / (kB T) grad_i U, with
    grad_2 U = -grad_1 U (PotentialDisplacementComputer.java:65-91)."""
    pot = orc.make_pot2(HARMONIC, SUB[case["subtype"]], case["k"])
    p1, p2 = np.array(case["p1"]), np.array(case["p2"])
    g1, _ = pair_eval(pot, case["r0"], p1 - p2)
    a = PDC["dt"] * PDC["D"][case["type"]] / (PDC["kB"] * PDC["T"])
    assert np.allclose(-a * g1, case["disp1"], rtol=0, atol=_pdc_tol(case))
    assert np.allclose(a * g1, case["disp2"], rtol=0, atol=_pdc_tol(case))


def pdc_spec(case):
    """Two particles of the case's type in a periodic box, only the case's
    potential on that type pair, radii r0/2 each, zero noise."""
    from readdygpu_amd.model import OrderTwoPotential, ParticleType, SystemSpec
    t = case["type"]
    spec = SystemSpec(box_lo=(-50.0, -50.0, -50.0), box_hi=(50.0, 50.0, 50.0), dt=PDC["dt"], kB=PDC["kB"],
                      T=PDC["T"], max_cutoff=10.0, seed=3)
    spec.types = [ParticleType(n, D=d, radius=0.5 * case["r0"]) for n, d in zip(("A", "B", "ABcmplx"), PDC["D"])]
    spec.pot2 = [OrderTwoPotential(HARMONIC, SUB[case["subtype"]], k=case["k"], pairs=[(t, t)])]
    xyz = np.array([case["p1"], case["p2"]], dtype=np.float32)
    return spec, xyz, np.array([t, t], dtype=np.int32)


@pytest.mark.parametrize("case", PDC["cases"])
def test_harmonic_pair_displacement_through_the_oracle_step(case):
    """The same vectors through the oracle's Euler-Maruyama step (orc_run) with
    zero noise: the step moves each particle by exactly its drift."""
    spec, xyz, types = pdc_spec(case)
    o = orc.OracleSystem(spec)
    x1, bad = o.run(xyz, types, 0, 1, noise=np.zeros((1, 3, 2), np.float32))
    assert bad == 0
    d = x1 - xyz.astype(np.float64)
    assert np.allclose(d[0], case["disp1"], rtol=0, atol=_pdc_tol(case))
    assert np.allclose(d[1], case["disp2"], rtol=0, atol=_pdc_tol(case))
