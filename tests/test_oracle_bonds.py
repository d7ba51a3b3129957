"""Pins the oracle's bonded pairs (group potentials): closed forms across the
periodic boundary, finite differences of the total energy, cells against
brute force, Newton's third law, and the rest-length override."""
import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import workloads
from readdygpu_amd.model import (ATTRACTIVE, HARMONIC, PIECEWISE_WELL, REPULSIVE, SPRING, WEAK_HARMONIC,
                                 OrderTwoPotential, ParticleType, SystemSpec)


def two_bead_spec(k=20.0, r0=None, kind=HARMONIC, sub=SPRING, length=0.0, depth=0.0):
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(100.0, 50.0, 50.0), dt=0.01)
    spec.types = [ParticleType("A", D=1.0, radius=1.0), ParticleType("B", D=1.0, radius=1.5)]
    spec.bonds = np.array([[0, 1]], np.int32)
    spec.bond_pots = [OrderTwoPotential(kind, sub, k, length=length, depth=depth)]
    spec.bond_r0 = None if r0 is None else [r0]
    return spec


def test_spring_bond_closed_form_across_the_boundary():
    spec = two_bead_spec(k=20.0)
    o = orc.OracleSystem(spec)
    # x = 1 and x = 97: minimum-image separation 4 across the x faces; r0 = 1 + 1.5
    xyz = np.array([[1.0, 10.0, 10.0], [97.0, 10.0, 10.0]])
    g, e = o.gradients(xyz, np.array([0, 1], np.int32), brute=True)
    assert e == pytest.approx(0.5 * 20.0 * (4.0 - 2.5) ** 2)
    # grad_0 U = k (r - r0) (x0 - x1)_mi / r, with (x0 - x1)_mi = +4 along x
    assert np.allclose(g[0], [20.0 * 1.5, 0, 0]) and np.allclose(g[1], -g[0])


def test_rest_length_override_and_spring_compression():
    spec = two_bead_spec(k=8.0, r0=5.0)
    o = orc.OracleSystem(spec)
    xyz = np.array([[10.0, 10.0, 10.0], [10.0, 13.0, 10.0]])
    g, e = o.gradients(xyz, np.array([0, 1], np.int32), brute=True)
    assert e == pytest.approx(0.5 * 8.0 * (3.0 - 5.0) ** 2)
    assert np.allclose(g[0], [0, 8.0 * 2.0, 0])           # compressed: pushed apart
    # attractive harmonic bonds act only beyond r0, repulsive only inside
    for sub, on in ((ATTRACTIVE, False), (REPULSIVE, True)):
        s2 = two_bead_spec(k=8.0, r0=5.0, sub=sub)
        g2, e2 = orc.OracleSystem(s2).gradients(xyz, np.array([0, 1], np.int32), brute=True)
        assert (e2 > 0) == on and (np.abs(g2).max() > 0) == on


def test_bond_has_no_cutoff():
    spec = two_bead_spec(k=1.0)
    spec.max_cutoff = 3.0
    xyz = np.array([[10.0, 10.0, 10.0], [30.0, 10.0, 10.0]])      # r = 20 >> max_cutoff
    g, e = orc.OracleSystem(spec).gradients(xyz, np.array([0, 1], np.int32), brute=True)
    assert e == pytest.approx(0.5 * (20.0 - 2.5) ** 2) and g[0][0] < 0


@pytest.mark.parametrize("kind,sub,k,length,depth", [(HARMONIC, SPRING, 30.0, 0, 0), (HARMONIC, ATTRACTIVE, 30.0, 0, 0),
                                                    (WEAK_HARMONIC, ATTRACTIVE, 4.0, 2.0, 0),
                                                    (PIECEWISE_WELL, ATTRACTIVE, 10.0, 2.0, 1.5)])
def test_total_gradient_is_energy_derivative(kind, sub, k, length, depth):
    spec, xyz, types = workloads.bonded_chains(n_chains=12, length=6, box=30.0, seed=3)
    spec.bond_pots = [OrderTwoPotential(kind, sub, k, length=length, depth=depth)]
    o = orc.OracleSystem(spec)
    x = xyz.astype(np.float64)
    g, _ = o.gradients(x, types, brute=True)
    rng = np.random.default_rng(1)
    h = 1e-6
    for i in rng.choice(len(x), 15, replace=False):
        for d in range(3):
            xp = x.copy(); xp[i, d] += h
            xm = x.copy(); xm[i, d] -= h
            fd = (o.gradients(xp, types, brute=True)[1] - o.gradients(xm, types, brute=True)[1]) / (2 * h)
            assert abs(fd - g[i, d]) <= 1e-5 * max(1.0, abs(g[i, d])), (i, d)


def test_bond_gradients_cells_equal_bruteforce_and_newton3():
    spec, xyz, types = workloads.bonded_chains(n_chains=200, length=8, box=40.0, seed=4)
    o = orc.OracleSystem(spec)
    gb, eb = o.gradients(xyz, types, brute=True)
    gc, ec = o.gradients(xyz, types, brute=False)
    assert np.allclose(gb, gc, atol=1e-9) and abs(eb - ec) < 1e-8 * max(1, abs(eb))
    assert np.abs(gb.sum(0)).max() < 1e-8 * max(1.0, np.abs(gb).max())
    # the bonds contribute: without them the gradient differs
    spec.bonds = None
    g0, _ = orc.OracleSystem(spec).gradients(xyz, types, brute=True)
    assert np.abs(g0 - gb).max() > 1.0
