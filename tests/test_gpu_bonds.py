"""Bonded groups on the GPU against the CPU oracle: gradients within the
FP32-vs-FP64 tolerance of the pair path (2e-5 * max|g| + 1e-5), bitwise
determinism across cell sizes and replica batching, bond bookkeeping through
reactions, and input checks."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import RdError, Simulation, workloads
from readdygpu_amd.model import ATTRACTIVE, HARMONIC, PIECEWISE_WELL, SPRING, WEAK_HARMONIC, OrderTwoPotential

pytestmark = pytest.mark.gpu


def chains(**kw):
    return workloads.bonded_chains(**{**dict(n_chains=300, length=8, box=40.0, seed=11), **kw})


def test_bond_gradients_match_oracle():
    spec, xyz, types = chains()
    # three bond potentials, one with a fixed rest length
    nb = len(spec.bonds)
    spec.bond_pots = [OrderTwoPotential(HARMONIC, SPRING, 50.0), OrderTwoPotential(WEAK_HARMONIC, ATTRACTIVE, 4.0,
                                                                                     length=2.0),
                      OrderTwoPotential(PIECEWISE_WELL, ATTRACTIVE, 10.0, length=2.0, depth=1.5)]
    spec.bond_pot = (np.arange(nb) % 3).astype(np.int32)
    spec.bond_r0 = [None, 2.5, None]
    with Simulation(spec, xyz, types) as sim:
        g, e = sim.gradients()
        sim.simulate(5)
        x5, t5 = sim.copy_pos_from_device()
        g5, e5 = sim.gradients()
    o = orc.OracleSystem(spec)
    for gg, ee, xx in ((g, e, xyz), (g5, e5, x5)):
        gr, er = o.gradients(xx.astype(np.float64), types, brute=True)
        tol = 2e-5 * np.abs(gr).max() + 1e-5
        assert np.abs(gg - gr).max() <= tol
        assert abs(ee - er) <= 1e-5 * max(1.0, abs(er))
    # some bonds cross the periodic boundary
    d = np.abs(xyz[spec.bonds[:, 0]] - xyz[spec.bonds[:, 1]]).max(1)
    assert (d > 20.0).any()


def test_bonded_trajectory_is_independent_of_the_cell_size():
    spec, xyz, types = chains()
    out = []
    for cell in (0.0, 5.0, 9.0):
        with Simulation(dataclasses.replace(spec, cell_size=cell), xyz, types) as sim:
            sim.simulate(30)
            out.append(sim.copy_pos_from_device()[0])
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_bonds_in_replicas_equal_separate_runs():
    spec, xyz, types = chains(n_chains=100)
    n = len(types)
    boxes = np.array([[0, 40, 0, 40, 0, 40], [0, 40, 0, 40, 0, 40], [0, 40, 0, 40, 0, 40]], float)
    counts = np.array([n, n, n])
    rng = np.random.default_rng(2)
    xs = [xyz, np.mod(xyz + rng.uniform(0, 40, 3), 40).astype(np.float32),
          np.mod(xyz + rng.uniform(0, 40, 3), 40).astype(np.float32)]
    xs = [np.minimum(x, np.float32(40) - np.float32(1e-3)) for x in xs]
    bs = dataclasses.replace(spec, bonds=np.concatenate([spec.bonds + r * n for r in range(3)]),
                             bond_pot=np.zeros(3 * len(spec.bonds), np.int32))
    with Simulation(bs, np.concatenate(xs), np.tile(types, 3), replicas=(boxes, counts)) as sim:
        sim.simulate(20)
        xb, _ = sim.copy_pos_from_device()
    for r in range(3):
        with Simulation(dataclasses.replace(spec, stream=r), xs[r], types) as one:
            one.simulate(20)
            xr, _ = one.copy_pos_from_device()
        assert np.array_equal(xb[r * n:(r + 1) * n], xr), f"replica {r}"


def test_stiff_bonds_hold_their_length():
    """Dimers with k = 200 kJ/mol/nm^2 (thermal spread sqrt(kT/k) = 0.11 nm)
    keep their rest length 2.2 nm over 2,000 steps; unbonded pairs drift."""
    spec, xyz, types = chains(n_chains=2000, length=2, box=80.0, k_bond=200.0, seed=12)
    spec.dt = 0.002
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(2000)
        x, _ = sim.copy_pos_from_device()
    d = x[spec.bonds[:, 0]] - x[spec.bonds[:, 1]]
    d -= 80.0 * np.round(d / 80.0)
    r = np.linalg.norm(d, axis=1)
    assert abs(r.mean() - 2.2) < 0.05 and r.max() < 3.5
    # bead 1 of chain c and bead 0 of chain c+1 are not bonded
    u = x[spec.bonds[1:, 0]] - x[spec.bonds[:-1, 1]]
    u -= 80.0 * np.round(u / 80.0)
    assert np.linalg.norm(u, axis=1).mean() > 10.0


def test_reactions_renumber_bonds():
    spec, xyz, types = chains(n_chains=50, length=4)
    with Simulation(spec, xyz, types) as sim:
        # delete particle 5 (inside chain 1) and create one particle
        cr = np.array([2, 0, -1, 0, 0, 0, 5, 0, 1, 20.0, 20.0, 20.0, -1], np.float32)
        sim.apply_reactions(cr)
        x, t = sim.copy_pos_from_device()
        g, _ = sim.gradients()
    n = len(types)
    keep = np.array([i for i in range(n) if i != 5])
    remap = -np.ones(n, int)
    remap[keep] = np.arange(n - 1)
    b = [(remap[a], remap[c]) for a, c in spec.bonds if a != 5 and c != 5]
    s2 = dataclasses.replace(spec, bonds=np.array(b, np.int32), bond_pot=np.zeros(len(b), np.int32))
    assert len(t) == n and t[-1] == 1
    gr, _ = orc.OracleSystem(s2).gradients(x.astype(np.float64), t, brute=True)
    assert np.abs(g - gr).max() <= 2e-5 * np.abs(gr).max() + 1e-5


def test_bad_bonds_are_rejected():
    spec, xyz, types = chains(n_chains=10, length=3)
    with Simulation(spec, xyz, types) as sim:
        p = [OrderTwoPotential(HARMONIC, SPRING, 1.0)]
        with pytest.raises(RdError):
            sim.set_bonds(np.array([[0, 0]]), None, p)               # self bond
        with pytest.raises(RdError):
            sim.set_bonds(np.array([[0, 1]]), np.array([3]), p)      # no such potential
        with pytest.raises(RdError):
            sim.set_bonds(np.array([[0, 1]]), None, p * 9)           # too many potentials
        sim.set_bonds(np.array([[0, 10_000]]), None, p)              # beyond the particle count
        with pytest.raises(RdError, match="beyond the particle count"):
            sim.simulate(1)
        sim.set_bonds(np.zeros((0, 2), np.int32), None, [])
        sim.simulate(2)
