"""The pair cases of the reference's ParticleConfigurationTest
(test/readdy/impl/sim/core/config/ParticleConfigurationTest.java:236-270
and :316-375), on the device.

Three type-0 particles of radius 1.4 nm at the positions of
test/testInputFiles/test_particleConfig/tplgy_coordinates.xml, in the box of
test/testInputFiles/test_default/param_global.xml ([-100, 100]^3).  The Java
test walks every pair its lattice offers: 3 pairs, the (0, 2) pair at
2.44949 nm.  It then creates a type-0 particle at (1, 2, 3), after which there
are 6 pairs at the distances listed in :363-375.  Here the full pair set is the
reaction-candidate set at a 3.5 nm reaction radius (every pair is closer), the
particle is created through the cReactions protocol, and the neighbour set
(d < 2.8 nm, the sum of the radii) is the subset of those distances below it.
"""
import numpy as np
import pytest

from readdygpu_amd import HARMONIC, REPULSIVE, OrderTwoPotential, ParticleType, Simulation, SystemSpec

pytestmark = pytest.mark.gpu

XYZ = np.array([[2, 2, 1.5], [1, 3, 4.5], [0, 1, 2.5]], dtype=np.float32)
EXPECTED = {(0, 2): 2.44949, (1, 2): 3.0, (0, 1): 3.316624, (0, 3): 1.802776, (1, 3): 1.802776, (2, 3): 1.5}


def pair_dists(pairs, x):
    return {(int(i), int(j)): float(np.linalg.norm(x[i] - x[j])) for i, j in pairs}


def test_pairs_before_and_after_creating_a_particle():
    spec = SystemSpec(box_lo=(-100.0, -100.0, -100.0), box_hi=(100.0, 100.0, 100.0), dt=1e-8, T=300.0, seed=1)
    spec.types = [ParticleType("A", D=1.0, radius=1.4)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0)]
    spec.reaction_radii = np.array([[3.5]])
    with Simulation(spec, XYZ, np.zeros(3, np.int32)) as sim:
        x, _ = sim.copy_pos_from_device()
        cand = pair_dists(sim.reaction_candidates(), x)
        assert sorted(cand) == [(0, 1), (0, 2), (1, 2)]
        assert cand[(0, 2)] == pytest.approx(2.44949, abs=1e-4)
        assert [tuple(p) for p in sim.neighbor_pairs()] == [(0, 2)]
        # createParticle(type 0, (1, 2, 3)) through cReactions: index -1, new type 0
        sim.apply_reactions(np.array([1, 100, 0, 1.0, 2.0, 3.0, -1], dtype=np.float32))
        x, t = sim.copy_pos_from_device()
        assert len(t) == 4 and t[3] == 0 and np.array_equal(x[3], [1.0, 2.0, 3.0])
        cand = pair_dists(sim.reaction_candidates(), x)
        near = [tuple(int(v) for v in p) for p in sim.neighbor_pairs()]
    assert sorted(cand) == sorted(EXPECTED)
    for pair, d in EXPECTED.items():
        assert cand[pair] == pytest.approx(d, abs=1e-4), pair
    assert near == sorted(p for p, d in EXPECTED.items() if d < 2.8)
