"""Known-answer vectors of the reference's JUnit tests and the step's error
reports, on the device through the C ABI.

* PotentialDisplacementComputerTest.java:245-305: the displacement of each
  particle of a harmonic pair (spring, repulsive, attractive) after one step
  with zero noise is its drift -dt D/(kB T) grad U.  The FP32 step rounds
  x - a g (|x| <= 6: ulp 4.8e-7), so the bound is 4 ulp of |x| beyond the
  vectors' own precision (1e-6 for the attractive vectors, given to 4-6
  digits).
* A particle that moves more than one box length in a step is still outside
  the box after the periodic correction: the step reports it
  (the reference loops the correction, CudaSimulation.cu:1857-1858; the
  oracle flags the same state).
* A pair gradient term beyond the fixed-point range is reported instead of
  being clamped silently.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import RdError, Simulation, workloads
from readdygpu_amd.model import ATTRACTIVE, HARMONIC, REPULSIVE, SPHERE, OrderOnePotential, OrderTwoPotential

from test_oracle_potentials import PDC, _pdc_tol, pdc_spec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", PDC["cases"])
def test_java_pair_displacements_on_the_device(case):
    spec, xyz, types = pdc_spec(case)
    with Simulation(spec, xyz, types) as sim:
        sim.set_noise_array(np.zeros((1, 3, 2), np.float32))
        sim.simulate(1)
        x1, t1 = sim.copy_pos_from_device()
    assert np.array_equal(t1, types)
    d = x1.astype(np.float64) - xyz.astype(np.float64)
    tol = _pdc_tol(case) + 4 * float(np.spacing(np.float32(6.0)))
    assert np.allclose(d[0], case["disp1"], rtol=0, atol=tol), (d[0], case["disp1"])
    assert np.allclose(d[1], case["disp2"], rtol=0, atol=tol), (d[1], case["disp2"])


def _far_jump_spec(k):
    """One particle 10 nm outside an attractive sphere of radius 10 in a
    100 nm box: the drift is a k 10 nm towards the centre."""
    spec, xyz, types = workloads.plain_10k(n=2)
    c = 50.0
    spec = dataclasses.replace(spec, pot1=[OrderOnePotential(SPHERE, ATTRACTIVE, k, origin=(c, c, c), radius=10.0)])
    xyz = np.array([[c + 20.0, c, c], [c, c, c]], np.float32)
    types = np.zeros(2, np.int32)
    return spec, xyz, types


def test_a_particle_moved_beyond_a_box_length_is_reported():
    spec, xyz, types = _far_jump_spec(1.0)
    a = spec.dt * spec.types[0].D / spec.kbt
    # drift 250 nm: x = 70 - 250 = -180, one correction gives -80, still outside
    spec, xyz, types = _far_jump_spec(25.0 / a)
    z = np.zeros((1, 3, 2), np.float32)
    xr, bad = orc.OracleSystem(spec).run(xyz, types, 0, 1, noise=z)
    assert bad == 1                    # the oracle flags it too
    with Simulation(spec, xyz, types) as sim:
        sim.set_noise_array(z)
        with pytest.raises(RdError, match="more than one box length"):
            sim.simulate(1)
    # 0.9 box lengths: x = -20, one correction brings it back to 80; not
    # reported, and the position equals the oracle's
    spec, xyz, types = _far_jump_spec(9.0 / a)
    xr, bad = orc.OracleSystem(spec).run(xyz, types, 0, 1, noise=z)
    assert bad == 0
    with Simulation(spec, xyz, types) as sim:
        sim.set_noise_array(z)
        sim.simulate(1)
        x1, _ = sim.copy_pos_from_device()
    assert np.abs(x1 - xr).max() < 1e-3


def test_fixed_point_range_overflow_is_reported():
    spec, xyz, types = workloads.plain_10k(n=2)
    xyz = np.array([[50.0, 50.0, 50.0], [50.5, 50.0, 50.0]], np.float32)   # r = 0.5 < r0 = 3
    types = np.zeros(2, np.int32)
    ok = dataclasses.replace(spec, pot2=[OrderTwoPotential(HARMONIC, REPULSIVE, k=1.0e8)])
    with Simulation(ok, xyz, types) as sim:              # |term| = 2.5e8 < 5e8: summed
        sim.set_noise_array(np.zeros((1, 3, 2), np.float32))
        g, _ = sim.gradients()
        assert abs(g[0, 0] - 2.5e8) < 1e2                # grad_0 U = k (r - r0)/r (x0 - x1)
    big = dataclasses.replace(spec, pot2=[OrderTwoPotential(HARMONIC, REPULSIVE, k=1.0e12)])
    with Simulation(big, xyz, types) as sim:
        with pytest.raises(RdError, match="fixed-point range"):
            sim.gradients()
    with Simulation(big, xyz, types) as sim:
        with pytest.raises(RdError, match="fixed-point range"):
            sim.simulate(1)
