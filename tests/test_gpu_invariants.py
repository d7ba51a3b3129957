"""The cell-sorted state stays consistent through every operation that
rebuilds it: steps, cell-size changes, reactions (with array growth), frame
rewinds, replica batches with reactions, bonds, checkpoint restore (every
particle inside its cell's slot range, every index once, cell starts rising
to N)."""
import dataclasses

import numpy as np
import pytest

from readdygpu_amd import Simulation, workloads

pytestmark = pytest.mark.gpu


def test_invariants_through_steps_reactions_and_rewinds(tmp_path):
    spec, xyz, types = workloads.plain_10k(n=4000)
    with Simulation(spec, xyz, types) as sim:
        assert sim.check_invariants() == 0
        sim.simulate(17)
        assert sim.check_invariants() == 0
        sim.frame_copy()
        sim.simulate_async(9)
        sim.frame_wait()
        sim.frame_rewind()
        assert sim.check_invariants() == 0
        e = [6] + [0, -1, 0, 0, 0, 7] + [0, -1, 0, 0, 0, 3999]
        for k in range(4):
            e += [0, k % 2, 10.0 + k, 20.0, 30.0, -1]
        sim.apply_reactions(np.array(e, np.float32))             # 4000 -> 4002 particles: arrays grow
        assert sim.info().n == 4002 and sim.check_invariants() == 0
        sim.set_launch_config(kernel=1, block=128)
        sim.simulate(5)
        assert sim.check_invariants() == 0
        p = str(tmp_path / "c.ckpt")
        sim.checkpoint(p)
        sim.simulate(3)
        sim.restore(p)
        assert sim.check_invariants() == 0
    for cell in (5.0, 9.0):
        with Simulation(dataclasses.replace(spec, cell_size=cell), xyz, types) as sim:
            sim.simulate(4)
            assert sim.check_invariants() == 0


def test_invariants_replicas_bonds_reactions():
    spec, boxes, counts, xyz, types = workloads.batched_replicas(n_replicas=6, seed=4, nmin=200, nmax=1500)
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.simulate(6)
        assert sim.check_invariants() == 0
        b = boxes[3]
        cr = np.array([2, 0, -1, 0, 0, 0, 1, 0, 1, b[0] + 1.0, b[2] + 1.0, b[4] + 1.0, -1], np.float32)
        sim.apply_reactions(cr, creation_replica=np.array([0, 3], np.int32))
        sim.simulate(3)
        assert sim.check_invariants() == 0
    bspec, bx, bt = workloads.bonded_chains(n_chains=100, length=5, box=35.0)
    with Simulation(bspec, bx, bt) as sim:
        sim.simulate(10)
        assert sim.check_invariants() == 0
