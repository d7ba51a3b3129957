"""Autotuner choices and their cache: a cell-size change on a running system
keeps the trajectory bit for bit (fixed-point pair sums), particles per lane
are a tuned dimension, and a cached choice is applied without timing."""
import json

import numpy as np
import pytest

from readdygpu_amd import RdError, Simulation, workloads

pytestmark = pytest.mark.gpu


def run(spec, xyz, types, steps, change=None, **launch):
    with Simulation(spec, xyz, types) as sim:
        if launch:
            sim.set_launch_config(**launch)
        sim.simulate(steps[0])
        if change is not None:
            change(sim)
        sim.simulate(steps[1])
        x, _ = sim.copy_pos_from_device()
        return x, sim.info()


def test_cell_size_change_keeps_the_trajectory():
    spec, xyz, types = workloads.plain_10k(n=6000)
    rc = float(spec.pair_range().max())
    ref, _ = run(spec, xyz, types, (7, 9))
    for cell in (1.3 * rc, 2.4 * rc, 0.0):
        x, info = run(spec, xyz, types, (7, 9), change=lambda s, c=cell: s.set_cell_size(c), ppt=2)
        assert info.step == 16
        if cell > 0:
            assert info.cell_size == pytest.approx(cell)
        assert np.array_equal(x, ref), cell


def test_tuning_cache_hit_applies_the_stored_choice(tmp_path):
    spec, xyz, types = workloads.plain_10k(n=6000)
    path = str(tmp_path / "tune.json")
    with Simulation(spec, xyz, types) as a:
        a.simulate(3)
        ms = a.autotune(trial_steps=2, cache=path)
        key = a.tune_key()
        lc_a, cell_a = a.launch_config(), a.info().cell_size
        assert a.info().step == 3
    db = json.load(open(path))
    assert db[key]["block"] == lc_a["used_block"] and db[key]["ppt"] == lc_a["used_ppt"]
    assert db[key]["cell_size"] == cell_a and db[key]["ms_per_step"] == ms
    ref, _ = run(spec, xyz, types, (3, 8))
    with Simulation(spec, xyz, types) as b:
        b.simulate(3)
        assert b.tune_key() == key
        assert b.autotune(trial_steps=2, cache=path) == ms          # hit: no timing
        lc_b = b.launch_config()
        assert (lc_b["used_block"], lc_b["used_ppt"]) == (lc_a["used_block"], lc_a["used_ppt"])
        assert b.info().cell_size == cell_a and b.info().step == 3
        b.simulate(8)
        x, _ = b.copy_pos_from_device()
    assert np.array_equal(x, ref)
    # another workload misses and is added next to the first
    spec2, xyz2, types2 = workloads.plain_10k(n=3000)
    with Simulation(spec2, xyz2, types2) as c:
        assert c.tune_key() != key
        c.autotune(trial_steps=2, cache=path)
    assert len(json.load(open(path))) == 2


def test_cell_size_too_fine_is_refused_and_the_lattice_kept():
    """A cell edge whose lattice would exceed the cell-index range is refused
    (RD_ERANGE); the previous lattice stays and the run continues bit for bit."""
    spec, xyz, types = workloads.plain_10k(n=3000)
    ref, _ = run(spec, xyz, types, (4, 6))

    def change(sim):
        before = sim.info().cell_size
        with pytest.raises(RdError):
            sim.set_cell_size(1e-4)
        assert sim.info().cell_size == before and sim.info().step == 4
    x, _ = run(spec, xyz, types, (4, 6), change=change)
    assert np.array_equal(x, ref)
