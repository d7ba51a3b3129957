"""rd_reaction_candidates_noalloc: the allocation-free variant of
rd_reaction_candidates (scratch in a fixed device array) returns the same
bit-exact candidate set, call after call, including a second call with a
smaller input that finds the scratch holding the first call's contents;
sizes beyond its scratch are refused with an error."""
import numpy as np
import pytest

from readdygpu_amd import RdError, Simulation, workloads

pytestmark = pytest.mark.gpu


def test_noalloc_candidates_equal_the_allocating_entry_on_two_calls():
    # first call: the 1M workload (A+B within 4 nm)
    spec, xyz, types = workloads.plain_1m()
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(5)
        ref = sim.reaction_candidates()
        got = sim.reaction_candidates_noalloc(len(ref) + 100)
        assert sim.last_count == len(ref) and len(ref) > 1000
        assert np.array_equal(got, ref)
        # a short buffer receives the first pairs of the sorted set and the full count
        part = sim.reaction_candidates_noalloc(100)
        assert sim.last_count == len(ref) and np.array_equal(part, ref[:100])
    # second call, smaller input, another handle: the scratch still holds the
    # first call's offsets and pairs
    spec, xyz, types = workloads.plain_1m(n=60_000, box=391.5)
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(3)
        ref = sim.reaction_candidates()
        got = sim.reaction_candidates_noalloc(len(ref) + 1)
        assert sim.last_count == len(ref) and len(ref) > 10
        assert np.array_equal(got, ref)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    """More than 2^20 candidate pairs: refused with RD_ERANGE; the handle and
    the allocating entry keep working."""
    spec, xyz, types = workloads.plain_10k(n=100_000)
    R = np.full((2, 2), 5.5)
    import dataclasses
    spec = dataclasses.replace(spec, box_hi=(126.0, 126.0, 126.0), reaction_radii=R)
    xyz = (xyz * 1.26).astype(np.float32)
    with Simulation(spec, xyz, types) as sim:
        ref = sim.reaction_candidates()
        assert len(ref) > (1 << 20)
        with pytest.raises(RdError, match="2\\^20 candidate pairs"):
            sim.reaction_candidates_noalloc(16)
        sim.simulate(1)
        assert len(sim.reaction_candidates()) > (1 << 20)
