"""Frame hand-off and device-side reactions.

The asynchronous frame copy returns what the synchronous copy returns; a
frame loop that runs the next frame's steps during the hand-off and rewinds
them when reactions arrive produces, bit for bit, the frames of the plain
sequential loop; the state after a device-side reaction compaction steps
exactly like a fresh handle built from the same particles."""
import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import RdError, Simulation, workloads

pytestmark = pytest.mark.gpu


def test_frame_copy_equals_copy_pos():
    spec, xyz, types = workloads.plain_10k(n=3000)
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(7)
        sim.frame_copy()
        xa, ta = sim.frame_wait()
        xb, tb = sim.copy_pos_from_device()
    assert np.array_equal(xa, xb) and np.array_equal(ta, tb)


def reactions_for(frame, n):
    """cReactions for a frame: frame 1 deletes 3 particles and creates 2,
    frame 3 deletes 1 and creates 4 (the particle count grows past the first
    allocation), other frames none."""
    if frame == 1:
        return np.array([4, 0, -1, 0, 0, 0, 5, 0, 1, 10.0, 20.0, 30.0, -1, 0, -1, 0, 0, 0, n - 1,
                         0, 0, 40.0, 50.0, 60.0, 17], np.float32)
    if frame == 3:
        e = [5, 0, -1, 0, 0, 0, 2]
        for k in range(4):
            e += [0, k % 2, 5.0 + k, 6.0, 7.0, -1]
        return np.array(e, np.float32)
    return None


def test_speculative_frames_equal_sequential_frames():
    spec, xyz, types = workloads.plain_10k(n=4000)
    spf, frames = 12, 5
    ref = []
    with Simulation(spec, xyz, types) as sim:
        for f in range(frames):
            sim.simulate(spf)
            x, t = sim.copy_pos_from_device()
            ref.append((x, t))
            cr = reactions_for(f, len(t))
            if cr is not None:
                sim.apply_reactions(cr)
    got = []
    rewinds = 0
    with Simulation(spec, xyz, types) as sim:
        sim.simulate_async(spf)
        for f in range(frames):
            sim.frame_copy()
            ahead = f + 1 < frames
            if ahead:
                sim.simulate_async(spf)          # next frame runs during the hand-off
            x, t = sim.frame_wait()
            got.append((x, t))
            cr = reactions_for(f, len(t))
            if cr is not None:
                if ahead:
                    sim.frame_rewind()
                    rewinds += 1
                sim.apply_reactions(cr)
                if ahead:
                    sim.simulate_async(spf)
        assert sim.info().step == frames * spf
    assert rewinds == 2
    for f in range(frames):
        assert np.array_equal(got[f][0], ref[f][0]) and np.array_equal(got[f][1], ref[f][1]), f"frame {f}"


def test_device_reactions_leave_a_complete_state():
    spec, xyz, types = workloads.plain_10k(n=2000)
    rng = np.random.default_rng(3)
    z = rng.normal(size=(4, 3, 2000 - 3 + 5)).astype(np.float32)
    with Simulation(spec, xyz, types, rdf=(10.0, 40)) as sim:
        sim.simulate(3)
        e = [8]
        for i in (11, 700, 1999):
            e += [0, -1, 0, 0, 0, i]
        for k in range(5):
            e += [0, k % 2, 10.0 + k, 20.0, 30.0, -1]
        sim.apply_reactions(np.array(e, np.float32))
        x1, t1 = sim.copy_pos_from_device()
        # the RDF normalisation uses the particle numbers after the reactions
        sim.call_rdf_calculation()
        cnt, frames = sim.copy_rdf_matrix()
        g = sim.normalize_rdf()
        o = orc.OracleSystem(spec)
        ref = o.rdf_counts(x1, t1, 10.0, 40, cells=False)
        assert np.array_equal(cnt, ref)
        npt = [int(np.sum(t1 == k)) for k in range(spec.n_types)]
        assert np.allclose(g, orc.rdf_normalize(spec.n_types, 40, 10.0, spec.volume(), npt, 1, ref), rtol=1e-12)
        sim.set_noise_array(z)
        sim.simulate(4)
        xa, ta = sim.copy_pos_from_device()
    assert len(t1) == 2002
    with Simulation(spec, x1, t1) as fresh:
        fresh.set_noise_array(z)
        fresh.simulate(4)
        xb, tb = fresh.copy_pos_from_device()
    assert np.array_equal(xa, xb) and np.array_equal(ta, tb)


def test_rewind_needs_a_frame_copy_of_the_same_particles():
    spec, xyz, types = workloads.plain_10k(n=500)
    with Simulation(spec, xyz, types) as sim:
        with pytest.raises(RdError):
            sim.frame_rewind()
        sim.frame_copy()
        sim.frame_wait()
        sim.apply_reactions(np.array([1, 0, -1, 0, 0, 0, 4], np.float32))
        with pytest.raises(RdError):
            sim.frame_rewind()
