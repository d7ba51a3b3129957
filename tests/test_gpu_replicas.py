"""Batched replicas: one handle holding independent periodic boxes.

Every replica must evolve bit for bit as the same system run alone (same box,
noise stream = replica index), and match the CPU oracle per replica:
bit-exact noise, cell ids, pair sets and RDF counts; gradients within the
FP32-vs-FP64 tolerance of the plain path (2e-5 * max|g| + 1e-5).
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import RdError, Simulation, workloads

pytestmark = pytest.mark.gpu


def small_batch(n_replicas=12, seed=7):
    """Uneven replicas (100..3000 particles) of the batched workload's system."""
    return workloads.batched_replicas(n_replicas=n_replicas, seed=seed, nmin=100, nmax=3000)


def replica(spec, boxes, counts, xyz, types, r):
    off = int(np.sum(counts[:r]))
    n = int(counts[r])
    b = boxes[r]
    s = dataclasses.replace(spec, box_lo=(b[0], b[2], b[4]), box_hi=(b[1], b[3], b[5]), stream=r)
    return s, xyz[off:off + n], types[off:off + n], off


@pytest.fixture(scope="module")
def batch():
    return small_batch()


@pytest.mark.parametrize("kernel", [1, 2])
def test_replicas_equal_separate_runs_bitwise(batch, kernel):
    """Per-workgroup items (1) and the resident-grid work queue (2)."""
    spec, boxes, counts, xyz, types = batch
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.set_launch_config(kernel=kernel)
        info = sim.info()
        assert info.n_replicas == len(counts) and info.n == int(counts.sum())
        sim.simulate(25)
        xb, tb = sim.copy_pos_from_device()
    for r in range(len(counts)):
        s, x, t, off = replica(spec, boxes, counts, xyz, types, r)
        with Simulation(s, x, t) as one:
            one.simulate(25)
            xr, tr = one.copy_pos_from_device()
        n = len(t)
        assert np.array_equal(xb[off:off + n], xr), f"replica {r}"
        assert np.array_equal(tb[off:off + n], tr)


def test_replica_noise_and_cell_ids_bitexact(batch):
    spec, boxes, counts, xyz, types = batch
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        z = sim.noise(3, 2)
        sim.simulate(4)
        x, t = sim.copy_pos_from_device()
        ids = sim.cell_ids()
    for r in range(len(counts)):
        s, _, _, off = replica(spec, boxes, counts, xyz, types, r)
        n = int(counts[r])
        ref = orc.noise(spec.seed, r, 3, 2, n)
        assert np.array_equal(z[:, :, off:off + n], ref), f"replica {r}"
        o = orc.OracleSystem(s)
        assert np.array_equal(ids[off:off + n], o.cell_index(x[off:off + n])[0]), f"replica {r}"


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_replica_gradients_and_pairs(batch, kernel):
    spec, boxes, counts, xyz, types = batch
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.set_launch_config(kernel=kernel)
        sim.simulate(3)
        x, t = sim.copy_pos_from_device()
        g, e = sim.gradients()
        p = sim.neighbor_pairs()
    eref = 0.0
    pref = []
    for r in range(len(counts)):
        s, _, _, off = replica(spec, boxes, counts, xyz, types, r)
        n = int(counts[r])
        o = orc.OracleSystem(s)
        gr, er = o.gradients(x[off:off + n].astype(np.float64), t[off:off + n], brute=True)
        tol = 2e-5 * np.abs(gr).max() + 1e-5
        assert np.abs(g[off:off + n] - gr).max() <= tol, f"replica {r}"
        eref += er
        pr = o.pairs(x[off:off + n], t[off:off + n], brute=True)
        pref.append(pr + off)
    assert abs(e - eref) <= 1e-5 * max(1.0, abs(eref))
    assert np.array_equal(p, np.concatenate(pref))


def test_replica_rdf_counts_bitexact(batch):
    spec, boxes, counts, xyz, types = batch
    R = len(counts)
    rmax, nb = 10.0, 50
    with Simulation(spec, xyz, types, replicas=(boxes, counts), rdf=(rmax, nb)) as sim:
        ref = np.zeros((R, 2, 2, nb), dtype=np.uint64)
        for f in range(2):
            sim.simulate(5)
            x, t = sim.copy_pos_from_device()
            sim.call_rdf_calculation()
            for r in range(R):
                s, _, _, off = replica(spec, boxes, counts, xyz, types, r)
                n = int(counts[r])
                ref[r] += orc.OracleSystem(s).rdf_counts(x[off:off + n], t[off:off + n], rmax, nb, cells=False)
        c, frames = sim.copy_rdf_matrix()
        g = sim.normalize_rdf()
    assert c.shape == (R, 2, 2, nb) and frames == 2
    assert np.array_equal(c, ref)
    for r in range(R):
        s, _, t0, _ = replica(spec, boxes, counts, xyz, types, r)
        npt = [int(np.sum(t0 == 0)), int(np.sum(t0 == 1))]
        gr = orc.rdf_normalize(2, nb, rmax, s.volume(), npt, 2, ref[r])
        assert np.allclose(g[r], gr, rtol=1e-12), f"replica {r}"


def test_replica_autotune_keeps_trajectory(batch):
    spec, boxes, counts, xyz, types = batch
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as a:
        a.simulate(10)
        xa, _ = a.copy_pos_from_device()
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as b:
        b.autotune(trial_steps=3)
        assert b.info().step == 0 and b.info().n_replicas == len(counts)
        b.simulate(10)
        xb, _ = b.copy_pos_from_device()
    assert np.array_equal(xa, xb)


def test_full_batched_workload_runs():
    """The 256-replica workload (about 2.6M particles): finite, inside each
    box, every replica keeps its particle count and types."""
    spec, boxes, counts, xyz, types = workloads.batched_replicas()
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.simulate(20)
        x, t = sim.copy_pos_from_device()
    assert np.array_equal(t, types)
    off = 0
    for r, n in enumerate(counts):
        xs = x[off:off + n]
        lo = boxes[r, 0::2].astype(np.float32)
        hi = boxes[r, 1::2].astype(np.float32)
        assert np.isfinite(xs).all()
        assert (xs >= lo).all() and (xs <= hi).all(), f"replica {r}"
        off += n


def test_replica_reactions_equal_separate_runs(batch):
    """Reactions in a handle of replicas: removals, creations (placed by
    creation_replica) and type changes give, replica for replica, the state
    and the following trajectory of the same reactions applied to each system
    alone."""
    spec, boxes, counts, xyz, types = batch
    R = len(counts)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    local = {}            # replica -> list of local entries (readdyId, newType, x, y, z, index)
    for r in range(R):
        b = boxes[r]
        e = []
        if r % 3 == 0:
            e.append([0, -1, 0, 0, 0, 0])                                  # remove local 0
        if r % 4 == 1:
            e.append([0, 1, b[0] + 1.0, b[2] + 2.0, b[4] + 3.0, -1])       # create one B
        if r % 5 == 2:
            e.append([0, 0, b[0] + 4.0, b[2] + 4.0, b[4] + 4.0, 3])        # type change of local 3
        if r == R - 1:
            e.append([0, -1, 0, 0, 0, int(counts[r]) - 1])                  # remove the last particle
        local[r] = e
    glob, rep = [], []
    for r in range(R):
        for x in local[r]:
            g = list(x)
            if g[5] != -1:
                g[5] += int(offs[r])
            glob.append(g)
            rep.append(r)
    cr = np.array([len(glob)] + [v for g in glob for v in g], np.float32)
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.simulate(5)
        sim.apply_reactions(cr, creation_replica=np.array(rep, np.int32))
        info = sim.info()
        x1, t1 = sim.copy_pos_from_device()
        sim.simulate(10)
        x2, t2 = sim.copy_pos_from_device()
    new_counts = []
    for r in range(R):
        s, x, t, _ = replica(spec, boxes, counts, xyz, types, r)
        with Simulation(s, x, t) as one:
            one.simulate(5)
            if local[r]:
                one.apply_reactions(np.array([len(local[r])] + [v for g in local[r] for v in g], np.float32))
            xa, ta = one.copy_pos_from_device()
            one.simulate(10)
            xb, tb = one.copy_pos_from_device()
        o = int(sum(new_counts))
        n = len(ta)
        new_counts.append(n)
        assert np.array_equal(x1[o:o + n], xa) and np.array_equal(t1[o:o + n], ta), f"replica {r}"
        assert np.array_equal(x2[o:o + n], xb) and np.array_equal(t2[o:o + n], tb), f"replica {r}"
    assert info.n == sum(new_counts) and info.n_replicas == R
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        with pytest.raises(RdError):             # a pure creation needs its replica
            sim.apply_reactions(np.array([1, 0, 1, 1.0, 1.0, 1.0, -1], np.float32))


def test_ragged_replicas_with_empty_and_wave_boundary_sizes(batch):
    """Replicas of 0, 1, 63, 64, 65, 0 and 130 particles (empty replicas, one
    particle, sizes on both sides of a 64-lane wave): each evolves bit for bit
    as the same box alone, and the pair sets and gradients match the oracle."""
    spec, boxes, _, _, _ = batch
    counts = np.array([0, 1, 63, 64, 65, 0, 130], dtype=np.int32)
    boxes = boxes[:len(counts)]
    rng = np.random.default_rng(11)
    xyz, types = [], []
    for r, c in enumerate(counts):
        b = boxes[r]
        lo, hi = np.array([b[0], b[2], b[4]]), np.array([b[1], b[3], b[5]])
        xyz.append((lo + rng.random((c, 3)) * (hi - lo) * 0.999).astype(np.float32))
        types.append(rng.integers(0, spec.n_types, c).astype(np.int32))
    xyz, types = np.concatenate(xyz), np.concatenate(types)
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.simulate(9)
        xb, tb = sim.copy_pos_from_device()
        g, e = sim.gradients()
        p = sim.neighbor_pairs()
    pref, eref = [], 0.0
    for r in range(len(counts)):
        s, x, t, off = replica(spec, boxes, counts, xyz, types, r)
        n = len(t)
        with Simulation(s, x, t) as one:
            one.simulate(9)
            xr, tr = one.copy_pos_from_device()
        assert np.array_equal(xb[off:off + n], xr), f"replica {r}"
        assert np.array_equal(tb[off:off + n], tr)
        if n == 0:
            continue
        o = orc.OracleSystem(s)
        gr, er = o.gradients(xb[off:off + n].astype(np.float64), tb[off:off + n], brute=True)
        assert np.abs(g[off:off + n] - gr).max() <= 2e-5 * np.abs(gr).max() + 1e-5, f"replica {r}"
        eref += er
        pref.append(o.pairs(xb[off:off + n], tb[off:off + n], brute=True) + off)
    assert abs(e - eref) <= 1e-5 * max(1.0, abs(eref))
    assert np.array_equal(p, np.concatenate(pref))
