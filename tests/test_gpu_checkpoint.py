"""Checkpoint / resume and failure reporting.

A run resumed from a snapshot continues bit for bit as the uninterrupted run
(positions, types, bonds, replica layout and the step counter that keys the
noise are restored); snapshots from another model or seed are refused; a
step that produces non-finite positions reports it."""
import dataclasses

import numpy as np
import pytest

from readdygpu_amd import RdError, Simulation, workloads
from readdygpu_amd.model import ATTRACTIVE, SPHERE, OrderOnePotential

pytestmark = pytest.mark.gpu


def test_resume_plain_run_bitwise(tmp_path):
    spec, xyz, types = workloads.plain_10k(n=3000)
    with Simulation(spec, xyz, types) as a:
        a.simulate(30)
        xa, ta = a.copy_pos_from_device()
    p = str(tmp_path / "s.ckpt")
    with Simulation(spec, xyz, types) as b:
        b.simulate(12)
        b.checkpoint(p)
    with Simulation(spec, xyz[:10], types[:10]) as c:      # another state, replaced by the snapshot
        c.simulate(3)
        c.restore(p)
        assert c.info().step == 12 and c.info().n == 3000
        c.simulate(18)
        xc, tc = c.copy_pos_from_device()
        assert c.info().step == 30
    assert np.array_equal(xa, xc) and np.array_equal(ta, tc)


def test_resume_replicas_with_bonds_bitwise(tmp_path):
    spec, xyz, types = workloads.bonded_chains(n_chains=60, length=6, box=30.0, seed=8)
    n = len(types)
    boxes = np.array([[0, 30, 0, 30, 0, 30], [0, 30, 0, 30, 0, 30]], float)
    counts = np.array([n, n])
    x2 = np.minimum(np.mod(xyz + 7.5, 30.0), np.float32(30.0) - np.float32(1e-3)).astype(np.float32)
    bs = dataclasses.replace(spec, bonds=np.concatenate([spec.bonds, spec.bonds + n]),
                             bond_pot=np.zeros(2 * len(spec.bonds), np.int32))
    X, T = np.concatenate([xyz, x2]), np.tile(types, 2)
    with Simulation(bs, X, T, replicas=(boxes, counts)) as a:
        a.simulate(25)
        xa, _ = a.copy_pos_from_device()
    p = str(tmp_path / "r.ckpt")
    with Simulation(bs, X, T, replicas=(boxes, counts)) as b:
        b.simulate(10)
        b.checkpoint(p)
    with Simulation(bs, X, T, replicas=(boxes, counts)) as c:
        c.restore(p)
        c.simulate(15)
        xc, _ = c.copy_pos_from_device()
    assert np.array_equal(xa, xc)


def test_snapshot_of_another_seed_or_file_is_refused(tmp_path):
    spec, xyz, types = workloads.plain_10k(n=500)
    p = str(tmp_path / "a.ckpt")
    with Simulation(spec, xyz, types) as a:
        a.simulate(2)
        a.checkpoint(p)
    with Simulation(dataclasses.replace(spec, seed=spec.seed + 1), xyz, types) as b:
        with pytest.raises(RdError, match="seed"):
            b.restore(p)
        q = tmp_path / "junk.ckpt"
        q.write_bytes(b"not a snapshot")
        with pytest.raises(RdError):
            b.restore(str(q))


def test_non_finite_positions_are_reported():
    """A confinement stiff enough to overflow FP32 makes the update
    non-finite; the step reports it instead of running on."""
    spec, xyz, types = workloads.plain_10k(n=200)
    spec = dataclasses.replace(spec, pot1=[OrderOnePotential(SPHERE, ATTRACTIVE, 3.0e38, origin=(50.0, 50.0, 50.0),
                                                             radius=1.0)])
    with Simulation(spec, xyz, types) as sim:
        with pytest.raises(RdError, match="non-finite"):
            sim.simulate(1)


def test_snapshot_whose_replica_sizes_do_not_add_up_is_refused(tmp_path):
    """The replica sizes stored in a snapshot must add up to its particle
    count: a crafted file whose sizes overrun the stored arrays (or wrap past
    2^31) is refused before any particle is read from it."""
    import struct
    spec, xyz, types = workloads.plain_10k(n=400)
    p = tmp_path / "a.ckpt"
    with Simulation(spec, xyz, types) as a:
        a.simulate(1)
        a.checkpoint(str(p))
    raw = bytearray(p.read_bytes())
    # magic 8, version 4, seed 8, step 8, nt 4, nrep 4, n 4; replica 0: lo, hi (6 doubles), stream 4, count 4
    off = 8 + 4 + 8 + 8 + 4 + 4 + 4 + 48 + 4
    assert struct.unpack_from("<I", raw, off)[0] == 400
    with Simulation(spec, xyz, types) as b:
        for bad in (401, 1 << 31, 0xFFFFFFFF):
            q = tmp_path / f"bad{bad}.ckpt"
            r = bytearray(raw)
            struct.pack_into("<I", r, off, bad)
            q.write_bytes(bytes(r))
            with pytest.raises(RdError, match="not a snapshot"):
                b.restore(str(q))
        # the handle is unchanged and still runs
        b.simulate(2)
        assert b.info().n == 400
