"""The learned-slab and crowded-cytosol workloads at their full sizes on the
device, checked against the oracle where it finishes in seconds and through
size-independent properties elsewhere.

learned_slab (2,000,000 particles, 512-piece composite):
  * the exact composite on 20,000 seeded particle positions against the FP64
    oracle (energy |e - e_ref| <= 1e-5 max(1, |e_ref|); gradient
    |g - g_ref| <= 1e-4 max|g_ref| + 1e-4 where the two lowest pieces differ
    by more than 1e-3 in energy: at a switch between pieces either gradient
    is correct);
  * the external pass of the step on all 2M slots equals that evaluation: the
    gradient with the composite minus the gradient without it is the
    composite gradient, to the rounding of the FP32 addition;
  * 20 exact and 20 surrogate steps keep every position finite and inside the
    box and every type count.
crowded_cytosol (64 replicas, 5.06M particles, 256-disk membrane on species 6-7):
  * bit-exact replica-local cell ids and A+B candidates against the oracle
    for the three smallest replicas;
  * one frame of A+B -> C through rd_apply_reactions_replicas at full size:
    per-replica and per-type counts, positions inside their boxes, the sorted
    state's invariants, and a step after it;
  * the surrogate's gain at the accuracy gates: force RMSE <= 1%, pooled RDF
    within 0.5% of the exact run, a faster external pass and >= 3x the plain
    form's particle-steps/s.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import Simulation, workloads

pytestmark = pytest.mark.gpu


def _inside(x, lo, hi):
    return bool(np.isfinite(x).all() and np.all(x >= np.asarray(lo, np.float32)) and
                np.all(x <= np.asarray(hi, np.float32)))


def test_learned_slab_full_size():
    spec, pieces, xyz, types = workloads.learned_slab()
    n = len(types)
    assert n == 2_000_000 and len(pieces) == 512
    rng = np.random.default_rng(44)
    idx = np.sort(rng.choice(n, 20_000, replace=False))
    with Simulation(spec, xyz, types) as sim:
        g0, _ = sim.gradients()
        sim.set_composite(pieces)
        g, e = sim.composite_eval(xyz[idx])
        gr, er, _, gap = orc.composite(pieces, xyz[idx].astype(np.float64))
        assert np.abs(e - er).max() <= 1e-5 * max(1.0, np.abs(er).max())
        sel = (gap > 1e-3 * np.maximum(1.0, er)) | (er == 0.0)
        assert sel.mean() > 0.99, sel.mean()
        assert np.abs(g[sel] - gr[sel]).max() <= 1e-4 * np.abs(gr).max() + 1e-4
        # the step's external pass over every slot
        g1, _ = sim.gradients()
        gc, _ = sim.composite_eval(xyz)
        diff = (g1.astype(np.float64) - g0) - gc
        assert np.abs(diff).max() <= 4 * np.spacing(np.abs(g1).max().astype(np.float32)), np.abs(diff).max()
        sim.simulate(20)
        x, t = sim.copy_pos_from_device()
        assert _inside(x, spec.box_lo, spec.box_hi)
        assert np.array_equal(np.bincount(t, minlength=3), np.bincount(types, minlength=3))
        # the surrogate, trained briefly on the exact composite, in the step
        sim.surrogate_init(seed=4)
        sim.surrogate_make_data(1 << 20, seed=4)
        curve = sim.surrogate_train(steps=200, batch=8192, lr=3e-3, lr_min=1e-4, seed=4)
        assert np.isfinite(curve).all() and curve[-20:].mean() < curve[:5].mean()
        sim.set_external_mode(True)
        sim.simulate(20)
        x2, t2 = sim.copy_pos_from_device()
        assert _inside(x2, spec.box_lo, spec.box_hi)
        assert np.array_equal(np.bincount(t2, minlength=3), np.bincount(types, minlength=3))
        assert sim.check_invariants() == 0


def _replica_spec(spec, b, r):
    import dataclasses
    return dataclasses.replace(spec, box_lo=(b[0], b[2], b[4]), box_hi=(b[1], b[3], b[5]), stream=r)


def test_crowded_cytosol_full_size():
    spec, boxes, counts, xyz, types, pieces, mtypes = workloads.crowded_cytosol()
    n = int(counts.sum())
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    with Simulation(spec, xyz, types, replicas=(boxes, counts)) as sim:
        sim.set_composite(pieces, types=mtypes)
        sim.simulate(5)
        x, t = sim.copy_pos_from_device()
        ids = sim.cell_ids()
        cand = sim.reaction_candidates()
        for r in np.argsort(counts)[:3]:
            a, m = int(offs[r]), int(counts[r])
            o = orc.OracleSystem(_replica_spec(spec, boxes[r], int(r)), cell=sim.info().cell_size)
            assert np.array_equal(ids[a:a + m], o.cell_index(x[a:a + m])[0]), r
            mine = cand[(cand[:, 0] >= a) & (cand[:, 0] < a + m)]
            assert np.all((mine[:, 1] >= a) & (mine[:, 1] < a + m))
            assert np.array_equal(mine - a, o.candidates(x[a:a + m], t[a:a + m])), r
        # one frame of A + B -> C (species 0 + 1 -> 2): A removed, B becomes C at the midpoint
        rng = np.random.default_rng(7)
        seen = np.zeros(n, bool)
        sel = []
        for i, j in cand[rng.permutation(len(cand))]:
            if not seen[i] and not seen[j]:
                seen[i] = seen[j] = True
                sel.append((i, j))
        sel = np.array(sel[:5000])
        k = len(sel)
        assert k > 1000
        a_idx = np.where(t[sel[:, 0]] == 0, sel[:, 0], sel[:, 1])
        b_idx = np.where(t[sel[:, 0]] == 0, sel[:, 1], sel[:, 0])
        rep = np.searchsorted(offs, a_idx, side="right") - 1
        lo, hi = boxes[rep][:, 0::2], boxes[rep][:, 1::2]
        L = hi - lo
        d = x[b_idx].astype(np.float64) - x[a_idx]
        d -= L * np.round(d / L)
        mid = np.minimum(np.mod(x[a_idx] + 0.5 * d - lo, L) + lo, hi - 1e-3)
        ent = np.zeros((k, 2, 6), np.float32)
        ent[:, 0, 1] = -1
        ent[:, 0, 5] = a_idx
        ent[:, 1, 1] = 2
        ent[:, 1, 2:5] = mid
        ent[:, 1, 5] = b_idx
        cr = np.concatenate([[2 * k], ent.reshape(-1)]).astype(np.float32)
        sim.apply_reactions(cr, creation_replica=np.repeat(rep, 2).astype(np.int32))
        x2, t2 = sim.copy_pos_from_device()
        assert len(t2) == n - k
        want = np.bincount(t, minlength=8).astype(np.int64)
        want[0] -= k
        want[1] -= k
        want[2] += k
        assert np.array_equal(np.bincount(t2, minlength=8), want)
        cnt2 = counts - np.bincount(rep, minlength=len(counts))
        offs2 = np.concatenate([[0], np.cumsum(cnt2)])
        for r in range(len(counts)):
            xr = x2[offs2[r]:offs2[r + 1]]
            assert _inside(xr, boxes[r][0::2], boxes[r][1::2]), r
        assert sim.check_invariants() == 0
        sim.simulate(3)
        x3, _ = sim.copy_pos_from_device()
        assert np.isfinite(x3).all() and sim.check_invariants() == 0


def _ext_pass_ms(sim, reps=10):
    import ctypes as C
    ms = C.c_double()
    assert sim.lib.rd_bench_ext(sim.h, reps, C.byref(ms)) == 0
    return ms.value


def _steps_per_s(sim, n, steps=40):
    import time
    sim.simulate(steps)                        # graphs captured, state warmed
    t0 = time.perf_counter()
    sim.simulate(steps)
    return n * steps / (time.perf_counter() - t0)


def test_cytosol_surrogate_gain_at_the_gates():
    """What the trained network gains on the crowded cytosol: at force RMSE
    <= 1% on fresh samples and an RDF pooled over the replicas within 0.5%
    (relative L2) of the exact run from the same start with the same noise
    (200 steps, a frame every 2 steps), the surrogate's external pass is
    faster than the exact pass, and the batched step with it delivers >= 3x
    the particle-steps/s of the plain form (each replica on its own handle
    with the exact membrane)."""
    import dataclasses
    spec, boxes, counts, xyz, types, pieces, mtypes = workloads.crowded_cytosol()
    n = int(counts.sum())
    lmax = float(boxes[:, 1].max())

    def handle(w=None, norm=None):
        sim = Simulation(spec, xyz, types, replicas=(boxes, counts), rdf=(20.0, 100))
        sim.set_composite(pieces, types=mtypes)
        if w is not None:
            sim.set_surrogate_weights(w)
            sim.set_surrogate_norm(*norm)
            sim.set_external_mode(True)
        return sim

    def pooled_rdf(sim):
        for _ in range(100):
            sim.simulate(2)
            sim.call_rdf_calculation()
        g = sim.normalize_rdf()
        return (g * counts[:, None, None, None]).sum(0)

    with handle() as ex:
        ex.surrogate_init(seed=5)
        ex.set_surrogate_norm((0.0, 0.0, 0.0), (lmax, lmax, lmax), 1.0)
        ex.surrogate_make_data(1 << 22, seed=5)
        ex.surrogate_train(steps=3000, batch=8192, lr=3e-3, lr_min=1e-5, seed=5)
        rmse = ex.surrogate_validate(400_000, seed=55)
        w, norm = ex.surrogate_weights(), ex.surrogate_norm()
        ms_exact = _ext_pass_ms(ex)
        g_exact = pooled_rdf(ex)
    with handle(w, norm) as su:
        ms_sur = _ext_pass_ms(su)
        g_sur = pooled_rdf(su)
        batched = _steps_per_s(su, n)
    dev = float(np.linalg.norm(g_sur - g_exact) / np.linalg.norm(g_exact))
    t_plain = 0.0
    off = 0
    for r, c in enumerate(counts):
        b = boxes[r]
        s1 = dataclasses.replace(spec, box_lo=(b[0], b[2], b[4]), box_hi=(b[1], b[3], b[5]), stream=r)
        with Simulation(s1, xyz[off:off + c], types[off:off + c]) as one:
            one.set_composite(pieces, types=mtypes)
            t_plain += int(c) * 40 / _steps_per_s(one, int(c))
        off += c
    plain = n * 40 / t_plain
    print(f"cytosol: force rel RMSE {rmse:.4f}, pooled RDF rel L2 {dev:.5f}, external pass exact {ms_exact:.3f} ms "
          f"surrogate {ms_sur:.3f} ms, batched+surrogate {batched:.3e} plain {plain:.3e} particle-steps/s")
    assert rmse <= 0.01
    assert dev <= 0.005
    assert ms_sur < ms_exact
    assert batched >= 3.0 * plain
