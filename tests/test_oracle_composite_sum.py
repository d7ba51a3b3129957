"""The learned-slab geometry with its pieces summed, as the reference's
orderOne adds every order-one potential of a particle's type
(CudaSimulation.cu:1049-1236): 512 attractive z-normal disks and cylinders
(k = 5 kJ/mol/nm^2) in a 2000 x 2000 x 400 nm slab.

What the oracle shows, at FP64:
* every disk adds k to d2U/dz2 everywhere (its axial term is k s^2 / 2 at any
  distance), so d2U/dz2 >= 256 k = 1280 kJ/mol/nm^2 at every point of the
  slab; the in-plane curvature far from the pieces approaches 512 k;
* the Euler-Maruyama step x' = x - a dU/dx (a = dt D / kT) maps a deviation
  from the minimum along z to (1 - a K) times itself, so it diverges when
  a K > 2: with dt = 0.05 ns, kT = 2.437 kJ/mol and the slab's D of 1.0, 0.5
  and 0.25 nm^2/ns, a K_zz is 26, 13 and 6.6;
* the oracle's own step (orc_run) throws the particles out of the box in one
  step at those D;
* below the stability bound the bowl pulls every particle into one point
  within a few steps: with the pair repulsion on, most pairs of a 2,000-
  particle sample are then within range (the neighbour search degenerates to
  all pairs).
"""
import dataclasses

import numpy as np

from oracle import oracle as orc
from readdygpu_amd import workloads


def _slab(n):
    spec, pieces, xyz, types = workloads.learned_slab(n=n)
    return spec, pieces, xyz, types


def _hessian(pieces, x, h=1e-3):
    H = np.zeros((3, 3))
    for d in range(3):
        xp = np.array(x, float); xm = np.array(x, float)
        xp[d] += h; xm[d] -= h
        gp, _ = orc.composite_sum(pieces, xp[None])
        gm, _ = orc.composite_sum(pieces, xm[None])
        H[:, d] = (gp[0] - gm[0]) / (2 * h)
    return 0.5 * (H + H.T)


def test_composite_sum_is_the_sum_of_the_pieces():
    """orc_composite_sum against the pieces one at a time (each piece's
    energy is pinned to 0.5 k dist^2 by brute-force distances in
    test_oracle_surrogate.py) and its gradient against central differences
    of its own energy."""
    spec, pieces, xyz, types = _slab(64)
    sub = pieces[::16]
    x = xyz[:12].astype(np.float64)
    g, e = orc.composite_sum(sub, x)
    for i in range(len(x)):
        parts = [orc.piece_eval(p, x[i]) for p in sub]
        assert abs(e[i] - sum(pe for pe, _ in parts)) <= 1e-9 * max(1.0, abs(e[i]))
        assert np.allclose(g[i], np.sum([pg for _, pg in parts], axis=0), rtol=1e-12, atol=1e-9)
        h = 1e-4
        for d in range(3):
            xp = x[i].copy(); xp[d] += h
            xm = x[i].copy(); xm[d] -= h
            fd = (orc.composite_sum(sub, xp[None])[1][0] - orc.composite_sum(sub, xm[None])[1][0]) / (2 * h)
            assert abs(fd - g[i, d]) <= 1e-5 * max(1.0, np.abs(g[i]).max()), (i, d)


def test_summed_slab_curvature_exceeds_the_euler_maruyama_bound():
    spec, pieces, xyz, types = _slab(64)
    k = pieces[0].k
    n_disk = sum(1 for p in pieces if p.kind == 1)
    assert n_disk == 256 and len(pieces) == 512 and k == 5.0
    rng = np.random.default_rng(3)
    pts = np.array(spec.box_lo) + rng.random((24, 3)) * (np.array(spec.box_hi) - np.array(spec.box_lo))
    kzz, kmin = [], []
    for x in pts:
        H = _hessian(pieces, x)
        kzz.append(H[2, 2])
        kmin.append(np.linalg.eigvalsh(H)[0])
    kzz = np.array(kzz)
    # closed form of the disks' share: n_disk * k everywhere
    assert np.all(kzz >= n_disk * k - 1e-3), kzz.min()
    assert np.all(np.array(kmin) > 0), "the summed potential is a convex bowl at every sample point"
    a_k = [spec.dt * t.D / spec.kbt * kzz.min() for t in spec.types]
    assert min(a_k) > 2.0, a_k
    # the largest D the EM step tolerates along z is far below the slab's
    d_max = 2.0 * spec.kbt / (spec.dt * kzz.max())
    assert d_max < 0.05 and d_max < min(t.D for t in spec.types) / 5, d_max


def test_summed_slab_throws_particles_out_in_one_step():
    spec, pieces, xyz, types = _slab(1000)
    sp = dataclasses.replace(spec, pot1=list(pieces), pot2=[])
    x1, bad = orc.OracleSystem(sp).run(xyz, types, 0, 1, noise=np.zeros((1, 3, len(types)), np.float32))
    assert bad == 1
    lo, hi = np.array(spec.box_lo), np.array(spec.box_hi)
    outside = np.any((x1 < lo) | (x1 > hi), axis=1)
    assert outside.mean() > 0.5, outside.mean()


def test_summed_slab_collapses_below_the_stability_bound():
    spec, pieces, xyz, types = _slab(2000)
    # stable step: a K_max < 1 along every axis (D far below the slab's)
    H = _hessian(pieces, np.array([1000.0, 1000.0, 200.0]))
    kmax = np.linalg.eigvalsh(H)[-1]
    D = 0.5 * spec.kbt / (spec.dt * 2.0 * kmax)
    sp = dataclasses.replace(spec, pot1=list(pieces),
                             types=[dataclasses.replace(t, D=D) for t in spec.types])
    o = orc.OracleSystem(sp)
    x, bad = o.run(xyz, types, 0, 40)
    assert bad == 0
    spread0 = np.linalg.norm(xyz - xyz.mean(0), axis=1).max()
    spread = np.linalg.norm(x - x.mean(0), axis=1).max()
    assert spread0 > 500.0 and spread < 5.0, (spread0, spread)
    # most pairs within the repulsion range: the cell list holds all particles in a few cells
    p = o.pairs(x, types, brute=True)
    n = len(types)
    assert len(p) > 0.5 * n * (n - 1) / 2, len(p)
