"""Seeded synthetic workloads.

plain_10k   10,000 particles, 2 species, 100 nm periodic box, harmonic
            repulsion k=10, dt=0.1 ns, 1,000 steps, RDF 0-20 nm / 200 bins.
plain_1m    1,000,000 particles, 4 species, 1000 nm box, repulsion + a
            piecewise harmonic well (depth 1, width 1) between species 0 and 1,
            spherical confinement R=480 nm, A+B candidates at 4 nm, dt=0.05 ns,
            500 steps, RDF 0-20 nm / 200 bins.

Values the workload descriptions leave open are fixed here: T = 293.15 K with
kB = 0.0083144621 kJ/mol/K (the reference GPU input's constants), sphere
k = 1 kJ/mol/nm^2, well inner stiffness 0 (the harmonic repulsion supplies it),
maxCutoff = 10 nm as in lib/ReaDDyGPU.cpp:101.
"""
from __future__ import annotations

import numpy as np

from .model import (ATTRACTIVE, CYLINDER, DISK, HARMONIC, PIECEWISE_WELL, REPULSIVE, SPHERE, SPRING,
                    OrderOnePotential, OrderTwoPotential, ParticleType, SystemSpec)


def uniform_positions(rng: np.random.Generator, n: int, lo, hi) -> np.ndarray:
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    x = (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)
    # keep float32 rounding inside [lo, hi)
    return np.minimum(np.maximum(x, lo.astype(np.float32)), np.nextafter(hi.astype(np.float32), -np.inf))


def plain_10k(n: int = 10_000, seed: int = 1, cell_size: float = 0.0):
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(100.0, 100.0, 100.0), dt=0.1, seed=seed,
                      cell_size=cell_size)
    spec.types = [ParticleType("A", D=1.0, radius=1.5), ParticleType("B", D=0.5, radius=2.0)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0)]
    rng = np.random.default_rng(seed)
    xyz = uniform_positions(rng, n, spec.box_lo, spec.box_hi)
    types = np.repeat(np.arange(2, dtype=np.int32), [n // 2, n - n // 2])
    return spec, xyz, types


PLAIN_10K = dict(steps=1000, rdf_every=100, rdf_range=20.0, rdf_bins=200)


def plain_1m(n: int = 1_000_000, seed: int = 2, cell_size: float = 0.0, box: float = 1000.0):
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(box, box, box), dt=0.05, seed=seed,
                      cell_size=cell_size)
    radii = np.linspace(1.0, 3.0, 4)
    Ds = np.linspace(0.1, 2.0, 4)
    spec.types = [ParticleType(f"S{i}", D=float(Ds[i]), radius=float(radii[i])) for i in range(4)]
    spec.pot2 = [
        OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0),
        OrderTwoPotential(PIECEWISE_WELL, ATTRACTIVE, k=0.0, length=1.0, depth=1.0, pairs=[(0, 1)]),
    ]
    c = 0.5 * box
    spec.pot1 = [OrderOnePotential(SPHERE, ATTRACTIVE, k=1.0, origin=(c, c, c), radius=0.48 * box)]
    R = np.zeros((4, 4))
    R[0, 1] = R[1, 0] = 4.0
    spec.reaction_radii = R
    rng = np.random.default_rng(seed)
    xyz = uniform_positions(rng, n, spec.box_lo, spec.box_hi)
    types = np.repeat(np.arange(4, dtype=np.int32), n // 4)
    if len(types) < n:
        types = np.concatenate([types, np.zeros(n - len(types), dtype=np.int32)])
    return spec, xyz, types


PLAIN_1M = dict(steps=500, rdf_range=20.0, rdf_bins=200)


def batched_replicas(n_replicas: int = 256, seed: int = 3, nmin: int = 500, nmax: int = 50_000,
                     phi: float = 0.05, cell_size: float = 0.0):
    """256 independent periodic boxes, particle counts log-uniform in
    [500, 50000], 2 species as plain_10k (radii 1.5/2.0 nm, D 1.0/0.5), harmonic
    repulsion k=10, dt=0.1 ns; each box edge makes the volume fraction 0.05.
    Returns (spec, boxes [R,6], counts [R], xyz [sum N,3], types [sum N]);
    replica r uses noise stream r."""
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0), dt=0.1, seed=seed, cell_size=cell_size)
    spec.types = [ParticleType("A", D=1.0, radius=1.5), ParticleType("B", D=0.5, radius=2.0)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0)]
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_replicas))).astype(np.int64)
    boxes = np.zeros((n_replicas, 6))
    xyz, types = [], []
    vol_a, vol_b = 4.0 / 3.0 * np.pi * 1.5 ** 3, 4.0 / 3.0 * np.pi * 2.0 ** 3
    for r, n in enumerate(counts):
        na = n // 2
        L = float(np.cbrt((na * vol_a + (n - na) * vol_b) / phi))
        boxes[r] = (0.0, L, 0.0, L, 0.0, L)
        xyz.append(uniform_positions(rng, int(n), (0.0, 0.0, 0.0), (L, L, L)))
        types.append(np.repeat(np.arange(2, dtype=np.int32), [na, n - na]))
    return spec, boxes, counts, np.concatenate(xyz), np.concatenate(types)


BATCHED = dict(steps=1000, rdf_range=20.0, rdf_bins=200)


def composite_pieces(rng: np.random.Generator, n_pieces: int, lo, hi, k: float = 5.0, rmin: float = 10.0,
                     rmax: float = 80.0):
    """Attractive z-normal pieces, alternating DISK and CYLINDER, centres
    uniform in [lo, hi), radii and cylinder heights uniform in [rmin, rmax]."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    pieces = []
    for i in range(n_pieces):
        c = lo + rng.random(3) * (hi - lo)
        r = float(rng.uniform(rmin, rmax))
        hgt = float(rng.uniform(rmin, rmax))
        kind = DISK if i % 2 == 0 else CYLINDER
        pieces.append(OrderOnePotential(kind, ATTRACTIVE, k=k, origin=tuple(float(v) for v in c), normal=(0.0, 0.0, 1.0),
                                        radius=r, height=hgt if kind == CYLINDER else 0.0))
    return pieces


def learned_slab(n: int = 2_000_000, seed: int = 4, n_pieces: int = 512, box=(2000.0, 2000.0, 400.0),
                 cell_size: float = 0.0):
    """2,000,000 particles of 3 species (radii 1.0/1.5/2.0 nm, D 1.0/0.5/0.25)
    in a 2000 x 2000 x 400 nm periodic slab; harmonic repulsion k=10; dt=0.05 ns;
    a composite external potential of 512 attractive z-normal pieces (disks and
    cylinders alternating, centres uniform in the slab, radii and heights
    uniform in [10, 80] nm, k=5) combined by nearest piece.
    Returns (spec, pieces, xyz, types)."""
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=tuple(float(b) for b in box), dt=0.05, seed=seed,
                      cell_size=cell_size)
    spec.types = [ParticleType("A", D=1.0, radius=1.0), ParticleType("B", D=0.5, radius=1.5),
                  ParticleType("C", D=0.25, radius=2.0)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0)]
    rng = np.random.default_rng(seed)
    pieces = composite_pieces(rng, n_pieces, spec.box_lo, spec.box_hi)
    xyz = uniform_positions(rng, n, spec.box_lo, spec.box_hi)
    types = (np.arange(n) % 3).astype(np.int32)
    return spec, pieces, xyz, types


LEARNED = dict(steps=500, train_samples=4_000_000)


def crowded_cytosol(n_replicas: int = 64, seed: int = 5, nmin: int = 20_000, nmax: int = 200_000, phi: float = 0.3,
                    n_pieces: int = 256, cell_size: float = 0.0):
    """64 periodic cubes, particle counts log-uniform in [20,000, 200,000],
    8 species (radii linspace(1, 5, 8) nm, D = 2 / radius nm^2/ns, assigned
    round-robin), volume fraction 0.3; harmonic repulsion k=10 plus two
    attractive piecewise wells (depth 1, width 1) for species pairs (0,1) and
    (2,3); A+B (species 0,1) reaction candidates at 3 nm; dt = 0.01 ns.
    Membrane: 256 attractive z-normal disks (k=5, radii uniform in [10, 80] nm)
    on the plane z = 0.4 x the smallest box edge, centres uniform over the
    largest box's x-y extent, combined by nearest piece; it acts on the two
    largest species (6, 7).
    Returns (spec, boxes [R,6], counts [R], xyz, types, pieces, membrane_types)."""
    radii = np.linspace(1.0, 5.0, 8)
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0), dt=0.01, seed=seed, cell_size=cell_size)
    spec.types = [ParticleType(f"S{i}", D=float(2.0 / radii[i]), radius=float(radii[i])) for i in range(8)]
    spec.pot2 = [
        OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0),
        OrderTwoPotential(PIECEWISE_WELL, ATTRACTIVE, k=0.0, length=1.0, depth=1.0, pairs=[(0, 1)]),
        OrderTwoPotential(PIECEWISE_WELL, ATTRACTIVE, k=0.0, length=1.0, depth=1.0, pairs=[(2, 3)]),
    ]
    R = np.zeros((8, 8))
    R[0, 1] = R[1, 0] = 3.0
    spec.reaction_radii = R
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_replicas))).astype(np.int64)
    vols = 4.0 / 3.0 * np.pi * radii ** 3
    boxes = np.zeros((n_replicas, 6))
    xyz, types = [], []
    for r, n in enumerate(counts):
        t = (np.arange(n) % 8).astype(np.int32)
        L = float(np.cbrt(vols[t].sum() / phi))
        boxes[r] = (0.0, L, 0.0, L, 0.0, L)
        xyz.append(uniform_positions(rng, int(n), (0.0, 0.0, 0.0), (L, L, L)))
        types.append(t)
    lmin, lmax = float(boxes[:, 1].min()), float(boxes[:, 1].max())
    zm = 0.4 * lmin
    pieces = []
    for i in range(n_pieces):
        c = (float(rng.uniform(0, lmax)), float(rng.uniform(0, lmax)), zm)
        pieces.append(OrderOnePotential(DISK, ATTRACTIVE, k=5.0, origin=c, normal=(0.0, 0.0, 1.0),
                                        radius=float(rng.uniform(10.0, 80.0))))
    return spec, boxes, counts, np.concatenate(xyz), np.concatenate(types), pieces, [6, 7]


CYTOSOL = dict(steps=1000, rdf_range=20.0, rdf_bins=100)


def bonded_chains(n_chains: int = 2000, length: int = 8, box: float = 120.0, seed: int = 6, k_bond: float = 50.0):
    """Linear chains of `length` beads (types alternate 0/1, radii 1.0/1.2 nm,
    D 1.0/0.8 nm^2/ns) joined by harmonic springs (k_bond, rest length = the
    pair's collision radius sum), with harmonic repulsion k=10 between all
    beads, in a periodic cube; dt = 0.01 ns.  Each chain starts as a random
    walk with steps uniform in [1.8, 2.8] nm (rest length 2.2 nm) from a
    uniform origin, wrapped into the box, so some bonds cross the periodic
    boundary.
    Returns (spec, xyz, types); the bonds are in spec.bonds."""
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(box, box, box), dt=0.01, seed=seed)
    spec.types = [ParticleType("A", D=1.0, radius=1.0), ParticleType("B", D=0.8, radius=1.2)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0)]
    rng = np.random.default_rng(seed)
    n = n_chains * length
    xyz = np.zeros((n, 3))
    for c in range(n_chains):
        p = rng.uniform(0.0, box, 3)
        for b in range(length):
            if b:
                v = rng.normal(size=3)
                p = p + rng.uniform(1.8, 2.8) * v / np.linalg.norm(v)
            xyz[c * length + b] = p
    xyz = np.mod(xyz, box).astype(np.float32)
    xyz = np.minimum(xyz, np.float32(box) - np.float32(1e-3))
    types = (np.arange(n) % 2).astype(np.int32)
    first = np.arange(n).reshape(n_chains, length)[:, :-1].reshape(-1)
    spec.bonds = np.stack([first, first + 1], 1).astype(np.int32)
    spec.bond_pot = np.zeros(len(first), np.int32)
    spec.bond_pots = [OrderTwoPotential(HARMONIC, SPRING, k=k_bond)]
    return spec, xyz, types
