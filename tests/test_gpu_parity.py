"""GPU parity of the plain BD step against the CPU oracle, through the C ABI.

Bit-exact: noise, cell ids, neighbour pairs, reaction candidates, RDF counts.
Tolerances (FP32 GPU vs FP64 oracle, same inputs and noise):
  gradients      |g - g_ref| <= 2e-5 * max|g_ref| + 1e-5
  one step       |x - x_ref| <= 1e-4 nm (minimum image); at 1M particles the
                 per-particle FP32 bound 4 ulp(|x|) + a (2^-20 |g| + 1e-3)
  20 steps       |x - x_ref| <= 2e-3 nm
  energy         relative 1e-5
"""
import numpy as np
import pytest

from oracle import oracle as orc
from readdygpu_amd import (ATTRACTIVE, BOX, CYLINDER, DISK, HARMONIC, PIECEWISE_WELL, REPULSIVE, SPHERE, SPRING,
                           WEAK_HARMONIC, OrderOnePotential, OrderTwoPotential, ParticleType, RdError, Simulation,
                           SystemSpec, workloads)

pytestmark = pytest.mark.gpu


def mi_err(a, b, spec):
    L = np.asarray(spec.box_hi, float) - np.asarray(spec.box_lo, float)
    d = np.abs(np.asarray(a, float) - np.asarray(b, float))
    return np.minimum(d, L - d).max() if d.size else 0.0


def rich_spec(seed=11, box=40.0, cell=0.0):
    """Every potential kind on one small system."""
    spec = SystemSpec(box_lo=(-20.0, 0.0, 5.0), box_hi=(-20.0 + box, box, 5.0 + box), dt=0.02, seed=seed,
                      cell_size=cell, max_cutoff=6.0)
    spec.types = [ParticleType("A", 1.0, 1.0), ParticleType("B", 0.5, 1.5, collision_radii=[1.2, 0.0, 0.0]),
                  ParticleType("C", 2.0, 0.8)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, 10.0),
                 OrderTwoPotential(PIECEWISE_WELL, ATTRACTIVE, 2.0, length=1.0, depth=1.0, pairs=[(0, 1)]),
                 OrderTwoPotential(WEAK_HARMONIC, ATTRACTIVE, 0.5, length=0.8, pairs=[(2, 2)]),
                 OrderTwoPotential(HARMONIC, ATTRACTIVE, 0.2, pairs=[(1, 2)]),
                 OrderTwoPotential(HARMONIC, SPRING, 0.1, pairs=[(0, 0)])]
    c = (0.0, 20.0, 25.0)
    spec.pot1 = [OrderOnePotential(SPHERE, ATTRACTIVE, 1.0, origin=c, radius=15.0),
                 OrderOnePotential(DISK, ATTRACTIVE, 0.3, origin=c, normal=(0.3, 0.2, 1.0), radius=8.0, types=[0]),
                 OrderOnePotential(CYLINDER, ATTRACTIVE, 0.5, origin=c, normal=(0, 1, 0), radius=10.0, height=12.0,
                                   types=[1]),
                 OrderOnePotential(CYLINDER, REPULSIVE, 0.5, origin=c, normal=(0, 0, 1), radius=3.0, height=6.0,
                                   types=[2]),
                 OrderOnePotential(BOX, ATTRACTIVE, 2.0, origin=(-18, 2, 7), extension=(36, 36, 36), types=[0, 2]),
                 OrderOnePotential(SPHERE, REPULSIVE, 0.7, origin=c, radius=2.0)]
    R = np.zeros((3, 3))
    R[0, 1] = R[1, 0] = 3.0
    spec.reaction_radii = R
    rng = np.random.default_rng(seed)
    xyz = workloads.uniform_positions(rng, 3000, spec.box_lo, spec.box_hi)
    types = rng.integers(0, 3, 3000).astype(np.int32)
    return spec, xyz, types


def dense_spec(n=6000, box=22.0, seed=9):
    """About 110 stencil candidates per particle: the flat kernel's candidate
    list spills over several LDS passes per wave."""
    spec = SystemSpec(box_lo=(0, 0, 0), box_hi=(box, box, box), dt=0.002, seed=seed)
    spec.types = [ParticleType("A", 1.0, 1.0), ParticleType("B", 0.5, 1.2)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, 10.0),
                 OrderTwoPotential(PIECEWISE_WELL, ATTRACTIVE, 1.0, length=0.5, depth=0.5, pairs=[(0, 1)])]
    rng = np.random.default_rng(seed)
    xyz = workloads.uniform_positions(rng, n, spec.box_lo, spec.box_hi)
    types = rng.integers(0, 2, n).astype(np.int32)
    return spec, xyz, types


@pytest.fixture(scope="module")
def c1():
    return workloads.plain_10k()


def test_noise_bitexact(c1):
    spec, xyz, types = c1
    with Simulation(spec, xyz, types) as sim:
        z = sim.noise(5, 3)
    assert np.array_equal(z, orc.noise(spec.seed, spec.stream, 5, 3, len(types)))


def test_cell_ids_bitexact(c1):
    spec, xyz, types = c1
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types) as sim:
        assert list(sim.info().lattice) == list(o.lattice())
        assert np.array_equal(sim.cell_ids(), o.cell_index(xyz)[0])
        sim.simulate(10)
        x, t = sim.copy_pos_from_device()
        assert np.array_equal(sim.cell_ids(), o.cell_index(x)[0])


@pytest.mark.parametrize("which", ["plain_10k", "rich", "dense", "denser"])
def test_gradients_and_energy_match_fp64(which, c1):
    """dense: 36 partners within range per particle on average (at most 68),
    denser: 84 (at most 127), so the 16-entry hit list of the force kernels
    fills and is evaluated over up to 5 and 8 passes."""
    if which == "plain_10k":
        spec, xyz, types = c1
    elif which == "rich":
        spec, xyz, types = rich_spec()
    else:
        spec, xyz, types = dense_spec() if which == "dense" else dense_spec(n=14000)
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types) as sim:
        g, e = sim.gradients()
    gr, er = o.gradients(xyz.astype(np.float64), types, brute=True)
    tol = 2e-5 * np.abs(gr).max() + 1e-5
    assert np.abs(g - gr).max() <= tol, np.abs(g - gr).max()
    assert abs(e - er) <= 1e-5 * max(1.0, abs(er))


@pytest.mark.parametrize("which", ["plain_10k", "rich"])
def test_steps_match_fp64_resynced(which, c1):
    """Each GPU step against one FP64 oracle step from the GPU's previous state."""
    spec, xyz, types = c1 if which == "plain_10k" else rich_spec()
    o = orc.OracleSystem(spec)
    worst = 0.0
    with Simulation(spec, xyz, types) as sim:
        x = xyz
        for step in range(20):
            sim.simulate(1)
            xg, _ = sim.copy_pos_from_device()
            xr, bad = o.run(x, types, step, 1)
            assert bad == 0
            worst = max(worst, mi_err(xg, xr, spec))
            x = xg
    print(f"{which}: worst one-step error over 20 steps {worst:.3e} nm")
    assert worst <= 1e-4


def test_free_run_matches_fp64_dilute(c1):
    """Dilute repulsive fluid: the FP32 trajectory stays within 2e-3 nm of the
    FP64 oracle over 20 steps (stiff systems amplify rounding by
    |1 - a*lambda_max(H)| per step and are checked step by step above)."""
    spec, xyz, types = c1
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(20)
        x20, _ = sim.copy_pos_from_device()
    xr20, _ = o.run(xyz, types, 0, 20)
    err = mi_err(x20, xr20, spec)
    print(f"20-step free-run max error {err:.3e} nm")
    assert err <= 2e-3


def test_step_is_deterministic_and_noise_array_equals_generator(c1):
    spec, xyz, types = c1
    with Simulation(spec, xyz, types) as a:
        a.simulate(7)
        xa, _ = a.copy_pos_from_device()
    with Simulation(spec, xyz, types) as b:
        b.simulate(7)
        xb, _ = b.copy_pos_from_device()
    assert np.array_equal(xa, xb)
    with Simulation(spec, xyz, types) as c:
        c.set_noise_array(orc.noise(spec.seed, spec.stream, 0, 7, len(types)))
        c.simulate(7)
        xc, _ = c.copy_pos_from_device()
        with pytest.raises(RdError):
            c.simulate(1)            # the array holds 7 steps only
    assert np.array_equal(xa, xc)


@pytest.mark.parametrize("which", ["plain_10k", "rich", "plain_1m_slice", "dense"])
def test_launch_configs_give_identical_trajectories(which, c1):
    """Three cell sizes (different sort orders inside cells), workgroup sizes,
    particles per lane, the work-queue form and the autotuner: bitwise the
    same positions and gradients (fixed-point pair sums)."""
    if which == "plain_10k":
        spec, xyz, types = c1
    elif which == "rich":
        spec, xyz, types = rich_spec()
    elif which == "dense":
        spec, xyz, types = dense_spec()
    else:
        spec, xyz, types = workloads.plain_1m(n=200_000, box=585.0)
    ref = None
    rc = float(spec.pair_range().max())
    shapes = [dict(kernel=1), dict(kernel=0, block=64), dict(kernel=1, block=512), dict(kernel=1, block=128),
              dict(kernel=1, block=128, ppt=2), dict(kernel=1, ppt=4), dict(kernel=0, block=64, ppt=2),
              dict(kernel=2), dict(kernel=2, block=64), dict(kernel=2, block=512)]
    for cell in (0.0, 1.3 * rc, 2.1 * rc):
        for shape in shapes:
            spec.cell_size = cell
            with Simulation(spec, xyz, types) as sim:
                sim.set_launch_config(**shape)
                sim.simulate(6)
                x, _ = sim.copy_pos_from_device()
                g, _ = sim.gradients()
                used = sim.launch_config()
            assert used["used_kernel"] == (shape.get("kernel") or 1), used
            assert used["used_block"] == (shape.get("block") or 256), used
            assert used["used_ppt"] == (shape.get("ppt") or 1), used
            if ref is None:
                ref = (x, g)
            assert np.array_equal(x, ref[0]), (cell, shape)
            assert np.array_equal(g, ref[1]), (cell, shape)
    spec.cell_size = 0.0
    with Simulation(spec, xyz, types) as sim:
        sim.autotune(trial_steps=3)
        assert sim.info().step == 0
        sim.simulate(6)
        x, _ = sim.copy_pos_from_device()
    assert np.array_equal(x, ref[0])


@pytest.mark.parametrize("which", ["plain_10k", "rich"])
def test_pairs_and_candidates_bitexact(which, c1):
    spec, xyz, types = c1 if which == "plain_10k" else rich_spec()
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(5)
        x, t = sim.copy_pos_from_device()
        p = sim.neighbor_pairs()
        c = sim.reaction_candidates()
    assert np.array_equal(p, o.pairs(x, t, brute=True))
    assert np.array_equal(c, o.candidates(x, t, brute=True))
    assert len(p) > 0


def test_rdf_counts_bitexact_and_normalised(c1):
    spec, xyz, types = c1
    o = orc.OracleSystem(spec)
    ref = np.zeros((2, 2, 200), dtype=np.uint64)
    with Simulation(spec, xyz, types, rdf=(20.0, 200)) as sim:
        for f in range(3):
            sim.simulate(10)
            x, t = sim.copy_pos_from_device()
            sim.call_rdf_calculation()
            ref += o.rdf_counts(x, t, 20.0, 200, cells=False)
        c, frames = sim.copy_rdf_matrix()
        g = sim.normalize_rdf()
    assert frames == 3
    assert np.array_equal(c, ref)
    gr = orc.rdf_normalize(2, 200, 20.0, spec.volume(), [5000, 5000], 3, ref)
    assert np.allclose(g, gr, rtol=1e-12)
    assert abs(g[:, :, 100:].mean() - 1.0) < 0.05     # dilute fluid: g -> 1 at range


@pytest.mark.parametrize("cell", [7.0, 6.5])
def test_rdf_counts_with_a_partial_top_cell(cell):
    """Lattice edges that leave a partial slab at the top of the box (it wraps
    into cell 0): particles crowded into that slab and into the cells whose
    stencils reach cell 0 or stop just short of it give the oracle's counts
    bit for bit, also where the kernel skips the minimum image."""
    box = 100.0
    spec = SystemSpec(box_lo=(0, 0, 0), box_hi=(box, box, box), dt=0.1, seed=5, cell_size=cell)
    spec.types = [ParticleType("A", 1.0, 1.5), ParticleType("B", 0.5, 2.0)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, 10.0)]
    rng = np.random.default_rng(21)
    L = int(box / cell)
    parts = [rng.random((3000, 3)) * box]
    for lo_, hi_ in ((L * cell, box), (2 * cell, 5 * cell), (0.0, cell)):
        q = rng.random((600, 3)) * box
        q[:, rng.integers(0, 3)] = lo_ + rng.random(600) * (hi_ - lo_)
        parts.append(q)
    xyz = np.clip(np.concatenate(parts), 0.0, np.nextafter(box, 0.0)).astype(np.float32)
    types = rng.integers(0, 2, len(xyz)).astype(np.int32)
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types, rdf=(20.0, 100)) as sim:
        assert sim.info().lattice[0] == L
        sim.call_rdf_calculation()
        c, _ = sim.copy_rdf_matrix()
    assert np.array_equal(c, o.rdf_counts(xyz, types, 20.0, 100, cells=False))


@pytest.mark.parametrize("nt,nbins", [(12, 200), (16, 170)])
def test_rdf_counts_with_large_histograms(nt, nbins):
    """Histograms of 115-170 KB: the LDS path with a raised LDS limit, or the
    direct global-atomic path past 160 KB; counts bit for bit the oracle's."""
    box = 40.0
    spec = SystemSpec(box_lo=(0, 0, 0), box_hi=(box, box, box), dt=0.1, seed=8)
    spec.types = [ParticleType(f"T{k}", 1.0, 0.5 + 0.05 * k) for k in range(nt)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, 10.0)]
    rng = np.random.default_rng(nt)
    xyz = (rng.random((2500, 3)) * box).astype(np.float32)
    types = rng.integers(0, nt, len(xyz)).astype(np.int32)
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types, rdf=(10.0, nbins)) as sim:
        sim.simulate(3)
        x, t = sim.copy_pos_from_device()
        sim.call_rdf_calculation()
        c, _ = sim.copy_rdf_matrix()
    assert c.shape == (nt, nt, nbins)
    assert np.array_equal(c, o.rdf_counts(x, t, 10.0, nbins, cells=False))


@pytest.mark.parametrize("box,cell,n", [(12.0, 4.2, 400), (9.0, 4.0, 200), (30.0, 9.5, 2000)])
def test_small_lattices_and_wide_cells(box, cell, n):
    """L < 3 (each cell once) and cells wider than the range."""
    spec = SystemSpec(box_lo=(0, 0, 0), box_hi=(box, box, box), dt=0.01, seed=3, cell_size=cell)
    spec.types = [ParticleType("A", 1.0, 1.0), ParticleType("B", 1.0, 0.9)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, 10.0)]
    R = np.array([[0.0, 3.0], [3.0, 0.0]])
    spec.reaction_radii = R
    rng = np.random.default_rng(n)
    xyz = workloads.uniform_positions(rng, n, spec.box_lo, spec.box_hi)
    types = rng.integers(0, 2, n).astype(np.int32)
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types, rdf=(box / 2 - 0.01, 30)) as sim:
        g, _ = sim.gradients()
        sim.simulate(3)
        x, t = sim.copy_pos_from_device()
        assert np.array_equal(sim.neighbor_pairs(), o.pairs(x, t, brute=True))
        assert np.array_equal(sim.reaction_candidates(), o.candidates(x, t, brute=True))
        sim.call_rdf_calculation()
        c, _ = sim.copy_rdf_matrix()
    assert np.array_equal(c, o.rdf_counts(x, t, np.float32(box / 2 - 0.01), 30, cells=False))
    gr, _ = o.gradients(xyz.astype(np.float64), types, brute=True)
    assert np.abs(g - gr).max() <= 2e-5 * np.abs(gr).max() + 1e-5


def test_edge_cases_empty_single_boundary():
    spec = SystemSpec(box_lo=(0, 0, 0), box_hi=(50, 50, 50), dt=0.1, seed=9)
    spec.types = [ParticleType("A", 1.0, 1.0)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, 10.0)]
    # empty system
    with Simulation(spec, np.zeros((0, 3), np.float32), np.zeros(0, np.int32)) as sim:
        sim.simulate(3)
        assert sim.info().n == 0 and len(sim.neighbor_pairs()) == 0
    # particles on the box faces (x == lo and x == hi) and in one cell
    xyz = np.array([[0, 0, 0], [50, 50, 50], [0.5, 0.5, 0.5], [49.9, 0.2, 25.0], [25, 25, 25], [25.5, 25, 25]],
                   dtype=np.float32)
    types = np.zeros(6, np.int32)
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types) as sim:
        assert np.array_equal(sim.cell_ids(), o.cell_index(xyz)[0])
        assert np.array_equal(sim.neighbor_pairs(), o.pairs(xyz, types, brute=True))
        g, _ = sim.gradients()
        gr, _ = o.gradients(xyz.astype(np.float64), types, brute=True)
        assert np.abs(g - gr).max() < 1e-4
        sim.simulate(2)
        x, _ = sim.copy_pos_from_device()
    xr, _ = o.run(xyz, types, 0, 2)
    assert mi_err(x, xr, spec) < 1e-4
    # out-of-box input is rejected as createNeighborList does (CudaSimulation.cu:590-593)
    with pytest.raises(RdError):
        Simulation(spec, np.array([[51.0, 1, 1]], np.float32), np.zeros(1, np.int32))


def test_reactions_protocol():
    spec, xyz, types = workloads.plain_10k(n=1000)
    with Simulation(spec, xyz, types) as sim:
        sim.simulate(2)
        x0, t0 = sim.copy_pos_from_device()
        # delete particles 3 and 10, create one B at (1,2,3), change type of 7 (delete + create)
        cr = np.array([3,
                       100, -1, 0, 0, 0, 3,
                       101, 1, 1.0, 2.0, 3.0, -1,
                       102, -1, 0, 0, 0, 10], dtype=np.float32)
        sim.apply_reactions(cr)
        x1, t1 = sim.copy_pos_from_device()
        keep = np.ones(1000, bool)
        keep[[3, 10]] = False
        assert len(t1) == 999
        assert np.array_equal(x1[:998], x0[keep]) and np.array_equal(t1[:998], t0[keep])
        assert np.allclose(x1[998], [1, 2, 3]) and t1[998] == 1
        sim.simulate(2)
        with pytest.raises(RdError):
            sim.apply_reactions(np.array([1, 0, -1, 0, 0, 0, 5000], np.float32))


def test_full_size_1m_properties():
    """1M particles: bit-exact sets against the oracle's cell-list path, and
    size-independent properties of the step."""
    spec, xyz, types = workloads.plain_1m()
    o = orc.OracleSystem(spec)
    with Simulation(spec, xyz, types, rdf=(20.0, 200)) as sim:
        assert np.array_equal(sim.cell_ids(), o.cell_index(xyz)[0])
        sim.simulate(20)
        x, t = sim.copy_pos_from_device()
        assert np.array_equal(np.bincount(t), np.bincount(types))
        assert np.array_equal(sim.cell_ids(), o.cell_index(x)[0])
        p = sim.neighbor_pairs()
        assert np.array_equal(p, o.pairs(x, t))
        c = sim.reaction_candidates()
        assert np.array_equal(c, o.candidates(x, t))
        sim.call_rdf_calculation()
        cnt, _ = sim.copy_rdf_matrix()
        assert np.array_equal(cnt, o.rdf_counts(x, t, 20.0, 200, cells=True))
        # one step from this state against the FP64 oracle (cell-list path)
        sim.simulate(1)
        x1, _ = sim.copy_pos_from_device()
    xr1, bad = o.run(x, t, 20, 1, nthreads=0)
    assert bad == 0
    # FP32 bound of one update x' = (x - a g) + b z against FP64 with the same
    # x and z: four roundings of |x'| (product, difference, noise product,
    # sum: 4 ulp) plus a times the FP32 error of the gradient (2^-20 |g|
    # relative, and 1e-3 absolute for the cancellation in k (r - r0) and
    # k (d - R) near the contact and the confinement surface)
    g, _ = o.gradients(x.astype(np.float64), t, nthreads=0)
    a = np.array([spec.dt * ty.D / spec.kbt for ty in spec.types])[t]
    L = np.asarray(spec.box_hi, float) - np.asarray(spec.box_lo, float)
    d = np.abs(x1.astype(np.float64) - xr1)
    d = np.minimum(d, L - d).max(axis=1)
    ulp = np.spacing(np.maximum(np.abs(x1), np.abs(xr1).astype(np.float32)).max(axis=1).astype(np.float32))
    tol = 4.0 * ulp.astype(np.float64) + a * (2.0 ** -20 * np.abs(g).max(axis=1) + 1e-3)
    worst = int(np.argmax(d / tol))
    assert np.all(d <= tol), (worst, d[worst], tol[worst], x1[worst], xr1[worst])
