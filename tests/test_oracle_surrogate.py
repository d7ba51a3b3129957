"""The oracle of the composite potential and of the MLP surrogate, pinned by
independent means: geometry (distance to a disk / solid cylinder by brute
force sampling), finite differences, and torch's FP64 autograd and BF16
rounding."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from readdygpu_amd import CYLINDER, DISK, workloads
from readdygpu_amd.model import ATTRACTIVE, OrderOnePotential


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e4, 1000), rng.normal(0, 1e-6, 1000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, -1.00390625, 65504.0, 1e38]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(orc.bf16(v), ref)
    assert all(orc.lib().orc_bf16(float(x)) == float(r) for x, r in zip(v[:200], ref[:200]))


def torch_mlp(w, xn, bf16=False):
    """The same network in torch FP64 (BF16 rounding at the kernels' points)."""
    w = torch.as_tensor(w, dtype=torch.float64)
    x = torch.as_tensor(np.asarray(xn, np.float32), dtype=torch.float64)

    def r(t):
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64) if bf16 else t

    W1 = r(w[0:192].view(64, 3)); b1 = w[192:256]
    W2 = r(w[256:4352].view(64, 64)); b2 = w[4352:4416]
    W3 = r(w[4416:8512].view(64, 64)); b3 = w[8512:8576]
    W4 = r(w[8576:8768].view(3, 64)); b4 = w[8768:8771]
    silu = torch.nn.functional.silu
    h1 = r(silu(x @ W1.T + b1))
    h2 = r(silu(h1 @ W2.T + b2))
    h3 = r(silu(h2 @ W3.T + b3))
    return h3 @ W4.T + b4


def rand_weights(rng):
    w = np.zeros(orc.MLP_NPARAM)
    w[0:192] = rng.uniform(-1.4, 1.4, 192)
    w[192:256] = rng.normal(0, 0.1, 64)
    w[256:4352] = rng.uniform(-0.3, 0.3, 4096)
    w[4352:4416] = rng.normal(0, 0.1, 64)
    w[4416:8512] = rng.uniform(-0.3, 0.3, 4096)
    w[8512:8576] = rng.normal(0, 0.1, 64)
    w[8576:8768] = rng.uniform(-0.3, 0.3, 192)
    w[8768:8771] = rng.normal(0, 0.1, 3)
    return w.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("bf16", [False, True])
def test_mlp_forward_matches_torch(bf16):
    rng = np.random.default_rng(1)
    w = rand_weights(rng)
    xn = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    y = orc.mlp_forward(w, xn, bf16)
    yt = torch_mlp(w, xn, bf16).numpy()
    # bf16 mode: an activation that lands on a rounding tie may round the
    # other way after FP64-vs-torch differences in the last bits
    tol = 1e-12 if not bf16 else 1e-2 * np.abs(yt).max()
    assert np.abs(y - yt).max() <= tol


def test_mlp_inference_form_is_the_network():
    """The inference kernels' form (layer 1 and the biases times -log2 e, W4
    times -ln 2, activations zs / (1 + 2^zs)) against torch's network: equal
    to rounding without BF16, and within BF16 noise of the FP64 network with
    the kernels' BF16 and FP32 roundings, as close as the plain BF16 form."""
    rng = np.random.default_rng(11)
    w = rand_weights(rng)
    xn = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    yt = torch_mlp(w, xn).numpy()
    yi = orc.mlp_forward(w, xn, False, inference=True)
    assert np.abs(yi - yt).max() <= 1e-12 * np.abs(yt).max()
    yi16 = orc.mlp_forward(w, xn, True, inference=True)
    y16 = orc.mlp_forward(w, xn, True)
    e_inf = np.linalg.norm(yi16 - yt) / np.linalg.norm(yt)
    e_bf = np.linalg.norm(y16 - yt) / np.linalg.norm(yt)
    assert 1e-5 < e_inf <= 1e-2 and e_inf <= 3 * e_bf, (e_inf, e_bf)


def test_mlp_gradient_matches_autograd_and_finite_differences():
    rng = np.random.default_rng(2)
    w = rand_weights(rng)
    xn = rng.uniform(-1, 1, (97, 3)).astype(np.float32)
    t = rng.normal(0, 1, (97, 3)).astype(np.float32)
    loss, g = orc.mlp_loss_grad(w, xn, t, False)
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    yt = torch_mlp(wt, xn)
    lt = ((yt - torch.as_tensor(t, dtype=torch.float64)) ** 2).sum() / len(xn)
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-12 * abs(loss)
    assert np.abs(g - wt.grad.numpy()).max() <= 1e-10 * np.abs(g).max()
    # central differences on a sample of parameters from every block
    idx = np.concatenate([rng.choice(np.arange(a, b), min(4, b - a), replace=False) for a, b in
                          [(0, 192), (192, 256), (256, 4352), (4352, 4416), (4416, 8512), (8512, 8576),
                           (8576, 8768), (8768, 8771)]])
    for i in idx:
        h = 1e-6
        wp = w.copy(); wp[i] += h
        wm = w.copy(); wm[i] -= h
        fd = (orc.mlp_loss_grad(wp, xn, t, False)[0] - orc.mlp_loss_grad(wm, xn, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_mlp_bf16_gradient_is_close_to_fp64():
    rng = np.random.default_rng(3)
    w = rand_weights(rng)
    xn = rng.uniform(-1, 1, (256, 3)).astype(np.float32)
    t = rng.normal(0, 1, (256, 3)).astype(np.float32)
    _, g64 = orc.mlp_loss_grad(w, xn, t, False)
    _, g16 = orc.mlp_loss_grad(w, xn, t, True)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert 1e-5 < rel < 3e-2       # BF16 rounding is there and stays small


def dist_to_piece_bruteforce(p, x, n=400):
    """Distance from x to a z-normal disk (filled circle) or solid cylinder,
    by dense sampling of the piece's boundary (points inside give 0)."""
    o = np.asarray(p.origin, float)
    v = np.asarray(x, float) - o
    r = np.hypot(v[0], v[1])
    hh = 0.5 * p.height if p.kind == CYLINDER else 0.0
    if p.kind == CYLINDER and r <= p.radius and abs(v[2]) <= hh:
        return 0.0
    th = np.linspace(0, 2 * np.pi, n, endpoint=False)
    rr = np.linspace(0, p.radius, n // 4)
    best = np.inf
    zs = [0.0] if p.kind == DISK else [-hh, hh]
    for z in zs:                                   # disk / caps (filled circles)
        R, T = np.meshgrid(rr, th)
        pts = np.stack([R * np.cos(T), R * np.sin(T), np.full_like(R, z)], -1).reshape(-1, 3)
        best = min(best, np.sqrt(((pts - v) ** 2).sum(1)).min())
    if p.kind == CYLINDER:                         # side wall
        Z, T = np.meshgrid(np.linspace(-hh, hh, n // 4), th)
        pts = np.stack([p.radius * np.cos(T), p.radius * np.sin(T), Z], -1).reshape(-1, 3)
        best = min(best, np.sqrt(((pts - v) ** 2).sum(1)).min())
    return best


def test_piece_energy_is_harmonic_in_the_distance_to_the_piece():
    rng = np.random.default_rng(4)
    pieces = [OrderOnePotential(DISK, ATTRACTIVE, 5.0, origin=(1.0, 2.0, 3.0), radius=7.0),
              OrderOnePotential(CYLINDER, ATTRACTIVE, 3.0, origin=(-2.0, 0.5, 1.0), radius=4.0, height=6.0)]
    for p in pieces:
        for _ in range(60):
            x = np.asarray(p.origin) + rng.uniform(-15, 15, 3)
            e, _ = orc.piece_eval(p, x)
            d = dist_to_piece_bruteforce(p, x)
            # sampling resolution ~ 2*pi*R/400 and R/100 on a surface
            assert abs(np.sqrt(2 * e / p.k) - d) <= 0.12, (p.kind, x)


def test_composite_gradient_is_the_derivative_of_the_nearest_piece_energy():
    rng = np.random.default_rng(5)
    pieces = workloads.composite_pieces(rng, 24, (0, 0, 0), (300, 300, 100))
    x = rng.uniform((0, 0, 0), (300, 300, 100), (400, 3))
    g, e, which, gap = orc.composite(pieces, x)
    assert which.min() >= 0
    h = 1e-5
    ok = 0
    for i in range(len(x)):
        if gap[i] < 1e-2:
            continue                      # near a switch between pieces
        fd = np.zeros(3)
        for d in range(3):
            xp = x[i].copy(); xp[d] += h
            xm = x[i].copy(); xm[d] -= h
            fd[d] = (orc.composite(pieces, xp[None])[1][0] - orc.composite(pieces, xm[None])[1][0]) / (2 * h)
        assert np.abs(fd - g[i]).max() <= 1e-4 * max(1.0, np.abs(g[i]).max()), i
        ok += 1
    assert ok > 300
    # the minimum is a minimum: no piece has a lower energy
    for i in range(0, len(x), 37):
        es = [orc.piece_eval(p, x[i])[0] for p in pieces]
        assert e[i] == min(es) and which[i] == int(np.argmin(es))
