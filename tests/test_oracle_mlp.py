"""Pins of the CPU oracle of the learned field surrogate (oracle/mlpref.c).

The surrogate has no counterpart in the reference, so nothing of the
reference binds it.  The oracle is checked against an independent library
evaluation instead: PyTorch (float64) modules and autograd for the forward
pass and the gradients, PyTorch's bfloat16 conversion for the rounding
points of the GPU-emulating mode, finite differences for the gradient, and
the closed-form properties of the features (periodicity) and of gelu.
"""
import numpy as np
import pytest
import torch

from oracle import tpref
from test_particle_amd import surrogate

L_BOX = 129.73


# the GPU kernels' activation, restated with numpy's IEEE half type: the
# FP32 input converted toward zero, every FMA (exact in FP64 for half
# operands) and the final difference rounded to nearest even in FP16
GELU_TAIL = [0.045379638671875, -0.1727294921875, 0.21728515625, 0.0158538818359375,
             -0.3017578125, 0.2413330078125, 0.018310546875, -0.08465576171875,
             0.020843505859375]


def _gelu_fp16(z):
    zf = np.asarray(z, np.float64).astype(np.float32)
    with np.errstate(over="ignore"):
        h = zf.astype(np.float16)
    away = np.abs(h.astype(np.float64)) > np.abs(zf.astype(np.float64))
    h = np.where(away, np.nextafter(h, np.float16(0)), h)
    s = (np.minimum(np.abs(h), np.float16(4)).astype(np.float64) * 0.5 - 1.0).astype(np.float16)
    p = np.full(h.shape, np.float16(GELU_TAIL[8]))
    for c in reversed(GELU_TAIL[:8]):
        p = (p.astype(np.float64) * s.astype(np.float64) + c).astype(np.float16)
    g = np.maximum(h, np.float16(0)).astype(np.float64) - p.astype(np.float64)
    return g.astype(np.float16).astype(np.float64)


def _torch_net(params, q, emulate=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = [torch.tensor(a, dtype=torch.float64)
                                      for a in surrogate.split_params(params)]
    qt = torch.tensor(np.asarray(q, dtype=np.float64))

    def bf(t):
        return t.to(torch.bfloat16).to(torch.float64) if emulate else t

    def f32(t):
        return t.to(torch.float32).to(torch.float64) if emulate else t

    def gelu(z):
        if emulate:
            return torch.tensor(_gelu_fp16(z.detach().numpy()).reshape(z.shape))
        return torch.nn.functional.gelu(z, approximate="tanh")

    def net(x):
        if emulate:
            xf = x.to(torch.float32)
            qf = qt.to(torch.float32)
            u = torch.addcmul(torch.addcmul(xf[:, :1] * qf[:, 0], xf[:, 1:2], qf[:, 1]),
                              xf[:, 2:3], qf[:, 2])
            r = (u - torch.round(u)).to(torch.float64)
        else:
            u = x @ qt.T
            r = u - torch.round(u)
        f = torch.cat([bf(torch.sin(2 * np.pi * r)), bf(torch.cos(2 * np.pi * r))], dim=1)

        def dense(h, W, b):
            # bf16 operands, FP64 sum, then the FP32 bias add of the kernels
            return f32(f32(h @ bf(W)) + f32(b)) if emulate else h @ W + b
        h1 = bf(gelu(dense(f, W1, b1)))
        h2 = bf(gelu(dense(h1, W2, b2)))
        h3 = f32(gelu(dense(h2, W3, b3)))
        return h3 @ W4 + b4
    return net


def _setup(n=16, seed=3):
    rng = np.random.default_rng(seed)
    params = surrogate.init_params(seed).astype(np.float64)
    q = surrogate.fourier_frequencies(L_BOX, sigma=3.0, seed=seed).astype(np.float64)
    x = rng.uniform(0.0, L_BOX, (n, 3))
    y = rng.normal(0.0, 1.0, (n, 3))
    return params, q, x, y


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-20, 200),
                        rng.normal(0, 1e20, 200), [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0]])
    v = v.astype(np.float32).astype(np.float64)
    ref = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(tpref.bf16_round(v), ref)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert tpref.bf16_round([1.0 + 2 ** -8])[0] == 1.0
    assert tpref.bf16_round([1.0 + 3 * 2 ** -8])[0] == 1.0 + 2 ** -6


def test_gelu_is_the_tanh_form_and_close_to_erf_gelu():
    from scipy.special import erf
    z = np.linspace(-12.0, 12.0, 4001)
    g, gd = tpref.gelu(z)
    tanh_form = 0.5 * z * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (z + 0.044715 * z ** 3)))
    assert np.allclose(g, tanh_form, rtol=1e-13, atol=1e-15)
    exact = 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    assert np.max(np.abs(g - exact)) < 1e-3
    h = 1e-6
    fd = (tpref.gelu(z + h)[0] - tpref.gelu(z - h)[0]) / (2 * h)
    assert np.allclose(gd, fd, rtol=1e-7, atol=1e-8)
    assert np.all(np.isfinite(tpref.gelu(np.array([-1e4, 1e4]))[0]))


def test_gpu_activation_is_the_fp16_tail_polynomial():
    """The kernels' gelu (oracle emulation with frexp-based FP16 rounding)
    equals numpy's half-precision restatement bit for bit, special values
    included, and stays within 2^-10 |z| + 1.1e-3 of the tanh form for
    |z| < 65504 (6.3e-4 for z <= 0); beyond, the FP16 input saturates at
    65504."""
    rng = np.random.default_rng(1)
    z = np.concatenate([np.linspace(-10.0, 10.0, 200001), rng.normal(0.0, 3.0, 100000),
                        rng.normal(0.0, 1e-4, 10000), np.linspace(-3000.0, 3000.0, 20001),
                        [0.0, -0.0, 4.0, -4.0, 65504.0, 65519.0, 65520.0, 1e5, -1e5, 6e-8,
                         -6e-8, 1e-9, 3e-5]])
    g = tpref.gelu_gpu(z)
    assert np.array_equal(g, _gelu_fp16(z))
    exact = tpref.gelu(z)[0]
    inside = np.abs(z) < 65504.0
    err = np.abs(g - exact)
    assert np.all(err[inside] <= 2.0 ** -10 * np.abs(z[inside]) + 1.1e-3)
    assert np.all(err[z <= 0] <= 6.3e-4)
    assert tpref.gelu_gpu([1e5])[0] == 65504.0 and abs(tpref.gelu_gpu([-1e5])[0]) <= 6.3e-4
    # a dropped tail term or a wrong coefficient shows at the 1e-2 level
    assert np.max(err[np.abs(z) < 4]) > 1e-4           # it is an approximation
    assert np.max(np.abs(g - np.maximum(z, 0))[np.abs(z) < 2]) > 0.1


def test_forward_matches_torch_fp64():
    params, q, x, _ = _setup()
    out = tpref.mlp_forward(params, q, x)
    ref = _torch_net(params, q)(torch.tensor(x)).detach().numpy()
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-13)


def test_forward_emulation_matches_torch_bf16_points():
    params, q, x, _ = _setup(n=64)
    out = tpref.mlp_forward(params, q, x, emulate=True)
    ref = _torch_net(params, q, emulate=True)(torch.tensor(x)).detach().numpy()
    # identical rounding points; only the order of the FP64 sums differs, so
    # a bf16 rounding can flip at a tie: compare to a tiny fraction of scale
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(out - ref)) < 1e-4 * scale
    # and the emulated network is close to the FP64 one (bf16 level)
    full = tpref.mlp_forward(params, q, x)
    assert np.max(np.abs(out - full)) < 3e-2 * np.max(np.abs(full))


def test_features_are_periodic():
    params, q, x, _ = _setup(n=8)
    q = np.rint(q * L_BOX) / L_BOX           # exact lattice vectors in FP64
    base = tpref.mlp_forward(params, q, x)
    for a in range(3):
        xs = x.copy()
        xs[:, a] += L_BOX
        # q = n / L with L_BOX: q . (x + L e_a) = q . x + n_a (integer)
        assert np.allclose(tpref.mlp_forward(params, q, xs), base, rtol=0, atol=1e-10)


def test_gradient_matches_torch_autograd():
    params, q, x, y = _setup(n=12)
    loss, g = tpref.mlp_loss_grad(params, q, x, y)
    pt = torch.tensor(params, requires_grad=True)
    parts = []
    o = 0
    for fin, fout in surrogate.LAYERS:
        parts.append(pt[o:o + fin * fout].reshape(fin, fout))
        o += fin * fout
        parts.append(pt[o:o + fout])
        o += fout
    W1, b1, W2, b2, W3, b3, W4, b4 = parts
    qt = torch.tensor(q)
    xt = torch.tensor(x)
    u = xt @ qt.T
    r = u - torch.round(u)
    f = torch.cat([torch.sin(2 * np.pi * r), torch.cos(2 * np.pi * r)], dim=1)
    act = lambda z: torch.nn.functional.gelu(z, approximate="tanh")  # noqa: E731
    out = act(act(act(f @ W1 + b1) @ W2 + b2) @ W3 + b3) @ W4 + b4
    tl = torch.mean((out - torch.tensor(y)) ** 2)
    tl.backward()
    assert abs(loss - tl.item()) <= 1e-13 * abs(tl.item())
    gr = pt.grad.numpy()
    assert np.allclose(g, gr, rtol=1e-10, atol=1e-14 * np.max(np.abs(gr)))


def test_gradient_finite_differences():
    params, q, x, y = _setup(n=6, seed=5)
    _, g = tpref.mlp_loss_grad(params, q, x, y)
    rng = np.random.default_rng(7)
    # indices in every parameter block
    bounds, o = [], 0
    for fin, fout in surrogate.LAYERS:
        bounds.append((o, o + fin * fout))
        o += fin * fout
        bounds.append((o, o + fout))
        o += fout
    idx = np.concatenate([rng.integers(a, b, 6) for a, b in bounds])
    for i in idx:
        h = 1e-5
        p1 = params.copy()
        p1[i] += h
        p2 = params.copy()
        p2[i] -= h
        fd = (tpref.mlp_loss_grad(p1, q, x, y)[0] - tpref.mlp_loss_grad(p2, q, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g[i]), 1e-4), (i, fd, g[i])


def test_periodic_modes_keep_amplitude_and_divergence_free():
    rows, period = surrogate.surrogate_field(200)
    k = rows[:, :3]
    n = k / (2 * np.pi / period)
    assert np.allclose(n, np.rint(n), atol=1e-9)
    assert np.all(np.any(np.rint(n) != 0, axis=1))
    assert np.max(np.abs(np.sum(k * rows[:, 3:6], axis=1))) < 1e-12 * np.max(np.abs(k))
    from test_particle_amd.fields import turbulence_modes
    orig = turbulence_modes(200)
    assert np.allclose(np.linalg.norm(rows[:, 3:6], axis=1),
                       np.linalg.norm(orig[:, 3:6], axis=1), rtol=1e-12)
    # the field is periodic in the box
    from test_particle_amd.fields import eval_modes_numpy
    pts = np.random.default_rng(0).uniform(0, period, (50, 3))
    b0 = eval_modes_numpy(rows, pts, 1.0)
    b1 = eval_modes_numpy(rows, pts + [period, -period, 2 * period], 1.0)
    assert np.allclose(b0, b1, atol=1e-9)


@pytest.mark.parametrize("sigma", [2.0, 6.0])
def test_fourier_frequencies_on_lattice(sigma):
    q = surrogate.fourier_frequencies(L_BOX, sigma=sigma, seed=1)
    n = q.astype(np.float64) * L_BOX
    assert q.shape == (128, 3)
    assert np.allclose(n, np.rint(n), atol=1e-3)
    assert np.all(np.any(np.rint(n) != 0, axis=1))
    assert 0.5 * sigma < np.std(np.rint(n)) < 1.5 * sigma


def test_logradial_frequencies_on_lattice():
    q = surrogate.fourier_frequencies_logradial(L_BOX, 16.0, seed=1)
    n = q.astype(np.float64) * L_BOX
    assert q.shape == (128, 3)
    assert np.allclose(n, np.rint(n), atol=1e-3)
    r = np.linalg.norm(np.rint(n), axis=1)
    assert np.all(r > 0) and r.max() <= 16.0 + 0.5 * np.sqrt(3.0)
    # log-uniform radii: about as many below 4 as between 4 and 16
    assert 0.3 < np.mean(r < 4.0) < 0.7


def test_spectral_tail_brackets():
    """The tail of one mode is its amplitude over |B| when the cut is below
    its lattice radius and zero above; over the whole table it is the rms
    fluctuation over |B|."""
    period = 2 * np.pi
    rows = np.array([[3.0, 0.0, 0.0, 0.0, 0.3, 0.0, 0.0],
                     [0.0, 0.0, 10.0, 0.2, 0.0, 0.0, 1.0]])
    tot = np.sqrt(0.5 * (0.09 + 0.04))
    bmag = np.sqrt(1.0 + 0.5 * (0.09 + 0.04))
    assert np.isclose(surrogate.spectral_tail(rows, period, 0.5), tot / bmag, rtol=1e-14)
    assert np.isclose(surrogate.spectral_tail(rows, period, 5.0), np.sqrt(0.02) / bmag,
                      rtol=1e-14)
    assert surrogate.spectral_tail(rows, period, 10.5) == 0.0
