"""The numpy reference of the learned-guess MLP (oracle/mlp_ref.py), pinned against
independent implementations: finite differences of its own loss, torch autograd and
torch.optim.Adam on the CPU, torch's BF16 conversion, and the erf form of GELU."""
import numpy as np
import pytest

import mlp_ref as R


def _case(n=64, seed=0):
    rng = np.random.default_rng(seed)
    from reprobus_amd.mlp import init_params

    flat = init_params(seed).astype(np.float64)
    flat += rng.normal(0, 0.01, flat.size) * (flat != 0)     # non-trivial biases stay zero
    P = R.split(flat)
    P["b1"][:] = rng.normal(0, 0.1, R.WIDTH)
    P["b4"][:R.NOUT] = rng.normal(0, 0.1, R.NOUT)
    X = rng.normal(0, 1, (n, R.NIN)) * 3 + 1
    T = rng.normal(0, 1, (n, R.NOUT)) * 2
    norm = (np.full(R.NIN, 1.0), np.full(R.NIN, 3.0), np.zeros(R.NOUT), np.full(R.NOUT, 2.0))
    return flat, X, T, norm


def test_backward_matches_finite_differences():
    flat, X, T, norm = _case()
    loss, g = R.loss_and_grads(flat, X, T, norm, use_bf16=False)
    rng = np.random.default_rng(3)
    live = np.ones(flat.size, bool)
    P = R.split(np.arange(flat.size))
    live[P["W1"][R.NIN:].ravel().astype(int)] = False       # padding rows: zero input
    live[P["W4"][:, R.NOUT:].ravel().astype(int)] = False
    live[P["b4"][R.NOUT:].ravel().astype(int)] = False
    for _ in range(4):
        v = rng.normal(0, 1, flat.size) * live
        eps = 1e-6
        lp, _ = R.loss_and_grads(flat + eps * v, X, T, norm, use_bf16=False)
        lm, _ = R.loss_and_grads(flat - eps * v, X, T, norm, use_bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ v) < 1e-6 * max(1.0, abs(fd)), (fd, g @ v)
    assert np.all(g[~live] == 0)


def test_backward_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    flat, X, T, norm = _case(n=48, seed=1)
    loss, g = R.loss_and_grads(flat, X, T, norm, use_bf16=False)
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in R.split(flat).items()}
    x = torch.zeros(X.shape[0], R.PAD, dtype=torch.float64)
    x[:, :R.NIN] = (torch.tensor(X) - 1.0) / 3.0
    a = x
    for l in (1, 2, 3):
        a = torch.nn.functional.gelu(a @ P[f"W{l}"] + P[f"b{l}"], approximate="tanh")
    y = (a @ P["W4"] + P["b4"])[:, :R.NOUT]
    lt = torch.mean((y - torch.tensor(T) / 2.0) ** 2)
    lt.backward()
    gt = np.concatenate([P[name].grad.numpy().ravel() for name, _ in R.SHAPES])
    assert abs(lt.item() - loss) < 1e-12 * max(1, loss)
    assert np.max(np.abs(gt - g)) < 1e-12 * max(1.0, np.max(np.abs(gt)))


def test_adam_matches_torch_optim():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    p0 = rng.normal(0, 1, 500)
    grads = [rng.normal(0, 1, 500) * 10.0 ** rng.uniform(-6, 0, 500) for _ in range(5)]
    pt = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(500), np.zeros(500)
    for t, gr in enumerate(grads, 1):
        pt.grad = torch.tensor(gr)
        opt.step()
        p, m, v = R.adam_step(p, gr, m, v, t, 1e-3)
    assert np.max(np.abs(pt.detach().numpy() - p)) < 1e-14


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    x = (rng.normal(0, 1, 100000) * 10.0 ** rng.uniform(-20, 20, 100000)).astype(np.float32)
    ties = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 2 ** -126], dtype=np.float32)
    x = np.concatenate([x, ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(R.bf16(x), ref)
    assert R.bf16(np.float32(1 + 2 ** -8)) == 1.0 and R.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6


def test_gelu_tanh_form_is_close_to_erf_form():
    from scipy.special import erf

    z = np.linspace(-8, 8, 4001)
    exact = 0.5 * z * (1 + erf(z / np.sqrt(2)))
    assert np.max(np.abs(R.gelu(z) - exact)) < 1e-3
    assert R.gelu(0.0) == 0.0
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - R.gelu_grad(z))) < 1e-8
