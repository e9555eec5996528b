"""The learned-guess MLP's HIP kernels (BF16 MFMA GEMMs, GELU, loss, backward, Adam,
features, guess) against the numpy reference oracle/mlp_ref.py, which rounds to BF16 at
the points the device does; remaining differences come from FP32 summation order."""
import numpy as np
import pytest

import mlp_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    from reprobus_amd.mlp import GuessMLP

    m = GuessMLP(0)
    yield m
    m.close()


def _case(n, seed):
    from reprobus_amd.mlp import init_params

    rng = np.random.default_rng(seed)
    flat = init_params(seed)
    P = R.split(flat)   # float64 views are copies; edit the float32 vector through offsets
    o = 0
    for name, shape in R.SHAPES:
        k = int(np.prod(shape))
        if name.startswith("b"):
            valid = R.NOUT if name == "b4" else k
            flat[o:o + valid] = rng.normal(0, 0.1, valid).astype(np.float32)
        o += k
    X = (rng.normal(0, 1, (n, R.NIN)) * 3 + 1).astype(np.float32)
    T = (rng.normal(0, 1, (n, R.NOUT)) * 2 - 0.5).astype(np.float32)
    norm = (np.full(R.NIN, 1.0, np.float32), np.full(R.NIN, 3.0, np.float32),
            np.full(R.NOUT, -0.5, np.float32), np.full(R.NOUT, 2.0, np.float32))
    return flat, X, T, norm


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_forward_matches_reference(net):
    import torch

    for n, seed in ((4096, 0), (1000, 1), (70, 2)):       # ragged batches: partial tiles
        flat, X, T, norm = _case(n, seed)
        net.set_params(flat)
        net.set_norm(*norm)
        Y = torch.empty((n, R.NOUT), dtype=torch.float32, device="cuda:0")
        net.forward(_dev(X), Y)
        torch.cuda.synchronize()
        ref = R.forward(flat, X, norm)
        err = np.abs(Y.cpu().numpy() - ref)
        scale = np.abs(ref).max()
        print(f"[mlp] forward n={n}: max abs err {err.max():.3e} (scale {scale:.2f})", flush=True)
        assert err.max() < 2e-2 * scale
        assert np.median(err) < 1e-4 * scale


def test_gradients_match_reference(net):
    import torch

    n = 8192
    flat, X, T, norm = _case(n, 3)
    net.set_params(flat)
    net.set_norm(*norm)
    g, loss = net.gradients(_dev(X), _dev(T))
    rloss, rg = R.loss_and_grads(flat, X, T, norm)
    print(f"[mlp] loss gpu {loss:.6f} ref {rloss:.6f}", flush=True)
    assert abs(loss - rloss) < 1e-4 * rloss
    Gg, Gr = R.split(g), R.split(rg)
    for name, _ in R.SHAPES:
        d = np.linalg.norm(Gg[name] - Gr[name]) / np.linalg.norm(Gr[name])
        print(f"[mlp] grad {name}: rel {d:.2e}", flush=True)
        assert d < 3e-3, name
    P = R.split(np.arange(R.NPARAM))
    assert np.all(g[P["W4"][:, R.NOUT:].ravel().astype(int)] == 0)
    assert np.all(g[P["b4"][R.NOUT:].astype(int)] == 0)


def test_adam_step_applies_its_own_gradient(net):
    import torch

    n = 2048
    flat, X, T, norm = _case(n, 4)
    net.set_params(flat)
    net.set_norm(*norm)
    m = np.zeros(R.NPARAM)
    v = np.zeros(R.NPARAM)
    p = flat.astype(np.float64)
    Xd, Td = _dev(X), _dev(T)
    for t in (1, 2, 3):
        g, _ = net.gradients(Xd, Td)
        net.train_step(Xd, Td, 1e-3)
        p, m, v = R.adam_step(p, g.astype(np.float64), m, v, t, 1e-3)
        got = net.get_params()
        assert np.max(np.abs(got - p)) < 1e-6, t
        p = got.astype(np.float64)


def test_training_reduces_loss(net):
    """A smooth target (a fixed random linear map through tanh) is learned."""
    import torch
    from reprobus_amd.mlp import init_params

    rng = np.random.default_rng(5)
    n = 8192
    A = rng.normal(0, 0.3, (R.NIN, R.NOUT))
    X = rng.normal(0, 1, (16 * n, R.NIN)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    net.set_params(init_params(1))
    net.set_norm(np.zeros(R.NIN), np.ones(R.NIN), np.zeros(R.NOUT), np.ones(R.NOUT))
    Xd, Td = _dev(X), _dev(T)
    first = net.train_step(Xd[:n], Td[:n], 1e-3, want_loss=True)
    for it in range(300):
        j = (it % 16) * n
        last = net.train_step(Xd[j:j + n], Td[j:j + n], 1e-3, want_loss=(it == 299))
    print(f"[mlp] training loss {first:.4f} -> {last:.5f}", flush=True)
    assert last < 0.2 * first


def test_features_and_guess(net):
    import torch
    from reprobus_amd import synth
    from reprobus_amd.chem import make_grid
    from reprobus_amd.mlp import LONG_SLOTS

    inp = synth.make_c1(seed=6, nlon=8, nlat=5)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    T, p, sza, q, h = (_dev(np.ascontiguousarray(a, dtype=np.float64))
                       for a in (inp.T, inp.p, inp.sza, inp.qj1, inp.hc))
    nb = g.nivbas * g.nlat * g.nlon
    X = torch.empty((nb, R.NIN), dtype=torch.float32, device="cuda:0")
    net.features(grid, inp.dt, T, p, sza, q, h, X)
    torch.cuda.synchronize()
    Xh = X.cpu().numpy()
    lg = lambda a: np.log(np.maximum(a, 1e-30)).reshape(a.shape[0], -1).T
    ref = np.concatenate([lg(inp.qj1[LONG_SLOTS]), lg(inp.hc),
                          inp.T.reshape(1, -1).T, np.log(inp.p).reshape(1, -1).T,
                          np.cos(np.radians(inp.sza)).reshape(1, -1).T,
                          lg(inp.qj1[[11]]), np.full((nb, 1), np.log(inp.dt))], axis=1)
    assert np.max(np.abs(Xh - ref.astype(np.float32)) / np.maximum(np.abs(ref), 1e-3)) < 1e-6
    # the guess (one fused launch: features, network, formula): its implied log
    # increments against the GEMM forward on the same features and against the reference;
    # the fused network sums each MFMA slab in another order, so BF16 roundings of
    # activations may differ: the forward tolerances apply
    flat, _, _, norm = _case(8, 7)
    net.set_params(flat)
    net.set_norm(*norm)
    Y = torch.empty((nb, R.NOUT), dtype=torch.float32, device="cuda:0")
    net.forward(X, Y)
    gq, gh = torch.zeros_like(q), torch.zeros_like(h)
    net.guess(grid, inp.dt, T, p, sza, q, h, gq, gh)
    torch.cuda.synchronize()
    y = Y.cpu().numpy().astype(np.float64)
    gqh, ghh = gq.cpu().numpy(), gh.cpu().numpy()
    x_long = np.maximum(inp.qj1[LONG_SLOTS].reshape(40, -1), 1e-30)
    x_short = np.maximum(inp.hc.reshape(15, -1), 1e-30)
    y_fused = np.concatenate([np.log(gqh[LONG_SLOTS].reshape(40, -1) / x_long),
                              np.log(ghh.reshape(15, -1) / x_short)]).T
    ref = R.forward(flat, Xh, norm)
    scale = np.abs(ref).max()
    for name, other in (("gemm forward", y), ("reference", ref)):
        err = np.abs(y_fused - other)
        print(f"[mlp] fused guess vs {name}: max {err.max():.3e} median {np.median(err):.3e} "
              f"(scale {scale:.2f})", flush=True)
        assert err.max() < 2e-2 * scale
        assert np.median(err) < 1e-4 * scale
    # finite everywhere (boxes of the ragged last workgroup included); copied slots exact
    assert np.all(np.isfinite(gqh)) and np.all(np.isfinite(ghh))
    for s in (10, 11, 42):
        assert np.array_equal(gqh[s], inp.qj1[s])


def test_bad_arguments(net):
    import torch
    from reprobus_amd._lib import RbError

    X = torch.zeros((4, R.NIN), dtype=torch.float32, device="cuda:0")
    with pytest.raises(RbError):
        net.train_step(X, X[:, :R.NOUT].contiguous(), 0.0)
    with pytest.raises(RbError):
        net.set_norm(np.zeros(R.NIN), np.zeros(R.NIN), np.zeros(R.NOUT), np.ones(R.NOUT))
