"""Surrogate MLP: the numpy oracle is pinned by central finite differences
(float64); the HIP forward / backward kernels are checked against the oracle
with the same bf16 operand rounding; training on the GPU must reduce the loss."""
import numpy as np
import pytest

from oracle import mlp_ref


def _net(nin, width, nh, nout, seed):
    rng = np.random.default_rng(seed)
    dims = mlp_ref.layer_dims(nin, width, nh, nout)
    p = np.concatenate([np.concatenate([rng.uniform(-1, 1, o * d) * np.sqrt(3.0 / d), rng.normal(0, 0.1, o)])
                        for d, o in dims])
    return dims, p.astype(np.float32).astype(np.float64)


def _data(n, nin, nout, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.3, 2.0, (n, nin))
    Y = np.tanh(X[:, :1] * 0.5) * rng.normal(1, 0.1, (1, nout)) + 0.1 * rng.normal(size=(n, nout))
    norm = (X.mean(0), X.std(0) + 0.1, Y.mean(0), Y.std(0) + 0.1)
    return X, Y, norm


def test_oracle_gradient_matches_finite_differences():
    dims, p = _net(5, 7, 2, 3, 1)
    X, Y, norm = _data(11, 5, 3, 2)
    loss, g = mlp_ref.loss_grad(p, dims, X, Y, norm, "f64")
    rng = np.random.default_rng(3)
    for q in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[q] = 1e-6
        fd = (mlp_ref.loss_grad(p + e, dims, X, Y, norm)[0] - mlp_ref.loss_grad(p - e, dims, X, Y, norm)[0]) / 2e-6
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])) + 1e-9, (q, fd, g[q])


def test_oracle_bf16_mode_is_close_to_f64():
    dims, p = _net(29, 64, 3, 65, 4)
    X, Y, norm = _data(300, 29, 65, 5)
    l64, g64 = mlp_ref.loss_grad(p, dims, X, Y, norm, "f64")
    l16, g16 = mlp_ref.loss_grad(p, dims, X, Y, norm, "bf16")
    assert abs(l16 - l64) / l64 < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2


@pytest.mark.gpu
def test_gpu_loss_grad_matches_oracle():
    import torch
    from cfdwavemaker_amd.mlp import Mlp
    for (nin, width, nh, nout, n) in ((29, 64, 3, 65, 257), (7, 96, 2, 5, 1000)):
        dims, p = _net(nin, width, nh, nout, 6)
        X, Y, norm = _data(n, nin, nout, 7)
        m = Mlp(nin, width, nh, nout)
        m.set_params(p)
        m.set_norm(*norm)
        Xd = torch.tensor(X, dtype=torch.float32, device="cuda")
        Yd = torch.tensor(Y, dtype=torch.float32, device="cuda")
        loss, g = m.loss_grad(Xd, Yd)
        normf = tuple(np.asarray(a, np.float32).astype(np.float64) for a in norm)
        lr, gr = mlp_ref.loss_grad(p, dims, X.astype(np.float32), Y.astype(np.float32), normf, "bf16")
        print(f"dims {dims} loss gpu {loss:.6e} oracle {lr:.6e}", flush=True)
        assert abs(loss - lr) / lr < 1e-4
        q = 0
        for d, o in dims:
            for sz in (d * o, o):
                a, b = g[q:q + sz], gr[q:q + sz]
                rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
                print(f"  block size {sz} rel grad err {rel:.2e}", flush=True)
                assert rel < 1e-4
                q += sz
        # forward in raw units
        out = m.forward(Xd).cpu().numpy()
        ref = mlp_ref.predict(p, dims, X.astype(np.float32), normf, "bf16")
        rel = np.abs(out - ref).max() / np.abs(ref).max()
        print(f"  forward rel err {rel:.2e}", flush=True)
        assert rel < 1e-3
        m.close()


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    import torch
    from cfdwavemaker_amd.mlp import Mlp
    rng = np.random.default_rng(8)
    n, nin, nout = 20000, 6, 4
    X = rng.uniform(-2, 2, (n, nin))
    Y = np.stack([np.sin(X[:, 0]) * X[:, 1], X[:, 2] ** 2, np.cos(X[:, 3] + X[:, 4]), X[:, 5]], 1)
    m = Mlp(nin, 128, 3, nout, seed=1)
    m.set_norm(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    Xd = torch.tensor(X, dtype=torch.float32, device="cuda")
    Yd = torch.tensor(Y, dtype=torch.float32, device="cuda")
    l0, _ = m.loss_grad(Xd[:4096], Yd[:4096])
    curve = m.train(Xd, Yd, steps=600, batch=2048, lr=2e-3, seed=2, report=100)
    l1, _ = m.loss_grad(Xd[:4096], Yd[:4096])
    print(f"loss before {l0:.3e} after {l1:.3e}; curve {curve[curve > 0]}", flush=True)
    assert l1 < 0.05 * l0
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(29, 256, 3, 65), (7, 96, 2, 5), (30, 24, 2, 40)])
def test_fused_inference_equals_layerwise(shape):
    """Batches of 16,384 rows or more run the whole network in one fused
    kernel; smaller ones layer by layer.  Both give the same outputs bit for
    bit (same operand rounding, k order and epilogue)."""
    import torch
    from cfdwavemaker_amd.mlp import Mlp
    nin, width, nh, nout = shape
    dims, p = _net(nin, width, nh, nout, 8)
    n = 20000
    X, Y, norm = _data(n, nin, nout, 9)
    m = Mlp(nin, width, nh, nout)
    m.set_params(p)
    m.set_norm(*norm)
    Xd = torch.tensor(X, dtype=torch.float32, device="cuda")
    big = m.forward(Xd).cpu().numpy()
    small = np.concatenate([m.forward(Xd[a:a + 5000]).cpu().numpy() for a in range(0, n, 5000)])
    assert np.all(np.isfinite(big))
    assert np.array_equal(big, small), float(np.max(np.abs(big - small)))
    m.close()
