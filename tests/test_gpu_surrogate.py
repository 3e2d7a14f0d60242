"""The MVC surrogate network on the GPU (cgs_mlp_*): forward, input
derivatives and parameter gradients against the numpy restatement
(oracle/mlp_oracle.py), training, and the surrogate binding against exact MVC.

Bars: weights within 5e-3 of the weights' spread of the oracle's bf16-rounded
float64 forward; input derivatives within 2% (Frobenius) of the oracle's
forward-mode tangents with the same bf16 roundings (and within 15% of the
float64 network's central differences); parameter gradients with cosine >= 0.999 and
relative error <= 3% per layer against the oracle's float64 backward (the GPU
rounds the backward activations to bf16); training lowers the loss at least
10x and the held-out squared error at least 10x below the untrained
network's; the surrogate binding's inside mask equals the exact one and every
deformed position is within 1e-3 of the cage diagonal of exact MVC's (points
whose residual fails the test take the exact path, so this holds whatever the
network's accuracy; the number of such fallbacks is printed)."""
import numpy as np
import pytest

import mlp_oracle as mo
from cindergs_amd.mvc import MvcGpu, icosphere, sample_inside
from cindergs_amd.surrogate import MvcSurrogate

pytestmark = pytest.mark.gpu


def _layout(net):
    return net.rows, net.cols, net.offsets


def test_forward_and_input_derivatives_match_oracle(gpu):
    net = MvcSurrogate(gpu, 642)
    params, fourier = net.get_params()
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    dx = gpu.to_device(x)
    w, g = net.infer(dx, x.shape[0])
    W, G = w.numpy(), g.numpy()
    ref = mo.weights(params, fourier, *_layout(net), x, 642, use_bf16=True)
    spread = np.abs(ref - ref.mean()).max()
    err = np.abs(W - ref).max()
    print(f"[case] surrogate forward: max|dw|={err:.2e} spread={spread:.2e}", flush=True)
    assert err <= 5e-3 * spread + 1e-6
    assert np.abs(W.sum(1) - 1).max() < 1e-4
    # dw/dx against the oracle's forward-mode tangents with the same bf16
    # roundings, and (loosely: bf16 hidden activations) against central
    # differences of the float64 network
    tan = mo.input_derivatives(params, fourier, *_layout(net), x, 642)
    rel_t = np.linalg.norm(G - tan) / np.linalg.norm(tan)
    print(f"[case] surrogate dw/dx vs bf16 tangents: relative error {rel_t:.3e}", flush=True)
    assert rel_t < 0.02
    h = 1e-4
    num = np.zeros((x.shape[0], 642, 3))
    for a in range(3):
        xp, xm = x.astype(np.float64).copy(), x.astype(np.float64).copy()
        xp[:, a] += h
        xm[:, a] -= h
        num[:, :, a] = (mo.weights(params, fourier, *_layout(net), xp, 642, use_bf16=False)
                        - mo.weights(params, fourier, *_layout(net), xm, 642, use_bf16=False)) / (2 * h)
    rel = np.linalg.norm(G - num) / np.linalg.norm(num)
    print(f"[case] surrogate dw/dx vs float64 differences: relative error {rel:.3e}", flush=True)
    assert rel < 0.15
    net.close()


def test_parameter_gradient_matches_oracle(gpu):
    net = MvcSurrogate(gpu, 642)
    params, fourier = net.get_params()
    rng = np.random.default_rng(2)
    n = 256
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    t = rng.dirichlet(np.ones(642), n).astype(np.float32)
    grad, loss = net.gradient(gpu.to_device(x), gpu.to_device(t), n)
    rloss, rgrad = mo.loss_and_grad(params, fourier, *_layout(net), x, t, 642, use_bf16=True)
    assert abs(loss - rloss) <= 1e-2 * rloss
    for l in range(5):
        o, r, c = net.offsets[l], net.rows[l], net.cols[l]
        a = grad[o:o + r * (c + 32)].astype(np.float64)
        b = rgrad[o:o + r * (c + 32)]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[case] layer {l}: cosine {cos:.6f} relative error {rel:.3e}", flush=True)
        assert cos >= 0.999 and rel <= 0.03
    net.close()


def test_training_and_surrogate_binding(gpu):
    cage = icosphere()
    mg = MvcGpu(gpu, cage)

    def inside_fn(p):
        d = gpu.to_device(p)
        w = mg.winding(d, p.shape[0])
        d.free()
        return w

    n_train = 65536
    xt = sample_inside(cage, n_train, seed=5, inside_fn=inside_fn, batch=1 << 17)
    dxt = gpu.to_device(xt)
    wt, _, _ = mg.bind(dxt, n_train, gradients=False)
    net = MvcSurrogate(gpu, cage.num_vertices)
    untrained = MvcSurrogate(gpu, cage.num_vertices, seed=7)
    losses = net.train(dxt, wt, n_train, steps=600, batch=8192, lr=1e-3)
    print(f"[case] training loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    assert np.all(np.isfinite(losses)) and losses[-1] * 10 <= losses[0]
    # fresh points plus a few outside
    xe = sample_inside(cage, 8192, seed=6, inside_fn=inside_fn, batch=1 << 15)
    xe = np.concatenate([xe, np.array([[0, 0, 2.0], [2.0, 0, 0]], np.float32)])
    dxe = gpu.to_device(xe)
    n = xe.shape[0]
    ew, eg, em = mg.bind(dxe, n)
    inside = em.numpy().astype(bool)
    W_exact = ew.numpy()[inside]
    mse = []
    for m in (net, untrained):
        w, _ = m.infer(dxe, n, gradients=False)
        mse.append(float(((w.numpy()[inside] - W_exact) ** 2).mean()))
    print(f"[case] held-out squared error: trained {mse[0]:.3e} untrained {mse[1]:.3e}", flush=True)
    assert mse[0] * 10 <= mse[1]
    sw, sg, sm, fb = net.bind(cage, dxe, n, residual_tol=2e-4)
    print(f"[case] surrogate binding: {fb} of {n} points took the exact path", flush=True)
    np.testing.assert_array_equal(sm.numpy(), em.numpy())
    rng = np.random.default_rng(8)
    Vd = (cage.vertices + rng.uniform(-0.1, 0.1, cage.vertices.shape)).astype(np.float32)
    pe, _ = mg.deform(ew, eg, em, dxe, Vd, n)
    ps, _ = mg.deform(sw, sg, sm, dxe, Vd, n)
    err = np.abs(ps.numpy() - pe.numpy()).max()
    print(f"[case] deformed position error {err:.3e} (diag {cage.diagonal():.3f})", flush=True)
    assert err <= 1e-3 * cage.diagonal()
    net.close()
    untrained.close()
