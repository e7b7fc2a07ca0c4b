"""HIP MLP kernels checked against a float64 numpy restatement of the same
network (forward, loss, parameter gradients), and a small training run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def numpy_mlp(params, dims, x, shift, scale):
    """Forward pass in float64; returns output and per-layer caches."""
    from math import sqrt
    from scipy.special import erf
    h = (x.astype(np.float64) - shift) * scale
    acts, pres, off = [h], [], 0
    for l, (i, o) in enumerate(dims):
        W = params[off:off + i * o].reshape(o, i).astype(np.float64)
        off += i * o
        b = params[off:off + o].astype(np.float64)
        off += o
        z = h @ W.T + b
        pres.append(z)
        h = 0.5 * z * (1 + erf(z / sqrt(2))) if l + 1 < len(dims) else z
        acts.append(h)
    return h, acts, pres


def numpy_grad(params, dims, x, t, shift, scale):
    from math import pi, sqrt
    from scipy.special import erf
    y, acts, pres = numpy_mlp(params, dims, x, shift, scale)
    n = y.size
    loss = np.sum((y - t) ** 2) / n
    dz = 2 * (y - t) / n
    grads, offs, off = [None] * len(dims), [], 0
    for (i, o) in dims:
        offs.append(off)
        off += i * o + o
    for l in range(len(dims) - 1, -1, -1):
        i, o = dims[l]
        W = params[offs[l]:offs[l] + i * o].reshape(o, i).astype(np.float64)
        gW = dz.T @ acts[l]
        gb = dz.sum(axis=0)
        grads[l] = np.concatenate([gW.ravel(), gb])
        if l == 0:
            break
        dh = dz @ W
        z = pres[l - 1]
        dz = dh * (0.5 * (1 + erf(z / sqrt(2))) + z * np.exp(-0.5 * z * z) / sqrt(2 * pi))
    return loss, np.concatenate(grads)


def test_mlp_forward_and_gradient_match_numpy(gpu):
    from cm_amd.warm_start import WarmStartMLP, FEATURE_SHIFT, FEATURE_SCALE
    net = WarmStartMLP(hidden=32, layers=3, seed=3)
    dims = [(9, 32), (32, 32), (32, 32), (32, 4)]
    rng = np.random.default_rng(0)
    n = 300  # not a multiple of the GEMM tiles
    x = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.1, 0.1, (n, 3)), rng.uniform(0, 0.04, (n, 1)),
                        rng.uniform(1, 5, (n, 1)), rng.uniform(0, 10, (n, 1))], axis=1).astype(np.float32)
    t = rng.standard_normal((n, 4)).astype(np.float32)
    p = net.parameters()
    y = net.forward(x)
    yref, _, _ = numpy_mlp(p, dims, x, FEATURE_SHIFT.astype(np.float64), FEATURE_SCALE.astype(np.float64))
    assert np.abs(y - yref).max() / np.abs(yref).max() < 1e-5
    loss, g = net.gradient(x, t)
    lref, gref = numpy_grad(p, dims, x, t, FEATURE_SHIFT.astype(np.float64), FEATURE_SCALE.astype(np.float64))
    assert abs(loss - lref) / lref < 1e-5
    assert np.linalg.norm(g - gref) / np.linalg.norm(gref) < 1e-4


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def numpy_mlp_bf16(params, dims, x, shift, scale):
    """The BF16 inference rule: FP32-normalised input, weights and hidden
    activations rounded to bf16, exact products, FP32 bias and GELU."""
    from math import sqrt
    from scipy.special import erf
    h = bf16((x.astype(np.float32) - shift.astype(np.float32)) * scale.astype(np.float32)).astype(np.float64)
    off = 0
    for l, (i, o) in enumerate(dims):
        W = bf16(params[off:off + i * o].reshape(o, i)).astype(np.float64)
        off += i * o
        b = params[off:off + o].astype(np.float64)
        off += o
        z = h @ W.T + b
        if l + 1 < len(dims):
            h = bf16(0.5 * z * (1 + erf(z / sqrt(2)))).astype(np.float64)
        else:
            h = z
    return h


def numpy_mlp_bf16_hidden(params, dims, x, shift, scale):
    """The BF16_HIDDEN inference rule: the first and last layers in FP32;
    between them, hidden activations and weights rounded to bf16, except the
    last hidden activation (the FP32 output layer's input)."""
    from math import sqrt
    from scipy.special import erf
    gelu = lambda z: 0.5 * z * (1 + erf(z / sqrt(2)))  # noqa: E731
    h = ((x.astype(np.float32) - shift.astype(np.float32)) * scale.astype(np.float32)).astype(np.float64)
    off = 0
    L = len(dims)
    for l, (i, o) in enumerate(dims):
        W = params[off:off + i * o].reshape(o, i).astype(np.float64)
        if 0 < l < L - 1:
            W = bf16(W).astype(np.float64)
        off += i * o
        b = params[off:off + o].astype(np.float64)
        off += o
        z = h @ W.T + b
        if l + 1 < L:
            h = gelu(z)
            if l + 2 < L:  # feeds a bf16 layer
                h = bf16(h).astype(np.float64)
        else:
            h = z
    return h


def test_bf16_bit_rounding_rule():
    """The emulation's rounding: nearest, ties to even, on the upper 16 bits."""
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    r = bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 4 * 2 ** -8)
    assert r[3] == 1.0 and r[4] == -2.5 and r[5] == 0.0


@pytest.mark.parametrize("hidden,layers,n", [(32, 3, 300), (128, 4, 5000)])
def test_mlp_bf16_inference_matches_emulation(gpu, hidden, layers, n):
    """BF16 inference on MFMA against the numpy emulation of the same rounding
    (only the FP32 accumulation order differs) and against the FP64 network."""
    from cm_amd import _lib
    from cm_amd.warm_start import WarmStartMLP, FEATURE_SHIFT, FEATURE_SCALE
    net = WarmStartMLP(hidden=hidden, layers=layers, seed=9)
    dims = [(9, hidden)] + [(hidden, hidden)] * (layers - 1) + [(hidden, 4)]
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.1, 0.1, (n, 3)), rng.uniform(0, 0.04, (n, 1)),
                        rng.uniform(1, 5, (n, 1)), rng.uniform(0, 10, (n, 1))], axis=1).astype(np.float32)
    p = net.parameters()
    y32 = net.forward(x)
    net.set_inference_precision(_lib.OC_BF16)
    y16 = net.forward(x)
    emu = numpy_mlp_bf16(p, dims, x, FEATURE_SHIFT, FEATURE_SCALE)
    exact, _, _ = numpy_mlp(p, dims, x, FEATURE_SHIFT.astype(np.float64), FEATURE_SCALE.astype(np.float64))
    scale_y = np.abs(exact).max()
    # same rounding rule: differences come from accumulation order and rare
    # one-ulp bf16 flips of a hidden activation
    assert np.abs(y16 - emu).max() <= 2e-2 * scale_y
    assert np.median(np.abs(y16 - emu)) <= 1e-3 * scale_y
    # bf16 against the FP64 network: bf16 resolution (2^-8), not a wrong layout
    assert np.abs(y16 - exact).max() <= 5e-2 * scale_y
    assert np.linalg.norm(y16 - exact) / np.linalg.norm(exact) < 2e-2
    # new parameters reach the bf16 weights; FP32 inference is unchanged
    p2 = p * np.float32(0.5)
    net.set_parameters(p2)
    emu2 = numpy_mlp_bf16(p2, dims, x, FEATURE_SHIFT, FEATURE_SCALE)
    assert np.abs(net.forward(x) - emu2).max() <= 2e-2 * np.abs(emu2).max()
    net.set_parameters(p)
    # bf16 hidden layers between FP32 first and last layers
    net.set_inference_precision(_lib.OC_BF16_HIDDEN)
    yh = net.forward(x)
    emu_h = numpy_mlp_bf16_hidden(p, dims, x, FEATURE_SHIFT, FEATURE_SCALE)
    assert np.abs(yh - emu_h).max() <= 2e-2 * scale_y
    assert np.median(np.abs(yh - emu_h)) <= 1e-3 * scale_y
    assert np.linalg.norm(yh - exact) / np.linalg.norm(exact) < 2e-2
    net.set_inference_precision(_lib.OC_FP32)
    assert np.array_equal(net.forward(x), y32)
    with pytest.raises(_lib.CMISSError):
        net.set_inference_precision(_lib.OC_FP64)
    net.destroy()


def test_mlp_training_reduces_loss(gpu):
    from cm_amd.warm_start import WarmStartMLP
    net = WarmStartMLP(hidden=64, layers=2, seed=5)
    rng = np.random.default_rng(1)
    n = 20000
    x = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.1, 0.1, (n, 3)), rng.uniform(0, 0.04, (n, 1)),
                        rng.uniform(1, 5, (n, 1)), rng.uniform(0, 10, (n, 1))], axis=1).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, 0]) * x[:, 1], x[:, 3] * 5 + x[:, 7] * 0.1, x[:, 2] ** 2, x[:, 8] * x[:, 4]],
                 axis=1).astype(np.float32)
    net.set_dataset(x, t)
    first = net.gradient(x[:4096], t[:4096])[0]
    last = net.train(steps=1500, batch=1024, learning_rate=3e-3, seed=2)
    assert last < 0.02 * first
