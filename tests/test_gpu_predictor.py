"""The learned initial iterate's HIP kernels (bf16 MFMA, FP32 accumulation)
against the CPU oracle (oracle/predictor.py), and its training on the GPU.

Bounds, stated once:
- forward: the oracle runs with the same bfloat16 roundings of x', W1, W2 and
  h1 and sums in FP64; the GPU sums 16 (pointwise) and 1728 (conv) products
  in FP32, so outputs agree to 1e-4 relative L2;
- gradient: the GPU also rounds dz2 to bfloat16 as the MFMA operand of both
  backward convolutions (relative 2^-9 per element), which the FP64 oracle
  does not; every parameter group's gradient agrees to 2e-2 relative L2 and
  the loss to 1e-5.
"""
import numpy as np
import pytest

from conftest import gpu_available, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _params(seed):
    from oracle import predictor as P
    rng = np.random.default_rng(seed)
    p = P.init_params(seed)
    p["W3"] = rng.standard_normal(P.SHAPES["W3"]) * 0.2
    p["b3"] = rng.standard_normal(P.SHAPES["b3"]) * 0.1
    p["b2"] = rng.standard_normal(P.SHAPES["b2"]) * 0.1
    # round-trip through float32, the GPU's master precision
    return {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}


def _features(rng, shape, outside=None):
    x = rng.standard_normal(shape + (16,)).astype(np.float32)
    x[..., 15] = 1.0
    if outside is not None:
        x[outside] = 0.0  # cells outside the domain: zero features, mask 0
    return x


@pytest.mark.parametrize("shape,margin", [((2, 16, 16, 16), 1), ((1, 9, 13, 11), 0), ((1, 12, 20, 24), 0)])
def test_gpu_predictor_forward_matches_oracle(shape, margin):
    from oracle import predictor as P
    from channel_flow_amd.solver import Predictor
    rng = np.random.default_rng(sum(shape))
    p = _params(3)
    outside = (slice(None), slice(0, 2)) if margin == 1 else None
    x = _features(rng, shape, outside)
    print(f"predictor forward {shape} margin {margin}", flush=True)
    with Predictor(dataset_crops=1, batch_crops=2) as pr:
        pr.set_params(P.pack(p))
        y = pr.forward(x, margin=margin)
    yo, _ = P.forward(p, x.astype(np.float64), margin=margin, bf=True)
    assert y.shape == yo.shape
    assert rel_l2(y, yo) <= 1e-4, rel_l2(y, yo)


def test_gpu_predictor_gradient_matches_oracle():
    from oracle import predictor as P
    from channel_flow_amd.solver import Predictor
    rng = np.random.default_rng(21)
    p = _params(4)
    x = _features(rng, (3, 16, 16, 16), (slice(2, 3), slice(None), slice(0, 3)))
    t = rng.standard_normal((3, 14, 14, 14, 5)).astype(np.float32)
    with Predictor(dataset_crops=1, batch_crops=4) as pr:
        pr.set_params(P.pack(p))
        g, loss = pr.gradient(x, t)
    y, cache = P.forward(p, x.astype(np.float64), margin=1, bf=True)
    go = P.backward(p, cache, y, t.astype(np.float64))
    assert abs(loss - P.loss(y, t)) <= 1e-5 * P.loss(y, t)
    gg = P.unpack(g)
    for k in P.ORDER:
        e = rel_l2(gg[k], go[k])
        print(f"gradient {k}: relative L2 {e:.2e}", flush=True)
        assert e <= 2e-2, (k, e)


def test_gpu_predictor_learns_a_teacher(tmp_path):
    """Training on the GPU (pack, forward, head, both backward convs,
    reduction, Adam): crops whose targets come from a fixed random network of
    the same shape.  The held-out loss falls well below the base guess's."""
    from oracle import predictor as P
    from channel_flow_amd.solver import Predictor
    rng = np.random.default_rng(31)
    teacher = _params(7)
    teacher["W3"] *= 3.0
    n = 48
    x = _features(rng, (n, 16, 16, 16))
    y, _ = P.forward(teacher, x.astype(np.float64), margin=1, bf=True)
    # the dataset channels the solver's capture writes: Q^n, d1, d2, target, mask
    D = np.zeros((n, 16, 16, 16, 24), np.float32)
    D[..., 0:15] = x[..., 0:15]
    D[:, 1:-1, 1:-1, 1:-1, 15:20] = y
    D[..., 20] = 1.0
    with Predictor(dataset_crops=n, batch_crops=8) as pr:
        pr._load_dataset(D)
        # identity scales and a zero base guess: x' = the features, y* = the target
        pr.set_norm(np.array([0] * 5 + [1] * 5 + [1] * 5 + [1] * 5 + [0, 0], np.float32))
        pr.init_params(5)
        l0, b0 = pr.eval(40, 8)
        losses = pr.train(40, steps=300, lr=2e-3, seed=3)
        l1, b1 = pr.eval(40, 8)
    print(f"teacher: held-out loss {l0:.3e} -> {l1:.3e} (base {b0:.3e}); train {losses[0]:.3e} -> "
          f"{losses[-1]:.3e}", flush=True)
    assert b0 == b1 and abs(l0 - b0) <= 1e-6 * b0  # the untrained network adds nothing
    assert np.all(np.isfinite(losses))
    assert l1 < 0.2 * b1
