"""The predictor's CPU oracle (oracle/predictor.py), pinned without the GPU:

- its hand-written backward pass against central finite differences of its
  own FP64 forward pass, for every parameter group, and a planted mistake
  (taps of the conv gradient mirrored) that the same check rejects;
- its 27-point convolution against torch.nn.functional.conv3d (CPU, FP64),
  which fixes the tap order, the zero padding and the channel orientation;
- its GELU against torch's erf-form GELU, and its bfloat16 rounding against
  torch's float32 -> bfloat16 conversion (round to nearest even).
"""
import numpy as np
import pytest

from oracle import predictor as P


def _params(seed):
    rng = np.random.default_rng(seed)
    p = P.init_params(seed)
    p["W3"] = rng.standard_normal(P.SHAPES["W3"]) * 0.3
    p["b3"] = rng.standard_normal(P.SHAPES["b3"]) * 0.1
    p["b2"] = rng.standard_normal(P.SHAPES["b2"]) * 0.1
    return p


def _inputs(seed, shape=(2, 5, 6, 7)):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape + (P.C_IN,))
    x[..., 15] = 1.0
    return x


def _fd(p, x, t, key, idx, h=1e-6):
    pp = {k: v.copy() for k, v in p.items()}
    pp[key][idx] += h
    lp = P.loss(P.forward(pp, x)[0], t)
    pp[key][idx] -= 2 * h
    lm = P.loss(P.forward(pp, x)[0], t)
    return (lp - lm) / (2 * h)


def test_backward_matches_finite_differences():
    p = _params(1)
    x = _inputs(2)
    y, cache = P.forward(p, x)
    t = np.random.default_rng(3).standard_normal(y.shape)
    g = P.backward(p, cache, y, t)
    rng = np.random.default_rng(4)
    for key in P.ORDER:
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in P.SHAPES[key])
            fd = _fd(p, x, t, key, idx)
            assert abs(fd - g[key][idx]) <= 1e-6 * max(abs(fd), 1e-3), (key, idx, fd, g[key][idx])


def test_planted_mirrored_taps_fail_the_check():
    """dW2 formed with the neighbour at p - o_t instead of p + o_t (a plausible
    sign slip) disagrees with the finite differences."""
    p = _params(5)
    x = _inputs(6)
    y, cache = P.forward(p, x)
    t = np.random.default_rng(7).standard_normal(y.shape)
    g = P.backward(p, cache, y, t)
    wrong = g["W2"][::-1].copy()  # tap t takes tap 26 - t's gradient: o_t -> -o_t
    worst = 0.0
    rng = np.random.default_rng(8)
    for _ in range(6):
        idx = (int(rng.integers(0, 27)),) + tuple(int(rng.integers(0, 64)) for _ in range(2))
        if idx[0] == 13:
            continue  # the centre tap is its own mirror
        fd = _fd(p, x, t, "W2", idx)
        worst = max(worst, abs(fd - wrong[idx]) / max(abs(fd), 1e-12))
    assert worst > 1e-3


def test_conv_matches_torch_conv3d():
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(9)
    p = _params(9)
    h1 = rng.standard_normal((2, 5, 6, 7, P.C_HID))
    # the oracle's conv on an h1 volume: forward with W1 = identity would need
    # 64 inputs, so call the shift-and-sum directly through forward's pieces
    inner = ((0, 5), (0, 6), (0, 7))
    z = np.broadcast_to(p["b2"], (2, 5, 6, 7, P.C_HID)).copy()
    for t, o in enumerate(P.taps()):
        z += P._shift(h1, o, inner, 0) @ p["W2"][t].T
    # torch: [n][C][D][H][W]; W2[t][out][in] -> weight[out][in][di+1][dj+1][dk+1]
    w = np.zeros((P.C_HID, P.C_HID, 3, 3, 3))
    for t, (di, dj, dk) in enumerate(P.taps()):
        w[:, :, di + 1, dj + 1, dk + 1] = p["W2"][t]
    zt = F.conv3d(torch.from_numpy(np.moveaxis(h1, -1, 1).copy()), torch.from_numpy(w),
                  torch.from_numpy(p["b2"].copy()), padding=1)
    zt = np.moveaxis(zt.numpy(), 1, -1)
    np.testing.assert_allclose(z, zt, rtol=1e-12, atol=1e-12)


def test_gelu_and_bf16_match_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    z = rng.standard_normal(10000) * 3
    np.testing.assert_allclose(P.gelu(z), torch.nn.functional.gelu(torch.from_numpy(z)).numpy(), rtol=1e-13,
                               atol=1e-15)
    # gelu' against a central difference
    h = 1e-6
    np.testing.assert_allclose(P.gelu_grad(z), (P.gelu(z + h) - P.gelu(z - h)) / (2 * h), rtol=1e-6, atol=1e-8)
    a = (rng.standard_normal(100000) * 10.0 ** rng.integers(-6, 6, 100000)).astype(np.float32)
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(P.bf16(a), want)


def test_forward_volume_margin_semantics():
    """margin 1 drops the outer cells of every volume; margin 0 keeps them
    with zeros outside: the inner outputs of a margin-0 run on a volume equal
    those of a margin-1 run on the same volume where no tap leaves it."""
    p = _params(11)
    x = _inputs(12, (1, 6, 7, 8))
    y1, _ = P.forward(p, x, margin=1)
    y0, _ = P.forward(p, x, margin=0)
    np.testing.assert_allclose(y0[:, 1:-1, 1:-1, 1:-1], y1, rtol=1e-13, atol=1e-13)
