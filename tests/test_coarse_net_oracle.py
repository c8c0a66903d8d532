"""The oracle of the learned coarse proposal (oracle/cnn.py), pinned by mathematics: the 3 x 3
convolution against scipy.signal.correlate, its two backward maps by adjointness, the correlation
volume against its definition summed point by point, the whole backward pass against central
finite differences, and the bfloat16 rounding against its bit-level definition. No GPU."""
import numpy as np
import pytest
from scipy.signal import correlate

from oracle import cnn as C


def test_conv3x3_is_a_cross_correlation_with_zero_padding():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 7, 9, 5))
    W = rng.standard_normal((4, 9, 5))
    b = rng.standard_normal(4)
    y = C.conv3x3(x, W, b)
    for n in range(2):
        for o in range(4):
            want = b[o] + sum(correlate(x[n, :, :, c], W[o, :, c].reshape(3, 3), mode="same") for c in range(5))
            np.testing.assert_allclose(y[n, :, :, o], want, rtol=1e-12, atol=1e-12)


def test_conv3x3_backward_maps_are_its_adjoints():
    """<conv(x), g> = <x, input_grad(g)> and = <W, weight_grad(g, x)> (conv is linear in x and W);
    a transposed or unflipped map fails these."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 8, 3))
    W = rng.standard_normal((5, 9, 3))
    g = rng.standard_normal((2, 6, 8, 5))
    lhs = float((C.conv3x3(x, W) * g).sum())
    assert abs(lhs - float((x * C.conv3x3_input_grad(g, W)).sum())) < 1e-9 * max(1.0, abs(lhs))
    dW, db = C.conv3x3_weight_grad(g, x)
    assert abs(lhs - float((W * dW).sum())) < 1e-9 * max(1.0, abs(lhs))
    np.testing.assert_allclose(db, g.reshape(-1, 5).sum(axis=0), rtol=1e-12)


def test_cost_volume_is_its_definition():
    rng = np.random.default_rng(2)
    h, w, R = 5, 7, 2
    n1, n2 = rng.standard_normal((h, w)), rng.standard_normal((h, w))
    cv = C.cost_volume(n1, n2, R)

    def at(a, y, x):
        return a[y, x] if 0 <= y < h and 0 <= x < w else 0.0
    D = 2 * R + 1
    for y in range(h):
        for x in range(w):
            for dy in range(-R, R + 1):
                for dx in range(-R, R + 1):
                    s = sum(at(n1, y + qy, x + qx) * at(n2, y + qy + dy, x + qx + dx)
                            for qy in (-1, 0, 1) for qx in (-1, 0, 1))
                    assert abs(cv[y, x, (dy + R) * D + dx + R] - s / 9.0) < 1e-12


def test_down8_and_standardise():
    img = np.arange(20 * 13, dtype=np.float64).reshape(20, 13)
    d = C.down8(img)
    assert d.shape == (3, 2)
    assert d[0, 0] == img[:8, :8].mean() and d[2, 1] == img[16:, 8:].mean()
    s = C.standardise(np.array([1.0, 2.0, 3.0, 4.0]))
    assert abs(s.mean()) < 1e-15 and abs((s ** 2).mean() - 1.0) < 1e-12
    assert (C.standardise(np.full(5, 2.5)) == 0).all()


def test_bf16_rounds_to_nearest_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert C.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert C.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert C.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7
    v = np.random.default_rng(3).standard_normal(1000).astype(np.float32)
    r = C.bf16(v)
    assert (np.abs(r - v) <= np.abs(v) * 2.0 ** -8).all()
    assert (r.astype(np.float32).view(np.uint32) & 0xFFFF == 0).all()


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    layers = []
    for ci, co, _ in C.LAYERS:
        layers.append((rng.standard_normal((co, 9, ci)) * scale * np.sqrt(2.0 / (9 * ci)),
                       rng.standard_normal(co) * 0.1))
    return C.flatten(layers)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_pass_matches_finite_differences(seed):
    """Every layer's weights and biases: the analytic gradient against central differences of the
    loss (float64, no rounding). A dropped ReLU mask, a wrong flip or a transposed weight fails."""
    rng = np.random.default_rng(10 + seed)
    nb, h8, w8 = 1, 6, 7
    a1, a2 = rng.standard_normal((nb, h8, w8)), rng.standard_normal((nb, h8, w8))
    target = rng.standard_normal((nb, h8, w8, 2))
    p = random_params(seed)
    loss, g, _ = C.gradient(p, a1, a2, target)
    offs = np.cumsum([0] + [co * 9 * ci + co for ci, co, _ in C.LAYERS])
    idx = []
    for k in range(len(C.LAYERS)):
        idx += list(rng.integers(offs[k], offs[k + 1], 6))
    for i in idx:
        e = 1e-5 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (C.gradient(pp, a1, a2, target)[0] - C.gradient(pm, a1, a2, target)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
