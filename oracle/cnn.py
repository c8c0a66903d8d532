"""FP64 numpy restatement of the learned coarse proposal (optflow_soil_amd/csrc/cnn.hip).

TEST INFRASTRUCTURE ONLY, like the rest of oracle/: the GPU kernels are checked against it. The
network has no counterpart in the reference (it stands in for the coarse levels of the POC
pyramid, rust/crates/optflow-core/src/poc.rs:694-766); what pins this restatement is mathematics:
the convolution against scipy.signal.correlate, the correlation volume against its definition
summed term by term, and the backward pass against central finite differences of the forward
pass (tests/test_coarse_net_oracle.py).

With bf16=True the forward and backward passes round to bfloat16 (nearest even) exactly where the
GPU kernels store bfloat16: the correlation volume, the weights, every hidden activation, the
output gradient and every hidden gradient; sums stay in float64.
"""
import numpy as np

CV_R = 6
CV_CH = (2 * CV_R + 1) ** 2
LAYERS = [(CV_CH, 32, True), (32, 32, True), (32, 32, True), (32, 32, True), (32, 32, True), (32, 2, False)]
LOSS_EPS = 0.01


def bf16(x):
    """float -> the nearest bfloat16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def n_params():
    return sum(co * 9 * ci + co for ci, co, _ in LAYERS)


def unflatten(p):
    """Flat parameters -> [(W [cout, 9, cin], b [cout])] per layer (the layout of include/ofx.h)."""
    p = np.asarray(p, dtype=np.float64)
    out, k = [], 0
    for ci, co, _ in LAYERS:
        W = p[k:k + co * 9 * ci].reshape(co, 9, ci)
        k += co * 9 * ci
        b = p[k:k + co]
        k += co
        out.append((W, b))
    assert k == p.size
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def down8(img):
    """Mean of each 8 x 8 block of pixels (the blocks at the right and bottom edges hold what
    exists)."""
    img = np.asarray(img, dtype=np.float64)
    h, w = img.shape
    h8, w8 = (h + 7) // 8, (w + 7) // 8
    out = np.empty((h8, w8))
    for i in range(h8):
        for j in range(w8):
            out[i, j] = img[8 * i:8 * i + 8, 8 * j:8 * j + 8].mean()
    return out


def standardise(a):
    """(a - mean) / sd over the map (population sd; a constant map gives 0)."""
    a = np.asarray(a, dtype=np.float64)
    m = a.mean()
    var = max((a * a).mean() - m * m, 0.0)
    return (a - m) * (1.0 / np.sqrt(var) if var > 1e-20 else 0.0)


def cost_volume(n1, n2, R=CV_R):
    """[h, w, (2R + 1)^2]: channel (dy + R)(2R + 1) + (dx + R) at p is the mean over the 3 x 3
    neighbourhood q of n1(p + q) n2(p + q + d), zero outside the maps."""
    h, w = n1.shape
    D = 2 * R + 1
    P1 = np.zeros((h + 2, w + 2))
    P1[1:-1, 1:-1] = n1
    P2 = np.zeros((h + 2 + 2 * R, w + 2 + 2 * R))
    P2[1 + R:1 + R + h, 1 + R:1 + R + w] = n2
    out = np.zeros((h, w, D * D))
    for dy in range(-R, R + 1):
        for dx in range(-R, R + 1):
            c = (dy + R) * D + (dx + R)
            acc = np.zeros((h, w))
            for qy in (-1, 0, 1):
                for qx in (-1, 0, 1):
                    a = P1[1 + qy:1 + qy + h, 1 + qx:1 + qx + w]
                    b = P2[1 + R + qy + dy:1 + R + qy + dy + h, 1 + R + qx + dx:1 + R + qx + dx + w]
                    acc += a * b
            out[:, :, c] = acc / 9.0
    return out


def conv3x3(x, W, b=None):
    """y[n, i, j, o] = b[o] + sum_{t, c} x[n, i + dy(t), j + dx(t), c] W[o, t, c], t = 3 (dy + 1) +
    (dx + 1), zero padding."""
    nb, h, w, ci = x.shape
    co = W.shape[0]
    P = np.zeros((nb, h + 2, w + 2, ci))
    P[:, 1:-1, 1:-1] = x
    y = np.zeros((nb, h, w, co))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += P[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] @ W[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3x3_input_grad(dy, W):
    """d loss / d x of conv3x3, given d loss / d y."""
    nb, h, w, co = dy.shape
    ci = W.shape[2]
    P = np.zeros((nb, h + 2, w + 2, co))
    P[:, 1:-1, 1:-1] = dy
    dx_ = np.zeros((nb, h, w, ci))
    for t in range(9):
        oy, ox = t // 3 - 1, t % 3 - 1
        # x[p + off(t)] contributed to y[p]: dx[q] += dy[q - off(t)] W[:, t, :]
        dx_ += P[:, 1 - oy:1 - oy + h, 1 - ox:1 - ox + w] @ W[:, t, :]
    return dx_


def conv3x3_weight_grad(dy, x):
    nb, h, w, ci = x.shape
    co = dy.shape[3]
    P = np.zeros((nb, h + 2, w + 2, ci))
    P[:, 1:-1, 1:-1] = x
    dW = np.zeros((co, 9, ci))
    D = dy.reshape(-1, co)
    for t in range(9):
        oy, ox = t // 3 - 1, t % 3 - 1
        dW[:, t, :] = D.T @ P[:, 1 + oy:1 + oy + h, 1 + ox:1 + ox + w].reshape(-1, ci)
    return dW, D.sum(axis=0)


def features(a1, a2):
    """The correlation volume of a batch of cell maps (nb, h8, w8) -> (nb, h8, w8, 169)."""
    return np.stack([cost_volume(standardise(x1), standardise(x2)) for x1, x2 in zip(a1, a2)])


def forward(params, cv, bf16_like_gpu=False):
    """Activations [cv, h1 .. h5] and the prediction (nb, h8, w8, 2) in cells."""
    r = bf16 if bf16_like_gpu else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unflatten(params)
    acts = [r(cv)]
    x = acts[0]
    for k, ((W, b), (_, _, relu)) in enumerate(zip(layers, LAYERS)):
        y = conv3x3(x, r(W), np.asarray(b, np.float32).astype(np.float64) if bf16_like_gpu else b)
        if relu:
            x = r(np.maximum(y, 0.0))
            acts.append(x)
        else:
            return acts, y


def loss_and_grad(pred, target):
    """Mean Charbonnier end-point error sqrt(ex^2 + ey^2 + eps^2) - eps and its gradient."""
    e = pred - target
    r = np.sqrt((e ** 2).sum(axis=-1, keepdims=True) + LOSS_EPS ** 2)
    n = e.shape[0] * e.shape[1] * e.shape[2]
    return float((r - LOSS_EPS).sum() / n), e / (r * n)


def gradient(params, a1, a2, target, bf16_like_gpu=False):
    """(loss, flat gradient, prediction) of one forward and backward pass."""
    r = bf16 if bf16_like_gpu else (lambda v: np.asarray(v, dtype=np.float64))
    cv = features(a1, a2)
    acts, pred = forward(params, cv, bf16_like_gpu)
    loss, dpred = loss_and_grad(pred, target)
    layers = unflatten(params)
    grads = [None] * len(layers)
    dy = r(dpred)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        dW, db = conv3x3_weight_grad(dy, acts[k])
        grads[k] = (dW, db)
        if k == 0:
            break
        dx = conv3x3_input_grad(dy, r(W))
        dy = r(dx * (acts[k] > 0))
    return loss, flatten(grads), pred
