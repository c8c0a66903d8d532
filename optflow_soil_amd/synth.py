"""Seeded synthetic speckle inputs of the benchmark workloads, with exactly known displacement.

The pattern is a sum of isotropic Gaussian dots (sigma uniform in [1.2, 2.5] px, amplitude
uniform in [0.3, 1.0], about 6000 dots per 256 x 256 px) normalised to [0, 1]. A frame is the
pattern sampled bilinearly at the inverse-mapped position of every pixel, plus Gaussian noise,
quantised to uint8; the flow engine reads it as v / 255 (python/optflow_ref/io.py:15-29).
Displacements are Lagrangian: material point X moves to x = X + u(X) (python/optflow_ref/synth.py:1-6).
"""
import numpy as np

DOTS_PER_PX = 6000.0 / (256 * 256)
MARGIN = 64


def make_pattern(h, w, seed=0, margin=MARGIN):
    """The speckle pattern rendered on [-margin, w + margin) x [-margin, h + margin), in [0, 1]."""
    rng = np.random.default_rng(seed)
    H, W = h + 2 * margin, w + 2 * margin
    n = int(round(DOTS_PER_PX * H * W))
    cx = rng.uniform(0, W, n)
    cy = rng.uniform(0, H, n)
    sig = rng.uniform(1.2, 2.5, n)
    amp = rng.uniform(0.3, 1.0, n)
    img = np.zeros(H * W)
    R = 10  # 4 sigma of the widest dot
    off = np.arange(-R, R + 1)
    chunk = 20000
    for s in range(0, n, chunk):
        x0 = np.floor(cx[s:s + chunk]).astype(np.int64)
        y0 = np.floor(cy[s:s + chunk]).astype(np.int64)
        xs = x0[:, None] + off[None, :]  # (k, 2R+1)
        ys = y0[:, None] + off[None, :]
        gx = np.exp(-0.5 * ((xs - cx[s:s + chunk, None]) / sig[s:s + chunk, None]) ** 2)
        gy = np.exp(-0.5 * ((ys - cy[s:s + chunk, None]) / sig[s:s + chunk, None]) ** 2)
        val = amp[s:s + chunk, None, None] * gy[:, :, None] * gx[:, None, :]
        ok = ((ys >= 0) & (ys < H))[:, :, None] & ((xs >= 0) & (xs < W))[:, None, :]
        idx = ys[:, :, None] * W + xs[:, None, :]
        img += np.bincount(idx[ok], weights=val[ok], minlength=H * W)
    img = img.reshape(H, W)
    lo, hi = img.min(), img.max()
    return (img - lo) / (hi - lo)


def _bilinear(img, x, y, margin):
    """Sample the pattern (pixel (0, 0) of the frame at img[margin, margin]) at frame coords."""
    H, W = img.shape
    xf = np.clip(x + margin, 0, W - 1.000001)
    yf = np.clip(y + margin, 0, H - 1.000001)
    x0 = np.floor(xf).astype(np.int64)
    y0 = np.floor(yf).astype(np.int64)
    tx, ty = xf - x0, yf - y0
    return ((1 - ty) * ((1 - tx) * img[y0, x0] + tx * img[y0, x0 + 1])
            + ty * ((1 - tx) * img[y0 + 1, x0] + tx * img[y0 + 1, x0 + 1]))


def render(pattern, field=None, noise=0.0, seed=0, margin=MARGIN, quantise=True):
    """Frame of size pattern - 2 margin; field(X, Y) -> (ux, uy) is inverted by fixed point."""
    H, W = pattern.shape
    h, w = H - 2 * margin, W - 2 * margin
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    X, Y = xx, yy
    if field is not None:
        for _ in range(30):  # x = X + u(X)  ->  X = x - u(X)
            ux, uy = field(X, Y)
            X, Y = xx - ux, yy - uy
    v = _bilinear(pattern, X, Y, margin)
    if noise > 0:
        v = v + np.random.default_rng(seed).normal(0.0, noise, v.shape)
    if not quantise:
        return v
    return np.clip(np.rint(v * 255.0), 0, 255).astype(np.uint8)


def as_float(frame_u8):
    return frame_u8.astype(np.float32) / np.float32(255.0)


# ---------------------------------------------------------------- workload fields
def stretch_translate(exx, t, center):
    """Uniform stretch exx along x about `center` plus translation t."""
    def f(X, Y):
        return exx * (X - center[0]) + t[0], np.zeros_like(Y) + t[1]
    return f


def bowl_and_band(h, w, depth=6.0, radius=300.0, slip=3.0, band_width=40.0, angle_deg=35.0):
    """Settlement bowl uy = -depth exp(-r^2 / (2 radius^2)) about the centre plus a band of the given
    width through the centre at angle_deg carrying a tangential slip (tanh profile)."""
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    a = np.deg2rad(angle_deg)
    tx, ty = np.cos(a), np.sin(a)     # along the band
    nx, ny = -np.sin(a), np.cos(a)    # across it

    def f(X, Y):
        r2 = (X - cx) ** 2 + (Y - cy) ** 2
        uy = -depth * np.exp(-r2 / (2 * radius ** 2))
        s = nx * (X - cx) + ny * (Y - cy)
        g = 0.5 * slip * np.tanh(s / (band_width / 2.0))
        return g * tx, uy + g * ty
    return f


def convex_polygon_mask(h, w, seed=0, n=8, coverage=0.6, search_step=1):
    """A seeded random convex n-gon around the centre scaled to cover about `coverage` of the frame.
    search_step > 1 measures the coverage on every search_step-th pixel while searching the scale
    (faster, the coverage then holds to about a percent)."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(0.8, 1.0, n)
    px, py = np.cos(ang) * rad, np.sin(ang) * rad
    # convex hull of the points (they are sorted by angle around the origin: gift-wrap for safety)
    pts = sorted(zip(px, py))

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])
    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    hull = np.array(lower[:-1] + upper[:-1])
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    u = (xx - (w - 1) / 2.0) / ((w - 1) / 2.0)
    v = (yy - (h - 1) / 2.0) / ((h - 1) / 2.0)

    def inside(scale, u=u, v=v):
        m = np.ones(u.shape, bool)
        for i in range(len(hull)):
            a, b = hull[i] * scale, hull[(i + 1) % len(hull)] * scale
            m &= (b[0] - a[0]) * (v - a[1]) - (b[1] - a[1]) * (u - a[0]) >= 0
        return m
    us, vs = u[::search_step, ::search_step], v[::search_step, ::search_step]
    lo, hi = 0.1, 2.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if inside(mid, us, vs).mean() < coverage:
            lo = mid
        else:
            hi = mid
    return inside(0.5 * (lo + hi))


def workload_c1(seed=0):
    """256 x 256, 5 frames: frame k stretched by exx = 0.005 k about (128, 128) and translated by
    (0.8 k, 0) px, noise sigma 0.005 (seed k), uint8. Returns (frames u8, truth(k) -> (ux, uy))."""
    h = w = 256
    pat = make_pattern(h, w, seed=seed)
    fields = [stretch_translate(0.005 * k, (0.8 * k, 0.0), (128.0, 128.0)) for k in range(5)]
    frames = [render(pat, None if k == 0 else fields[k], noise=0.005, seed=k) for k in range(5)]
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    return frames, (lambda k: fields[k](xx, yy))


def workload_c2(size=2048, seed=0):
    """A size x size pair with the settlement bowl and shear band, noise sigma 0.005, a convex 8-gon
    ROI of ~60 %. Returns (I1 u8, I2 u8, roi bool, truth (ux, uy))."""
    h = w = size
    pat = make_pattern(h, w, seed=seed)
    f = bowl_and_band(h, w, radius=300.0 * size / 2048.0, band_width=40.0)
    a = render(pat, None, noise=0.005, seed=1)
    b = render(pat, f, noise=0.005, seed=2)
    roi = convex_polygon_mask(h, w, seed=seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    return a, b, roi, f(xx, yy)


C4_SIZES = ((512, 512), (768, 1024), (1024, 1024), (1536, 1536), (1024, 2048))  # (h, w)


def affine_field(exx, eyy, exy, rot_deg, t, center):
    """u(X) = (E + R - I)(X - c) + t: a small strain (exx, eyy, exy), a rotation and a translation."""
    r = np.deg2rad(rot_deg)
    a11, a12 = exx + (np.cos(r) - 1.0), exy - np.sin(r)
    a21, a22 = exy + np.sin(r), eyy + (np.cos(r) - 1.0)

    def f(X, Y):
        dx, dy = X - center[0], Y - center[1]
        return a11 * dx + a12 * dy + t[0], a21 * dx + a22 * dy + t[1]
    f.coeffs = (a11, a12, a21, a22, (float(center[0]), float(center[1])), (float(t[0]), float(t[1])))
    return f


def render_affine(pattern, ox, oy, h, w, f, noise=0.0, seed=0):
    """An h x w frame of an affine field f (affine_field) over the pattern's window at (ox, oy):
    the same steps as render, with the inverse map X = c + M^-1 (x - c - t) in closed form
    (M = I + A) instead of the fixed point."""
    a11, a12, a21, a22, (cx, cy), (tx, ty) = f.coeffs
    det = (1 + a11) * (1 + a22) - a12 * a21
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    rx, ry = xx - cx - tx, yy - cy - ty
    X = cx + ((1 + a22) * rx - a12 * ry) / det
    Y = cy + (-a21 * rx + (1 + a11) * ry) / det
    v = _bilinear(pattern, X + ox, Y + oy, 0)
    if noise > 0:
        v = v + np.random.default_rng(seed).normal(0.0, noise, v.shape)
    return np.clip(np.rint(v * 255.0), 0, 255).astype(np.uint8)


def workload_c4(n_pairs=64, seed=0):
    """The batched uneven workload: n_pairs distinct pairs, sizes drawn uniformly from C4_SIZES.
    Each pair is its own window of one 4096 x 4096 speckle pattern (a seeded place), under its own
    random affine field (strains within +-2 %, rotation +-1 deg, translation +-8 px), with its own
    noise (sigma 0.005) and a convex 8-gon ROI of its own covering 40-90 %. Returns the list of
    (I1 u8, I2 u8, roi)."""
    rng = np.random.default_rng(seed)
    pat = make_pattern(4096, 4096, seed=seed + 4000, margin=0)
    out = []
    for k in range(n_pairs):
        h, w = C4_SIZES[int(rng.integers(len(C4_SIZES)))]
        ox = int(rng.integers(64, 4096 - w - 64))
        oy = int(rng.integers(64, 4096 - h - 64))
        f = affine_field(*rng.uniform(-0.02, 0.02, 3), rng.uniform(-1.0, 1.0), rng.uniform(-8.0, 8.0, 2),
                         ((w - 1) / 2.0, (h - 1) / 2.0))
        a = np.clip(np.rint((pat[oy:oy + h, ox:ox + w]
                             + np.random.default_rng(10 * k + 1).normal(0.0, 0.005, (h, w))) * 255.0), 0, 255)
        b = render_affine(pat, ox, oy, h, w, f, noise=0.005, seed=10 * k + 2)
        roi = convex_polygon_mask(h, w, seed=k, coverage=float(rng.uniform(0.4, 0.9)), search_step=8)
        out.append((a.astype(np.uint8), b, roi))
    return out


def growing_bowl_and_band_frames_gpu(n, nf, seed=0):
    """The long-sequence workload's frames, rendered with torch on the GPU in f64 (the steps of
    render): frame k carries s = k / (nf - 1) times the bowl-and-band field of an n x n frame (the
    inverse of x = X + u(X) by 30 fixed-point steps, a bilinear sample of the pattern, edges clamped
    as _bilinear), then a brightness drift of -3 % .. +3 %, Gaussian noise of sigma 0.01 (torch's
    generator, seeded per frame) and 8-bit quantisation. Yields host f32 frames one at a time, and
    the truth of the last frame is bowl_and_band scaled by 1."""
    import torch
    dev = torch.device("cuda:0")
    pat = make_pattern(n, n, seed=seed)
    P = torch.from_numpy(pat).to(dev)
    H, W = P.shape
    m = (H - n) // 2
    yy, xx = torch.meshgrid(torch.arange(n, dtype=torch.float64, device=dev),
                            torch.arange(n, dtype=torch.float64, device=dev), indexing="ij")
    c = (n - 1) / 2.0
    a = np.deg2rad(35.0)
    tx, ty, nx, ny = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)

    def field(X, Y, s):
        r2 = (X - c) ** 2 + (Y - c) ** 2
        uy = -6.0 * torch.exp(-r2 / (2 * 300.0 ** 2))
        g = 0.5 * 3.0 * torch.tanh((nx * (X - c) + ny * (Y - c)) / 20.0)
        return s * g * tx, s * (uy + g * ty)
    for k in range(nf):
        s = k / (nf - 1)
        X, Y = xx, yy
        if k:
            for _ in range(30):
                ux, uy = field(X, Y, s)
                X, Y = xx - ux, yy - uy
        xf = torch.clamp(X + m, 0, W - 1.000001)
        yf = torch.clamp(Y + m, 0, H - 1.000001)
        x0, y0 = torch.floor(xf).long(), torch.floor(yf).long()
        fx, fy = xf - x0, yf - y0
        v = ((1 - fy) * ((1 - fx) * P[y0, x0] + fx * P[y0, x0 + 1])
             + fy * ((1 - fx) * P[y0 + 1, x0] + fx * P[y0 + 1, x0 + 1]))
        gen = torch.Generator(device=dev)
        gen.manual_seed(100 + k)
        v = v * (0.97 + 0.06 * s) + 0.01 * torch.randn(v.shape, generator=gen, device=dev, dtype=torch.float64)
        q = torch.clamp(torch.round(v * 255.0), 0, 255)
        yield (q.float() / 255.0).cpu().numpy()


def gaussian_bumps(h, w, n=6, amp=40.0, scale=(150.0, 600.0), seed=0):
    """A sum of n Gaussian-bump vector fields u = a exp(-|X - c|^2 / (2 s^2)) with centres uniform over
    the frame, scales s uniform in `scale` and vectors a of random direction, rescaled together so
    that the largest |u| over the frame is `amp` px."""
    rng = np.random.default_rng(seed)
    cx, cy = rng.uniform(0, w, n), rng.uniform(0, h, n)
    s = rng.uniform(scale[0], scale[1], n)
    ang = rng.uniform(0, 2 * np.pi, n)
    mag = rng.uniform(0.3, 1.0, n)
    ax, ay = mag * np.cos(ang), mag * np.sin(ang)

    def raw(X, Y):
        ux, uy = np.zeros_like(X), np.zeros_like(Y)
        for k in range(n):
            g = np.exp(-((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) / (2.0 * s[k] ** 2))
            ux += ax[k] * g
            uy += ay[k] * g
        return ux, uy
    yy, xx = np.mgrid[0:h:4, 0:w:4].astype(np.float64)
    ux, uy = raw(xx, yy)
    k = amp / float(np.sqrt(ux ** 2 + uy ** 2).max())

    def f(X, Y):
        ux, uy = raw(X, Y)
        return k * ux, k * uy
    return f


def workload_c5(seed=0):
    """A 4000 x 3000 pair with large smooth motion: six Gaussian bumps, |u| up to 40 px, noise sigma
    0.01, full frame. Returns (I1 u8, I2 u8, truth (ux, uy))."""
    h, w = 3000, 4000
    pat = make_pattern(h, w, seed=seed)
    f = gaussian_bumps(h, w, seed=seed + 7)
    a = render(pat, None, noise=0.01, seed=1)
    b = render(pat, f, noise=0.01, seed=2)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    return a, b, f(xx, yy)
