// Coarse stage of the oracle (test infrastructure only): windowed, band-limited phase-only
// correlation on a grid over a Gaussian pyramid, outlier cleaning and fill. Restates
// python/optflow_ref/poc.py:18-239; the grid end for odd windows follows
// rust/crates/optflow-core/src/poc.rs:244-256.
#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace orc {

namespace {

double hann(int i, int n) { return 0.5 - 0.5 * std::cos(2 * M_PI * i / n); }

double parabola(double m, double c, double p) {
    double den = m - 2.0 * c + p;
    return den == 0 ? 0.0 : 0.5 * (m - p) / den;
}

}  // namespace

// poc.py:39-54: shift d with b(x) ~ a(x - d)
void poc_shift(const double *a, const double *b, int h, int w, double *dx, double *dy, double *peak) {
    const size_t n = (size_t)h * w;
    double ma = 0, mb = 0;
    for (size_t i = 0; i < n; i++) { ma += a[i]; mb += b[i]; }
    ma /= n;
    mb /= n;
    std::vector<cd> A(n), B(n);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            double wv = hann(y, h) * hann(x, w);
            A[i] = (a[i] - ma) * wv;
            B[i] = (b[i] - mb) * wv;
        }
    for (int y = 0; y < h; y++) fft_line(&A[(size_t)y * w], w, false), fft_line(&B[(size_t)y * w], w, false);
    std::vector<cd> col(h);
    for (auto *M : {&A, &B})
        for (int x = 0; x < w; x++) {
            for (int y = 0; y < h; y++) col[y] = (*M)[(size_t)y * w + x];
            fft_line(col.data(), h, false);
            for (int y = 0; y < h; y++) (*M)[(size_t)y * w + x] = col[y];
        }
    double lpsum = 0;
    for (int y = 0; y < h; y++) {
        int ky = (y <= (h - 1) / 2) ? y : y - h;
        bool oky = std::fabs((double)ky / h) <= 0.25;
        for (int x = 0; x < w; x++) {
            int kx = (x <= (w - 1) / 2) ? x : x - w;
            bool okx = std::fabs((double)kx / w) <= 0.25;
            size_t i = (size_t)y * w + x;
            cd c = B[i] * std::conj(A[i]);
            c /= std::abs(c) + 1e-12;
            double lp = (oky && okx) ? 1.0 : 0.0;
            lpsum += lp;
            A[i] = c * lp;
        }
    }
    const double lpm = lpsum / n;
    std::vector<cd> &C = A;
    for (int y = 0; y < h; y++) fft_line(&C[(size_t)y * w], w, true);
    for (int x = 0; x < w; x++) {
        for (int y = 0; y < h; y++) col[y] = C[(size_t)y * w + x];
        fft_line(col.data(), h, true);
        for (int y = 0; y < h; y++) C[(size_t)y * w + x] = col[y];
    }
    std::vector<double> r(n);
    const int sh = (h + 1) / 2, sw = (w + 1) / 2;  // numpy.fft.fftshift
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            r[(size_t)y * w + x] = C[(size_t)((y + sh) % h) * w + (x + sw) % w].real() / lpm;
    // np.argmax: first maximum, a NaN wins
    size_t best = 0;
    for (size_t i = 0; i < n; i++) {
        if (std::isnan(r[i])) { best = i; break; }
        if (r[i] > r[best]) best = i;
    }
    int iy = (int)(best / w), ix = (int)(best % w);
    iy = std::min(std::max(iy, 1), h - 2);
    ix = std::min(std::max(ix, 1), w - 2);
    double sx = parabola(r[(size_t)iy * w + ix - 1], r[(size_t)iy * w + ix], r[(size_t)iy * w + ix + 1]);
    double sy = parabola(r[(size_t)(iy - 1) * w + ix], r[(size_t)iy * w + ix], r[(size_t)(iy + 1) * w + ix]);
    *dx = ix - w / 2 + sx;
    *dy = iy - h / 2 + sy;
    *peak = r[(size_t)iy * w + ix];
}

namespace {

int pow2_floor(int n) {
    if (n < 1) return 0;
    int p = 1;
    while (p * 2 <= n) p *= 2;
    return p;
}

// rust poc.rs:244-256 (equal to poc.py:61-64 for even windows)
std::vector<int> grid_axis(int n, int win, int step) {
    std::vector<int> out;
    int stop = n + 1 - (win - win / 2);
    int start = win / 2;
    if (start >= stop) return out;
    int count = (stop - start - 1) / step + 1;
    for (int k = 0; k < count; k++) out.push_back(start + k * step);
    return out;
}

struct Grid {
    std::vector<int> xs, ys;
    std::vector<double> ux, uy, peak;  // ny x nx
};

std::vector<double> window(const double *img, int W, int y0, int x0, int wy, int wx) {
    std::vector<double> o((size_t)wy * wx);
    for (int y = 0; y < wy; y++)
        for (int x = 0; x < wx; x++) o[(size_t)y * wx + x] = img[(size_t)(y0 + y) * W + x0 + x];
    return o;
}

// poc.py:67-92
Grid poc_grid(const double *I1, const double *I2, int H, int W, const std::vector<int> &xs,
              const std::vector<int> &ys, int wy, int wx, const std::vector<double> *ix,
              const std::vector<double> *iy, int iters = 3) {
    Grid g;
    g.xs = xs;
    g.ys = ys;
    const int ny = (int)ys.size(), nx = (int)xs.size();
    g.ux.assign((size_t)ny * nx, 0);
    g.uy.assign((size_t)ny * nx, 0);
    g.peak.assign((size_t)ny * nx, 0);
    const int hy = wy / 2, hx = wx / 2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < ny * nx; k++) {
        const int i = k / nx, j = k % nx;
        const int y = ys[i], x = xs[j];
        double dx = ix ? (*ix)[k] : 0.0, dy = iy ? (*iy)[k] : 0.0;
        std::vector<double> a = window(I1, W, y - hy, x - hx, wy, wx);
        double pk = 0.0;
        for (int t = 0; t < iters; t++) {
            const long ox = (long)std::nearbyint(dx), oy = (long)std::nearbyint(dy);
            const long y0 = y - hy + oy, x0 = x - hx + ox;
            if (y0 < 0 || x0 < 0 || y0 + wy > H || x0 + wx > W) { pk = 0.0; break; }
            std::vector<double> b = window(I2, W, (int)y0, (int)x0, wy, wx);
            double sx, sy;
            poc_shift(a.data(), b.data(), wy, wx, &sx, &sy, &pk);
            dx = ox + sx;
            dy = oy + sy;
            if (std::fabs(sx) < 0.5 && std::fabs(sy) < 0.5) break;
        }
        g.ux[k] = dx;
        g.uy[k] = dy;
        g.peak[k] = pk;
    }
    return g;
}

// poc.py:95-140
Grid clean(const Grid &g, double peak_min, double thresh, const std::vector<double> *fx,
           const std::vector<double> *fy, int max_pass = 100) {
    const int ny = (int)g.ys.size(), nx = (int)g.xs.size();
    const size_t n = (size_t)ny * nx;
    std::vector<double> ux = g.ux, uy = g.uy;
    std::vector<uint8_t> valid(n), out;
    for (size_t k = 0; k < n; k++) valid[k] = g.peak[k] >= peak_min;
    static const int offs[8][2] = {{-1, -1}, {-1, 0}, {-1, 1}, {0, -1}, {0, 1}, {1, -1}, {1, 0}, {1, 1}};
    auto neighbours = [&](int i, int j, const std::vector<uint8_t> &ok, int *nb) {
        int c = 0;
        for (auto &o : offs) {
            int a = i + o[0], b = j + o[1];
            if (a >= 0 && a < ny && b >= 0 && b < nx && ok[(size_t)a * nx + b]) nb[c++] = a * nx + b;
        }
        return c;
    };
    out = valid;
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            size_t k = (size_t)i * nx + j;
            if (!valid[k]) continue;
            int nb[8];
            int c = neighbours(i, j, valid, nb);
            if (c < 3) continue;
            for (auto *comp : {&ux, &uy}) {
                std::vector<double> vals(c), dev(c);
                for (int q = 0; q < c; q++) vals[q] = (*comp)[nb[q]];
                double med = median(vals);
                for (int q = 0; q < c; q++) dev[q] = std::fabs(vals[q] - med);
                double res = median(dev);
                if (std::fabs((*comp)[k] - med) / (res + 0.1) > thresh) out[k] = 0;
            }
        }
    bool any = false;
    for (auto v : out) any |= v != 0;
    Grid r;
    r.xs = g.xs;
    r.ys = g.ys;
    if (!any) {
        r.ux = fx ? *fx : std::vector<double>(n, 0.0);
        r.uy = fy ? *fy : std::vector<double>(n, 0.0);
        r.peak.assign(n, 0.0);
        for (size_t k = 0; k < n; k++) r.peak[k] = g.peak[k] * 0;
        return r;
    }
    for (int pass = 0; pass < max_pass; pass++) {
        bool all = true;
        for (auto v : out) all &= v != 0;
        if (all) break;
        std::vector<uint8_t> nw = out;
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) {
                size_t k = (size_t)i * nx + j;
                if (out[k]) continue;
                int nb[8];
                int c = neighbours(i, j, out, nb);
                if (c) {
                    std::vector<double> vx(c), vy(c);
                    for (int q = 0; q < c; q++) { vx[q] = ux[nb[q]]; vy[q] = uy[nb[q]]; }
                    ux[k] = median(vx);
                    uy[k] = median(vy);
                    nw[k] = 1;
                }
            }
        out = nw;
    }
    r.ux = ux;
    r.uy = uy;
    r.peak.resize(n);
    for (size_t k = 0; k < n; k++) r.peak[k] = valid[k] ? g.peak[k] : 0.0;
    return r;
}

// poc.py:143-161: bilinear interpolation of the grid field, clamped to the grid
void grid_frac(double q, const std::vector<int> &nodes, double &t, int &i0) {
    if (nodes.size() == 1) { t = 0; i0 = 0; return; }
    double f = (q - nodes[0]) / (double)(nodes[1] - nodes[0]);
    f = std::min(std::max(f, 0.0), (double)nodes.size() - 1);
    i0 = std::min((int)std::floor(f), (int)nodes.size() - 2);
    t = f - i0;
}

void grid_sample(const Grid &g, const std::vector<double> &gux, const std::vector<double> &guy, double xq,
                 double yq, double &ox, double &oy) {
    double tx, ty;
    int x0, y0;
    grid_frac(xq, g.xs, tx, x0);
    grid_frac(yq, g.ys, ty, y0);
    const int nx = (int)g.xs.size(), ny = (int)g.ys.size();
    int x1 = std::min(x0 + 1, nx - 1), y1 = std::min(y0 + 1, ny - 1);
    auto bil = [&](const std::vector<double> &a) {
        return (1 - ty) * ((1 - tx) * a[(size_t)y0 * nx + x0] + tx * a[(size_t)y0 * nx + x1]) +
               ty * ((1 - tx) * a[(size_t)y1 * nx + x0] + tx * a[(size_t)y1 * nx + x1]);
    };
    ox = bil(gux);
    oy = bil(guy);
}

std::vector<double> downsample(const std::vector<double> &img, int h, int w, int &nh, int &nw) {
    std::vector<double> b((size_t)h * w);
    gauss_blur(img.data(), h, w, 1.0, b.data());
    nh = (h + 1) / 2;
    nw = (w + 1) / 2;
    std::vector<double> o((size_t)nh * nw);
    for (int y = 0; y < nh; y++)
        for (int x = 0; x < nw; x++) o[(size_t)y * nw + x] = b[(size_t)(2 * y) * w + 2 * x];
    return o;
}

}  // namespace

// the grid positions along one axis (public form of grid_axis for the other oracle files)
std::vector<int> grid_positions(int n, int win, int step) { return grid_axis(n, win, step); }

// poc.py:172-187
int level_window(int h, int w, int win, int &wy, int &wx) {
    int lim_h = pow2_floor(2 * h / 3), lim_w = pow2_floor(2 * w / 3);
    if (win <= lim_h && win <= lim_w) { wy = wx = win; return 1; }
    if (h <= w) {
        wy = std::min(win, lim_h);
        wx = std::min(lim_w, std::max(win, win * win / std::max(wy, 1)));
    } else {
        wx = std::min(win, lim_w);
        wy = std::min(lim_h, std::max(win, win * win / std::max(wx, 1)));
    }
    return std::min(wy, wx) >= 16;
}

// poc.py:190-195
int n_levels(double max_disp, int win, int h, int w) {
    double v = std::ceil(std::log2(std::max(max_disp, 1e-9) / (win / 4.0)));
    int L = (int)std::max(0.0, v);
    int wy, wx;
    while (L > 0 && !level_window(h >> L, w >> L, win, wy, wx)) L--;
    return L;
}

// poc.py:216-239 plus the dense start of track.py:131-137
void coarse_dense(const double *I1, const double *I2, int H, int W, const orc_cfg &c, const double *init_x,
                  const double *init_y, double *gx_out, double *gy_out, std::vector<double> *grid_ux,
                  std::vector<double> *grid_uy, std::vector<double> *grid_peak) {
    std::vector<int> xs = grid_axis(W, c.poc_win, c.grid_step), ys = grid_axis(H, c.poc_win, c.grid_step);
    const int L = n_levels(c.max_disp, c.poc_win, H, W);
    std::vector<std::vector<double>> pa(1, std::vector<double>(I1, I1 + (size_t)H * W));
    std::vector<std::vector<double>> pb(1, std::vector<double>(I2, I2 + (size_t)H * W));
    std::vector<int> ph(1, H), pw(1, W);
    for (int l = 0; l < L; l++) {
        int nh, nw;
        pa.push_back(downsample(pa.back(), ph.back(), pw.back(), nh, nw));
        pb.push_back(downsample(pb.back(), ph.back(), pw.back(), nh, nw));
        ph.push_back(nh);
        pw.push_back(nw);
    }
    Grid g;
    bool have = false;
    for (int lev = L; lev >= 0; lev--) {
        const int h = ph[lev], w = pw[lev];
        int wy = c.poc_win, wx = c.poc_win;
        std::vector<int> gxs = xs, gys = ys;
        if (lev > 0) {
            level_window(h, w, c.poc_win, wy, wx);
            gxs = grid_axis(w, wx, c.grid_step);
            gys = grid_axis(h, wy, c.grid_step);
        }
        const size_t n = gxs.size() * gys.size();
        std::vector<double> ix(n), iy(n);
        bool use_init = false;
        if (have) {
            for (size_t i = 0; i < gys.size(); i++)
                for (size_t j = 0; j < gxs.size(); j++) {
                    double ox, oy;
                    grid_sample(g, g.ux, g.uy, gxs[j] / 2.0, gys[i] / 2.0, ox, oy);
                    ix[i * gxs.size() + j] = 2.0 * ox;
                    iy[i * gxs.size() + j] = 2.0 * oy;
                }
            use_init = true;
        } else if (init_x) {
            const int sc = 1 << lev;
            for (size_t i = 0; i < gys.size(); i++)
                for (size_t j = 0; j < gxs.size(); j++) {
                    size_t p = (size_t)(gys[i] * sc) * W + gxs[j] * sc;
                    ix[i * gxs.size() + j] = init_x[p] / sc;
                    iy[i * gxs.size() + j] = init_y[p] / sc;
                }
            use_init = true;
        }
        Grid raw = poc_grid(pa[lev].data(), pb[lev].data(), h, w, gxs, gys, wy, wx, use_init ? &ix : nullptr,
                            use_init ? &iy : nullptr);
        g = clean(raw, c.poc_peak_min, c.median_thresh, use_init ? &ix : nullptr, use_init ? &iy : nullptr);
        have = true;
    }
    if (grid_ux) { *grid_ux = g.ux; *grid_uy = g.uy; *grid_peak = g.peak; }
    if (!gx_out) return;
    std::vector<double> cx, cy;
    if (init_x) {
        const size_t nx = g.xs.size();
        cx.resize(g.ux.size());
        cy.resize(g.uy.size());
        for (size_t i = 0; i < g.ys.size(); i++)
            for (size_t j = 0; j < nx; j++) {
                size_t p = (size_t)g.ys[i] * W + g.xs[j];
                cx[i * nx + j] = g.ux[i * nx + j] - init_x[p];
                cy[i * nx + j] = g.uy[i * nx + j] - init_y[p];
            }
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            double ox, oy;
            if (init_x) {
                grid_sample(g, cx, cy, x, y, ox, oy);
                gx_out[p] = init_x[p] + ox;
                gy_out[p] = init_y[p] + oy;
            } else {
                grid_sample(g, g.ux, g.uy, x, y, ox, oy);
                gx_out[p] = ox;
                gy_out[p] = oy;
            }
        }
}

}  // namespace orc

using namespace orc;

extern "C" void orc_poc_shift(const double *a, const double *b, int h, int w, double *dx, double *dy,
                              double *peak) {
    poc_shift(a, b, h, w, dx, dy, peak);
}

extern "C" int orc_level_window(int h, int w, int win, int *wy, int *wx) {
    return level_window(h, w, win, *wy, *wx);
}

extern "C" int orc_n_levels(double max_disp, int win, int h, int w) { return n_levels(max_disp, win, h, w); }

extern "C" void orc_grid_dims(int h, int w, const orc_cfg *cfg, int *nx, int *ny) {
    *nx = (int)grid_axis(w, cfg->poc_win, cfg->grid_step).size();
    *ny = (int)grid_axis(h, cfg->poc_win, cfg->grid_step).size();
}

extern "C" void orc_coarse_dense(const double *i1, const double *i2, int h, int w, const orc_cfg *cfg,
                                 const double *init_ux, const double *init_uy, double *ux, double *uy) {
    coarse_dense(i1, i2, h, w, *cfg, init_ux, init_uy, ux, uy, nullptr, nullptr, nullptr);
}

extern "C" int orc_coarse_flow(const double *i1, const double *i2, int h, int w, const orc_cfg *cfg,
                               const double *init_ux, const double *init_uy, double *gux, double *guy,
                               double *peak) {
    std::vector<double> a, b, c;
    coarse_dense(i1, i2, h, w, *cfg, init_ux, init_uy, nullptr, nullptr, &a, &b, &c);
    std::memcpy(gux, a.data(), a.size() * sizeof(double));
    std::memcpy(guy, b.data(), b.size() * sizeof(double));
    std::memcpy(peak, c.data(), c.size() * sizeof(double));
    return 0;
}
