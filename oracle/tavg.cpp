// Temporal averaging and frame conversion of the oracle (test infrastructure only).
//
// Restates rust/crates/optflow-cli/src/tavg.rs:23-207 (python/optflow_ref/tavg.py:11-96): the
// robust affine fit to the coarse POC grid, the alignment of a neighbour frame to the centre frame
// by the cubic B-spline warp (the Rust's warp; the Python's map_coordinates agrees away from the
// border), and the f64 mean in the order centre, then the others. And
// rust/crates/optflow-cli/src/io.rs:44-52, :259-281 (python/optflow_ref/io.py:15-54): grey levels
// in f64 (8-bit / 255, 16-bit / 65535, 0.299 R + 0.587 G + 0.114 B left to right) and the box
// downscale that crops the remainder (row sums first, in order).
#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace orc {

namespace {

constexpr size_t MIN_POINTS = 6;  // tavg.rs:20

// np.median of a non-empty vector
double median_of(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return (n % 2) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// tavg.rs:35-57: 3 x 3 (augmented) elimination with partial pivoting, the pivot being the last
// of equal largest magnitudes (Rust's max_by); false when a pivot is below 1e-12
bool solve3(double a[3][4], double x[3]) {
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int i = c + 1; i < 3; i++)
            if (std::fabs(a[i][c]) >= std::fabs(a[p][c])) p = i;
        if (std::fabs(a[p][c]) < 1e-12) return false;
        if (p != c)
            for (int j = 0; j < 4; j++) std::swap(a[c][j], a[p][j]);
        for (int i = c + 1; i < 3; i++) {
            const double f = a[i][c] / a[c][c];
            for (int j = c; j < 4; j++) a[i][j] -= f * a[c][j];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double s = a[i][3];
        for (int j = i + 1; j < 3; j++) s -= a[i][j] * x[j];
        x[i] = s / a[i][i];
    }
    return true;
}

// tavg.rs:62-82: least squares v = c0 + c1 x + c2 y over the selected points, in coordinates
// centred on their mean, returned in absolute coordinates
bool lstsq(const std::vector<double> &xs, const std::vector<double> &ys, const std::vector<double> &v,
           const std::vector<uint8_t> &sel, double c[3]) {
    double mx = 0, my = 0, n = 0;
    for (size_t i = 0; i < xs.size(); i++)
        if (sel[i]) {
            mx += xs[i];
            my += ys[i];
            n += 1;
        }
    mx /= n;
    my /= n;
    double m[3][4] = {{0}};
    for (size_t i = 0; i < xs.size(); i++) {
        if (!sel[i]) continue;
        const double row[3] = {1.0, xs[i] - mx, ys[i] - my};
        for (int r = 0; r < 3; r++) {
            for (int k = 0; k < 3; k++) m[r][k] += row[r] * row[k];
            m[r][3] += row[r] * v[i];
        }
    }
    double s[3];
    if (!solve3(m, s)) return false;
    c[0] = s[0] - s[1] * mx - s[2] * my;
    c[1] = s[1];
    c[2] = s[2];
    return true;
}

}  // namespace

// tavg.rs:87-110: 4 passes; after each fit the points whose residual exceeds
// max(3 x 1.4826 MAD, 0.05 px) are dropped for the next one; false with fewer than 6 left
bool fit_affine(const std::vector<double> &xs, const std::vector<double> &ys, const std::vector<double> &ux,
                const std::vector<double> &uy, double cx[3], double cy[3]) {
    const size_t n = xs.size();
    std::vector<uint8_t> sel(n, 1);
    bool have = false;
    for (int pass = 0; pass < 4; pass++) {
        size_t cnt = 0;
        for (uint8_t v : sel) cnt += v;
        if (cnt < MIN_POINTS) return false;
        double ax[3], ay[3];
        if (!lstsq(xs, ys, ux, sel, ax) || !lstsq(xs, ys, uy, sel, ay)) return false;
        std::copy(ax, ax + 3, cx);
        std::copy(ay, ay + 3, cy);
        have = true;
        std::vector<double> r(n), rs;
        for (size_t i = 0; i < n; i++) {
            r[i] = std::hypot(ax[0] + ax[1] * xs[i] + ax[2] * ys[i] - ux[i], ay[0] + ay[1] * xs[i] + ay[2] * ys[i] - uy[i]);
            if (sel[i]) rs.push_back(r[i]);
        }
        const double lim = std::max(3.0 * (1.4826 * median_of(rs)), 0.05);
        for (size_t i = 0; i < n; i++) sel[i] = r[i] <= lim;
    }
    return have;
}

// tavg.rs:118-163: `other` resampled onto the grid of `centre` (out, f32 values), or false when no
// motion could be estimated. mode 1: the affine fit to the level-0 POC grid points with a peak
// above 0 (inside `mask` if given); mode 2: the dense coarse field.
bool align_to(const double *centre, const double *other, int h, int w, int mode, const uint8_t *mask,
              const orc_cfg &c, double *out) {
    const size_t n = (size_t)w * h;
    if (std::min(w, h) < c.poc_win) return false;  // no POC grid fits (tavg.rs:290-296)
    std::vector<double> ux(n), uy(n);
    if (mode == 2) {
        std::vector<double> gx(n), gy(n);
        coarse_dense(centre, other, h, w, c, nullptr, nullptr, gx.data(), gy.data(), nullptr, nullptr, nullptr);
        for (size_t i = 0; i < n; i++) {  // stored as f32 images (tavg.rs:128-135)
            ux[i] = (double)(float)gx[i];
            uy[i] = (double)(float)gy[i];
        }
    } else {
        std::vector<double> gux, guy, gpk;
        coarse_dense(centre, other, h, w, c, nullptr, nullptr, nullptr, nullptr, &gux, &guy, &gpk);
        const std::vector<int> gxs = grid_positions(w, c.poc_win, c.grid_step);
        const std::vector<int> gys = grid_positions(h, c.poc_win, c.grid_step);
        std::vector<double> xs, ys, vx, vy;
        for (size_t iy = 0; iy < gys.size(); iy++)
            for (size_t ix = 0; ix < gxs.size(); ix++) {
                const size_t j = iy * gxs.size() + ix;
                // the grid is f32 in the Rust (poc.rs Grid)
                if ((float)gpk[j] > 0.0f && (!mask || mask[(size_t)gys[iy] * w + gxs[ix]])) {
                    xs.push_back(gxs[ix]);
                    ys.push_back(gys[iy]);
                    vx.push_back((double)(float)gux[j]);
                    vy.push_back((double)(float)guy[j]);
                }
            }
        if (xs.size() < MIN_POINTS) return false;
        double cx[3], cy[3];
        if (!fit_affine(xs, ys, vx, vy, cx, cy)) return false;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const size_t i = (size_t)y * w + x;
                ux[i] = (double)(float)(cx[0] + cx[1] * x + cx[2] * y);
                uy[i] = (double)(float)(cy[0] + cy[1] * x + cy[2] * y);
            }
    }
    std::vector<double> cf(other, other + n);
    bspline_real(cf, h, w);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            out[i] = (double)(float)sample_at(cf.data(), h, w, x + ux[i], y + uy[i]);
        }
    return true;
}

}  // namespace orc

using namespace orc;

extern "C" int orc_fit_affine(const double *xs, const double *ys, const double *ux, const double *uy, long n,
                              double *cx, double *cy) {
    return fit_affine(std::vector<double>(xs, xs + n), std::vector<double>(ys, ys + n),
                      std::vector<double>(ux, ux + n), std::vector<double>(uy, uy + n), cx, cy)
               ? 1
               : 0;
}

extern "C" int orc_align_to(const double *centre, const double *other, int h, int w, int mode, const uint8_t *mask,
                            const orc_cfg *cfg, double *out) {
    return align_to(centre, other, h, w, mode, mask, *cfg, out) ? 1 : 0;
}

// tavg.rs:168-207: frames is nf x h x w (f32 values); mode 0 plain, 1 affine, 2 dense; with
// keep_unaligned a neighbour whose motion cannot be estimated enters unaligned. out: the mean as
// f32 values; returns the number of frames in it.
extern "C" int orc_tavg_window(const double *frames, int nf, int centre, int h, int w, int mode, const uint8_t *mask,
                               int keep_unaligned, const orc_cfg *cfg, double *out) {
    const size_t n = (size_t)w * h;
    const double *c = frames + (size_t)centre * n;
    std::vector<std::vector<double>> used;
    for (int i = 0; i < nf; i++) {
        if (i == centre) continue;
        const double *f = frames + (size_t)i * n;
        if (mode == 0) {
            used.emplace_back(f, f + n);
            continue;
        }
        std::vector<double> a(n);
        if (align_to(c, f, h, w, mode, mask, *cfg, a.data()))
            used.push_back(std::move(a));
        else if (keep_unaligned)
            used.emplace_back(f, f + n);
    }
    if (used.empty()) {
        std::copy(c, c + n, out);
        return 1;
    }
    const double cnt = (double)(used.size() + 1);
    for (size_t i = 0; i < n; i++) {
        double s = c[i];
        for (const auto &u : used) s += u[i];
        out[i] = (double)(float)(s / cnt);
    }
    return (int)used.size() + 1;
}

// io.rs:44-52, :259-281: pixels (h x w x channels, 8 or 16 bit) -> f64 grey levels, box
// downscaled by `factor` (the remainder cropped): out is (h / factor) x (w / factor)
extern "C" void orc_to_gray(const void *pixels, int bits, int channels, int h, int w, int factor, double *out) {
    const size_t n = (size_t)w * h;
    std::vector<double> g(n);
    const double scale = bits == 16 ? 65535.0 : 255.0;
    for (size_t i = 0; i < n; i++) {
        auto at = [&](int ch) {
            const size_t k = i * channels + ch;
            return bits == 16 ? ((const uint16_t *)pixels)[k] / scale : ((const uint8_t *)pixels)[k] / scale;
        };
        g[i] = channels >= 3 ? 0.299 * at(0) + 0.587 * at(1) + 0.114 * at(2) : at(0);
    }
    if (factor <= 1) {
        std::copy(g.begin(), g.end(), out);
        return;
    }
    const int ow = w / factor, oh = h / factor;
    const double norm = (double)factor * factor;
    for (int j = 0; j < oh; j++)
        for (int i = 0; i < ow; i++) {
            double sum = 0.0;
            for (int dy = 0; dy < factor; dy++) {
                double row = 0.0;
                const size_t base = (size_t)(j * factor + dy) * w + (size_t)i * factor;
                for (int dx = 0; dx < factor; dx++) row += g[base + dx];
                sum += row;
            }
            out[(size_t)j * ow + i] = sum / norm;
        }
}
