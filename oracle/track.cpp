// Pair and sequence pipeline of the oracle (test infrastructure only): coarse stage, fine stage,
// guard and strain (python/optflow_ref/track.py:45-146), the chained sequence with reference
// updates, composition and link confidence (track.py:231-328), and the hybrid sequence
// (rust/crates/optflow-core/src/track/hybrid.rs, track.py:331-505).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace orc {

void refine(const double *i1, const double *i2, int h, int w, const double *ux0, const double *uy0,
            const orc_cfg &c, double *ux_out, double *uy_out, double *conf_out, double *resid_out,
            double *energy_out);
void coarse_dense(const double *I1, const double *I2, int H, int W, const orc_cfg &c, const double *init_x,
                  const double *init_y, double *gx_out, double *gy_out, std::vector<double> *grid_ux,
                  std::vector<double> *grid_uy, std::vector<double> *grid_peak);
void displacement_gradient(const double *ux, const double *uy, int h, int w, int win, double *dudx,
                           double *dudy, double *dvdx, double *dvdy);
void strain_fields(const double *dudx, const double *dudy, const double *dvdx, const double *dvdy, size_t n,
                   double *exx, double *eyy, double *exy, double *gmax, double *vol);

namespace {

struct Pair {
    std::vector<double> ux, uy, conf, raw_ux, raw_uy;
    std::vector<double> exx, eyy, exy, gmax, vol;
    double t_coarse = 0, t_fine = 0, t_strain = 0;
};

// track.py:56-108 without the crack options: validity, guard band, strain
void finish(const double *ux, const double *uy, const double *conf, int h, int w, const orc_cfg &c,
            const uint8_t *mask, Pair &out) {
    Cfg C(c);
    const size_t n = (size_t)h * w;
    std::vector<uint8_t> low(n), bad(n);
    for (size_t i = 0; i < n; i++) {
        low[i] = conf[i] < c.conf_min;
        bad[i] = low[i] || (mask && !mask[i]);
    }
    const double guard = C.edge_guard_px();
    if (guard > 0) {
        std::vector<uint8_t> src(n);
        large_components(low.data(), h, w, C.edge_guard_min_area_px(), src.data());
        bool any = false;
        for (size_t i = 0; i < n; i++) {
            if (mask && !mask[i]) src[i] = 1;
            any |= src[i] != 0;
        }
        if (any) {
            std::vector<uint8_t> notsrc(n);
            for (size_t i = 0; i < n; i++) notsrc[i] = !src[i];
            std::vector<double> d(n);
            edt(notsrc.data(), h, w, d.data(), nullptr);
            for (size_t i = 0; i < n; i++) bad[i] |= d[i] <= guard;
        }
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int bd = std::min(std::min(x, y), std::min(w - 1 - x, h - 1 - y));
                if (bd < guard) bad[(size_t)y * w + x] = 1;
            }
    }
    out.ux.resize(n);
    out.uy.resize(n);
    for (size_t i = 0; i < n; i++) {
        out.ux[i] = bad[i] ? NAN : ux[i];
        out.uy[i] = bad[i] ? NAN : uy[i];
    }
    std::vector<double> g[4];
    for (auto &v : g) v.resize(n);
    displacement_gradient(out.ux.data(), out.uy.data(), h, w, c.strain_win, g[0].data(), g[1].data(),
                          g[2].data(), g[3].data());
    out.exx.resize(n); out.eyy.resize(n); out.exy.resize(n); out.gmax.resize(n); out.vol.resize(n);
    strain_fields(g[0].data(), g[1].data(), g[2].data(), g[3].data(), n, out.exx.data(), out.eyy.data(),
                  out.exy.data(), out.gmax.data(), out.vol.data());
}

// ---------------------------------------------------------------- tiled fine stage
// rust/crates/optflow-core/src/tile.rs:17-146 (geometry) and :225-500 (refine_tiled_from_on,
// join_tiles): cores of `tile` px that partition the frame, each refined on its extension (the core
// plus tile_margin, clipped) against the window of image 2 that the extension's start field
// reaches; only the cores are kept, and the confidence is formed afterwards from the joined
// energy (median over the whole frame) and the joined field (border rule).
struct Rect {
    int x0, y0, x1, y1;
    long area() const { return (long)(x1 - x0) * (y1 - y0); }
};
struct Tile {
    Rect core, ext;
};

// tile.rs:38-48
void tile_axis(int n, int t, int m, int k, int &c0, int &c1, int &e0, int &e1) {
    c0 = k * t;
    c1 = std::min(n, c0 + t);
    e0 = std::max(0, c0 - m);
    e1 = std::min(n, c1 + m);
}

// tile.rs:53-76: row by row
std::vector<Tile> tiles(int w, int h, int t, int m) {
    std::vector<Tile> out;
    const int nx = (w + t - 1) / t, ny = (h + t - 1) / t;
    for (int ky = 0; ky < ny; ky++)
        for (int kx = 0; kx < nx; kx++) {
            Tile tl;
            tile_axis(w, t, m, kx, tl.core.x0, tl.core.x1, tl.ext.x0, tl.ext.x1);
            tile_axis(h, t, m, ky, tl.core.y0, tl.core.y1, tl.ext.y0, tl.ext.y1);
            out.push_back(tl);
        }
    return out;
}

// tile.rs:80-95
int tile_margin(const orc_cfg &c) {
    const Cfg C(c);
    return std::max({64, (int)std::ceil(4.0 * c.sigma_w), (int)std::ceil(4.0 * c.sigma_s),
                     (int)std::ceil(std::max(C.edge_guard_px(), 0.0)), c.strain_win});
}

// tile.rs:112-146: the extension moved by the finite range of its start field, widened by the
// margin, clipped (never empty)
Rect displaced_window(const Rect &ext, const double *ux, const double *uy, int ew, int eh, int w, int h, int m) {
    auto range = [&](const double *u, double &lo, double &hi) {
        lo = INFINITY;
        hi = -INFINITY;
        for (long i = 0; i < (long)ew * eh; i++) {
            // the start field is f32 in the reference
            const double v = (double)(float)u[i];
            if (std::isfinite(v)) {
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
        }
        if (lo > hi) lo = hi = 0.0;
    };
    auto axis = [&](int first, int end, double lo, double hi, int n, int &a, int &b) {
        double fa = std::floor(first + lo) - m;
        double fb = std::ceil((end - 1) + hi) + 1.0 + m;
        fa = std::min(std::max(fa, 0.0), (double)(n - 1));
        a = (int)fa;
        fb = std::min(std::max(fb, (double)(a + 1)), (double)n);
        b = (int)fb;
    };
    double lx, hx, ly, hy;
    range(ux, lx, hx);
    range(uy, ly, hy);
    Rect r;
    axis(ext.x0, ext.x1, lx, hx, w, r.x0, r.x1);
    axis(ext.y0, ext.y1, ly, hy, h, r.y0, r.y1);
    return r;
}

std::vector<double> crop(const double *a, int w, const Rect &r) {
    const int rw = r.x1 - r.x0;
    std::vector<double> out((size_t)rw * (r.y1 - r.y0));
    for (int y = r.y0; y < r.y1; y++)
        std::copy(a + (size_t)y * w + r.x0, a + (size_t)y * w + r.x1, out.begin() + (size_t)(y - r.y0) * rw);
    return out;
}

// The fine stage of a pair from the dense start (gx, gy): whole frame, or tiled when the frame has
// a side over c.tile (track.rs:580-598)
void fine_stage(const double *i1, const double *i2, int h, int w, const double *gx, const double *gy,
                const orc_cfg &c, double *ux, double *uy, double *conf) {
    if (c.tile <= 0 || std::max(w, h) <= c.tile) {
        refine(i1, i2, h, w, gx, gy, c, ux, uy, conf, nullptr, nullptr);
        return;
    }
    const int m = tile_margin(c);
    const std::vector<Tile> ts = tiles(w, h, c.tile, m);
    std::vector<size_t> order(ts.size());
    for (size_t k = 0; k < ts.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ts[a].ext.area() > ts[b].ext.area(); });
    const size_t n = (size_t)w * h;
    std::vector<double> energy(n), resid(n);
    for (size_t k : order) {
        const Tile &t = ts[k];
        const int ew = t.ext.x1 - t.ext.x0, eh = t.ext.y1 - t.ext.y0;
        const std::vector<double> sx = crop(gx, w, t.ext), sy = crop(gy, w, t.ext);
        const Rect win = displaced_window(t.ext, sx.data(), sy.data(), ew, eh, w, h, m);
        const std::vector<double> a = crop(i1, w, t.ext), b = crop(i2, w, win);
        const size_t ne = (size_t)ew * eh;
        std::vector<double> tux(ne), tuy(ne), tconf(ne), tres(ne), ten(ne);
        refine_shifted(a.data(), eh, ew, b.data(), win.y1 - win.y0, win.x1 - win.x0, t.ext.x0 - win.x0,
                       t.ext.y0 - win.y0, sx.data(), sy.data(), c, tux.data(), tuy.data(), tconf.data(), tres.data(),
                       ten.data());
        for (int y = t.core.y0; y < t.core.y1; y++)
            for (int x = t.core.x0; x < t.core.x1; x++) {
                const size_t l = (size_t)(y - t.ext.y0) * ew + (x - t.ext.x0), i = (size_t)y * w + x;
                ux[i] = tux[l];
                uy[i] = tuy[l];
                resid[i] = tres[l];
                energy[i] = ten[l];
            }
    }
    confidence_of(energy.data(), resid.data(), ux, uy, h, w, c, conf);
}

// track.py:115-146
int pair(const double *i1, const double *i2, int h, int w, const uint8_t *mask, const double *ix,
         const double *iy, const orc_cfg &c, Pair &out) {
    const size_t n = (size_t)h * w;
    if (mask) {
        bool any = false;
        for (size_t i = 0; i < n; i++) any |= mask[i] != 0;
        if (!any) { set_error("ROI mask is empty"); return 1; }
    }
    if (std::min(h, w) < c.poc_win) { set_error("image smaller than poc_win"); return 1; }
    double t0 = omp_get_wtime();
    std::vector<double> gx(n), gy(n);
    coarse_dense(i1, i2, h, w, c, ix, iy, gx.data(), gy.data(), nullptr, nullptr, nullptr);
    double t1 = omp_get_wtime();
    out.raw_ux.resize(n);
    out.raw_uy.resize(n);
    out.conf.resize(n);
    fine_stage(i1, i2, h, w, gx.data(), gy.data(), c, out.raw_ux.data(), out.raw_uy.data(), out.conf.data());
    double t2 = omp_get_wtime();
    finish(out.raw_ux.data(), out.raw_uy.data(), out.conf.data(), h, w, c, mask, out);
    double t3 = omp_get_wtime();
    out.t_coarse = t1 - t0;
    out.t_fine = t2 - t1;
    out.t_strain = t3 - t2;
    return 0;
}

void store(const Pair &p, size_t n, orc_result *r) {
    auto put = [&](double *dst, const std::vector<double> &src) {
        if (dst) std::memcpy(dst, src.data(), n * sizeof(double));
    };
    put(r->ux, p.ux); put(r->uy, p.uy); put(r->conf, p.conf);
    put(r->raw_ux, p.raw_ux); put(r->raw_uy, p.raw_uy);
    put(r->exx, p.exx); put(r->eyy, p.eyy); put(r->exy, p.exy); put(r->gmax, p.gmax); put(r->vol, p.vol);
    r->t_coarse = p.t_coarse;
    r->t_fine = p.t_fine;
    r->t_strain = p.t_strain;
}

// track.py:231-242: bilinear sample with NaN propagation, NaN outside the image
double sample_bilinear(const double *a, int h, int w, double x, double y) {
    bool out_of = !(std::isfinite(x) && std::isfinite(y)) || x < 0 || x > w - 1 || y < 0 || y > h - 1;
    if (out_of) return NAN;
    double xc = std::min(std::max(x, 0.0), (double)w - 1), yc = std::min(std::max(y, 0.0), (double)h - 1);
    int x0 = std::min((int)std::floor(xc), w - 2), y0 = std::min((int)std::floor(yc), h - 2);
    double tx = xc - x0, ty = yc - y0;
    const double *r0 = a + (size_t)y0 * w, *r1 = r0 + w;
    return (1 - ty) * ((1 - tx) * r0[x0] + tx * r0[x0 + 1]) + ty * ((1 - tx) * r1[x0] + tx * r1[x0 + 1]);
}

// track.py:245-250
void compose(const std::vector<double> &ax, const std::vector<double> &ay, const std::vector<double> &bx,
             const std::vector<double> &by, int h, int w, std::vector<double> &ox, std::vector<double> &oy) {
    const size_t n = (size_t)h * w;
    std::vector<double> rx(n), ry(n);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            double px = x + ax[i], py = y + ay[i];
            rx[i] = ax[i] + sample_bilinear(bx.data(), h, w, px, py);
            ry[i] = ay[i] + sample_bilinear(by.data(), h, w, px, py);
        }
    ox.swap(rx);
    oy.swap(ry);
}

// track.py:253-257
std::vector<double> conf_at(const std::vector<double> &conf, const std::vector<double> &ux,
                            const std::vector<double> &uy, int h, int w) {
    std::vector<double> o((size_t)h * w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            double v = sample_bilinear(conf.data(), h, w, x + ux[i], y + uy[i]);
            o[i] = std::isfinite(v) ? v : 0.0;
        }
    return o;
}

double border_distance(double x, double y, int w, int h) {
    return std::min(std::min(x, y), std::min(w - 1 - x, h - 1 - y));
}

// track.py:317-328 with phaseflow.reportable_flow
double mean_conf(const Pair &leg, const std::vector<double> &ux, const std::vector<double> &uy, const uint8_t *mask,
                 int h, int w, const orc_cfg &c) {
    const size_t n = (size_t)h * w;
    const double B = Cfg(c).edge_guard_px() > 0 ? 4.0 * c.sigma_w : 0.0;
    std::vector<double> rep(n);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            rep[i] = (border_distance(x, y, w, h) >= B &&
                      border_distance(x + leg.raw_ux[i], y + leg.raw_uy[i], w, h) >= B) ? 1.0 : 0.0;
        }
    double s = 0;
    size_t cnt = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            double X = x + ux[i], Y = y + uy[i];
            double cv = sample_bilinear(leg.conf.data(), h, w, X, Y);
            double rv = sample_bilinear(rep.data(), h, w, X, Y);
            if (std::isfinite(cv) && rv > 0.5 && (!mask || mask[i])) { s += cv; cnt++; }
        }
    return cnt ? s / cnt : 0.0;
}

// track.py:260-315 (seq_mode chain)
int sequence_chain(const double *frames, int nf, int h, int w, const uint8_t *mask, const orc_cfg &c,
                   orc_result *res) {
    const size_t n = (size_t)h * w;
    const double *ref = frames, *prev = frames;
    int ref_idx = 0;
    std::vector<double> u0x(n, 0.0), u0y(n, 0.0), c0r(n, 1.0);
    Pair prev_res;
    bool have_prev = false;
    for (int k = 1; k < nf; k++) {
        const double *fk = frames + (size_t)k * n;
        Pair leg;
        int rc = pair(ref, fk, h, w, nullptr, have_prev ? prev_res.raw_ux.data() : nullptr,
                      have_prev ? prev_res.raw_uy.data() : nullptr, c, leg);
        if (rc) return rc;
        double cm = mean_conf(leg, u0x, u0y, mask, h, w, c);
        if (cm < c.ref_update_conf && ref_idx != k - 1) {
            std::vector<double> ca = conf_at(prev_res.conf, u0x, u0y, h, w);
            for (size_t i = 0; i < n; i++) c0r[i] = std::min(c0r[i], ca[i]);
            compose(u0x, u0y, prev_res.raw_ux, prev_res.raw_uy, h, w, u0x, u0y);
            ref = prev;
            ref_idx = k - 1;
            leg = Pair();
            rc = pair(ref, fk, h, w, nullptr, nullptr, nullptr, c, leg);
            if (rc) return rc;
            cm = mean_conf(leg, u0x, u0y, mask, h, w, c);
        }
        std::vector<double> cx, cy;
        compose(u0x, u0y, leg.raw_ux, leg.raw_uy, h, w, cx, cy);
        std::vector<double> ca = conf_at(leg.conf, u0x, u0y, h, w), conf0(n);
        for (size_t i = 0; i < n; i++) conf0[i] = std::min(c0r[i], ca[i]);
        Pair out;
        finish(cx.data(), cy.data(), conf0.data(), h, w, c, mask, out);
        out.conf = conf0;
        out.raw_ux = cx;
        out.raw_uy = cy;
        out.t_coarse = leg.t_coarse;
        out.t_fine = leg.t_fine;
        out.t_strain = leg.t_strain;
        store(out, n, &res[k - 1]);
        res[k - 1].reference = ref_idx;
        res[k - 1].mean_conf = cm;
        prev = fk;
        prev_res = leg;
        have_prev = true;
    }
    return 0;
}


// ---------------------------------------------------------------- seq_mode "hybrid"
// rust/crates/optflow-core/src/track/hybrid.rs (python/optflow_ref/track.py:331-505)

// Sum over the (2r + 1)^2 window, zero outside the image, by running sums along rows then columns
// (hybrid.rs:67-112, the order of its additions).
std::vector<double> box_sum(const std::vector<double> &a, int w, int h, int r) {
    std::vector<double> rows(a.size()), out(a.size());
    for (int y = 0; y < h; y++) {
        const double *in = a.data() + (size_t)y * w;
        double *o = rows.data() + (size_t)y * w;
        double s = 0.0;
        for (int x = 0; x < std::min(r, w); x++) s += in[x];
        for (int x = 0; x < w; x++) {
            if (x + r < w) s += in[x + r];
            if (x > r) s -= in[x - r - 1];
            o[x] = s;
        }
    }
    std::vector<double> acc(w, 0.0);
    for (int y = 0; y < std::min(r, h); y++)
        for (int x = 0; x < w; x++) acc[x] += rows[(size_t)y * w + x];
    for (int y = 0; y < h; y++) {
        if (y + r < h)
            for (int x = 0; x < w; x++) acc[x] += rows[(size_t)(y + r) * w + x];
        if (y > r)
            for (int x = 0; x < w; x++) acc[x] -= rows[(size_t)(y - r - 1) * w + x];
        std::copy(acc.begin(), acc.end(), out.begin() + (size_t)y * w);
    }
    return out;
}

// m c = b for two right-hand sides, Gaussian elimination with partial pivoting (the first row of
// largest |pivot| wins ties as Rust's max_by keeps the last maximum); false if singular
// (hybrid.rs:114-145).
bool solve3(double m[3][3], double b[3][2], double x[3][2]) {
    double a[3][5];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) a[i][j] = m[i][j];
        a[i][3] = b[i][0];
        a[i][4] = b[i][1];
    }
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int i = c + 1; i < 3; i++)
            if (!(std::fabs(a[i][c]) < std::fabs(a[p][c]))) p = i;  // max_by: the last of equal maxima
        if (a[p][c] == 0.0 || !std::isfinite(a[p][c])) return false;
        for (int j = 0; j < 5; j++) std::swap(a[c][j], a[p][j]);
        for (int i = c + 1; i < 3; i++) {
            const double f = a[i][c] / a[c][c];
            for (int j = c; j < 5; j++) a[i][j] -= f * a[c][j];
        }
    }
    for (int k = 0; k < 2; k++)
        for (int i = 2; i >= 0; i--) {
            double s = a[i][3 + k];
            for (int j = i + 1; j < 3; j++) s -= a[i][j] * x[j][k];
            x[i][k] = s / a[i][i];
        }
    return true;
}

// hybrid.rs:147-260 (_local_affine_agreement, track.py:331-380)
std::vector<uint8_t> local_affine_agreement(int w, int h, const std::vector<double> &ux, const std::vector<double> &uy,
                                            const std::vector<uint8_t> &support, const std::vector<double> &tx,
                                            const std::vector<double> &ty, double radius, long min_support,
                                            double tol) {
    const size_t n = (size_t)w * h;
    const int r = std::max<long>(1, (long)std::nearbyint(radius));  // round half to even
    const double ridge = 1e-3 * (double)(r * r);
    std::vector<double> ux0(n), uy0(n);
    for (size_t i = 0; i < n; i++) {
        ux0[i] = support[i] ? ux[i] : 0.0;
        uy0[i] = support[i] ? uy[i] : 0.0;
    }
    std::vector<uint8_t> sel = support, ok_fit(n, 0);
    std::vector<double> ax(n, NAN), ay(n, NAN), s1;
    auto xs = [&](size_t i) { return (double)(i % w); };
    auto ys = [&](size_t i) { return (double)(i / w); };
    for (int pass = 0; pass < 2; pass++) {
        std::vector<double> wt(n);
        for (size_t i = 0; i < n; i++) wt[i] = sel[i] ? 1.0 : 0.0;
        auto prod = [&](auto f) {
            std::vector<double> v(n);
            for (size_t i = 0; i < n; i++) v[i] = f(i);
            return box_sum(v, w, h, r);
        };
        s1 = box_sum(wt, w, h, r);
        const auto sx = prod([&](size_t i) { return wt[i] * xs(i); });
        const auto sy = prod([&](size_t i) { return wt[i] * ys(i); });
        const auto sxx = prod([&](size_t i) { return wt[i] * xs(i) * xs(i); });
        const auto sxy = prod([&](size_t i) { return wt[i] * xs(i) * ys(i); });
        const auto syy = prod([&](size_t i) { return wt[i] * ys(i) * ys(i); });
        std::vector<double> rhs[6];
        for (int k = 0; k < 2; k++) {
            const auto &comp = k ? uy0 : ux0;
            rhs[3 * k] = prod([&](size_t i) { return wt[i] * comp[i]; });
            rhs[3 * k + 1] = prod([&](size_t i) { return wt[i] * xs(i) * comp[i]; });
            rhs[3 * k + 2] = prod([&](size_t i) { return wt[i] * ys(i) * comp[i]; });
        }
#pragma omp parallel for schedule(static)
        for (long li = 0; li < (long)n; li++) {
            const size_t i = (size_t)li;
            if (s1[i] < 3.0) {
                ax[i] = ay[i] = NAN;
                ok_fit[i] = 0;
                continue;
            }
            const double x = xs(i), y = ys(i);
            const double sxc = sx[i] - x * s1[i], syc = sy[i] - y * s1[i];
            const double sxxc = sxx[i] - 2.0 * x * sx[i] + x * x * s1[i];
            const double sxyc = sxy[i] - x * sy[i] - y * sx[i] + x * y * s1[i];
            const double syyc = syy[i] - 2.0 * y * sy[i] + y * y * s1[i];
            double m[3][3] = {{s1[i], sxc, syc}, {sxc, sxxc + ridge * s1[i], sxyc}, {syc, sxyc, syyc + ridge * s1[i]}};
            double b[3][2], c[3][2];
            for (int k = 0; k < 2; k++) {
                const double su = rhs[3 * k][i];
                b[0][k] = su;
                b[1][k] = rhs[3 * k + 1][i] - x * su;
                b[2][k] = rhs[3 * k + 2][i] - y * su;
            }
            if (solve3(m, b, c)) {
                ax[i] = c[0][0];
                ay[i] = c[0][1];
            } else {
                ax[i] = ay[i] = NAN;
            }
            ok_fit[i] = 1;
        }
        if (pass == 0)
            for (size_t i = 0; i < n; i++)
                sel[i] = support[i] && std::hypot(ux[i] - ax[i], uy[i] - ay[i]) <= 3.0 * tol;
    }
    std::vector<uint8_t> ok(n);
    for (size_t i = 0; i < n; i++)
        ok[i] = s1[i] + 0.5 >= (double)min_support && ok_fit[i] && std::hypot(tx[i] - ax[i], ty[i] - ay[i]) <= tol;
    return ok;
}

// hybrid.rs:262-540 (track.py:383-505)
int sequence_hybrid(const double *frames, int nf, int h, int w, const uint8_t *mask, const orc_cfg &c,
                    orc_result *res) {
    const size_t n = (size_t)h * w;
    std::vector<uint8_t> roi(n, 1);
    if (mask)
        for (size_t i = 0; i < n; i++) roi[i] = mask[i] != 0;
    size_t n_roi = 0;
    for (size_t i = 0; i < n; i++) n_roi += roi[i];
    const double *ref = frames, *prev = frames;
    int ref_idx = 0;
    std::vector<double> u0x(n, 0.0), u0y(n, 0.0), c0r(n, 1.0);
    std::vector<uint8_t> anch(n, 1), good_prev(n, 1);
    std::vector<double> upx(n, 0.0), upy(n, 0.0);
    std::vector<long> nfb(n, 0);
    Pair prev_res;
    bool have_prev = false;
    std::vector<double> prev_adj_x, prev_adj_y;
    bool have_prev_adj = false;
    const long cap = (long)c.fb_max + 1;
    const double conf_min = c.conf_min, radius = Cfg(c).fb_radius_px();
    const bool gate_on = c.fb_min_support > 0;
    for (int k = 1; k < nf; k++) {
        const double *fk = frames + (size_t)k * n;
        Pair leg;
        int rc = pair(ref, fk, h, w, nullptr, have_prev ? prev_res.raw_ux.data() : nullptr,
                      have_prev ? prev_res.raw_uy.data() : nullptr, c, leg);
        if (rc) return rc;
        double cm = mean_conf(leg, u0x, u0y, mask, h, w, c);
        if (cm < c.ref_update_conf && ref_idx != k - 1) {
            // c0K is replaced (not min-accumulated) where the pixel was direct-good at k - 1, 0 elsewhere
            const std::vector<double> cp = conf_at(prev_res.conf, u0x, u0y, h, w);
            for (size_t i = 0; i < n; i++) c0r[i] = good_prev[i] ? cp[i] : 0.0;
            anch = good_prev;
            u0x = upx;
            u0y = upy;
            ref = prev;
            ref_idx = k - 1;
            leg = Pair();
            rc = pair(ref, fk, h, w, nullptr, nullptr, nullptr, c, leg);
            if (rc) return rc;
            cm = mean_conf(leg, u0x, u0y, mask, h, w, c);
        }
        std::vector<double> cx, cy;
        compose(u0x, u0y, leg.raw_ux, leg.raw_uy, h, w, cx, cy);
        const std::vector<double> leg_c = conf_at(leg.conf, u0x, u0y, h, w);
        std::vector<double> link(n);
        std::vector<uint8_t> direct_ok(n), good(n), cand_re(n), need(n);
        const bool reacq_ok = ref_idx + 1 < k;
        bool any_need = false;
        for (size_t i = 0; i < n; i++) {
            link[i] = anch[i] ? std::min(c0r[i], leg_c[i]) : leg_c[i];
            direct_ok[i] = std::isfinite(cx[i]) && std::isfinite(cy[i]) && link[i] >= conf_min;
            good[i] = direct_ok[i] && anch[i];
            cand_re[i] = direct_ok[i] && !anch[i] && roi[i] && reacq_ok;
            need[i] = !good[i] && roi[i] && std::isfinite(upx[i]) && std::isfinite(upy[i]) && nfb[i] < (long)c.fb_max;
            any_need |= need[i] != 0;
        }
        double t_coarse = leg.t_coarse, t_fine = leg.t_fine;
        std::vector<double> ax, ay, ca(n, 0.0), araw_x, araw_y;
        bool have_adj = false;
        if (any_need) {
            const std::vector<double> *rx, *ry, *rconf;
            Pair adj;
            if (ref_idx == k - 1) {
                rx = &leg.raw_ux;
                ry = &leg.raw_uy;
                rconf = &leg.conf;
            } else {
                rc = pair(prev, fk, h, w, nullptr, have_prev_adj ? prev_adj_x.data() : nullptr,
                          have_prev_adj ? prev_adj_y.data() : nullptr, c, adj);
                if (rc) return rc;
                t_coarse += adj.t_coarse;
                t_fine += adj.t_fine;
                rx = &adj.raw_ux;
                ry = &adj.raw_uy;
                rconf = &adj.conf;
            }
            std::vector<double> cax, cay;
            compose(upx, upy, *rx, *ry, h, w, cax, cay);
            ca = conf_at(*rconf, upx, upy, h, w);
            ax.resize(n);
            ay.resize(n);
            for (size_t i = 0; i < n; i++) {
                const bool fin = std::isfinite(cax[i]) && std::isfinite(cay[i]);
                ax[i] = fin ? cax[i] : upx[i];
                ay[i] = fin ? cay[i] : upy[i];
            }
            araw_x = *rx;
            araw_y = *ry;
            have_adj = true;
        } else {
            ax = upx;
            ay = upy;
        }
        std::vector<uint8_t> has_fb(n), re_nofb(n), re(n), fb_cand(n), fb_ok(n);
        bool any_re_nofb = false;
        for (size_t i = 0; i < n; i++) {
            has_fb[i] = need[i] && ca[i] >= conf_min;
            re_nofb[i] = cand_re[i] && !has_fb[i];
            any_re_nofb |= re_nofb[i] && roi[i];
        }
        auto masked = [&](const std::vector<double> &v) {
            std::vector<double> o(n);
            for (size_t i = 0; i < n; i++) o[i] = good[i] ? v[i] : NAN;
            return o;
        };
        auto support = [&]() {
            std::vector<uint8_t> s(n);
            for (size_t i = 0; i < n; i++) s[i] = good[i] && roi[i];
            return s;
        };
        if (gate_on && any_re_nofb) {
            const auto gate0 = local_affine_agreement(w, h, masked(cx), masked(cy), support(), cx, cy, radius,
                                                      c.fb_min_support, c.fb_tol);
            for (size_t i = 0; i < n; i++) re_nofb[i] = re_nofb[i] && gate0[i];
        }
        for (size_t i = 0; i < n; i++) {
            const bool agree = std::hypot(cx[i] - ax[i], cy[i] - ay[i]) <= c.fb_tol;
            re[i] = (cand_re[i] && has_fb[i] && agree) || re_nofb[i];
            if (re[i]) {
                c0r[i] = leg_c[i];
                anch[i] = 1;
                good[i] = 1;
            }
        }
        bool any_cand = false;
        for (size_t i = 0; i < n; i++) {
            nfb[i] = good[i] ? 0 : std::min(nfb[i] + 1, cap);
            fb_cand[i] = !good[i] && need[i] && nfb[i] <= (long)c.fb_max && ca[i] >= conf_min;
            any_cand |= fb_cand[i] && roi[i];
        }
        if (gate_on && any_cand) {
            const auto gate = local_affine_agreement(w, h, masked(cx), masked(cy), support(), ax, ay, radius,
                                                     c.fb_min_support, c.fb_tol);
            for (size_t i = 0; i < n; i++) fb_ok[i] = fb_cand[i] && gate[i];
        } else {
            fb_ok = fb_cand;
        }
        std::vector<double> ux(n), uy(n), conf(n), conf_in(n);
        long nmax = 0;
        size_t n_fb_ok = 0, n_rej = 0, n_re = 0;
        for (size_t i = 0; i < n; i++) {
            ux[i] = good[i] ? cx[i] : ax[i];
            uy[i] = good[i] ? cy[i] : ay[i];
            conf[i] = good[i] ? link[i] : ca[i] * std::max(0.0, 1.0 - (double)nfb[i] / (double)cap);
            // the low mask of finish: neither direct-good nor an accepted fallback
            conf_in[i] = (good[i] || fb_ok[i]) ? 1.0 : -1.0;
            nmax = std::max(nmax, nfb[i]);
            if (roi[i]) {
                n_fb_ok += fb_ok[i];
                n_rej += fb_cand[i] && !fb_ok[i];
                n_re += re[i];
            }
        }
        Pair out;
        finish(ux.data(), uy.data(), conf_in.data(), h, w, c, mask, out);
        out.conf = conf;
        out.raw_ux = ux;
        out.raw_uy = uy;
        out.t_coarse = t_coarse;
        out.t_fine = t_fine;
        out.t_strain = leg.t_strain;
        store(out, n, &res[k - 1]);
        if (res[k - 1].fb)
            for (size_t i = 0; i < n; i++) res[k - 1].fb[i] = (int32_t)nfb[i];
        res[k - 1].reference = ref_idx;
        res[k - 1].mean_conf = cm;
        res[k - 1].n_fb_max = (int)nmax;
        res[k - 1].fallback_frac = (double)n_fb_ok / (double)n_roi;
        res[k - 1].fb_gate_rejected = (double)n_rej / (double)n_roi;
        res[k - 1].reacquired = (double)n_re / (double)n_roi;
        upx = ux;
        upy = uy;
        good_prev = good;
        prev_adj_x = araw_x;
        prev_adj_y = araw_y;
        have_prev_adj = have_adj;
        prev = fk;
        prev_res = leg;
        have_prev = true;
    }
    return 0;
}

}  // namespace

}  // namespace orc

using namespace orc;

extern "C" int orc_track_pair(const double *i1, const double *i2, int h, int w, const uint8_t *mask,
                              const double *init_ux, const double *init_uy, const orc_cfg *cfg,
                              orc_result *res) {
    Pair p;
    int rc = pair(i1, i2, h, w, mask, init_ux, init_uy, *cfg, p);
    if (rc) return rc;
    store(p, (size_t)h * w, res);
    res->reference = 0;
    res->mean_conf = 0;
    return 0;
}

extern "C" int orc_track_sequence(const double *frames, int n, int h, int w, const uint8_t *mask,
                                  const orc_cfg *cfg, orc_result *res) {
    if (n < 1) { set_error("no frames"); return 1; }
    if (cfg->seq_hybrid) return sequence_hybrid(frames, n, h, w, mask, *cfg, res);
    return sequence_chain(frames, n, h, w, mask, *cfg, res);
}

extern "C" void orc_box_sum(const double *a, int h, int w, int r, double *out) {
    std::vector<double> v(a, a + (size_t)h * w);
    const auto o = box_sum(v, w, h, r);
    std::copy(o.begin(), o.end(), out);
}

extern "C" void orc_local_affine_agreement(const double *ux, const double *uy, const uint8_t *support, const double *tx,
                                           const double *ty, int h, int w, double radius, long min_support, double tol,
                                           uint8_t *ok) {
    const size_t n = (size_t)h * w;
    const auto o = local_affine_agreement(w, h, std::vector<double>(ux, ux + n), std::vector<double>(uy, uy + n),
                                          std::vector<uint8_t>(support, support + n), std::vector<double>(tx, tx + n),
                                          std::vector<double>(ty, ty + n), radius, min_support, tol);
    std::copy(o.begin(), o.end(), ok);
}

extern "C" void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
