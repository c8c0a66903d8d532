// Pair and sequence pipeline of the oracle (test infrastructure only): coarse stage, fine stage,
// guard and strain (python/optflow_ref/track.py:45-146), and the chained sequence with reference
// updates, composition and link confidence (track.py:231-328).
#include <omp.h>

#include <algorithm>
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
    refine(i1, i2, h, w, gx.data(), gy.data(), c, out.raw_ux.data(), out.raw_uy.data(), out.conf.data(),
           nullptr, nullptr);
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
    if (cfg->seq_hybrid) { set_error("hybrid sequences are not in the oracle yet"); return 1; }
    return sequence_chain(frames, n, h, w, mask, *cfg, res);
}

extern "C" void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
