// Gabor bank, reference phase and the fine stage (refine) for the oracle (test infrastructure
// only). Restates python/optflow_ref/phaseflow.py:73-511 in double precision, with the Rust
// port's Cholesky solve (rust/crates/optflow-core/src/phaseflow.rs:29-34, :394-455) and its
// optional pre-warp freeze (rust/crates/optflow-core/src/prewarp.rs:48-115).
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "common.hpp"

namespace orc {

void confidence_of(const double *energy, const double *resid, const double *ux, const double *uy, int h, int w,
                   const orc_cfg &c, double *conf_out);

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

namespace {

double fftfreq(int i, int n) {
    // numpy.fft.fftfreq(n)[i]
    int k = (i <= (n - 1) / 2) ? i : i - n;
    return (double)k / n;
}

// phaseflow.py:73-89: F planes hp x wp, scales outer, orientations inner
void gabor_bank(int hp, int wp, const orc_cfg &c, std::vector<double> &bank) {
    Cfg C(c);
    const int F = C.n_filters();
    bank.assign((size_t)F * hp * wp, 0.0);
    const double s_hp = c.dc_cut * c.lam * C.max_scale() / (2 * M_PI);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < hp; y++) {
        const double fy = fftfreq(y, hp) * 2 * M_PI;
        for (int x = 0; x < wp; x++) {
            const double fx = fftfreq(x, wp) * 2 * M_PI;
            const double hpf = 1.0 - std::exp(-(fx * fx + fy * fy) * s_hp * s_hp / 2);
            int f = 0;
            for (int si = 0; si < c.n_scales; si++) {
                const double w0 = 2 * M_PI / (c.lam * c.scales[si]);
                const double s = c.bandwidth * w0;
                for (int k = 0; k < c.n_orient; k++, f++) {
                    const double th = M_PI * k / c.n_orient;
                    const double ax = fx - w0 * std::cos(th), ay = fy - w0 * std::sin(th);
                    bank[((size_t)f * hp + y) * wp + x] = hpf * std::exp(-(ax * ax + ay * ay) / (2 * s * s));
                }
            }
        }
    }
}

int reflect(int i, int n) {
    // numpy.pad mode="reflect" for a pad narrower than the image
    if (i < 0) return -i;
    if (i >= n) return 2 * (n - 1) - i;
    return i;
}

std::vector<cd> padded_spectrum(const double *img, int h, int w) {
    const int hp = h + 2 * PAD, wp = w + 2 * PAD;
    std::vector<cd> a((size_t)hp * wp);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < hp; y++)
        for (int x = 0; x < wp; x++)
            a[(size_t)y * wp + x] = img[(size_t)reflect(y - PAD, h) * w + reflect(x - PAD, w)];
    fft2(a, hp, wp, false);
    return a;
}

struct RefPack {
    int F = 0;
    std::vector<cd> r;                 // [f][pixel]
    std::vector<double> gx, gy, amp;   // [f][pixel]
};

// phaseflow.py:100-111
void reference_phase(const double *i1, int h, int w, const orc_cfg &c, const std::vector<double> &bank,
                     RefPack &rp) {
    const int hp = h + 2 * PAD, wp = w + 2 * PAD;
    const int F = Cfg(c).n_filters();
    const size_t n = (size_t)h * w, np_ = (size_t)hp * wp;
    rp.F = F;
    rp.r.resize(F * n);
    rp.gx.resize(F * n);
    rp.gy.resize(F * n);
    rp.amp.resize(F * n);
    std::vector<cd> S = padded_spectrum(i1, h, w);
    std::vector<cd> P(np_), PX(np_), PY(np_);
    for (int f = 0; f < F; f++) {
        const double *G = &bank[f * np_];
#pragma omp parallel for schedule(static)
        for (int y = 0; y < hp; y++) {
            const double fy = fftfreq(y, hp) * 2 * M_PI;
            for (int x = 0; x < wp; x++) {
                const double fx = fftfreq(x, wp) * 2 * M_PI;
                size_t i = (size_t)y * wp + x;
                cd p = S[i] * G[i];
                P[i] = p;
                PX[i] = p * cd(0, fx);
                PY[i] = p * cd(0, fy);
            }
        }
        fft2(P, hp, wp, true);
        fft2(PX, hp, wp, true);
        fft2(PY, hp, wp, true);
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t j = (size_t)(y + PAD) * wp + (x + PAD);
                cd r = P[j], rx = PX[j], ry = PY[j];
                double a2 = std::norm(r) + 1e-12;
                size_t o = f * n + (size_t)y * w + x;
                rp.r[o] = r;
                rp.gx[o] = (std::conj(r) * rx).imag() / a2;
                rp.gy[o] = (std::conj(r) * ry).imag() / a2;
                rp.amp[o] = std::sqrt(a2);
            }
    }
}

// phaseflow.py:114-120: sum_q W(q - p) (qx - px)^px (qy - py)^py a(q), edges replicated
void window_moment(const double *a, int h, int w, double sigma, int px, int py, double *out,
                   std::vector<double> &tmp) {
    std::vector<double> k = gauss_kernel(sigma);
    const int r = (int)k.size() / 2;
    std::vector<double> ky(k.size()), kx(k.size());
    for (int i = -r; i <= r; i++) {
        ky[i + r] = k[i + r] * std::pow((double)i, py);
        kx[i + r] = k[i + r] * std::pow((double)i, px);
    }
    tmp.resize((size_t)h * w);
    correlate1d(a, h, w, ky.data(), (int)ky.size(), 0, NEAREST, tmp.data());
    correlate1d(tmp.data(), h, w, kx.data(), (int)kx.size(), 1, NEAREST, out);
}

struct Unknown { int comp, px, py; };

std::vector<Unknown> shape_unknowns(int order) {
    // phaseflow.py:123-129
    std::vector<std::pair<int, int>> pw = {{0, 0}, {1, 0}, {0, 1}};
    if (order == 2) { pw.push_back({2, 0}); pw.push_back({1, 1}); pw.push_back({0, 2}); }
    std::vector<Unknown> u = {{0, 0, 0}, {1, 0, 0}};
    for (size_t i = 1; i < pw.size(); i++)
        for (int c = 0; c < 2; c++) u.push_back({c, pw[i].first, pw[i].second});
    return u;
}

// Cholesky solve of the n x n system in place (lower triangle of m used); false unless every
// pivot is finite and positive and the solution finite (phaseflow.rs:394-455).
bool cholesky(double *m, double *b, int n) {
    for (int j = 0; j < n; j++) {
        double d = m[j * n + j];
        for (int k = 0; k < j; k++) d -= m[j * n + k] * m[j * n + k];
        if (!(std::isfinite(d) && d > 0)) return false;
        double l = std::sqrt(d);
        m[j * n + j] = l;
        for (int i = j + 1; i < n; i++) {
            double s = m[i * n + j];
            for (int k = 0; k < j; k++) s -= m[i * n + k] * m[j * n + k];
            m[i * n + j] = s / l;
        }
    }
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= m[i * n + k] * b[k];
        b[i] = s / m[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < n; k++) s -= m[k * n + i] * b[k];
        b[i] = s / m[i * n + i];
    }
    for (int i = 0; i < n; i++)
        if (!std::isfinite(b[i])) return false;
    return true;
}

}  // namespace

// phaseflow.py:132-188 (local_solve): sums = gxx, gxy, gyy, gxc, gyc, gcc, gw
void local_solve(const double *sums, const double *ux, const double *uy, int h, int w, double sigma,
                 int order, double *oux, double *ouy, double *energy, double *resid, double *slope) {
    const size_t N = (size_t)h * w;
    std::vector<Unknown> unk = shape_unknowns(order);
    const int n = (int)unk.size();
    // moment maps needed: key (source, px, py); sources 0..2 products, 3..4 rhs, 5 gcc, 6 gw
    std::map<std::tuple<int, int, int>, int> index;
    auto need = [&](int src, int px, int py) {
        auto key = std::make_tuple(src, px, py);
        if (!index.count(key)) { int id = (int)index.size(); index[key] = id; }
        return index[key];
    };
    std::vector<int> midx(n * n), bidx(n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int a = std::min(unk[i].comp, unk[j].comp), b = std::max(unk[i].comp, unk[j].comp);
            int src = (a == 0 && b == 0) ? 0 : (a == 0 ? 1 : 2);
            midx[i * n + j] = need(src, unk[i].px + unk[j].px, unk[i].py + unk[j].py);
        }
    for (int i = 0; i < n; i++) bidx[i] = need(3 + unk[i].comp, unk[i].px, unk[i].py);
    const int icc = need(5, 0, 0), iw = need(6, 0, 0);
    std::vector<double> maps(index.size() * N), tmp;
    for (auto &kv : index) {
        int src, px, py;
        std::tie(src, px, py) = kv.first;
        window_moment(sums + src * N, h, w, sigma, px, py, &maps[kv.second * N], tmp);
    }
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)N; p++) {
        double M[144], b[12], b0[12];
        for (int i = 0; i < n * n; i++) M[i] = maps[midx[i] * N + p];
        for (int i = 0; i < n; i++) b[i] = maps[bidx[i] * N + p];
        const double e = M[0] + M[n + 1];
        const double det2 = M[0] * M[n + 1] - M[1] * M[1];
        energy[p] = e;
        bool ok = det2 > APERTURE_TOL * e * e;
        if (!ok) {
            for (int i = 0; i < n * n; i++) M[i] = 0;
            for (int i = 0; i < n; i++) { M[i * n + i] = 1; b[i] = 0; }
        }
        double tr = 0;
        for (int i = 0; i < n; i++) tr += M[i * n + i];
        for (int i = 0; i < n; i++) M[i * n + i] += RIDGE * tr / n;
        for (int i = 0; i < n; i++) b0[i] = b[i];
        ok = cholesky(M, b, n) && ok;
        if (ok) {
            double tb = 0;
            for (int i = 0; i < n; i++) tb += b[i] * b0[i];
            double r = maps[icc * N + p] - tb;
            resid[p] = std::sqrt((r > 0 ? r : (std::isnan(r) ? r : 0.0)) / (maps[iw * N + p] + 1e-30));
            oux[p] = b[0];
            ouy[p] = b[1];
            for (int i = 0; i < 4; i++) slope[p * 4 + i] = b[2 + i];
        } else {
            resid[p] = M_PI;
            oux[p] = ux[p];
            ouy[p] = uy[p];
            for (int i = 0; i < 4; i++) slope[p * 4 + i] = 0;
        }
    }
    // energy threshold against the median window energy (phaseflow.py:149)
    const double thr = 1e-9 * median(std::vector<double>(energy, energy + N)) + 1e-30;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)N; p++)
        if (!(energy[p] > thr)) {
            resid[p] = M_PI;
            oux[p] = ux[p];
            ouy[p] = uy[p];
            for (int i = 0; i < 4; i++) slope[p * 4 + i] = 0;
        }
}

namespace {

// _confidence before the border rule (phaseflow.py:320-323)
void confidence_raw(const double *energy, const double *resid, size_t n, double phase_tol, double *conf) {
    const double scale = 0.25 * median(std::vector<double>(energy, energy + n)) + 1e-30;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) {
        double t = std::min(1.0, energy[i] / scale);
        double q = resid[i] / phase_tol;
        conf[i] = t * std::exp(-q * q);
    }
}

// phaseflow.py:326-355 with the low mask given (rust prewarp.rs LowFreeze supplies it)
void prewarp(const double *ux, const double *uy, const double *slope, const std::vector<uint8_t> &low,
             int h, int w, const orc_cfg &c, double *sx, double *sy) {
    const size_t n = (size_t)h * w;
    std::vector<double> fx(ux, ux + n), fy(uy, uy + n);
    bool any_low = false;
    for (size_t i = 0; i < n; i++) if (low[i]) { any_low = true; break; }
    if (any_low) {
        std::vector<uint8_t> notlow(n);
        for (size_t i = 0; i < n; i++) notlow[i] = !low[i];
        std::vector<double> d(n);
        edt(notlow.data(), h, w, d.data(), nullptr);
        const double R = Cfg(c).auto_edge_guard_px();
        std::vector<uint8_t> ok(n);
        size_t nok = 0;
        for (size_t i = 0; i < n; i++) { ok[i] = d[i] > R; nok += ok[i]; }
        if (nok > 0 && nok < n) {
            std::vector<uint8_t> notok(n);
            for (size_t i = 0; i < n; i++) notok[i] = !ok[i];
            std::vector<int32_t> idx(n);
            edt(notok.data(), h, w, nullptr, idx.data());
#pragma omp parallel for schedule(static)
            for (long i = 0; i < (long)n; i++) {
                if (ok[i]) continue;
                const int j = idx[i];
                const double dx = (double)(i % w) - (j % w), dy = (double)(i / w) - (j / w);
                const double *g = slope + 4 * (size_t)j;
                fx[i] = ux[j] + g[0] * dx + g[2] * dy;
                fy[i] = uy[j] + g[1] * dx + g[3] * dy;
            }
        }
    }
    gauss_blur(fx.data(), h, w, c.sigma_s, sx);
    gauss_blur(fy.data(), h, w, c.sigma_s, sy);
}

double border_distance(double x, double y, int w, int h) {
    return std::min(std::min(x, y), std::min(w - 1 - x, h - 1 - y));
}

}  // namespace

void gabor_bank_public(int hp, int wp, const orc_cfg &c, std::vector<double> &bank) { gabor_bank(hp, wp, c, bank); }

// phaseflow.py:427-511 without the crack options; image 2 may be any window (h2 x w2) of the
// second frame, pixel (x, y) of image 1 sitting at (x + shx, y + shy) in it before the field
// displaces it (rust/crates/optflow-core/src/phaseflow.rs:1416-1439, 1441-1456).
void refine_shifted(const double *i1, int h, int w, const double *i2, int h2, int w2, long shx, long shy,
                    const double *ux0, const double *uy0, const orc_cfg &c, double *ux_out, double *uy_out,
                    double *conf_out, double *resid_out, double *energy_out) {
    Cfg C(c);
    const int hp = h + 2 * PAD, wp = w + 2 * PAD;
    const size_t n = (size_t)h * w, np_ = (size_t)hp * wp;
    const int F = C.n_filters();
    std::vector<double> bank;
    gabor_bank(hp, wp, c, bank);
    RefPack rp;
    reference_phase(i1, h, w, c, bank, rp);
    std::vector<double> ux(ux0, ux0 + n), uy(uy0, uy0 + n);
    std::vector<double> resid(n, M_PI), energy(n, 1.0), slope(4 * n, 0.0);
    std::vector<double> nux(n), nuy(n);
    std::vector<double> c2(i2, i2 + (size_t)h2 * w2);
    bspline_real(c2, h2, w2);
    const double fsx = (double)shx, fsy = (double)shy;
    std::vector<double> sx(n), sy(n), warped(n), conf(n);
    std::vector<std::vector<cd>> resp(F);
    std::vector<double> sums(7 * n);
    std::vector<uint8_t> low(n), kept;
    bool have_kept = false;
    const int freeze = c.prewarp_freeze_at;
    for (int it = 0; it < c.outer_iters; it++) {
        if (it == 0) {
            gauss_blur(ux.data(), h, w, c.sigma_s, sx.data());
            gauss_blur(uy.data(), h, w, c.sigma_s, sy.data());
        } else {
            if (freeze >= 1 && it > freeze && have_kept) {
                low = kept;
            } else {
                confidence_raw(energy.data(), resid.data(), n, c.phase_tol, conf.data());
                for (size_t i = 0; i < n; i++) low[i] = conf[i] < c.conf_min;
                if (freeze >= 1 && it == freeze) { kept = low; have_kept = true; }
            }
            prewarp(ux.data(), uy.data(), slope.data(), low, h, w, c, sx.data(), sy.data());
        }
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t i = (size_t)y * w + x;
                warped[i] = sample_at(c2.data(), h2, w2, ((double)x + fsx) + sx[i], ((double)y + fsy) + sy[i]);
            }
        std::vector<cd> S2 = padded_spectrum(warped.data(), h, w);
        for (int f = 0; f < F; f++) {
            resp[f].resize(np_);
            const double *G = &bank[f * np_];
            for (size_t i = 0; i < np_; i++) resp[f][i] = S2[i] * G[i];
            fft2(resp[f], hp, wp, true);
            bspline_complex(resp[f], hp, wp);
        }
        std::vector<double> csx(sx), csy(sy);
        bspline_real(csx, h, w);
        bspline_real(csy, h, w);
        for (int in = 0; in < c.inner_iters; in++) {
#pragma omp parallel for schedule(static)
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    const size_t i = (size_t)y * w + x;
                    const double tx = x + ux[i], ty = y + uy[i];
                    double px = tx - sx[i], py = ty - sy[i];
                    for (int k = 0; k < 2; k++) {
                        double nx = tx - sample_at(csx.data(), h, w, px, py);
                        double ny = ty - sample_at(csy.data(), h, w, px, py);
                        px = nx;
                        py = ny;
                    }
                    double gxx = 0, gxy = 0, gyy = 0, gxc = 0, gyc = 0, gcc = 0, gw = 0;
                    for (int f = 0; f < F; f++) {
                        const size_t o = f * n + i;
                        cd r2 = sample_at(resp[f].data(), hp, wp, px + PAD, py + PAD);
                        cd r1 = rp.r[o];
                        cd prod = r1 * std::conj(r2);
                        double dphi = std::atan2(prod.imag(), prod.real());
                        double wk = rp.amp[o] * std::abs(r2);
                        double gx = rp.gx[o], gy = rp.gy[o];
                        double cc = dphi + gx * ux[i] + gy * uy[i];
                        gxx += wk * gx * gx; gxy += wk * gx * gy; gyy += wk * gy * gy;
                        gxc += wk * gx * cc; gyc += wk * gy * cc; gcc += wk * cc * cc;
                        gw += wk;
                    }
                    sums[0 * n + i] = gxx; sums[1 * n + i] = gxy; sums[2 * n + i] = gyy;
                    sums[3 * n + i] = gxc; sums[4 * n + i] = gyc; sums[5 * n + i] = gcc;
                    sums[6 * n + i] = gw;
                }
            local_solve(sums.data(), ux.data(), uy.data(), h, w, c.sigma_w, c.shape_order, nux.data(),
                        nuy.data(), energy.data(), resid.data(), slope.data());
            ux.swap(nux);
            uy.swap(nuy);
        }
    }
    confidence_of(energy.data(), resid.data(), ux.data(), uy.data(), h, w, c, conf_out);
    for (size_t i = 0; i < n; i++) {
        ux_out[i] = ux[i];
        uy_out[i] = uy[i];
        if (resid_out) resid_out[i] = resid[i];
        if (energy_out) energy_out[i] = energy[i];
    }
}

// phaseflow.py:320-323 and the border rule of phaseflow.rs:1617-1757: the clamped confidence of
// a whole field (its texture scale the median energy of all of it), 0 where the pixel or its
// match lies within B of the border
void confidence_of(const double *energy, const double *resid, const double *ux, const double *uy, int h, int w,
                   const orc_cfg &c, double *conf_out) {
    const size_t n = (size_t)h * w;
    std::vector<double> conf(n);
    confidence_raw(energy, resid, n, c.phase_tol, conf.data());
    const double B = Cfg(c).edge_guard_px() > 0 ? 4.0 * c.sigma_w : 0.0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            bool rep = border_distance(x, y, w, h) >= B && border_distance(x + ux[i], y + uy[i], w, h) >= B;
            double cl = std::min(std::max(conf[i], 0.0), 1.0);
            if (std::isnan(conf[i])) cl = conf[i];
            conf_out[i] = rep ? cl : 0.0;
        }
}

void refine(const double *i1, const double *i2, int h, int w, const double *ux0, const double *uy0,
            const orc_cfg &c, double *ux_out, double *uy_out, double *conf_out, double *resid_out,
            double *energy_out) {
    refine_shifted(i1, h, w, i2, h, w, 0, 0, ux0, uy0, c, ux_out, uy_out, conf_out, resid_out, energy_out);
}

}  // namespace orc

using namespace orc;

extern "C" const char *orc_last_error(void) { return g_err.c_str(); }

extern "C" int orc_n_filters(const orc_cfg *cfg) { return Cfg(*cfg).n_filters(); }

extern "C" void orc_gabor_bank(int hp, int wp, const orc_cfg *cfg, double *bank) {
    std::vector<double> b;
    gabor_bank(hp, wp, *cfg, b);
    std::memcpy(bank, b.data(), b.size() * sizeof(double));
}

extern "C" void orc_reference_phase(const double *i1, int h, int w, const orc_cfg *cfg, double *r_re,
                                    double *r_im, double *gx, double *gy, double *amp) {
    std::vector<double> bank;
    gabor_bank(h + 2 * PAD, w + 2 * PAD, *cfg, bank);
    RefPack rp;
    reference_phase(i1, h, w, *cfg, bank, rp);
    for (size_t i = 0; i < rp.r.size(); i++) {
        r_re[i] = rp.r[i].real();
        r_im[i] = rp.r[i].imag();
        gx[i] = rp.gx[i];
        gy[i] = rp.gy[i];
        if (amp) amp[i] = rp.amp[i];
    }
}

extern "C" void orc_local_solve(const double *sums, const double *ux, const double *uy, int h, int w,
                                double sigma, int order, double *oux, double *ouy, double *energy,
                                double *resid, double *slope) {
    local_solve(sums, ux, uy, h, w, sigma, order, oux, ouy, energy, resid, slope);
}

extern "C" int orc_refine(const double *i1, const double *i2, int h, int w, const double *ux0,
                          const double *uy0, const orc_cfg *cfg, double *ux, double *uy, double *conf,
                          double *resid, double *energy) {
    if (h <= PAD || w <= PAD) {
        set_error("refine: image too small for the reflect pad");
        return 1;
    }
    refine(i1, i2, h, w, ux0, uy0, *cfg, ux, uy, conf, resid, energy);
    return 0;
}

extern "C" int orc_refine_shifted(const double *i1, int h, int w, const double *i2, int h2, int w2, long shx,
                                  long shy, const double *ux0, const double *uy0, const orc_cfg *cfg, double *ux,
                                  double *uy, double *conf, double *resid, double *energy) {
    if (h <= PAD || w <= PAD) {
        set_error("refine: image too small for the reflect pad");
        return 1;
    }
    if (h2 < 1 || w2 < 1) {
        set_error("refine: empty image i2");
        return 1;
    }
    refine_shifted(i1, h, w, i2, h2, w2, shx, shy, ux0, uy0, *cfg, ux, uy, conf, resid, energy);
    return 0;
}
