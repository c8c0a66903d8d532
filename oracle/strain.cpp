// Strain of the oracle (test infrastructure only): NaN-aware quadratic least-squares
// (Savitzky-Golay) displacement gradient over win x win windows and the Green-Lagrange fields.
// Restates python/optflow_ref/strain.py:61-112 (full window) with the Rust port's exact integer
// window test and Cholesky pivot rule (rust/crates/optflow-core/src/strain.rs:37-108).
#include <algorithm>

#include "common.hpp"

namespace orc {

namespace {
// strain.py:10: quadratic basis 1, x, y, x^2, xy, y^2 as (power of x, power of y)
const int BASIS[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};

// strain.rs:64-108: Cholesky with relative pivot test, two right-hand sides
bool solve6(double m[6][6], double rhs[6][2], double out[6][2]) {
    double l[6][6] = {{0}};
    for (int i = 0; i < 6; i++)
        for (int j = 0; j <= i; j++) {
            double s = m[i][j];
            for (int k = 0; k < j; k++) s -= l[i][k] * l[j][k];
            if (i == j) {
                if (std::isnan(s) || s <= 1e-12 * m[i][i]) return false;
                l[i][i] = std::sqrt(s);
            } else {
                l[i][j] = s / l[j][j];
            }
        }
    for (int col = 0; col < 2; col++) {
        double y[6];
        for (int i = 0; i < 6; i++) {
            double s = rhs[i][col];
            for (int k = 0; k < i; k++) s -= l[i][k] * y[k];
            y[i] = s / l[i][i];
        }
        for (int i = 5; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < 6; k++) s -= l[k][i] * out[k][col];
            out[i][col] = s / l[i][i];
        }
    }
    return true;
}
}  // namespace

void displacement_gradient(const double *ux, const double *uy, int h, int w, int win, double *dudx,
                           double *dudy, double *dvdx, double *dvdy) {
    const size_t n = (size_t)h * w;
    const int r = win / 2;
    const int klen = 2 * r + 1;
    std::vector<double> m(n), u0(n), v0(n);
    for (size_t i = 0; i < n; i++) {
        bool ok = std::isfinite(ux[i]) && std::isfinite(uy[i]);
        m[i] = ok;
        u0[i] = ok ? ux[i] : 0.0;
        v0[i] = ok ? uy[i] : 0.0;
    }
    // moments of m up to total power 4, of u0 and v0 up to power 2 (strain.py:7-9)
    auto kern = [&](int p) {
        std::vector<double> k(klen);
        for (int t = -r; t <= r; t++) k[t + r] = std::pow((double)t, p);
        return k;
    };
    std::vector<double> tmp(n);
    auto moment = [&](const std::vector<double> &a, int px, int py, std::vector<double> &out) {
        std::vector<double> ky = kern(py), kx = kern(px);
        out.resize(n);
        correlate1d(a.data(), h, w, ky.data(), klen, 0, CONSTANT, tmp.data());
        correlate1d(tmp.data(), h, w, kx.data(), klen, 1, CONSTANT, out.data());
    };
    std::vector<double> mm[5][5];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            int a = BASIS[i][0] + BASIS[j][0], b = BASIS[i][1] + BASIS[j][1];
            if (mm[a][b].empty()) moment(m, a, b, mm[a][b]);
        }
    std::vector<double> mu[6], mv[6];
    for (int i = 0; i < 6; i++) {
        moment(u0, BASIS[i][0], BASIS[i][1], mu[i]);
        moment(v0, BASIS[i][0], BASIS[i][1], mv[i]);
    }
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)n; p++) {
        double count = mm[0][0][p];
        // strain.rs:37-46: count >= win^2 / 2 in exact integers
        bool ok = 2.0 * std::nearbyint(count) >= (double)win * win;
        double M[6][6], R[6][2], C[6][2];
        for (int i = 0; i < 6; i++) {
            for (int j = 0; j < 6; j++) M[i][j] = mm[BASIS[i][0] + BASIS[j][0]][BASIS[i][1] + BASIS[j][1]][p];
            R[i][0] = mu[i][p];
            R[i][1] = mv[i][p];
        }
        if (ok) ok = solve6(M, R, C);
        if (ok) {
            dudx[p] = C[1][0];
            dudy[p] = C[2][0];
            dvdx[p] = C[1][1];
            dvdy[p] = C[2][1];
        } else {
            dudx[p] = dudy[p] = dvdx[p] = dvdy[p] = NAN;
        }
    }
}

// strain.py:100-112
void strain_fields(const double *dudx, const double *dudy, const double *dvdx, const double *dvdy, size_t n,
                   double *exx, double *eyy, double *exy, double *gmax, double *vol) {
    for (size_t i = 0; i < n; i++) {
        double a = dudx[i], b = dudy[i], c = dvdx[i], d = dvdy[i];
        double Exx = a + 0.5 * (a * a + c * c);
        double Eyy = d + 0.5 * (b * b + d * d);
        double Exy = 0.5 * (b + c) + 0.5 * (a * b + c * d);
        exx[i] = Exx;
        eyy[i] = Eyy;
        exy[i] = Exy;
        double hd = (Exx - Eyy) / 2;
        gmax[i] = 2.0 * std::sqrt(hd * hd + Exy * Exy);
        vol[i] = (1 + a) * (1 + d) - b * c - 1.0;
    }
}

}  // namespace orc

extern "C" void orc_displacement_gradient(const double *ux, const double *uy, int h, int w, int win,
                                          double *dudx, double *dudy, double *dvdx, double *dvdy) {
    orc::displacement_gradient(ux, uy, h, w, win, dudx, dudy, dvdx, dvdy);
}
