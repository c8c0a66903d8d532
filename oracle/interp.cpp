// Cubic B-spline interpolation, separable correlation and Gaussian blur for the oracle
// (test infrastructure only). Follows python/optflow_ref/interp.py:13-91.
#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace orc {

namespace {
const double Z = std::sqrt(3.0) - 2.0;  // interp.py:8
// interp.py:9: ceil(log(1e-12) / log|z|) = 21
const int HORIZON = (int)std::ceil(std::log(1e-12) / std::log(std::fabs(std::sqrt(3.0) - 2.0)));

// In-place prefilter of n samples at stride s (interp.py:12-26). T is double or complex.
template <class T>
void prefilter_line(T *c, int n, long s) {
    if (n == 1) return;
    const double z = Z;
    for (int i = 0; i < n; i++) c[i * s] *= 6.0;
    int h = std::min(n, HORIZON);
    T acc = T(0);
    double zk = 1.0;
    for (int k = 0; k < h; k++) {
        acc += zk * c[k * s];
        zk *= z;
    }
    c[0] = acc;
    for (int i = 1; i < n; i++) c[i * s] += z * c[(i - 1) * s];
    c[(n - 1) * s] = (z / (z * z - 1.0)) * (c[(n - 1) * s] + z * c[(n - 2) * s]);
    for (int i = n - 2; i >= 0; i--) c[i * s] = z * (c[(i + 1) * s] - c[i * s]);
}

template <class T>
void bspline_2d(T *c, int h, int w) {
    // axis 0 (down each column), then axis 1 (along each row): interp.py:29-34
#pragma omp parallel
    {
        std::vector<T> col(h);
#pragma omp for schedule(static)
        for (int x = 0; x < w; x++) {
            for (int y = 0; y < h; y++) col[y] = c[(size_t)y * w + x];
            prefilter_line(col.data(), h, 1);
            for (int y = 0; y < h; y++) c[(size_t)y * w + x] = col[y];
        }
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) prefilter_line(c + (size_t)y * w, w, 1);
}
}  // namespace

void bspline_real(std::vector<double> &c, int h, int w) { bspline_2d(c.data(), h, w); }
void bspline_complex(std::vector<cd> &c, int h, int w) { bspline_2d(c.data(), h, w); }

int mirror(long i, int n) {
    if (n == 1) return 0;
    long p = 2L * (n - 1);
    i = std::labs(i) % p;
    return (int)(i >= n ? p - i : i);
}

void weights(double t, double wv[4]) {
    double t2 = t * t, t3 = t2 * t;
    double u = 1 - t;
    wv[0] = u * u * u / 6;
    wv[1] = (3 * t3 - 6 * t2 + 4) / 6;
    wv[2] = (-3 * t3 + 3 * t2 + 3 * t + 1) / 6;
    wv[3] = t3 / 6;
}

template <class T>
T sample_at(const T *coeffs, int h, int w, double x, double y) {
    // interp.py:52-69: clamp, 4 mirrored taps per axis, sum over x inside, then over y
    x = std::min(std::max(x, 0.0), w - 1.0);
    y = std::min(std::max(y, 0.0), h - 1.0);
    if (std::isnan(x) || std::isnan(y)) return T(NAN);
    long x0 = (long)std::floor(x), y0 = (long)std::floor(y);
    double wx[4], wy[4];
    weights(x - x0, wx);
    weights(y - y0, wy);
    int xi[4];
    for (int k = 0; k < 4; k++) xi[k] = mirror(x0 + k - 1, w);
    T out = T(0);
    for (int j = 0; j < 4; j++) {
        const T *row = coeffs + (size_t)mirror(y0 + j - 1, h) * w;
        T acc = wx[0] * row[xi[0]];
        for (int k = 1; k < 4; k++) acc = acc + wx[k] * row[xi[k]];
        out += wy[j] * acc;
    }
    return out;
}
template double sample_at<double>(const double *, int, int, double, double);
template cd sample_at<cd>(const cd *, int, int, double, double);

std::vector<double> gauss_kernel(double sigma) {
    int r = (int)(4.0 * sigma + 0.5);
    std::vector<double> k(2 * r + 1);
    double s = 0;
    for (int i = -r; i <= r; i++) {
        double t = i / sigma;
        k[i + r] = std::exp(-0.5 * t * t);
        s += k[i + r];
    }
    for (auto &v : k) v /= s;
    return k;
}

void correlate1d(const double *in, int h, int w, const double *k, int klen, int axis, Mode mode,
                 double *out) {
    const int r = klen / 2;
    if (axis == 0) {
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++) {
            double *o = out + (size_t)y * w;
            for (int x = 0; x < w; x++) o[x] = 0.0;
            for (int j = -r; j <= r; j++) {
                int yy = y + j;
                const double kj = k[j + r];
                if (yy < 0 || yy >= h) {
                    if (mode == CONSTANT) continue;
                    yy = yy < 0 ? 0 : h - 1;
                }
                const double *src = in + (size_t)yy * w;
                for (int x = 0; x < w; x++) o[x] += kj * src[x];
            }
        }
    } else {
#pragma omp parallel
        {
            std::vector<double> padded(w + 2 * r);
#pragma omp for schedule(static)
            for (int y = 0; y < h; y++) {
                const double *src = in + (size_t)y * w;
                for (int x = -r; x < w + r; x++) {
                    double v;
                    if (x < 0) v = mode == CONSTANT ? 0.0 : src[0];
                    else if (x >= w) v = mode == CONSTANT ? 0.0 : src[w - 1];
                    else v = src[x];
                    padded[x + r] = v;
                }
                double *o = out + (size_t)y * w;
                for (int x = 0; x < w; x++) {
                    double s = 0;
                    const double *p = &padded[x];
                    for (int j = 0; j < klen; j++) s += k[j] * p[j];
                    o[x] = s;
                }
            }
        }
    }
}

void gauss_blur(const double *in, int h, int w, double sigma, double *out) {
    // interp.py:80-91: correlate along axis 0 then axis 1, edges replicated
    std::vector<double> k = gauss_kernel(sigma);
    std::vector<double> tmp((size_t)h * w);
    correlate1d(in, h, w, k.data(), (int)k.size(), 0, NEAREST, tmp.data());
    correlate1d(tmp.data(), h, w, k.data(), (int)k.size(), 1, NEAREST, out);
}

double median(std::vector<double> v) {
    size_t n = v.size();
    if (n == 0) return NAN;
    for (double x : v)
        if (std::isnan(x)) return NAN;
    size_t mid = n / 2;
    std::nth_element(v.begin(), v.begin() + mid, v.end());
    double upper = v[mid];
    if (n % 2) return upper;
    double lower = *std::max_element(v.begin(), v.begin() + mid);
    return (lower + upper) / 2.0;
}

}  // namespace orc

using namespace orc;

extern "C" void orc_bspline_coeffs(const double *in, int h, int w, double *out) {
    std::vector<double> c(in, in + (size_t)h * w);
    bspline_real(c, h, w);
    std::memcpy(out, c.data(), c.size() * sizeof(double));
}

extern "C" void orc_bspline_coeffs_c(const double *in, int h, int w, double *out) {
    std::vector<cd> c((size_t)h * w);
    for (size_t i = 0; i < c.size(); i++) c[i] = cd(in[2 * i], in[2 * i + 1]);
    bspline_complex(c, h, w);
    for (size_t i = 0; i < c.size(); i++) { out[2 * i] = c[i].real(); out[2 * i + 1] = c[i].imag(); }
}

extern "C" void orc_sample(const double *coeffs, int h, int w, const double *x, const double *y,
                           long n, double *out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) out[i] = sample_at(coeffs, h, w, x[i], y[i]);
}

extern "C" void orc_warp(const double *img, int h, int w, const double *ux, const double *uy,
                         double *out) {
    std::vector<double> c(img, img + (size_t)h * w);
    bspline_real(c, h, w);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            out[i] = sample_at(c.data(), h, w, x + ux[i], y + uy[i]);
        }
}

extern "C" void orc_gauss_blur(const double *in, int h, int w, double sigma, double *out) {
    gauss_blur(in, h, w, sigma, out);
}

extern "C" double orc_median(const double *v, long n) {
    return median(std::vector<double>(v, v + n));
}
