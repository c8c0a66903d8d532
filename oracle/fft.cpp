// FP64 FFT of any length for the oracle (test infrastructure only).
//
// numpy.fft convention (python/optflow_ref/phaseflow.py:102-107, poc.py:44-48): forward
// X[k] = sum_n x[n] exp(-2 pi i k n / N), inverse with the + sign and a 1/N factor. The
// transform is a recursive decimation-in-time over the prime factors of N; a prime factor is
// combined by a direct DFT of that radix, so every length works without Bluestein. The Rust
// port calls rustfft 6.4.1 (rust/crates/optflow-core/src/fft.rs:40-48); both are exact DFTs up
// to rounding, which the tests pin against a naive DFT and numpy.
#include <map>
#include <mutex>

#include "common.hpp"

namespace orc {

namespace {

struct Plan {
    int n = 0;
    std::vector<int> factors;  // prime factors, 4 first where possible
    std::vector<cd> tw;        // exp(-2 pi i e / n), e in [0, n)
};

const Plan &plan_for(int n) {
    static std::mutex mu;
    static std::map<int, Plan *> plans;
    std::lock_guard<std::mutex> lock(mu);
    auto it = plans.find(n);
    if (it != plans.end()) return *it->second;
    Plan *p = new Plan;
    p->n = n;
    int m = n;
    while (m % 4 == 0) { p->factors.push_back(4); m /= 4; }
    for (int f = 2; m > 1;) {
        if (m % f == 0) { p->factors.push_back(f); m /= f; }
        else f = (f == 2) ? 3 : f + 2;
    }
    p->tw.resize(n);
    for (int e = 0; e < n; e++) {
        double a = -2.0 * M_PI * (double)e / (double)n;
        p->tw[e] = cd(std::cos(a), std::sin(a));
    }
    plans[n] = p;
    return *p;
}

// out[0..n) = DFT of in[0], in[stride], ... (length n); the twiddle of the length-n sub-problem is
// tw[e * (N / n)].
void rec(const cd *in, cd *out, int n, long stride, const Plan &P, size_t level, bool inverse,
         std::vector<cd> &tmp) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = P.factors[level];
    int m = n / p;
    for (int r = 0; r < p; r++) rec(in + r * stride, out + (long)r * m, m, stride * p, P, level + 1, inverse, tmp);
    const int scale = P.n / n;
    cd buf[64];
    std::vector<cd> big;
    cd *v = buf;
    if (p > 64) { big.resize(p); v = big.data(); }
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            long e = ((long)r * k % n) * scale;
            cd t = P.tw[e % P.n];
            if (inverse) t = std::conj(t);
            v[r] = out[(long)r * m + k] * t;
        }
        for (int q = 0; q < p; q++) {
            cd s = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * q % p) * (P.n / p);
                cd t = P.tw[e];
                if (inverse) t = std::conj(t);
                s += v[r] * t;
            }
            tmp[q] = s;
        }
        for (int q = 0; q < p; q++) out[k + (long)m * q] = tmp[q];
    }
}

}  // namespace

void fft_line(cd *x, int n, bool inverse) {
    if (n <= 1) return;
    const Plan &P = plan_for(n);
    std::vector<cd> src(x, x + n), tmp(std::max(n, 64));
    rec(src.data(), x, n, 1, P, 0, inverse, tmp);
    if (inverse) {
        double s = 1.0 / n;
        for (int i = 0; i < n; i++) x[i] *= s;
    }
}

void fft2(std::vector<cd> &a, int h, int w, bool inverse) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) fft_line(&a[(size_t)y * w], w, inverse);
#pragma omp parallel
    {
        std::vector<cd> col(h);
#pragma omp for schedule(static)
        for (int x = 0; x < w; x++) {
            for (int y = 0; y < h; y++) col[y] = a[(size_t)y * w + x];
            fft_line(col.data(), h, inverse);
            for (int y = 0; y < h; y++) a[(size_t)y * w + x] = col[y];
        }
    }
}

}  // namespace orc

extern "C" void orc_fft1(const double *in, int n, int inverse, double *out) {
    std::vector<orc::cd> a(n);
    for (int i = 0; i < n; i++) a[i] = orc::cd(in[2 * i], in[2 * i + 1]);
    orc::fft_line(a.data(), n, inverse != 0);
    for (int i = 0; i < n; i++) { out[2 * i] = a[i].real(); out[2 * i + 1] = a[i].imag(); }
}

extern "C" void orc_fft2(const double *in, int h, int w, int inverse, double *out) {
    std::vector<orc::cd> a((size_t)h * w);
    for (size_t i = 0; i < a.size(); i++) a[i] = orc::cd(in[2 * i], in[2 * i + 1]);
    orc::fft2(a, h, w, inverse != 0);
    for (size_t i = 0; i < a.size(); i++) { out[2 * i] = a[i].real(); out[2 * i + 1] = a[i].imag(); }
}
