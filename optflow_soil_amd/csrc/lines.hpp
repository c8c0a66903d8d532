// Device helpers for work on whole lines held in LDS: complex arithmetic, Stockham FFT passes
// and the segmented cubic B-spline prefilter.
#pragma once
#include <hip/hip_runtime.h>

namespace ofx {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 mul_minus_i(float2 a) { return make_float2(a.y, -a.x); }  // -i a
__device__ __forceinline__ float2 mul_plus_i(float2 a) { return make_float2(-a.y, a.x); }   // +i a

// ---------------------------------------------------------------- forward DFT kernels of small radix
// y_q = sum_t x_t exp(-2 pi i q t / R)
__device__ __forceinline__ void dft2(float2 *x) {
    float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}
__device__ __forceinline__ void dft4(float2 *x) {
    float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    float2 c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, mul_minus_i(d));
    x[3] = cadd(b, mul_plus_i(d));
}
__device__ __forceinline__ void dft3(float2 *x) {
    const float c = -0.5f, s = -0.86602540378443864676f;  // exp(-2 pi i / 3)
    float2 t = cadd(x[1], x[2]), d = csub(x[1], x[2]);
    float2 m = make_float2(x[0].x + c * t.x, x[0].y + c * t.y);
    float2 isd = make_float2(-s * d.y, s * d.x);  // i s d
    x[0] = cadd(x[0], t);
    x[1] = cadd(m, isd);
    x[2] = csub(m, isd);
}
__device__ __forceinline__ void dft5(float2 *x) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    float2 x0 = x[0];
    float2 p1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    float2 p2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    float2 q1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    float2 q2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    x[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    x[1] = cadd(p1, mul_minus_i(q1));
    x[4] = cadd(p1, mul_plus_i(q1));
    x[2] = cadd(p2, mul_minus_i(q2));
    x[3] = cadd(p2, mul_plus_i(q2));
}
__device__ __forceinline__ void dft8(float2 *x) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {x[0], x[2], x[4], x[6]};
    float2 o[4] = {x[1], x[3], x[5], x[7]};
    dft4(e);
    dft4(o);
    // o_q *= exp(-2 pi i q / 8)
    o[1] = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = mul_minus_i(o[2]);
    o[3] = make_float2(h * (-o[3].x + o[3].y), h * (-o[3].y - o[3].x));
    for (int q = 0; q < 4; q++) {
        x[q] = cadd(e[q], o[q]);
        x[q + 4] = csub(e[q], o[q]);
    }
}

template <int R>
__device__ __forceinline__ void dft_small(float2 *x) {
    if constexpr (R == 2) dft2(x);
    else if constexpr (R == 3) dft3(x);
    else if constexpr (R == 4) dft4(x);
    else if constexpr (R == 5) dft5(x);
    else if constexpr (R == 8) dft8(x);
}

// One self-sorting Stockham pass (radix R, p = product of the radices already applied) of `lines`
// lines of length n stored line-major at src / dst (line stride n). Butterfly i of a line reads
// src[i + q m] (m = n / R), applies W_{pR}^{q k} (k = i mod p), transforms, and writes
// dst[(i / p) p R + k + q p].
template <int R>
__device__ __forceinline__ void stockham_small(const float2 *src, float2 *dst, int n, int lines, int p,
                                               const float2 *__restrict__ tw, int tid, int nt) {
    const int m = n / R;
    const int step = n / (p * R);
    for (int item = tid; item < m * lines; item += nt) {
        const int line = item / m, i = item - line * m;
        const float2 *s = src + line * n;
        float2 *d = dst + line * n;
        const int k = i % p;
        float2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = s[i + q * m];
#pragma unroll
        for (int q = 1; q < R; q++) x[q] = cmul(x[q], tw[q * k * step]);
        dft_small<R>(x);
        const int j = (i / p) * p * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) d[j + q * p] = x[q];
    }
}

// The same pass for any radix R, one output per thread: output o = sum_t x_t W_{pR}^{t (o mod pR)}
// with x_t = src[i + t m], i = (o / pR) p + o mod p.
__device__ __forceinline__ void stockham_generic(const float2 *src, float2 *dst, int n, int lines, int p, int R,
                                                 const float2 *__restrict__ tw, int tid, int nt) {
    const int m = n / R, pR = p * R, step = n / pR;
    for (int item = tid; item < n * lines; item += nt) {
        const int line = item / n, o = item - line * n;
        const float2 *s = src + line * n;
        const int r = o % pR;
        const int i = (o / pR) * p + (o % p);
        float ax = 0.f, ay = 0.f;
        int e = 0;  // t * r mod pR
        for (int t = 0; t < R; t++) {
            float2 w = tw[e * step];
            float2 v = s[i + t * m];
            ax += v.x * w.x - v.y * w.y;
            ay += v.x * w.y + v.y * w.x;
            e += r;
            if (e >= pR) e -= pR;
        }
        dst[line * n + o] = make_float2(ax, ay);
    }
}

// Forward DFT of `lines` lines of length n in LDS buffer a (line-major); b is scratch of the same
// size. Returns the buffer holding the result. Ends with a barrier.
__device__ __forceinline__ float2 *fft_lines_lds(float2 *a, float2 *b, int n, int lines, const int *fac,
                                                 int nfac, const float2 *__restrict__ tw, int tid, int nt) {
    int p = 1;
    for (int s = 0; s < nfac; s++) {
        const int R = fac[s];
        switch (R) {
            case 2: stockham_small<2>(a, b, n, lines, p, tw, tid, nt); break;
            case 3: stockham_small<3>(a, b, n, lines, p, tw, tid, nt); break;
            case 4: stockham_small<4>(a, b, n, lines, p, tw, tid, nt); break;
            case 5: stockham_small<5>(a, b, n, lines, p, tw, tid, nt); break;
            case 8: stockham_small<8>(a, b, n, lines, p, tw, tid, nt); break;
            default: stockham_generic(a, b, n, lines, p, R, tw, tid, nt); break;
        }
        __syncthreads();
        float2 *t = a;
        a = b;
        b = t;
        p *= R;
    }
    return a;
}

// ---------------------------------------------------------------- cubic B-spline prefilter of a line
// Coefficients of python/optflow_ref/interp.py:12-26 (x 6, causal start over a 21-sample horizon,
// causal and anti-causal recursions with z = sqrt(3) - 2, mirror end), evaluated with the line
// split into segments of SEG samples: each segment starts its recursion WARM samples early from
// zero, which leaves an error of z^WARM (< 2e-14) relative to the sequential recursion, far
// below float32 storage. Arithmetic is double; the result is multiplied by `scale`.
constexpr int PF_SEG = 32;
constexpr int PF_WARM = 24;
constexpr int PF_HORIZON = 21;  // ceil(log(1e-12) / log|z|)

template <class T>
struct Dbl;
template <>
struct Dbl<float> {
    double v;
    __device__ Dbl() : v(0) {}
    __device__ explicit Dbl(float a) : v(a) {}
    __device__ Dbl(double a, int) : v(a) {}
    __device__ Dbl operator+(const Dbl &o) const { return Dbl(v + o.v, 0); }
    __device__ Dbl operator-(const Dbl &o) const { return Dbl(v - o.v, 0); }
    __device__ Dbl operator*(double s) const { return Dbl(v * s, 0); }
    __device__ float out(double s) const { return (float)(v * s); }
};
template <>
struct Dbl<float2> {
    double x, y;
    __device__ Dbl() : x(0), y(0) {}
    __device__ explicit Dbl(float2 a) : x(a.x), y(a.y) {}
    __device__ Dbl(double a, double b) : x(a), y(b) {}
    __device__ Dbl operator+(const Dbl &o) const { return Dbl(x + o.x, y + o.y); }
    __device__ Dbl operator-(const Dbl &o) const { return Dbl(x - o.x, y - o.y); }
    __device__ Dbl operator*(double s) const { return Dbl(x * s, y * s); }
    __device__ float2 out(double s) const { return make_float2((float)(x * s), (float)(y * s)); }
};

// `lines` lines of length n, line-major with stride n. src: input; mid: scratch (distinct from
// src); dst: output, which may alias src (src is only read before the barrier between passes).
template <class T>
__device__ void prefilter_lines(const T *src, T *mid, T *dst, int n, int lines, double scale, int tid, int nt) {
    const double z = -0.26794919243112270647;  // sqrt(3) - 2
    if (n == 1) {
        for (int l = tid; l < lines; l += nt) dst[l] = Dbl<T>(src[l]).out(scale);
        __syncthreads();
        return;
    }
    const int nseg = (n + PF_SEG - 1) / PF_SEG;
    // causal pass: mid[i] = 6 src[i] + z mid[i-1], mid[0] = sum_{k<h} z^k 6 src[k]
    for (int item = tid; item < nseg * lines; item += nt) {
        const int line = item / nseg, j = item - line * nseg;
        const T *sl = src + (size_t)line * n;
        T *ml = mid + (size_t)line * n;
        const int a = j * PF_SEG, b = min(a + PF_SEG, n);
        Dbl<T> acc;
        int i;
        if (a <= PF_WARM) {
            const int h = min(n, PF_HORIZON);
            double zk = 1.0;
            for (int k = 0; k < h; k++) {
                acc = acc + Dbl<T>(sl[k]) * (6.0 * zk);
                zk *= z;
            }
            if (a == 0) ml[0] = acc.out(1.0);
            i = 1;
        } else {
            i = a - PF_WARM;
        }
        for (; i < a; i++) acc = Dbl<T>(sl[i]) * 6.0 + acc * z;
        for (i = max(a, 1); i < b; i++) {
            acc = Dbl<T>(sl[i]) * 6.0 + acc * z;
            ml[i] = acc.out(1.0);
        }
    }
    __syncthreads();
    // anti-causal pass: dst[n-1] = z / (z^2 - 1) (mid[n-1] + z mid[n-2]), dst[i] = z (dst[i+1] - mid[i])
    const double zend = z / (z * z - 1.0);
    for (int item = tid; item < nseg * lines; item += nt) {
        const int line = item / nseg, j = item - line * nseg;
        const T *ml = mid + (size_t)line * n;
        T *dl = dst + (size_t)line * n;
        const int a = j * PF_SEG, b = min(a + PF_SEG, n);
        Dbl<T> acc;
        int i;
        if (b + PF_WARM >= n) {
            acc = (Dbl<T>(ml[n - 1]) + Dbl<T>(ml[n - 2]) * z) * zend;
            if (b == n) dl[n - 1] = acc.out(scale);
            i = n - 2;
        } else {
            i = b + PF_WARM - 1;
        }
        for (; i >= b; i--) acc = (acc - Dbl<T>(ml[i])) * z;
        for (i = min(b - 1, n - 2); i >= a; i--) {
            acc = (acc - Dbl<T>(ml[i])) * z;
            dl[i] = acc.out(scale);
        }
    }
    __syncthreads();
}

}  // namespace ofx
