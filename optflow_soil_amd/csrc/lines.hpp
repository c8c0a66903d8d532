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

// ---------------------------------------------------------------- LDS line layout
// With SH > 0, element e of an LDS line buffer lives at e + (e >> SH): one spare slot every 2^SH
// elements, so that passes writing at strides of 2..16 elements spread over the LDS banks.
template <int SH>
__device__ __forceinline__ int lp(int e) {
    if constexpr (SH > 0) return e + (e >> SH);
    else return e;
}
template <int SH>
__host__ __device__ constexpr int lp_len(int n) {
    return SH > 0 ? n + (n >> SH) + 1 : n;
}

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

// y_q = sum_t x_t W_R^{q t} in registers for an odd prime R, with W_R^m = w[m] (m < R) taken from
// the line's twiddle table. The pairs t, R - t share cos and flip sin: with a_t = x_t + x_{R-t} and
// b_t = x_t - x_{R-t} (t = 1 .. (R-1)/2), P_q = x_0 + sum_t a_t cos(2 pi q t / R) and
// Q_q = sum_t b_t sin(2 pi q t / R) give y_q = P_q - i Q_q and y_{R-q} = P_q + i Q_q: (R-1)^2 real
// multiply-adds instead of the (R-1)^2 complex ones of the direct sum.
template <int R>
__device__ __forceinline__ void dft_prime(float2 *x, const float2 *w) {
    constexpr int H = (R - 1) / 2;
    float2 a[H + 1], b[H + 1];
#pragma unroll
    for (int t = 1; t <= H; t++) {
        a[t] = cadd(x[t], x[R - t]);
        b[t] = csub(x[t], x[R - t]);
    }
    float2 y0 = x[0];
#pragma unroll
    for (int t = 1; t <= H; t++) y0 = cadd(y0, a[t]);
#pragma unroll
    for (int q = 1; q <= H; q++) {
        float px = x[0].x, py = x[0].y, qx = 0.f, qy = 0.f;
#pragma unroll
        for (int t = 1; t <= H; t++) {
            const float2 c = w[(q * t) % R];  // (cos, -sin) of 2 pi q t / R
            px = fmaf(a[t].x, c.x, px);
            py = fmaf(a[t].y, c.x, py);
            qx = fmaf(b[t].x, -c.y, qx);
            qy = fmaf(b[t].y, -c.y, qy);
        }
        x[q] = make_float2(px + qy, py - qx);
        x[R - q] = make_float2(px - qy, py + qx);
    }
    x[0] = y0;
}

// exp(-2 pi i m / N) for N = 12 and 16, as (cos, -sin)
template <int N>
struct TwConst;
template <>
struct TwConst<16> {
    static constexpr float c[16] = {1.f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                                    0.f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
                                    -1.f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f,
                                    0.f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f};
    static constexpr int Q = 4;  // c[(m + Q) mod N] is -sin(2 pi m / N)
};
template <>
struct TwConst<12> {
    static constexpr float c[12] = {1.f, 0.86602540378443864676f, 0.5f, 0.f, -0.5f, -0.86602540378443864676f,
                                    -1.f, -0.86602540378443864676f, -0.5f, 0.f, 0.5f, 0.86602540378443864676f};
    static constexpr int Q = 3;
};

// x * exp(-2 pi i m / N) for a compile-time m: exact for the multiples of N / 4
template <int N, int M>
__device__ __forceinline__ float2 tw_mul(float2 x) {
    constexpr int m = M % N;
    if constexpr (m == 0) return x;
    else if constexpr (4 * m == N) return mul_minus_i(x);
    else if constexpr (2 * m == N) return make_float2(-x.x, -x.y);
    else if constexpr (4 * m == 3 * N) return mul_plus_i(x);
    else {
        constexpr float c = TwConst<N>::c[m], sn = TwConst<N>::c[(m + TwConst<N>::Q) % N];  // cos, -sin
        return make_float2(x.x * c - x.y * sn, x.x * sn + x.y * c);
    }
}

template <int R>
__device__ __forceinline__ void dft_small(float2 *x);

// DFT of length N = N1 N2 from DFTs of N1 and N2 (Cooley-Tukey, twiddles in between):
// X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_N^(n2 k1) sum_n1 W_N1^(n1 k1) x[N2 n1 + n2].
template <int N, int N2I, int N1, int K = 0>
__device__ __forceinline__ void twiddle_row(const float2 (&t)[N1], float2 (&row)[N1]) {
    if constexpr (K < N1) {
        row[K] = tw_mul<N, N2I * K>(t[K]);
        twiddle_row<N, N2I, N1, K + 1>(t, row);
    }
}
template <int N1, int N2, int N2I = 0>
__device__ __forceinline__ void inner_columns(const float2 *x, float2 (&y)[N2][N1]) {
    if constexpr (N2I < N2) {
        float2 t[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) t[n1] = x[N2 * n1 + N2I];
        dft_small<N1>(t);
        twiddle_row<N1 * N2, N2I, N1>(t, y[N2I]);
        inner_columns<N1, N2, N2I + 1>(x, y);
    }
}
template <int N1, int N2>
__device__ __forceinline__ void dft_composite(float2 *x) {
    float2 y[N2][N1];
    inner_columns<N1, N2>(x, y);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
        float2 t[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++) t[n2] = y[n2][k1];
        dft_small<N2>(t);
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) x[k1 + N1 * k2] = t[k2];
    }
}

template <int R>
__device__ __forceinline__ void dft_small(float2 *x) {
    if constexpr (R == 2) dft2(x);
    else if constexpr (R == 3) dft3(x);
    else if constexpr (R == 4) dft4(x);
    else if constexpr (R == 5) dft5(x);
    else if constexpr (R == 8) dft8(x);
    else if constexpr (R == 12) dft_composite<4, 3>(x);
    else if constexpr (R == 16) dft_composite<4, 4>(x);
}

// One self-sorting Stockham pass (radix R, p = product of the radices already applied) of `lines`
// lines of length n stored line-major at src / dst (line stride n). Butterfly i of a line reads
// src[i + q m] (m = n / R), applies W_{pR}^{q k} (k = i mod p), transforms, and writes
// dst[(i / p) p R + k + q p].
template <int R, int SH = 0>
__device__ __forceinline__ void stockham_small(const float2 *src, float2 *dst, int n, int lines, int p,
                                               const float2 *__restrict__ tw, int tid, int nt) {
    const int m = n / R;
    const int step = n / (p * R);
    for (int item = tid; item < m * lines; item += nt) {
        const int line = item / m, i = item - line * m;
        const int base = line * n;
        const int k = i % p;
        float2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = src[lp<SH>(base + i + q * m)];
#pragma unroll
        for (int q = 1; q < R; q++) x[q] = cmul(x[q], tw[q * k * step]);
        dft_small<R>(x);
        const int j = (i / p) * p * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) dst[lp<SH>(base + j + q * p)] = x[q];
    }
}

// Stockham pass for a prime radix R in {7, 11, 13} with the butterfly in registers.
template <int R, int SH = 0>
__device__ __forceinline__ void stockham_prime(const float2 *src, float2 *dst, int n, int lines, int p,
                                               const float2 *__restrict__ tw, int tid, int nt) {
    const int m = n / R;
    const int step = n / (p * R);
    float2 wr[R];
#pragma unroll
    for (int q = 0; q < R; q++) wr[q] = tw[q * (n / R)];
    for (int item = tid; item < m * lines; item += nt) {
        const int line = item / m, i = item - line * m;
        const int base = line * n;
        const int k = i % p;
        float2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = src[lp<SH>(base + i + q * m)];
#pragma unroll
        for (int q = 1; q < R; q++) x[q] = cmul(x[q], tw[q * k * step]);
        dft_prime<R>(x, wr);
        const int j = (i / p) * p * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) dst[lp<SH>(base + j + q * p)] = x[q];
    }
}

// The same pass for any radix R, one output per thread: output o = sum_t x_t W_{pR}^{t (o mod pR)}
// with x_t = src[i + t m], i = (o / pR) p + o mod p.
template <int SH = 0>
__device__ __forceinline__ void stockham_generic(const float2 *src, float2 *dst, int n, int lines, int p, int R,
                                                 const float2 *__restrict__ tw, int tid, int nt) {
    const int m = n / R, pR = p * R, step = n / pR;
    for (int item = tid; item < n * lines; item += nt) {
        const int line = item / n, o = item - line * n;
        const int base = line * n;
        const int r = o % pR;
        const int i = (o / pR) * p + (o % p);
        float ax = 0.f, ay = 0.f;
        int e = 0;  // t * r mod pR
        for (int t = 0; t < R; t++) {
            float2 w = tw[e * step];
            float2 v = src[lp<SH>(base + i + t * m)];
            ax += v.x * w.x - v.y * w.y;
            ay += v.x * w.y + v.y * w.x;
            e += r;
            if (e >= pR) e -= pR;
        }
        dst[lp<SH>(base + o)] = make_float2(ax, ay);
    }
}

// i mod p for i < 2^16, p < 2^16, from magic = floor(2^32 / p) + 1 (0 for p = 1): the quotient
// umulhi(i, magic) is exact because i (magic p - 2^32) < 2^32.
__host__ __device__ inline unsigned int mod_magic(unsigned int p) {
    return p <= 1 ? 0u : (unsigned int)((0x100000000ull / p) + 1);
}
__device__ __forceinline__ int mod_by(int i, int p, unsigned int magic) {
    return p == 1 ? 0 : i - (int)__umulhi((unsigned int)i, magic) * p;
}

// One Stockham pass of radix R over a single line in LDS (p = product of the radices already
// applied, magic = mod_magic(p)). Butterfly i reads src[i + q m] (m = n / R), applies
// W_{pR}^{q k} (k = i mod p), transforms, and writes dst[i R - k (R - 1) + q p]
// (= (i / p) p R + k + q p).
template <int R, int SH>
__device__ __forceinline__ void stockham_line(const float2 *src, float2 *dst, int n, int p, unsigned int magic,
                                              const float2 *tw, int tid, int nt) {
    const int m = n / R;
    const int step = n / (p * R);
    float2 wr[R];
    if constexpr (R == 7 || R == 11 || R == 13) {
#pragma unroll
        for (int q = 0; q < R; q++) wr[q] = tw[q * m];
    }
    for (int i = tid; i < m; i += nt) {
        const int k = mod_by(i, p, magic);
        float2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = src[lp<SH>(i + q * m)];
        const int ks = k * step;
#pragma unroll
        for (int q = 1; q < R; q++) x[q] = cmul(x[q], tw[q * ks]);
        if constexpr (R == 7 || R == 11 || R == 13) dft_prime<R>(x, wr);
        else dft_small<R>(x);
        const int j = i * R - k * (R - 1);
#pragma unroll
        for (int q = 0; q < R; q++) dst[lp<SH>(j + q * p)] = x[q];
    }
}

// First Stockham pass (p = 1, no twiddles) of one line whose element e is load(e), written to dst.
template <int R, int SH, class L>
__device__ __forceinline__ void first_pass_small(L load, float2 *dst, int n, const float2 *tw, int tid, int nt) {
    const int m = n / R;
    float2 wr[R];
    if constexpr (R == 7 || R == 11 || R == 13) {
#pragma unroll
        for (int q = 0; q < R; q++) wr[q] = tw[q * m];
    }
    for (int i = tid; i < m; i += nt) {
        float2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = load(i + q * m);
        if constexpr (R == 7 || R == 11 || R == 13) dft_prime<R>(x, wr);
        else dft_small<R>(x);
#pragma unroll
        for (int q = 0; q < R; q++) dst[lp<SH>(i * R + q)] = x[q];
    }
}

// Forward DFT of one line of length n whose element e is load(e): the first pass reads straight from
// the loader (global memory) into registers, the others run in LDS buffers a / b with the twiddle
// table tw (LDS). magic[s] = mod_magic(fac[0] ... fac[s-1]). Returns the buffer holding the
// result; ends with a barrier. Radices without a register butterfly load the line into a first.
template <int SH, class L>
__device__ __forceinline__ float2 *fft_line_from(L load, float2 *a, float2 *b, int n, const int *fac,
                                                 const unsigned int *magic, int nfac, const float2 *tw, int tid,
                                                 int nt) {
    const int R0 = nfac > 0 ? fac[0] : 1;
    int s0 = 1;
    switch (R0) {
        case 2: first_pass_small<2, SH>(load, b, n, tw, tid, nt); break;
        case 3: first_pass_small<3, SH>(load, b, n, tw, tid, nt); break;
        case 4: first_pass_small<4, SH>(load, b, n, tw, tid, nt); break;
        case 5: first_pass_small<5, SH>(load, b, n, tw, tid, nt); break;
        case 7: first_pass_small<7, SH>(load, b, n, tw, tid, nt); break;
        case 8: first_pass_small<8, SH>(load, b, n, tw, tid, nt); break;
        case 12: first_pass_small<12, SH>(load, b, n, tw, tid, nt); break;
        case 16: first_pass_small<16, SH>(load, b, n, tw, tid, nt); break;
        case 11: first_pass_small<11, SH>(load, b, n, tw, tid, nt); break;
        case 13: first_pass_small<13, SH>(load, b, n, tw, tid, nt); break;
        default:
            for (int e = tid; e < n; e += nt) a[lp<SH>(e)] = load(e);
            s0 = 0;
            break;
    }
    __syncthreads();
    if (s0 == 1) {
        float2 *t = a;
        a = b;
        b = t;
    }
    int p = s0 == 1 ? R0 : 1;
    for (int s = s0; s < nfac; s++) {
        const int R = fac[s];
        const unsigned int mg = magic[s];
        switch (R) {
            case 2: stockham_line<2, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 3: stockham_line<3, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 4: stockham_line<4, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 5: stockham_line<5, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 8: stockham_line<8, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 12: stockham_line<12, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 7: stockham_line<7, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 11: stockham_line<11, SH>(a, b, n, p, mg, tw, tid, nt); break;
            case 13: stockham_line<13, SH>(a, b, n, p, mg, tw, tid, nt); break;
            default: stockham_generic<SH>(a, b, n, 1, p, R, tw, tid, nt); break;
        }
        __syncthreads();
        float2 *t = a;
        a = b;
        b = t;
        p *= R;
    }
    return a;
}

// ---------------------------------------------------------------- compile-time line transforms
// For a line length N = 16 R1 R2 ... whose later strides are multiples of 16, the padded LDS index
// lp(i + q M) of a butterfly's q-th element is lp(i) + q M 17 / 16, so every read and write of a
// pass is a constant offset from one base address, and the pass's twiddles W_{P R}^{q k} come from
// a table laid out [k][q - 1] (values gathered from the line's table, so the arithmetic is that of
// the runtime passes). NT threads per line.
template <int R>
constexpr bool prime_radix() {
    return R == 7 || R == 11 || R == 13 || R == 17 || R == 19;
}

template <int R, int N, int P, int SH, int NT>
__device__ __forceinline__ void ct_pass(const float2 *src, float2 *dst, const float2 *__restrict__ T,
                                        const float2 *__restrict__ tw_line, int tid) {
    constexpr int M = N / R;
    static_assert(SH == 4 && M % 16 == 0 && P % 16 == 0, "strides of whole 16-element groups");
    constexpr int SM = M + M / 16, SP = P + P / 16;
    float2 wr[R];
    if constexpr (prime_radix<R>()) {
#pragma unroll
        for (int q = 0; q < R; q++) wr[q] = tw_line[q * M];
    }
#pragma unroll
    for (int i0 = 0; i0 < M; i0 += NT) {
        const int i = i0 + tid;
        if (M % NT != 0 && i >= M) break;
        const int k = i % P;
        float2 x[R];
        const float2 *sp = src + lp<SH>(i);
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = sp[q * SM];
        const float2 *t = T + k * (R - 1);
#pragma unroll
        for (int q = 1; q < R; q++) x[q] = cmul(x[q], t[q - 1]);
        if constexpr (prime_radix<R>()) dft_prime<R>(x, wr);
        else dft_small<R>(x);
        float2 *dp = dst + lp<SH>((i / P) * P * R + k);
#pragma unroll
        for (int q = 0; q < R; q++) dp[q * SP] = x[q];
    }
}

// The passes after the first: radix R at stage P, then the rest; returns the buffer holding the
// result. Each pass ends with a barrier.
template <int N, int P, int SH, int NT, int R, int... Rest>
__device__ __forceinline__ float2 *ct_passes(float2 *src, float2 *dst, const float2 *__restrict__ T,
                                             const float2 *__restrict__ tw_line, int tid) {
    ct_pass<R, N, P, SH, NT>(src, dst, T, tw_line, tid);
    __syncthreads();
    if constexpr (sizeof...(Rest) == 0) return dst;
    else return ct_passes<N, P * R, SH, NT, Rest...>(dst, src, T + P * (R - 1), tw_line, tid);
}

template <int... Rs>
constexpr int radix_product() {
    return (1 * ... * Rs);
}

// The transform of a length-N line (N = 16 Rs...) whose element e is load(e): a radix-16 first pass
// from the loader, then the passes Rs; T holds the tables of the later passes one after the other.
// Returns the buffer holding the result; ends with a barrier.
template <int N, int SH, int NT, int... Rs, class L>
__device__ __forceinline__ float2 *fft_line_ct(L load, float2 *a, float2 *b, const float2 *__restrict__ T,
                                               const float2 *__restrict__ tw_line, int tid) {
    constexpr int M0 = N / 16;
    static_assert(N == 16 * radix_product<Rs...>(), "N = 16 Rs...");
#pragma unroll
    for (int i0 = 0; i0 < M0; i0 += NT) {
        const int i = i0 + tid;
        if (M0 % NT != 0 && i >= M0) break;
        float2 x[16];
#pragma unroll
        for (int q = 0; q < 16; q++) x[q] = load(i + q * M0);
        dft_small<16>(x);
        float2 *dp = a + 17 * i;  // lp(16 i + q) = 17 i + q
#pragma unroll
        for (int q = 0; q < 16; q++) dp[q] = x[q];
    }
    __syncthreads();
    return ct_passes<N, 16, SH, NT, Rs...>(a, b, T, tw_line, tid);
}

// The same passes on one LDS buffer per line: a pass reads every butterfly of the thread into
// registers, waits for the workgroup, then writes its results over the buffer. Two barriers per
// pass instead of one, but half the LDS, so twice the lines share a CU.
template <int R, int N, int P, int SH, int NT>
__device__ __forceinline__ void ct_pass_inplace(float2 *buf, const float2 *__restrict__ T,
                                                const float2 *__restrict__ tw_line, int tid) {
    constexpr int M = N / R, IT = (M + NT - 1) / NT;
    static_assert(SH == 4 && M % 16 == 0 && P % 16 == 0, "strides of whole 16-element groups");
    constexpr int SM = M + M / 16, SP = P + P / 16;
    float2 wr[R];
    if constexpr (prime_radix<R>()) {
#pragma unroll
        for (int q = 0; q < R; q++) wr[q] = tw_line[q * M];
    }
    float2 x[IT][R];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = it * NT + tid;
        if (M % NT != 0 && i >= M) break;
        const int k = i % P;
        const float2 *sp = buf + lp<SH>(i);
#pragma unroll
        for (int q = 0; q < R; q++) x[it][q] = sp[q * SM];
        const float2 *t = T + k * (R - 1);
#pragma unroll
        for (int q = 1; q < R; q++) x[it][q] = cmul(x[it][q], t[q - 1]);
        if constexpr (prime_radix<R>()) dft_prime<R>(x[it], wr);
        else dft_small<R>(x[it]);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = it * NT + tid;
        if (M % NT != 0 && i >= M) break;
        const int k = i % P;
        float2 *dp = buf + lp<SH>((i / P) * P * R + k);
#pragma unroll
        for (int q = 0; q < R; q++) dp[q * SP] = x[it][q];
    }
    __syncthreads();
}

template <int N, int P, int SH, int NT, int R, int... Rest>
__device__ __forceinline__ void ct_passes_inplace(float2 *buf, const float2 *__restrict__ T,
                                                  const float2 *__restrict__ tw_line, int tid) {
    ct_pass_inplace<R, N, P, SH, NT>(buf, T, tw_line, tid);
    if constexpr (sizeof...(Rest) > 0) ct_passes_inplace<N, P * R, SH, NT, Rest...>(buf, T + P * (R - 1), tw_line, tid);
}

// fft_line_ct on the one buffer `a`; the result is left in `a`. Ends with a barrier. FROM_LDS: the
// loader reads `a` itself, so every thread's loads finish before the first pass writes.
template <int N, int SH, int NT, bool FROM_LDS, int... Rs, class L>
__device__ __forceinline__ void fft_line_ct_inplace(L load, float2 *a, const float2 *__restrict__ T,
                                                    const float2 *__restrict__ tw_line, int tid) {
    constexpr int M0 = N / 16;
    static_assert(N == 16 * radix_product<Rs...>(), "N = 16 Rs...");
    static_assert(!FROM_LDS || M0 <= NT, "a loader reading the buffer needs one butterfly per thread");
#pragma unroll
    for (int i0 = 0; i0 < M0; i0 += NT) {
        const int i = i0 + tid;
        const bool on = M0 % NT == 0 || i < M0;
        float2 x[16];
        if (on) {
#pragma unroll
            for (int q = 0; q < 16; q++) x[q] = load(i + q * M0);
            dft_small<16>(x);
        }
        if constexpr (FROM_LDS) __syncthreads();
        if (on) {
            float2 *dp = a + 17 * i;  // lp(16 i + q) = 17 i + q
#pragma unroll
            for (int q = 0; q < 16; q++) dp[q] = x[q];
        }
    }
    __syncthreads();
    ct_passes_inplace<N, 16, SH, NT, Rs...>(a, T, tw_line, tid);
}

// Stockham pass with compile-time radix R, line length N and p = P over `lines` lines (line stride N):
// every index operation is a shift or a mask when N and P are powers of two.
template <int R, int N, int P>
__device__ __forceinline__ void stockham_ct(const float2 *src, float2 *dst, int lines, const float2 *tw, int tid,
                                            int nt) {
    constexpr int M = N / R, STEP = N / (P * R);
    for (int item = tid; item < M * lines; item += nt) {
        const int line = item / M, i = item - line * M;
        const int base = line * N;
        const int k = i % P;
        float2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = src[base + i + q * M];
        if constexpr (P > 1) {
#pragma unroll
            for (int q = 1; q < R; q++) x[q] = cmul(x[q], tw[q * k * STEP]);
        }
        dft_small<R>(x);
        const int j = (i / P) * P * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) dst[base + j + q * P] = x[q];
    }
}

// Forward DFT of `lines` lines of length 64 (radix 8, 8) in LDS buffer a; b is scratch. Returns
// the buffer holding the result (a). Ends with a barrier.
__device__ __forceinline__ float2 *fft64_lines_lds(float2 *a, float2 *b, int lines, const float2 *tw, int tid, int nt) {
    stockham_ct<8, 64, 1>(a, b, lines, tw, tid, nt);
    __syncthreads();
    stockham_ct<8, 64, 8>(b, a, lines, tw, tid, nt);
    __syncthreads();
    return a;
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
            case 7: stockham_prime<7>(a, b, n, lines, p, tw, tid, nt); break;
            case 11: stockham_prime<11>(a, b, n, lines, p, tw, tid, nt); break;
            case 13: stockham_prime<13>(a, b, n, lines, p, tw, tid, nt); break;
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
// zero, which leaves an error of z^WARM (7e-10) relative to the sequential recursion, far below
// float32 storage (6e-8). Arithmetic is double; the result is multiplied by `scale`.
// 17 = 1 mod 16: the lanes of a wavefront walking neighbouring segments of an LDS line of 8-byte
// elements start 136 bytes apart, i.e. on different banks (a multiple of 16 would put every lane on
// the same bank).
constexpr int PF_SEG = 17;
constexpr int PF_WARM = 16;
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

// Single-precision accumulator with the interface of Dbl: the recursions are contractive (|z| < 1),
// so float arithmetic stays within about two ulp of the exact coefficients.
struct Sgl2 {
    float x, y;
    __device__ Sgl2() : x(0.f), y(0.f) {}
    __device__ explicit Sgl2(float2 a) : x(a.x), y(a.y) {}
    __device__ Sgl2(float a, float b) : x(a), y(b) {}
    __device__ Sgl2 operator+(const Sgl2 &o) const { return Sgl2(x + o.x, y + o.y); }
    __device__ Sgl2 operator-(const Sgl2 &o) const { return Sgl2(x - o.x, y - o.y); }
    __device__ Sgl2 operator*(double s) const { return Sgl2(x * (float)s, y * (float)s); }
    __device__ float2 out(double s) const { return make_float2(x * (float)s, y * (float)s); }
};

// `lines` lines of length n, line-major with stride n. src: input; mid: scratch (distinct from
// src); dst: output, which may alias src (src is only read before the barrier between passes).
// A: the accumulator (Dbl<T>: double arithmetic; Sgl2: float).
template <class T, int SH = 0, class A = Dbl<T>>
__device__ void prefilter_lines(const T *src, T *mid, T *dst, int n, int lines, double scale, int tid, int nt) {
    const double z = -0.26794919243112270647;  // sqrt(3) - 2
    if (n == 1) {
        for (int l = tid; l < lines; l += nt) dst[lp<SH>(l)] = A(src[lp<SH>(l)]).out(scale);
        __syncthreads();
        return;
    }
    const int nseg = (n + PF_SEG - 1) / PF_SEG;
    // causal pass: mid[i] = 6 src[i] + z mid[i-1], mid[0] = sum_{k<h} z^k 6 src[k]
    for (int item = tid; item < nseg * lines; item += nt) {
        const int line = item / nseg, j = item - line * nseg;
        const int base = line * n;
        auto sl = [&](int i) { return src[lp<SH>(base + i)]; };
        auto mset = [&](int i, const A &v) { mid[lp<SH>(base + i)] = v.out(1.0); };
        const int a = j * PF_SEG, b = min(a + PF_SEG, n);
        A acc;
        int i;
        if (a <= PF_WARM) {
            const int h = min(n, PF_HORIZON);
            double zk = 1.0;
            for (int k = 0; k < h; k++) {
                acc = acc + A(sl(k)) * (6.0 * zk);
                zk *= z;
            }
            if (a == 0) mset(0, acc);
            i = 1;
        } else {
            i = a - PF_WARM;
        }
#pragma unroll 4
        for (; i < a; i++) acc = A(sl(i)) * 6.0 + acc * z;
#pragma unroll 4
        for (i = max(a, 1); i < b; i++) {
            acc = A(sl(i)) * 6.0 + acc * z;
            mset(i, acc);
        }
    }
    __syncthreads();
    // anti-causal pass: dst[n-1] = z / (z^2 - 1) (mid[n-1] + z mid[n-2]), dst[i] = z (dst[i+1] - mid[i])
    const double zend = z / (z * z - 1.0);
    for (int item = tid; item < nseg * lines; item += nt) {
        const int line = item / nseg, j = item - line * nseg;
        const int base = line * n;
        auto ml = [&](int i) { return mid[lp<SH>(base + i)]; };
        auto dset = [&](int i, const A &v) { dst[lp<SH>(base + i)] = v.out(scale); };
        const int a = j * PF_SEG, b = min(a + PF_SEG, n);
        A acc;
        int i;
        if (b + PF_WARM >= n) {
            acc = (A(ml(n - 1)) + A(ml(n - 2)) * z) * zend;
            if (b == n) dset(n - 1, acc);
            i = n - 2;
        } else {
            i = b + PF_WARM - 1;
        }
#pragma unroll 4
        for (; i >= b; i--) acc = (acc - A(ml(i))) * z;
#pragma unroll 4
        for (i = min(b - 1, n - 2); i >= a; i--) {
            acc = (acc - A(ml(i))) * z;
            dset(i, acc);
        }
    }
    __syncthreads();
}

// prefilter_lines for one complex line of compile-time length N held in `buf`, in place, with the
// float accumulator (Sgl2). Each pass first runs every segment's warm-up (which reads the samples of
// neighbouring segments) and keeps only its running value; after a barrier each thread walks its own
// segment, reading each sample just before it overwrites it, so no second buffer is needed. The
// arithmetic is that of prefilter_lines<float2, SH, Sgl2>, so the coefficients are the same bits.
// N > 1 samples; the N / 17 segments need as many threads.
template <int SH>
__device__ void prefilter_line_inplace(float2 *buf, int N, double scale, int tid, int nt) {
    const double z = -0.26794919243112270647;  // sqrt(3) - 2
    const int NSEG = (N + PF_SEG - 1) / PF_SEG;
    using A = Sgl2;
    auto at = [&](int i) { return buf[lp<SH>(i)]; };
    const int j = tid;
    const bool mine = j < NSEG;
    const int a = j * PF_SEG, b = min(a + PF_SEG, N);
    // causal pass: mid[i] = 6 src[i] + z mid[i-1], mid[0] = sum_{k<h} z^k 6 src[k]
    A acc;
    if (mine) {
        if (a <= PF_WARM) {
            const int H = N < PF_HORIZON ? N : PF_HORIZON;
            double zk = 1.0;
            for (int k = 0; k < H; k++) {
                acc = acc + A(at(k)) * (6.0 * zk);
                zk *= z;
            }
            for (int i = 1; i < a; i++) acc = A(at(i)) * 6.0 + acc * z;
        } else {
#pragma unroll 4
            for (int i = a - PF_WARM; i < a; i++) acc = A(at(i)) * 6.0 + acc * z;
        }
    }
    __syncthreads();
    if (mine) {
        if (a == 0) buf[lp<SH>(0)] = acc.out(1.0);
        for (int i = max(a, 1); i < b; i++) {
            acc = A(at(i)) * 6.0 + acc * z;
            buf[lp<SH>(i)] = acc.out(1.0);
        }
    }
    __syncthreads();
    // anti-causal pass: dst[n-1] = z / (z^2 - 1) (mid[n-1] + z mid[n-2]), dst[i] = z (dst[i+1] - mid[i])
    const double zend = z / (z * z - 1.0);
    A bcc;
    if (mine) {
        int i;
        if (b + PF_WARM >= N) {
            bcc = (A(at(N - 1)) + A(at(N - 2)) * z) * zend;
            i = N - 2;
        } else {
            i = b + PF_WARM - 1;
        }
#pragma unroll 4
        for (; i >= b; i--) bcc = (bcc - A(at(i))) * z;
    }
    __syncthreads();
    if (mine) {
        if (b == N) buf[lp<SH>(N - 1)] = bcc.out(scale);
        for (int i = min(b - 1, N - 2); i >= a; i--) {
            bcc = (bcc - A(at(i))) * z;
            buf[lp<SH>(i)] = bcc.out(scale);
        }
    }
    __syncthreads();
    (void)nt;
}

}  // namespace ofx
