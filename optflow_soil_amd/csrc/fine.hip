// Fine stage kernels: spline prefilter of real planes, Gaussian smoothing of the displacement,
// warp of the second image, reference pack, phase accumulation, window moments with the per-pixel
// least-squares solve, energy threshold, exact median, confidence.
//
// Reference: python/optflow_ref/phaseflow.py:100-511 (the specification) and its Rust port
// rust/crates/optflow-core/src/phaseflow.rs (storage f32, arithmetic f64).
#include <cmath>

#include "fine.hpp"
#include "lines.hpp"
#include "sample.hpp"

namespace ofx {

namespace {

constexpr double ZP = -0.26794919243112270647;  // sqrt(3) - 2

// ------------------------------------------------------------------ B-spline prefilter, real planes
// Along y: one thread per column walks the column (coalesced across the wavefront), sequential
// recursion in double (interp.py:12-26).
__global__ void k_prefilter_cols(float *c, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w || h == 1) return;
    const int hz = min(h, PF_HORIZON);
    double acc = 0, zk = 1;
    for (int k = 0; k < hz; k++) {
        acc += zk * 6.0 * (double)c[(size_t)k * w + x];
        zk *= ZP;
    }
    double prev = acc;
    c[x] = (float)acc;
    for (int y = 1; y < h; y++) {
        prev = 6.0 * (double)c[(size_t)y * w + x] + ZP * prev;
        c[(size_t)y * w + x] = (float)prev;
    }
    const double cm2 = (double)c[(size_t)(h - 2) * w + x];
    double d = (ZP / (ZP * ZP - 1.0)) * (prev + ZP * cm2);
    c[(size_t)(h - 1) * w + x] = (float)d;
    for (int y = h - 2; y >= 0; y--) {
        d = ZP * (d - (double)c[(size_t)y * w + x]);
        c[(size_t)y * w + x] = (float)d;
    }
}

// Along x: one workgroup per row, segmented recursion in LDS.
__global__ void __launch_bounds__(256) k_prefilter_rows(float *c, int w) {
    extern __shared__ float smem[];
    float *a = smem, *b = smem + w;
    float *row = c + (size_t)blockIdx.x * w;
    for (int x = threadIdx.x; x < w; x += blockDim.x) a[x] = row[x];
    __syncthreads();
    prefilter_lines<float>(a, b, a, w, 1, 1.0, threadIdx.x, blockDim.x);
    for (int x = threadIdx.x; x < w; x += blockDim.x) row[x] = a[x];
}

// ------------------------------------------------------------------ Gaussian smoothing
// Separable correlation with replicated edges (interp.py:80-91): y pass double -> double, x pass
// double -> float (the pre-warp is stored as float, phaseflow.rs:1538-1555).
__global__ void k_blur_y(const double *__restrict__ in, double *__restrict__ out, int w, int h,
                         const double *__restrict__ k, int r) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    double s = 0;
    for (int j = -r; j <= r; j++) {
        const int yy = min(max(y + j, 0), h - 1);
        s += k[j + r] * in[(size_t)yy * w + x];
    }
    out[(size_t)y * w + x] = s;
}

__global__ void k_blur_x(const double *__restrict__ in, float *__restrict__ out, int w, const double *__restrict__ k,
                         int r) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const double *row = in + (size_t)y * w;
    double s = 0;
    for (int j = -r; j <= r; j++) s += k[j + r] * row[min(max(x + j, 0), w - 1)];
    out[(size_t)y * w + x] = (float)s;
}

// ------------------------------------------------------------------ warp of image 2
// warped(X) = I2(X + s(X)) through the spline of I2 (phaseflow.py:455, interp.py:72-77).
__global__ void k_warp(const float *__restrict__ c2, const float *__restrict__ sx, const float *__restrict__ sy,
                       float *__restrict__ out, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const Taps tx = taps_at((double)x + (double)sx[i], w), ty = taps_at((double)y + (double)sy[i], h);
    out[i] = sample_real(c2, w, tx, ty);
}

// ------------------------------------------------------------------ reference pack
// Per pixel and filter: r, and the phase gradient Im(conj(r) r_x) / (|r|^2 + 1e-12) (likewise y)
// (phaseflow.py:100-111; gabor.rs:230-269). Stored as float4 (re, im, gx, gy); the amplitude
// sqrt(|r|^2 + 1e-12) is recomputed where it is used.
__global__ void k_refpack(const float2 *__restrict__ r, const float2 *__restrict__ rx, const float2 *__restrict__ ry,
                          float4 *__restrict__ out, int w, int h, int wp) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t j = (size_t)(y + PAD) * wp + (x + PAD);
    const float2 a = r[j], bx = rx[j], by = ry[j];
    const double re = a.x, im = a.y;
    const double a2 = re * re + im * im + 1e-12;
    const double gx = (re * (double)bx.y - im * (double)bx.x) / a2;
    const double gy = (re * (double)by.y - im * (double)by.x) / a2;
    out[(size_t)y * w + x] = make_float4(a.x, a.y, (float)gx, (float)gy);
}

// ------------------------------------------------------------------ accumulation
// For every pixel: the position X' in the pre-warped image with X' + s(X') = X + u (two fixed-point
// steps, phaseflow.py:469-472), the bicubic response of every filter there, the phase difference
// against the reference and the seven sums of the local solve (phaseflow.py:473-483):
// gxx, gxy, gyy, gxc, gyc, gcc, w with c = dphi + g . u. Samples and the phase are float; c and
// the sums are double because c carries g . u (tens of rad) whose square enters the residual.
__global__ void __launch_bounds__(256) k_accumulate(const double *__restrict__ ux, const double *__restrict__ uy,
                                                    const float *__restrict__ sx, const float *__restrict__ sy,
                                                    const float *__restrict__ csx, const float *__restrict__ csy,
                                                    const float2 *__restrict__ resp, const float4 *__restrict__ refp,
                                                    double *__restrict__ sums, int w, int h, int wp, int hp, int F) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h, np = (size_t)wp * hp;
    const size_t i = (size_t)y * w + x;
    const double u = ux[i], v = uy[i];
    const double tx = x + u, ty = y + v;
    double px = tx - (double)sx[i], py = ty - (double)sy[i];
    for (int it = 0; it < 2; it++) {
        const Taps ax = taps_at(px, w), ay = taps_at(py, h);
        const double nx = tx - (double)sample_real(csx, w, ax, ay);
        const double ny = ty - (double)sample_real(csy, w, ax, ay);
        px = nx;
        py = ny;
    }
    const Taps qx = taps_at(px + PAD, wp), qy = taps_at(py + PAD, hp);
    double gxx = 0, gxy = 0, gyy = 0, gxc = 0, gyc = 0, gcc = 0, gw = 0;
    for (int f = 0; f < F; f++) {
        const float2 r2 = sample_cplx(resp + (size_t)f * np, wp, qx, qy);
        const float4 p = refp[(size_t)f * n + i];
        // r1 * conj(r2)
        const float re = p.x * r2.x + p.y * r2.y;
        const float im = p.y * r2.x - p.x * r2.y;
        const double dphi = (double)atan2f(im, re);
        const double a1 = sqrt((double)p.x * p.x + (double)p.y * p.y + 1e-12);
        const double wk = a1 * (double)hypotf(r2.x, r2.y);
        const double gx = p.z, gy = p.w;
        const double c = dphi + gx * u + gy * v;
        const double wgx = wk * gx, wgy = wk * gy;
        gxx += wgx * gx;
        gxy += wgx * gy;
        gyy += wgy * gy;
        gxc += wgx * c;
        gyc += wgy * c;
        gcc += wk * c * c;
        gw += wk;
    }
    sums[0 * n + i] = gxx;
    sums[1 * n + i] = gxy;
    sums[2 * n + i] = gyy;
    sums[3 * n + i] = gxc;
    sums[4 * n + i] = gyc;
    sums[5 * n + i] = gcc;
    sums[6 * n + i] = gw;
}

// ------------------------------------------------------------------ window moments (order 1)
// M(key, px, py) = sum_q W(q - p) (qx - px)^px (qy - py)^py a(q), Gaussian W, replicated edges
// (phaseflow.py:114-120), as a y pass (weights k t^py) then an x pass (k t^px).
// y pass outputs, 15 planes: gxx t^0,1,2 | gxy t^0,1,2 | gyy t^0,1,2 | gxc t^0,1 | gyc t^0,1 | gcc | w
__global__ void k_moments_y(const double *__restrict__ sums, double *__restrict__ out, int w, int h,
                            const double *__restrict__ kt, int r) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    const int klen = 2 * r + 1;
    double acc[15];
#pragma unroll
    for (int q = 0; q < 15; q++) acc[q] = 0;
    for (int j = -r; j <= r; j++) {
        const int yy = min(max(y + j, 0), h - 1);
        const size_t i = (size_t)yy * w + x;
        const double k0 = kt[j + r], k1 = kt[klen + j + r], k2 = kt[2 * klen + j + r];
        double s[7];
#pragma unroll
        for (int q = 0; q < 7; q++) s[q] = sums[q * n + i];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            acc[3 * q + 0] += k0 * s[q];
            acc[3 * q + 1] += k1 * s[q];
            acc[3 * q + 2] += k2 * s[q];
        }
        acc[9] += k0 * s[3];
        acc[10] += k1 * s[3];
        acc[11] += k0 * s[4];
        acc[12] += k1 * s[4];
        acc[13] += k0 * s[5];
        acc[14] += k0 * s[6];
    }
    const size_t i = (size_t)y * w + x;
#pragma unroll
    for (int q = 0; q < 15; q++) out[q * n + i] = acc[q];
}

// x pass of the moments, then the 6 x 6 system of the affine shape function per pixel
// (phaseflow.py:132-161; Cholesky as rust phaseflow.rs:394-455): unknowns
// (ux, uy, dux/dx, duy/dx, dux/dy, duy/dy).
__global__ void __launch_bounds__(256) k_moments_x_solve(const double *__restrict__ ym, const double *__restrict__ ux_in,
                                                         const double *__restrict__ uy_in, double *__restrict__ ux_out,
                                                         double *__restrict__ uy_out, double *__restrict__ energy,
                                                         double *__restrict__ resid, double *__restrict__ slope, int w,
                                                         int h, const double *__restrict__ kt, int r) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    const int klen = 2 * r + 1;
    // mp[prod][a][b]: moment of product prod (xx, xy, yy) with x power a, y power b (a + b <= 2)
    double mp[3][3][3];
    double mr[2][2][2];  // rhs (gxc, gyc), a + b <= 1
    double mcc = 0, mw = 0;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) mp[p][a][b] = 0;
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) mr[p][a][b] = 0;
    const double *row = ym + (size_t)y * w;
    for (int j = -r; j <= r; j++) {
        const int xx = min(max(x + j, 0), w - 1);
        const double k0 = kt[j + r], k1 = kt[klen + j + r], k2 = kt[2 * klen + j + r];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const double v0 = row[(3 * p + 0) * n + xx];  // y power 0
            const double v1 = row[(3 * p + 1) * n + xx];  // y power 1
            const double v2 = row[(3 * p + 2) * n + xx];  // y power 2
            mp[p][0][0] += k0 * v0;
            mp[p][1][0] += k1 * v0;
            mp[p][2][0] += k2 * v0;
            mp[p][0][1] += k0 * v1;
            mp[p][1][1] += k1 * v1;
            mp[p][0][2] += k0 * v2;
        }
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const double v0 = row[(9 + 2 * p) * n + xx], v1 = row[(10 + 2 * p) * n + xx];
            mr[p][0][0] += k0 * v0;
            mr[p][1][0] += k1 * v0;
            mr[p][0][1] += k0 * v1;
        }
        mcc += k0 * row[13 * n + xx];
        mw += k0 * row[14 * n + xx];
    }
    // unknown i -> (component, x power, y power)
    const int uc[6] = {0, 1, 0, 1, 0, 1}, ua[6] = {0, 0, 1, 1, 0, 0}, ub[6] = {0, 0, 0, 0, 1, 1};
    double M[6][6], bv[6], b0[6];
#pragma unroll
    for (int a = 0; a < 6; a++) {
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const int prod = (uc[a] == 0 && uc[b] == 0) ? 0 : ((uc[a] == 1 && uc[b] == 1) ? 2 : 1);
            M[a][b] = mp[prod][ua[a] + ua[b]][ub[a] + ub[b]];
        }
        bv[a] = mr[uc[a]][ua[a]][ub[a]];
    }
    const size_t i = (size_t)y * w + x;
    const double e = M[0][0] + M[1][1];
    const double det2 = M[0][0] * M[1][1] - M[0][1] * M[0][1];
    energy[i] = e;
    bool ok = det2 > APERTURE_TOL * e * e;
    if (!ok) {
#pragma unroll
        for (int a = 0; a < 6; a++) {
#pragma unroll
            for (int b = 0; b < 6; b++) M[a][b] = (a == b) ? 1.0 : 0.0;
            bv[a] = 0;
        }
    }
    double tr = 0;
#pragma unroll
    for (int a = 0; a < 6; a++) tr += M[a][a];
#pragma unroll
    for (int a = 0; a < 6; a++) {
        M[a][a] += RIDGE * tr / 6.0;
        b0[a] = bv[a];
    }
    // Cholesky (lower triangle in place), then L y = b, L^T x = y
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = M[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) d -= M[j][k] * M[j][k];
        ok = ok && isfinite(d) && d > 0;
        const double l = sqrt(d);
        M[j][j] = l;
#pragma unroll
        for (int a = j + 1; a < 6; a++) {
            double s = M[a][j];
#pragma unroll
            for (int k = 0; k < j; k++) s -= M[a][k] * M[j][k];
            M[a][j] = s / l;
        }
    }
#pragma unroll
    for (int a = 0; a < 6; a++) {
        double s = bv[a];
#pragma unroll
        for (int k = 0; k < a; k++) s -= M[a][k] * bv[k];
        bv[a] = s / M[a][a];
    }
#pragma unroll
    for (int a = 5; a >= 0; a--) {
        double s = bv[a];
#pragma unroll
        for (int k = a + 1; k < 6; k++) s -= M[k][a] * bv[k];
        bv[a] = s / M[a][a];
    }
#pragma unroll
    for (int a = 0; a < 6; a++) ok = ok && isfinite(bv[a]);
    double *sl = slope + 4 * i;
    if (ok) {
        double tb = 0;
#pragma unroll
        for (int a = 0; a < 6; a++) tb += bv[a] * b0[a];
        const double rr = mcc - tb;
        resid[i] = sqrt((rr > 0 ? rr : (isnan(rr) ? rr : 0.0)) / (mw + 1e-30));
        ux_out[i] = bv[0];
        uy_out[i] = bv[1];
        sl[0] = bv[2];
        sl[1] = bv[3];
        sl[2] = bv[4];
        sl[3] = bv[5];
    } else {
        resid[i] = M_PI;
        ux_out[i] = ux_in[i];
        uy_out[i] = uy_in[i];
        sl[0] = sl[1] = sl[2] = sl[3] = 0;
    }
}

// Energy threshold (phaseflow.py:149): pixels with energy <= 1e-9 median + 1e-30 (or NaN) keep the
// input displacement and get resid = pi.
__global__ void k_energy_reset(const double *__restrict__ energy, const double *__restrict__ med,
                               const double *__restrict__ ux_in, const double *__restrict__ uy_in, double *ux,
                               double *uy, double *resid, double *slope, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double thr = 1e-9 * med[0] + 1e-30;
    if (!(energy[i] > thr)) {
        ux[i] = ux_in[i];
        uy[i] = uy_in[i];
        resid[i] = M_PI;
        slope[4 * i] = slope[4 * i + 1] = slope[4 * i + 2] = slope[4 * i + 3] = 0;
    }
}

// ------------------------------------------------------------------ exact median (np.median)
// Radix select over the order-preserving 64-bit keys of the doubles, 8 bits per pass, histograms
// privatised in LDS. State: prefix (known high bits), rank still to find, NaN count.
struct SelectState {
    unsigned long long prefix;
    unsigned long long rank;
    unsigned int nan_count;
    unsigned int pad;
    unsigned long long below_count;  // elements with key < the selected key
    unsigned long long below_max;    // largest key below it
};

__device__ __forceinline__ unsigned long long order_key(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ double from_key(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    return __longlong_as_double((long long)b);
}

__global__ void k_select_init(SelectState *st, unsigned int *hist, unsigned long long rank) {
    st->prefix = 0;
    st->rank = rank;
    st->nan_count = 0;
    st->below_count = 0;
    st->below_max = 0;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
}

__global__ void __launch_bounds__(256) k_select_hist(const double *__restrict__ v, size_t n, int shift,
                                                     SelectState *st, unsigned int *hist) {
    __shared__ unsigned int lh[256];
    lh[threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long prefix = st->prefix;
    const int hi = shift + 8;  // bits above the current digit must equal the prefix
    unsigned int nan_local = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double x = v[i];
        if (shift == 56 && isnan(x)) nan_local++;
        const unsigned long long k = order_key(x);
        if (hi >= 64 || (k >> hi) == (prefix >> hi)) atomicAdd(&lh[(k >> shift) & 255], 1u);
    }
    if (nan_local) atomicAdd(&st->nan_count, nan_local);
    __syncthreads();
    if (lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

__global__ void k_select_pick(SelectState *st, unsigned int *hist, int shift) {
    if (threadIdx.x != 0) return;
    unsigned long long rank = st->rank, cum = 0;
    int d = 0;
    for (; d < 256; d++) {
        if (cum + hist[d] > rank) break;
        cum += hist[d];
    }
    if (d == 256) d = 255;
    st->prefix |= (unsigned long long)d << shift;
    st->rank = rank - cum;
    for (int i = 0; i < 256; i++) hist[i] = 0;
}

__global__ void __launch_bounds__(256) k_select_below(const double *__restrict__ v, size_t n, SelectState *st) {
    const unsigned long long key = st->prefix;
    unsigned long long cnt = 0, mx = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const unsigned long long k = order_key(v[i]);
        if (k < key) {
            cnt++;
            mx = k > mx ? k : mx;
        }
    }
    // wavefront reduction, then one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_down(cnt, off);
        unsigned long long o = __shfl_down(mx, off);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0 && cnt) {
        atomicAdd(&st->below_count, cnt);
        atomicMax(&st->below_max, mx);
    }
}

__global__ void k_select_finish(const SelectState *st, unsigned long long k, int even, double *out) {
    if (threadIdx.x != 0) return;
    if (st->nan_count) {
        out[0] = NAN;
        return;
    }
    const double upper = from_key(st->prefix);
    if (!even) {
        out[0] = upper;
        return;
    }
    const double lower = (st->below_count >= k) ? from_key(st->below_max) : upper;
    out[0] = (lower + upper) / 2.0;
}

// ------------------------------------------------------------------ confidence
// conf = min(1, E / (0.25 median(E) + 1e-30)) exp(-(resid / phase_tol)^2) (phaseflow.py:320-323).
// low = conf < conf_min (before the border rule, for the pre-warp). With final != 0 also the
// border rule (phaseflow.py:297-317) and the clip to [0, 1] into conf_out.
__global__ void k_confidence(const double *__restrict__ energy, const double *__restrict__ resid,
                             const double *__restrict__ med, const double *__restrict__ ux,
                             const double *__restrict__ uy, double phase_tol, double conf_min, double B, int w, int h,
                             uint8_t *__restrict__ low, float *__restrict__ conf_out, double *__restrict__ conf64) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const double scale = 0.25 * med[0] + 1e-30;
    const double t = fmin(1.0, energy[i] / scale);
    const double q = resid[i] / phase_tol;
    const double c = t * exp(-q * q);
    if (low) low[i] = c < conf_min;
    if (conf_out) {
        const double X = x + ux[i], Y = y + uy[i];
        const double b1 = fmin(fmin((double)x, (double)y), fmin((double)(w - 1 - x), (double)(h - 1 - y)));
        const double b2 = fmin(fmin(X, Y), fmin((double)(w - 1) - X, (double)(h - 1) - Y));
        const bool rep = b1 >= B && b2 >= B;  // NaN compares false: outside
        const double cl = isnan(c) ? c : fmin(fmax(c, 0.0), 1.0);
        const double v = rep ? cl : 0.0;
        conf_out[i] = (float)v;
        if (conf64) conf64[i] = v;
    }
}

}  // namespace

// ------------------------------------------------------------------ host wrappers
static dim3 grid2(int w, int h) { return dim3((w + 255) / 256, h); }

void prefilter_real(float *c, int w, int h, hipStream_t s) {
    if (h > 1) hipLaunchKernelGGL(k_prefilter_cols, dim3((w + 255) / 256), dim3(256), 0, s, c, w, h);
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_prefilter_rows, dim3(h), dim3(256), 2 * w * sizeof(float), s, c, w);
    LAUNCHCHECK();
}

void blur(const double *in, double *tmp, float *out, int w, int h, const double *taps, int r, hipStream_t s) {
    hipLaunchKernelGGL(k_blur_y, grid2(w, h), dim3(256), 0, s, in, tmp, w, h, taps, r);
    hipLaunchKernelGGL(k_blur_x, grid2(w, h), dim3(256), 0, s, tmp, out, w, taps, r);
    LAUNCHCHECK();
}

void warp(const float *c2, const float *sx, const float *sy, float *out, int w, int h, hipStream_t s) {
    hipLaunchKernelGGL(k_warp, grid2(w, h), dim3(256), 0, s, c2, sx, sy, out, w, h);
    LAUNCHCHECK();
}

void refpack(const float2 *r, const float2 *rx, const float2 *ry, float4 *out, int w, int h, int wp, hipStream_t s) {
    hipLaunchKernelGGL(k_refpack, grid2(w, h), dim3(256), 0, s, r, rx, ry, out, w, h, wp);
    LAUNCHCHECK();
}

void accumulate(const double *ux, const double *uy, const float *sx, const float *sy, const float *csx,
                const float *csy, const float2 *resp, const float4 *refp, double *sums, int w, int h, int F,
                hipStream_t s) {
    const int wp = w + 2 * PAD, hp = h + 2 * PAD;
    hipLaunchKernelGGL(k_accumulate, grid2(w, h), dim3(256), 0, s, ux, uy, sx, sy, csx, csy, resp, refp, sums, w, h,
                       wp, hp, F);
    LAUNCHCHECK();
}

void moments_solve(const double *sums, double *ym, const double *ux_in, const double *uy_in, double *ux_out,
                   double *uy_out, double *energy, double *resid, double *slope, int w, int h, const double *kt,
                   int r, hipStream_t s) {
    hipLaunchKernelGGL(k_moments_y, grid2(w, h), dim3(256), 0, s, sums, ym, w, h, kt, r);
    hipLaunchKernelGGL(k_moments_x_solve, grid2(w, h), dim3(256), 0, s, ym, ux_in, uy_in, ux_out, uy_out, energy,
                       resid, slope, w, h, kt, r);
    LAUNCHCHECK();
}

void energy_reset(const double *energy, const double *med, const double *ux_in, const double *uy_in, double *ux,
                  double *uy, double *resid, double *slope, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_energy_reset, dim3((n + 255) / 256), dim3(256), 0, s, energy, med, ux_in, uy_in, ux, uy,
                       resid, slope, n);
    LAUNCHCHECK();
}

size_t median_scratch_bytes() { return sizeof(SelectState) + 256 * sizeof(unsigned int); }

void median(const double *v, size_t n, void *scratch, double *out, hipStream_t s) {
    SelectState *st = (SelectState *)scratch;
    unsigned int *hist = (unsigned int *)((char *)scratch + sizeof(SelectState));
    const unsigned long long k = n / 2;
    hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, s, st, hist, k);
    const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256);
    for (int shift = 56; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(k_select_hist, dim3(blocks), dim3(256), 0, s, v, n, shift, st, hist);
        hipLaunchKernelGGL(k_select_pick, dim3(1), dim3(64), 0, s, st, hist, shift);
    }
    const int even = (n % 2 == 0) ? 1 : 0;
    if (even) hipLaunchKernelGGL(k_select_below, dim3(blocks), dim3(256), 0, s, v, n, st);
    hipLaunchKernelGGL(k_select_finish, dim3(1), dim3(64), 0, s, st, k, even, out);
    LAUNCHCHECK();
}

void confidence(const double *energy, const double *resid, const double *med, const double *ux, const double *uy,
                double phase_tol, double conf_min, double B, int w, int h, uint8_t *low, float *conf_out,
                double *conf64, hipStream_t s) {
    hipLaunchKernelGGL(k_confidence, grid2(w, h), dim3(256), 0, s, energy, resid, med, ux, uy, phase_tol, conf_min, B,
                       w, h, low, conf_out, conf64);
    LAUNCHCHECK();
}

}  // namespace ofx
