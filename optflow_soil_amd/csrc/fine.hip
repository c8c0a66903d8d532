// Fine stage kernels: spline prefilter of real planes, Gaussian smoothing of the displacement,
// warp of the second image, reference pack, phase accumulation, window moments with the per-pixel
// least-squares solve, energy threshold, exact median, confidence.
//
// Reference: python/optflow_ref/phaseflow.py:100-511 (the specification) and its Rust port
// rust/crates/optflow-core/src/phaseflow.rs (storage f32, arithmetic f64).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "fine.hpp"
#include "lines.hpp"
#include "sample.hpp"

namespace ofx {

namespace {

constexpr double ZP = -0.26794919243112270647;  // sqrt(3) - 2

// ------------------------------------------------------------------ B-spline prefilter, real planes
// Along y: one thread per (column, segment of PF_SEG rows), coalesced across the wavefront; each
// segment starts its recursion PF_WARM rows early (error z^PF_WARM, as the LDS form) in double.
// Causal pass src -> mid, anti-causal pass mid -> dst (dst may be src).
// One (column, segment) of the causal pass, src -> mid. An interior segment (a full PF_SEG rows,
// its warm-up inside the frame) issues all PF_WARM + PF_SEG loads before the recursion uses them;
// the loads of a walk in a loop would each wait for the one before.
__device__ __forceinline__ void causal_segment(const float *__restrict__ src, float *__restrict__ mid, int w, int h,
                                               int x, int j) {
    const int a = j * PF_SEG, b = min(a + PF_SEG, h);
    if (a > PF_WARM && b - a == PF_SEG) {
        constexpr int NV = PF_WARM + PF_SEG;
        float v[NV];
        const float *p = src + (size_t)(a - PF_WARM) * w + x;
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = p[(size_t)k * w];
        double acc = 0;
#pragma unroll
        for (int k = 0; k < PF_WARM; k++) acc = 6.0 * (double)v[k] + ZP * acc;
        float *q = mid + (size_t)a * w + x;
#pragma unroll
        for (int k = 0; k < PF_SEG; k++) {
            acc = 6.0 * (double)v[PF_WARM + k] + ZP * acc;
            q[(size_t)k * w] = (float)acc;
        }
        return;
    }
    double acc = 0;
    int i;
    if (a <= PF_WARM) {
        const int hz = min(h, PF_HORIZON);
        double zk = 1;
        for (int k = 0; k < hz; k++) {
            acc += zk * 6.0 * (double)src[(size_t)k * w + x];
            zk *= ZP;
        }
        if (a == 0) mid[x] = (float)acc;
        i = 1;
    } else {
        i = a - PF_WARM;
    }
    for (; i < a; i++) acc = 6.0 * (double)src[(size_t)i * w + x] + ZP * acc;
    for (i = max(a, 1); i < b; i++) {
        acc = 6.0 * (double)src[(size_t)i * w + x] + ZP * acc;
        mid[(size_t)i * w + x] = (float)acc;
    }
}

// One (column, segment) of the anti-causal pass, mid -> dst, interior segments with their loads
// issued first as above.
__device__ __forceinline__ void anti_segment(const float *__restrict__ mid, float *__restrict__ dst, int w, int h,
                                             int x, int j) {
    const int a = j * PF_SEG, b = min(a + PF_SEG, h);
    if (b + PF_WARM < h && b - a == PF_SEG) {
        constexpr int NV = PF_SEG + PF_WARM;
        float v[NV];  // v[k] is row a + k
        const float *p = mid + (size_t)a * w + x;
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = p[(size_t)k * w];
        double acc = 0;
#pragma unroll
        for (int k = NV - 1; k >= PF_SEG; k--) acc = ZP * (acc - (double)v[k]);
        float *q = dst + (size_t)a * w + x;
#pragma unroll
        for (int k = PF_SEG - 1; k >= 0; k--) {
            acc = ZP * (acc - (double)v[k]);
            q[(size_t)k * w] = (float)acc;
        }
        return;
    }
    double acc = 0;
    int i;
    if (b + PF_WARM >= h) {
        acc = (ZP / (ZP * ZP - 1.0)) * ((double)mid[(size_t)(h - 1) * w + x] + ZP * (double)mid[(size_t)(h - 2) * w + x]);
        if (b == h) dst[(size_t)(h - 1) * w + x] = (float)acc;
        i = h - 2;
    } else {
        i = b + PF_WARM - 1;
    }
    for (; i >= b; i--) acc = ZP * (acc - (double)mid[(size_t)i * w + x]);
    for (i = min(b - 1, h - 2); i >= a; i--) {
        acc = ZP * (acc - (double)mid[(size_t)i * w + x]);
        dst[(size_t)i * w + x] = (float)acc;
    }
}

__global__ void k_prefilter_cols_causal(const float *__restrict__ src, float *__restrict__ mid, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < w) causal_segment(src, mid, w, h, x, blockIdx.y);
}

__global__ void k_prefilter_cols_anti(const float *__restrict__ mid, float *__restrict__ dst, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < w) anti_segment(mid, dst, w, h, x, blockIdx.y);
}

// Two planes per launch (blockIdx.z): the spline of each pre-warp component from the component
// itself, src -> mid -> dst, so the copy into dst comes for free.
struct PlanePair {
    const float *src[2];
    float *mid[2], *dst[2];
};
__global__ void k_prefilter_cols_causal2(PlanePair P, int w, int h) {
    const int z = blockIdx.z;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < w) causal_segment(z ? P.src[1] : P.src[0], z ? P.mid[1] : P.mid[0], w, h, x, blockIdx.y);
}

__global__ void k_prefilter_cols_anti2(PlanePair P, int w, int h) {
    const int z = blockIdx.z;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < w) anti_segment(z ? P.mid[1] : P.mid[0], z ? P.dst[1] : P.dst[0], w, h, x, blockIdx.y);
}

// Along x: one workgroup per row y, the rows y of both planes side by side in LDS as two lines of
// one segmented recursion (twice the segments of one row for the 256 threads to walk).
__global__ void __launch_bounds__(256) k_prefilter_rows2(PlanePair P, int w) {
    extern __shared__ float smem[];
    constexpr int SH = 4;
    float *a = smem, *b = smem + lp_len<SH>(2 * w);
    float *row0 = P.dst[0] + (size_t)blockIdx.x * w, *row1 = P.dst[1] + (size_t)blockIdx.x * w;
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
        a[lp<SH>(x)] = row0[x];
        a[lp<SH>(w + x)] = row1[x];
    }
    __syncthreads();
    prefilter_lines<float, SH>(a, b, a, w, 2, 1.0, threadIdx.x, blockDim.x);
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
        row0[x] = a[lp<SH>(x)];
        row1[x] = a[lp<SH>(w + x)];
    }
}

// Along x: one workgroup per row, segmented recursion in LDS.
__global__ void __launch_bounds__(256) k_prefilter_rows(float *c, int w) {
    extern __shared__ float smem[];
    constexpr int SH = 4;
    float *a = smem, *b = smem + lp_len<SH>(w);
    float *row = c + (size_t)blockIdx.x * w;
    for (int x = threadIdx.x; x < w; x += blockDim.x) a[lp<SH>(x)] = row[x];
    __syncthreads();
    prefilter_lines<float, SH>(a, b, a, w, 1, 1.0, threadIdx.x, blockDim.x);
    for (int x = threadIdx.x; x < w; x += blockDim.x) row[x] = a[lp<SH>(x)];
}

// ------------------------------------------------------------------ Gaussian smoothing
// Separable correlation with replicated edges (interp.py:80-91): y pass double -> double, x pass
// double -> float (the pre-warp is stored as float, phaseflow.rs:1538-1555).
struct BT {
    double k[2 * 64 + 1];
};
__global__ void k_blur_y(const double *__restrict__ in, double *__restrict__ out, int w, int h, const BT bt, int r) {
    const double *k = bt.k;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    double s = 0;
    for (int j = -r; j <= r; j++) {
        const int yy = min(max(y + j, 0), h - 1);
        s += k[j + r] * in[(size_t)yy * w + x];
    }
    out[(size_t)y * w + x] = s;
}

// Register-blocked y pass for a compile-time radius R (sigma_s = 4 gives R = 16): one thread per
// (column, BLUR_ROWS output rows); each of the BLUR_ROWS + 2R input rows is read once and added
// into every output row it reaches, taps in the order of the sequential sum.
constexpr int BLUR_ROWS = 16;

template <int R>
__global__ void __launch_bounds__(256) k_blur_y_rb(const double *__restrict__ in, double *__restrict__ out, int w,
                                                   int h, const BT bt) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y0 = blockIdx.y * BLUR_ROWS;
    if (x >= w) return;
    double acc[BLUR_ROWS];
#pragma unroll
    for (int q = 0; q < BLUR_ROWS; q++) acc[q] = 0;
#pragma unroll
    for (int j = 0; j < BLUR_ROWS + 2 * R; j++) {
        const int yy = min(max(y0 - R + j, 0), h - 1);
        const double v = in[(size_t)yy * w + x];
#pragma unroll
        for (int q = 0; q < BLUR_ROWS; q++) {
            const int t = j - q;
            if (t >= 0 && t <= 2 * R) acc[q] += bt.k[t] * v;
        }
    }
#pragma unroll
    for (int q = 0; q < BLUR_ROWS; q++)
        if (y0 + q < h) out[(size_t)(y0 + q) * w + x] = acc[q];
}

// x pass from an LDS row segment of 256 + 2r values (replicated edges)
__global__ void __launch_bounds__(256) k_blur_x_lds(const double *__restrict__ in, float *__restrict__ out, int w,
                                                    const BT bt, int r) {
    __shared__ double seg[256 + 2 * 64];
    const int x0 = blockIdx.x * 256, y = blockIdx.y;
    const double *row = in + (size_t)y * w;
    for (int c = threadIdx.x; c < 256 + 2 * r; c += blockDim.x) seg[c] = row[min(max(x0 - r + c, 0), w - 1)];
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= w) return;
    double s = 0;
    for (int j = 0; j <= 2 * r; j++) s += bt.k[j] * seg[threadIdx.x + j];
    out[(size_t)y * w + x] = (float)s;
}

// The x pass for a compile-time radius R, four adjacent outputs per thread (1024 per workgroup):
// each of a thread's 4 + 2R inputs is read from LDS once and added into the outputs it reaches,
// taps in the order of the sequential sum. Column c of the segment sits at (c mod 4) Q + c / 4, so
// the 64 lanes of a wavefront read 64 consecutive words.
template <int R>
__global__ void __launch_bounds__(256) k_blur_x_rb(const double *__restrict__ in, float *__restrict__ out, int w,
                                                   const BT bt) {
    constexpr int P = 1024, NB = 4, Q = (P + 2 * R + 3) / 4;
    __shared__ double seg[4 * Q];
    const int x0 = blockIdx.x * P, y = blockIdx.y;
    const double *row = in + (size_t)y * w;
    for (int c = threadIdx.x; c < P + 2 * R; c += 256)
        seg[(c & 3) * Q + (c >> 2)] = row[min(max(x0 - R + c, 0), w - 1)];
    __syncthreads();
    double acc[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) acc[q] = 0;
#pragma unroll
    for (int j = 0; j < NB + 2 * R; j++) {
        const double v = seg[(j & 3) * Q + threadIdx.x + (j >> 2)];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int t = j - q;
            if (t >= 0 && t <= 2 * R) acc[q] += bt.k[t] * v;
        }
    }
    float *o = out + (size_t)y * w;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int x = x0 + NB * threadIdx.x + q;
        if (x < w) o[x] = (float)acc[q];
    }
}

// ------------------------------------------------------------------ warp of image 2
// warped(X) = I2((x + shx) + s_x, (y + shy) + s_y) from the spline c2 of image 2 (w2 x h2), the
// window of the second frame displaced by the shift (phaseflow.rs:1416-1439; a zero shift is
// phaseflow.py:455 with interp.py:72-77).
__global__ void k_warp(const float *__restrict__ c2, int w2, int h2, double shx, double shy,
                       const float *__restrict__ sx, const float *__restrict__ sy, float *__restrict__ out, int w,
                       int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const Taps tx = taps_at(((double)x + shx) + (double)sx[i], w2), ty = taps_at(((double)y + shy) + (double)sy[i], h2);
    out[i] = sample_real(c2, w2, tx, ty);
}

// ------------------------------------------------------------------ reference pack
// Per pixel and filter: r, and the phase gradient Im(conj(r) r_x) / (|r|^2 + 1e-12) (likewise y)
// (phaseflow.py:100-111; gabor.rs:230-269). Stored as float4 (re, im, gx, gy); the amplitude
// sqrt(|r|^2 + 1e-12) is recomputed where it is used.
__global__ void k_refpack(const float2 *__restrict__ resp, float4 *__restrict__ out, int w, int h, int wp, int F,
                          size_t np) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, f = blockIdx.z;
    if (x >= w) return;
    const size_t j = (size_t)(y + PAD) * wp + (x + PAD);
    // filter f's planes r, rx, ry: planes f, F + f, 2F + f of the 3F inverses
    const float2 a = resp[f * np + j], bx = resp[(F + f) * np + j], by = resp[(2 * F + f) * np + j];
    const double re = a.x, im = a.y;
    const double a2 = re * re + im * im + 1e-12;
    const double gx = (re * (double)bx.y - im * (double)bx.x) / a2;
    const double gy = (re * (double)by.y - im * (double)by.x) / a2;
    out[(size_t)f * w * h + (size_t)y * w + x] = make_float4(a.x, a.y, (float)gx, (float)gy);
}

// ------------------------------------------------------------------ accumulation
// For every pixel: the position X' in the pre-warped image with X' + s(X') = X + u (two fixed-point
// steps, phaseflow.py:469-472), the bicubic response of every filter there, the phase difference
// against the reference and the seven sums of the local solve (phaseflow.py:473-483):
// gxx, gxy, gyy, gxc, gyc, gcc, w with c = dphi + g . u. Samples and the phase are float; c and
// the sums are double because c carries g . u (tens of rad) whose square enters the residual.
// Tiled accumulation: a 16 x 16 pixel tile finds the bounding box of its 4 x 4 response taps, and
// if it holds at most ACC_BOX nodes (smooth fields: about (16 + 3 + |spread of u - s|)^2), every
// filter's box is staged in LDS with coalesced loads (double-buffered: the box of filter f + 1 is
// fetched while filter f is evaluated) and the taps are read from LDS. Tiles whose taps spread
// wider, or reach the mirrored border of the padded plane, gather their taps from global memory.
constexpr int ACC_T = 16;
constexpr int ACC_BOX = 1600;  // nodes per staged box (40 x 40)

// The weight of a filter: the reference amplitude sqrt(|r|^2 + 1e-12) (stored as f32,
// gabor.rs:260-269) times the sample's modulus |s| (phaseflow.rs:1268-1277). With re + i im =
// r conj(s), re^2 + im^2 = |r|^2 |s|^2, so the product is one square root,
// sqrt(re^2 + im^2 + 1e-12 |s|^2), in f32 (a few ulp from the product of the two f32 roots).
__device__ __forceinline__ float weight_of(float re, float im, float s2) {
    return sqrtf(fmaf(re, re, fmaf(im, im, 1e-12f * s2)));
}

// One filter of a pixel (phaseflow.rs:1247-1401): the bicubic sample of the response box bb
// (tap offsets o), the phase difference against the reference pack p = (re, im, gx, gy), the
// weight, c = dphi + gx u + gy v and the seven sums s = gxx, gxy, gyy, gxc, gyc, gcc, gw.
__device__ __forceinline__ void acc_filter(const float2 *bb, const int (&o)[16], const Taps &qx, const Taps &qy,
                                           const float4 p, double u, double v, double (&s)[7]) {
    float r2x = 0.f, r2y = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float ax = 0.f, ay = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float2 t = bb[o[4 * j + k]];
            ax = fmaf(qx.w[k], t.x, ax);
            ay = fmaf(qx.w[k], t.y, ay);
        }
        r2x = fmaf(qy.w[j], ax, r2x);
        r2y = fmaf(qy.w[j], ay, r2y);
    }
    const float re = p.x * r2x + p.y * r2y;
    const float im = p.y * r2x - p.x * r2y;
    const double dphi = (double)atan2f(im, re);
    const double wk = (double)weight_of(re, im, fmaf(r2x, r2x, r2y * r2y));
    const double gx = p.z, gy = p.w;
    const double c = dphi + gx * u + gy * v;
    const double wgx = wk * gx, wgy = wk * gy;
    s[0] += wgx * gx;
    s[1] += wgx * gy;
    s[2] += wgy * gy;
    s[3] += wgx * c;
    s[4] += wgy * c;
    s[5] += wk * c * c;
    s[6] += wk;
}

// Stages a box of nodes in 16-byte pairs with the load-to-LDS path of gfx950: pair p = tid + 256 q
// (plane offset poff[q], p < npairs) lands at dst + 2 p, lane l of a wavefront writing 16 bytes at
// the wavefront's base (wbase = its first thread) + 16 l.
template <int PQ>
__device__ __forceinline__ void stage_pairs(const float2 *plane, float2 *dst, const int (&poff)[PQ], int npairs,
                                            int wbase) {
#pragma unroll
    for (int q = 0; q < PQ; q++)
        if ((int)threadIdx.x + 256 * q < npairs)
            __builtin_amdgcn_global_load_lds(
                (const void *)(plane + poff[q]),
                (__attribute__((address_space(3))) void *)(dst + 2 * (256 * q + wbase)), 16, 0, 0);
}

// Workgroups are dealt to the chip's 8 XCDs in turn (dispatch index L runs on XCD L mod 8) and
// each XCD has its own L2. Each row of tiles is split into 8 chunks of c = ceil(gx / 8) adjacent
// tiles, chunk x going to XCD x: neighbouring tiles, whose staged boxes overlap and share cache
// lines, meet in the same L2, while all XCDs still walk the same rows of the frame together.
// The grid has 8 c slots per row; slots past the last column return at once.
__device__ __forceinline__ void xcd_tile(int L, int gx, int &tx, int &ty) {
    const int x = L & 7, s = L >> 3, c = (gx + 7) >> 3;
    ty = s / c;
    tx = x * c + (s - ty * c);
}

// Six workgroups per CU: the two staged boxes (25.6 KB of LDS) allow six, and 80 VGPRs (81 without
// the hint) let six wavefronts share each SIMD.
template <bool DMA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_accumulate_tiled(const double *__restrict__ ux, const double *__restrict__ uy,
                                                          const float *__restrict__ sx, const float *__restrict__ sy,
                                                          const float *__restrict__ csx, const float *__restrict__ csy,
                                                          const float2 *__restrict__ resp,
                                                          const float4 *__restrict__ refp, double *__restrict__ sums,
                                                          int w, int h, int wp, int hp, int nf) {
    __shared__ int red[4][8];
    __shared__ int geom[4];
    const int lx = threadIdx.x & (ACC_T - 1), ly = threadIdx.x >> 4;
    int tx, ty;
    xcd_tile(blockIdx.x, (w + ACC_T - 1) / ACC_T, tx, ty);
    if (tx * ACC_T >= w) return;  // a padding slot: the whole workgroup leaves before any barrier
    const int x = tx * ACC_T + lx, y = ty * ACC_T + ly;
    const bool inside = x < w && y < h;
    const size_t n = (size_t)w * h, np = (size_t)wp * hp;
    const size_t i = inside ? (size_t)y * w + x : 0;
    double u = 0, v = 0, px = 0, py = 0;
    if (inside) {
        u = ux[i];
        v = uy[i];
        const double tx = x + u, ty = y + v;
        px = tx - (double)sx[i];
        py = ty - (double)sy[i];
        for (int it = 0; it < 2; it++) {
            const Taps ax = taps_at(px, w), ay = taps_at(py, h);
            const double nx = tx - (double)sample_real(csx, w, ax, ay);
            const double ny = ty - (double)sample_real(csy, w, ax, ay);
            px = nx;
            py = ny;
        }
    }
    const Taps qx = taps_at(px + PAD, wp), qy = taps_at(py + PAD, hp);
    // tap range of this pixel; a mirrored tap (border of the padded plane) sends the tile to the
    // global path
    int x0 = qx.idx[0], x1 = qx.idx[3], y0 = qy.idx[0], y1 = qy.idx[3];
    bool plain = qx.idx[1] == x0 + 1 && qx.idx[2] == x0 + 2 && x1 == x0 + 3 && qy.idx[1] == y0 + 1 &&
                 qy.idx[2] == y0 + 2 && y1 == y0 + 3;
    if (!inside) {
        x0 = y0 = 0x7fffffff;
        x1 = y1 = -1;
        plain = true;
    }
    // block min / max of the tap ranges (wavefront shuffles, then the 4 wavefronts)
    int mnx = x0, mxx = x1, mny = y0, mxy = y1, pl = plain ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) {
        mnx = min(mnx, __shfl_xor(mnx, off));
        mxx = max(mxx, __shfl_xor(mxx, off));
        mny = min(mny, __shfl_xor(mny, off));
        mxy = max(mxy, __shfl_xor(mxy, off));
        pl = min(pl, __shfl_xor(pl, off));
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = mnx;
        red[wv][1] = mxx;
        red[wv][2] = mny;
        red[wv][3] = mxy;
        red[wv][4] = pl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = red[0][0], b = red[0][1], c = red[0][2], d = red[0][3], e = red[0][4];
        for (int k = 1; k < 4; k++) {
            a = min(a, red[k][0]);
            b = max(b, red[k][1]);
            c = min(c, red[k][2]);
            d = max(d, red[k][3]);
            e = min(e, red[k][4]);
        }
        // DMA: the box starts on an even column and has an even width, so that it is staged in
        // 16-byte pairs of nodes; its extra column must stay inside the padded row
        const int bx = DMA ? (a & ~1) : a, bw = DMA ? ((b - bx + 2) & ~1) : b - a + 1, bh = d - c + 1;
        geom[0] = bx;
        geom[1] = c;
        geom[2] = bw;
        geom[3] = (e && bw > 0 && bh > 0 && (long)bw * bh <= ACC_BOX && (!DMA || bx + bw <= wp)) ? bh : -1;
    }
    __syncthreads();
    const int bx0 = geom[0], by0 = geom[1], bw = geom[2], bh = geom[3];
    // the seven sums gxx, gxy, gyy, gxc, gyc, gcc, gw
    double sm[7] = {0, 0, 0, 0, 0, 0, 0};
    if (bh > 0) {
        const int bn = bw * bh;
        // tap offsets inside the box (pixels outside the frame evaluate nothing)
        const int ox = inside ? x0 - bx0 : 0, oy = inside ? y0 - by0 : 0;
        int o[16];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int k = 0; k < 4; k++) o[4 * j + k] = (oy + j) * bw + ox + k;
        // the reference pack of filter f + 1 is loaded unconditionally (a pixel outside the frame
        // reads pixel 0), so that no register copy waits for it before filter f is evaluated
        float4 pc = refp[i];
        if constexpr (DMA) {
            // Staged by the load-to-LDS path of gfx950: pair p = tid + 256 q of the box (nodes
            // 2 p, 2 p + 1, one row) goes from the plane straight to LDS, lane l of a wavefront
            // writing 16 bytes at the wavefront's base + 16 l, so no register holds the next box.
            // The pair offsets in a plane are the same for every filter; the plane and the pairs
            // are 16-byte aligned (even padded width, even box column and width). The two boxes
            // are separate LDS arrays and the filter loop runs in pairs, so that the compiler sees
            // that the staging of one box does not touch the box being read and does not wait for
            // it there.
            __shared__ float2 box0[ACC_BOX], box1[ACC_BOX];
            constexpr int PQ = (ACC_BOX / 2 + 255) / 256;
            const int hw = bw >> 1, npairs = bn >> 1, wbase = threadIdx.x & ~63;
            int poff[PQ];
#pragma unroll
            for (int q = 0; q < PQ; q++) {
                const int p = threadIdx.x + 256 * q;
                const int r = p / hw, c = p - r * hw;
                poff[q] = p < npairs ? (by0 + r) * wp + bx0 + 2 * c : 0;
            }
            stage_pairs<PQ>(resp, box0, poff, npairs, wbase);
            for (int f = 0; f < nf; f += 2) {
                // filter f from box0 while filter f + 1 is staged into box1, then filter f + 1
                // from box1 while filter f + 2 is staged into box0. Before each: every load of
                // this thread done (the box to read and the pack of its filter), then every
                // thread's (that box complete, the other one no longer read).
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                float4 pn = pc;
                if (f + 1 < nf) {
                    stage_pairs<PQ>(resp + (size_t)(f + 1) * np, box1, poff, npairs, wbase);
                    pn = refp[(size_t)(f + 1) * n + i];
                }
                if (inside) acc_filter(box0, o, qx, qy, pc, u, v, sm);
                pc = pn;
                if (f + 1 < nf) {
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                    if (f + 2 < nf) {
                        stage_pairs<PQ>(resp + (size_t)(f + 2) * np, box0, poff, npairs, wbase);
                        pn = refp[(size_t)(f + 2) * n + i];
                    }
                    if (inside) acc_filter(box1, o, qx, qy, pc, u, v, sm);
                    pc = pn;
                }
            }
        } else {
            __shared__ float2 box[2][ACC_BOX];
            // offsets in a response plane of this thread's box elements e = tid + 256 q, the same
            // for every filter (a plane holds fewer than 2^31 nodes: the sides are at most 7744)
            int boff[ACC_BOX / 256 + 1];
#pragma unroll
            for (int q = 0; q < ACC_BOX / 256 + 1; q++) {
                const int e = threadIdx.x + 256 * q;
                const int r = e / bw, c = e - r * bw;
                boff[q] = e < bn ? (by0 + r) * wp + bx0 + c : 0;
            }
#pragma unroll
            for (int q = 0; q < ACC_BOX / 256 + 1; q++) {
                const int e = threadIdx.x + 256 * q;
                if (e < bn) box[0][e] = resp[boff[q]];
            }
            for (int f = 0; f < nf; f++) {
                __syncthreads();
                float2 nx[ACC_BOX / 256 + 1];
                float4 pn = pc;
                if (f + 1 < nf) {
                    const float2 *plane = resp + (size_t)(f + 1) * np;
#pragma unroll
                    for (int q = 0; q < ACC_BOX / 256 + 1; q++) {
                        const int e = threadIdx.x + 256 * q;
                        if (e < bn) nx[q] = plane[boff[q]];
                    }
                    pn = refp[(size_t)(f + 1) * n + i];
                }
                if (inside) acc_filter(box[f & 1], o, qx, qy, pc, u, v, sm);
                if (f + 1 < nf) {
                    float2 *dst = box[(f + 1) & 1];
#pragma unroll
                    for (int q = 0; q < ACC_BOX / 256 + 1; q++) {
                        const int e = threadIdx.x + 256 * q;
                        if (e < bn) dst[e] = nx[q];
                    }
                }
                pc = pn;
            }
        }
    } else if (inside) {
        for (int f = 0; f < nf; f++) {
            const float2 r2 = sample_cplx(resp + (size_t)f * np, wp, qx, qy);
            const float4 p = refp[(size_t)f * n + i];
            const float re = p.x * r2.x + p.y * r2.y;
            const float im = p.y * r2.x - p.x * r2.y;
            const double dphi = (double)atan2f(im, re);
            const double wk = (double)weight_of(re, im, fmaf(r2.x, r2.x, r2.y * r2.y));
            const double gx = p.z, gy = p.w;
            const double c = dphi + gx * u + gy * v;
            const double wgx = wk * gx, wgy = wk * gy;
            sm[0] += wgx * gx;
            sm[1] += wgx * gy;
            sm[2] += wgy * gy;
            sm[3] += wgx * c;
            sm[4] += wgy * c;
            sm[5] += wk * c * c;
            sm[6] += wk;
        }
    }
    if (!inside) return;
#pragma unroll
    for (int k = 0; k < 7; k++) sums[k * n + i] = sm[k];
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

__device__ __forceinline__ void solve_pixel(const double (&mp)[3][3][3], const double (&mr)[2][2][2], double mcc,
                                            double mw, size_t i, const double *__restrict__ ux_in,
                                            const double *__restrict__ uy_in, double *__restrict__ ux_out,
                                            double *__restrict__ uy_out, double *__restrict__ energy,
                                            double *__restrict__ resid, double *__restrict__ slope);

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
    solve_pixel(mp, mr, mcc, mw, (size_t)y * w + x, ux_in, uy_in, ux_out, uy_out, energy, resid, slope);
}

// The 6 x 6 system of the affine shape function at pixel i from its moments (phaseflow.py:132-161;
// Cholesky as rust phaseflow.rs:394-455): unknowns (ux, uy, dux/dx, duy/dx, dux/dy, duy/dy).
__device__ __forceinline__ void solve_pixel(const double (&mp)[3][3][3], const double (&mr)[2][2][2], double mcc,
                                            double mw, size_t i, const double *__restrict__ ux_in,
                                            const double *__restrict__ uy_in, double *__restrict__ ux_out,
                                            double *__restrict__ uy_out, double *__restrict__ energy,
                                            double *__restrict__ resid, double *__restrict__ slope) {
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
    // Cholesky (lower triangle in place), then L y = b, L^T x = y; one reciprocal square root per
    // column (the reciprocal of its diagonal, within an ulp or two of 1 / sqrt) instead of a
    // division per element; the diagonal itself is never used again. A diagonal that is not
    // positive gives a NaN or infinite reciprocal, so the result is not finite and the pixel falls
    // back, as with the divisions.
    double il[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = M[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) d -= M[j][k] * M[j][k];
        ok = ok && isfinite(d) && d > 0;
        il[j] = rsqrt(d);
#pragma unroll
        for (int a = j + 1; a < 6; a++) {
            double s = M[a][j];
#pragma unroll
            for (int k = 0; k < j; k++) s -= M[a][k] * M[j][k];
            M[a][j] = s * il[j];
        }
    }
#pragma unroll
    for (int a = 0; a < 6; a++) {
        double s = bv[a];
#pragma unroll
        for (int k = 0; k < a; k++) s -= M[a][k] * bv[k];
        bv[a] = s * il[a];
    }
#pragma unroll
    for (int a = 5; a >= 0; a--) {
        double s = bv[a];
#pragma unroll
        for (int k = a + 1; k < 6; k++) s -= M[k][a] * bv[k];
        bv[a] = s * il[a];
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

// Tiled forms of the two moment passes (radius r <= MT_MAXR). y pass: a 64-column x 64-row output
// tile, one input plane at a time staged with its (64 + 2r)-row halo in LDS; each thread covers one
// column and 16 rows. x pass: a 256-pixel row segment, all 15 y-pass planes with the 2r-column halo
// in LDS, then the solve.
constexpr int MT_MAXR = 32;
constexpr int MT_TW = 64, MT_TH = 64;
// the window weights k t^p (p = 0, 1, 2) as a kernel argument, read through the scalar cache
struct KT {
    double k[3][2 * MT_MAXR + 1];
};

__global__ void __launch_bounds__(256) k_moments_y_tiled(const double *__restrict__ sums, double *__restrict__ out,
                                                         int w, int h, const KT kt, int r) {
    extern __shared__ double tile[];  // (MT_TH + 2r) x MT_TW
    const int x0 = blockIdx.x * MT_TW, y0 = blockIdx.y * MT_TH;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const size_t n = (size_t)w * h;
    const int klen = 2 * r + 1, rows = MT_TH + 2 * r;
    const int x = x0 + tx;
    // input plane q feeds y powers 0..npow-1 into outputs base .. base + npow - 1
    const int npow[7] = {3, 3, 3, 2, 2, 1, 1}, base[7] = {0, 3, 6, 9, 11, 13, 14};
    for (int q = 0; q < 7; q++) {
        const double *src = sums + q * n;
        __syncthreads();
        for (int idx = threadIdx.x; idx < rows * MT_TW; idx += blockDim.x) {
            const int ry = idx / MT_TW, cx = idx - ry * MT_TW;
            const int yy = min(max(y0 - r + ry, 0), h - 1), xx = min(x0 + cx, w - 1);
            tile[idx] = src[(size_t)yy * w + xx];
        }
        __syncthreads();
        if (x >= w) continue;
        for (int k = 0; k < MT_TH / 4; k++) {
            const int ly = ty + 4 * k, y = y0 + ly;
            if (y >= h) break;
            double a0 = 0, a1 = 0, a2 = 0;
            const double *col = tile + ly * MT_TW + tx;  // input row y - r at col[0]
            if (npow[q] == 3) {
                for (int j = 0; j < klen; j++) {
                    const double v = col[j * MT_TW];
                    a0 += kt.k[0][j] * v;
                    a1 += kt.k[1][j] * v;
                    a2 += kt.k[2][j] * v;
                }
            } else if (npow[q] == 2) {
                for (int j = 0; j < klen; j++) {
                    const double v = col[j * MT_TW];
                    a0 += kt.k[0][j] * v;
                    a1 += kt.k[1][j] * v;
                }
            } else {
                for (int j = 0; j < klen; j++) a0 += kt.k[0][j] * col[j * MT_TW];
            }
            const size_t i = (size_t)y * w + x;
            out[(size_t)base[q] * n + i] = a0;
            if (npow[q] > 1) out[(size_t)(base[q] + 1) * n + i] = a1;
            if (npow[q] > 2) out[(size_t)(base[q] + 2) * n + i] = a2;
        }
    }
}

// RC > 0: the radius fixed at compile time (r = RC), so that the tap loop unrolls and every LDS
// read takes an immediate offset; RC = 0: any r <= 32 from the argument.
template <int RC>
__global__ void __launch_bounds__(256) k_moments_x_solve_tiled(const double *__restrict__ ym,
                                                               const double *__restrict__ ux_in,
                                                               const double *__restrict__ uy_in,
                                                               double *__restrict__ ux_out, double *__restrict__ uy_out,
                                                               double *__restrict__ energy, double *__restrict__ resid,
                                                               double *__restrict__ slope, int w, int h, const KT kt,
                                                               int r_arg) {
    extern __shared__ double seg[];  // 15 planes x (256 + 2r)
    const int r = RC > 0 ? RC : r_arg;
    const int x0 = blockIdx.x * 256, y = blockIdx.y;
    const int cols = 256 + 2 * r, klen = 2 * r + 1;
    const size_t n = (size_t)w * h;
    for (int idx = threadIdx.x; idx < 15 * cols; idx += blockDim.x) {
        const int q = idx / cols, c = idx - q * cols;
        const int xx = min(max(x0 - r + c, 0), w - 1);
        seg[idx] = ym[(size_t)q * n + (size_t)y * w + xx];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= w) return;
    double mp[3][3][3], mr[2][2][2], mcc = 0, mw = 0;
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
    const double *s0 = seg + threadIdx.x;  // column x - r
#pragma unroll (RC > 0 ? 11 : 1)
    for (int j = 0; j < klen; j++) {
        const double k0 = kt.k[0][j], k1 = kt.k[1][j], k2 = kt.k[2][j];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const double v0 = s0[(3 * p + 0) * cols + j], v1 = s0[(3 * p + 1) * cols + j], v2 = s0[(3 * p + 2) * cols + j];
            mp[p][0][0] += k0 * v0;
            mp[p][1][0] += k1 * v0;
            mp[p][2][0] += k2 * v0;
            mp[p][0][1] += k0 * v1;
            mp[p][1][1] += k1 * v1;
            mp[p][0][2] += k0 * v2;
        }
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const double v0 = s0[(9 + 2 * p) * cols + j], v1 = s0[(10 + 2 * p) * cols + j];
            mr[p][0][0] += k0 * v0;
            mr[p][1][0] += k1 * v0;
            mr[p][0][1] += k0 * v1;
        }
        mcc += k0 * s0[13 * cols + j];
        mw += k0 * s0[14 * cols + j];
    }
    solve_pixel(mp, mr, mcc, mw, (size_t)y * w + x, ux_in, uy_in, ux_out, uy_out, energy, resid, slope);
}

// Register-blocked forms for a compile-time radius R (the default sigma_w = 4 gives R = 16).
// y pass: one thread per (column, 16 consecutive output rows, input plane); the 16 + 2R input rows
// are read once straight from global memory (coalesced across the wavefront) and each is added
// into every output row it reaches.
//
// The weights stay in scalar registers: the Gaussian taps k_t = kt.k[0][t] are symmetric
// (k_t = k_{2R-t} exactly), so R + 1 values serve all 2R + 1 taps, and the weights k_t (t - R)^p of
// the powers p = 1, 2 are not loaded at all. With a_j = j - R - C for input row j of the block and
// c_q = q - C for output row q (C = 8), t - R = a_j - c_q, so with T_p = sum_j k_{j-q} a_j^p v_j,
//   sum_t k_t (t - R) v = T1 - c_q T0,   sum_t k_t (t - R)^2 v = T2 - 2 c_q T1 + c_q^2 T0.
// The row factors a_j^p multiply each input value once; the multiply-adds per output are as many
// as with the three weight rows, which need 6R + 3 scalar register pairs and spill to vector lanes.
constexpr int RB_ROWS = 16;
constexpr int RB_C = 8;

template <int R, int NP>
__device__ __forceinline__ void y_block(const double *__restrict__ src, double *__restrict__ out, size_t n, int w,
                                        int h, int x, int y0, const KT &kt) {
    double acc[RB_ROWS][NP];
#pragma unroll
    for (int q = 0; q < RB_ROWS; q++)
#pragma unroll
        for (int p = 0; p < NP; p++) acc[q][p] = 0;
#pragma unroll
    for (int j = 0; j < RB_ROWS + 2 * R; j++) {
        const int yy = min(max(y0 - R + j, 0), h - 1);
        const double v = src[(size_t)yy * w + x];
        const double a = (double)(j - R - RB_C);
        double vp[NP];
        vp[0] = v;
        if constexpr (NP > 1) vp[1] = a * v;
        if constexpr (NP > 2) vp[2] = (a * a) * v;
#pragma unroll
        for (int q = 0; q < RB_ROWS; q++) {
            const int t = j - q;  // tap index 0 .. 2R of output row y0 + q
            if (t >= 0 && t <= 2 * R) {
                const double k = kt.k[0][t <= R ? t : 2 * R - t];
#pragma unroll
                for (int p = 0; p < NP; p++) acc[q][p] = fma(k, vp[p], acc[q][p]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RB_ROWS; q++) {
        const int y = y0 + q;
        if (y < h) {
            const double c = (double)(q - RB_C);
            const size_t i = (size_t)y * w + x;
            out[i] = acc[q][0];
            if constexpr (NP > 1) out[n + i] = fma(-c, acc[q][0], acc[q][1]);
            if constexpr (NP > 2) out[2 * n + i] = fma(c * c, acc[q][0], fma(-2 * c, acc[q][1], acc[q][2]));
        }
    }
}

template <int R>
__global__ void __launch_bounds__(256) k_moments_y_rb(const double *__restrict__ sums, double *__restrict__ out, int w,
                                                      int h, const KT kt) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y0 = blockIdx.y * RB_ROWS, q = blockIdx.z;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    // plane q of the sums feeds the y powers 0 .. NP - 1 of outputs base ..
    const double *src = sums + (size_t)q * n;
    switch (q) {
        case 0: y_block<R, 3>(src, out + 0 * n, n, w, h, x, y0, kt); break;
        case 1: y_block<R, 3>(src, out + 3 * n, n, w, h, x, y0, kt); break;
        case 2: y_block<R, 3>(src, out + 6 * n, n, w, h, x, y0, kt); break;
        case 3: y_block<R, 2>(src, out + 9 * n, n, w, h, x, y0, kt); break;
        case 4: y_block<R, 2>(src, out + 11 * n, n, w, h, x, y0, kt); break;
        case 5: y_block<R, 1>(src, out + 13 * n, n, w, h, x, y0, kt); break;
        default: y_block<R, 1>(src, out + 14 * n, n, w, h, x, y0, kt); break;
    }
}

// x pass for a compile-time radius R, register-blocked along x, then the solve. A workgroup takes a
// 256-pixel row segment:
//  1. the 15 y-pass planes of the segment and its 2R-column halo are staged in LDS;
//  2. each wavefront takes a fixed set of planes, balanced by the x powers they feed (7, 7, 6 and 6
//     moments), and each lane 4 consecutive pixels: the lane reads its 4 + 2R inputs of a plane once
//     and adds each into the 4 outputs it reaches, in the row-factor form of y_block (the x powers
//     recombined per output from T_p = sum_t k_t a^p v);
//  3. the 26 moments of every pixel go to LDS, over the staging area;
//  4. each thread solves one pixel.
// A lane's inputs sit at columns 4 l + j. Column c of a plane is stored at (c mod 4) XB_YQ + c / 4,
// so the 64 lanes read 64 consecutive words; moment m of pixel p is stored at
// m XB_P + (p mod 4) 64 + p / 4, so the lanes write consecutive words in step 3 and, solving pixel
// 4 l + wave in step 4, read consecutive words.
constexpr int XB_P = 256, XB_B = 4, XB_C = 2;

template <int R>
struct XbLayout {
    static constexpr int YQ = (XB_P + 2 * R + 3) / 4;
    static constexpr int YS = 4 * YQ;
    // the moments go to LDS in two rounds over the staging area: 19, then 7
    static constexpr int WORDS = (15 * YS > 19 * XB_P) ? 15 * YS : 19 * XB_P;
};

template <int R, int NP, int ROW>
__device__ __forceinline__ void x_plane(const double *__restrict__ ys, int lane, const KT &kt,
                                        double (&res)[7][XB_B]) {
    constexpr int YQ = XbLayout<R>::YQ;
    double acc[XB_B][NP];
#pragma unroll
    for (int q = 0; q < XB_B; q++)
#pragma unroll
        for (int p = 0; p < NP; p++) acc[q][p] = 0;
#pragma unroll
    for (int j = 0; j < XB_B + 2 * R; j++) {
        const double v = ys[(j & 3) * YQ + lane + (j >> 2)];
        const double a = (double)(j - R - XB_C);
        double vp[NP];
        vp[0] = v;
        if constexpr (NP > 1) vp[1] = a * v;
        if constexpr (NP > 2) vp[2] = (a * a) * v;
#pragma unroll
        for (int q = 0; q < XB_B; q++) {
            const int t = j - q;  // tap 0 .. 2R of output q
            if (t >= 0 && t <= 2 * R) {
                const double k = kt.k[0][t <= R ? t : 2 * R - t];
#pragma unroll
                for (int p = 0; p < NP; p++) acc[q][p] = fma(k, vp[p], acc[q][p]);
            }
        }
        // Without this, the compiler orders the multiply-adds chain by chain and keeps every
        // input and its row factors live at once (over 200 VGPRs); touching the accumulators
        // every two inputs keeps the input-major order.
        if (j & 1) {
#pragma unroll
            for (int q = 0; q < XB_B; q++)
#pragma unroll
                for (int p = 0; p < NP; p++) asm volatile("" : "+v"(acc[q][p]));
        }
    }
#pragma unroll
    for (int q = 0; q < XB_B; q++) {
        const double c = (double)(q - XB_C);
        res[ROW][q] = acc[q][0];
        if constexpr (NP > 1) res[ROW + 1][q] = fma(-c, acc[q][0], acc[q][1]);
        if constexpr (NP > 2) res[ROW + 2][q] = fma(c * c, acc[q][0], fma(-2 * c, acc[q][1], acc[q][2]));
    }
}

template <int NP, int ROW>
__device__ __forceinline__ void x_put(double *__restrict__ mom, int base, int lane, const double (&res)[7][XB_B]) {
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int q = 0; q < XB_B; q++) mom[(base + p) * XB_P + q * 64 + lane] = res[ROW + p][q];
}

template <int R>
__global__ void __launch_bounds__(256) k_moments_x_rb(const double *__restrict__ ym, const double *__restrict__ ux_in,
                                                      const double *__restrict__ uy_in, double *__restrict__ ux_out,
                                                      double *__restrict__ uy_out, double *__restrict__ energy,
                                                      double *__restrict__ resid, double *__restrict__ slope, int w,
                                                      int h, const KT kt) {
    using L = XbLayout<R>;
    __shared__ double lds[L::WORDS];
    const int x0 = blockIdx.x * XB_P, y = blockIdx.y;
    const size_t n = (size_t)w * h;
    // stage: thread t takes column t of every plane and, for t < 2R, column 256 + t; the plane
    // offset is uniform, so the addresses cost no per-plane vector arithmetic. Every load is issued
    // before the first is waited for.
    {
        static_assert(2 * R <= XB_P, "the halo columns are taken by the first 2R threads");
        const int c0 = threadIdx.x, c1 = threadIdx.x + XB_P;
        const bool two = c1 < XB_P + 2 * R;
        const double *row = ym + (size_t)y * w;
        const double *p0 = row + min(max(x0 - R + c0, 0), w - 1);
        const double *p1 = row + min(max(x0 - R + (two ? c1 : c0), 0), w - 1);
        double v0[15], v1[15];
#pragma unroll
        for (int q = 0; q < 15; q++) v0[q] = p0[(size_t)q * n];
        if (two) {
#pragma unroll
            for (int q = 0; q < 15; q++) v1[q] = p1[(size_t)q * n];
        }
        double *d0 = lds + (c0 & 3) * L::YQ + (c0 >> 2), *d1 = lds + (c1 & 3) * L::YQ + (c1 >> 2);
#pragma unroll
        for (int q = 0; q < 15; q++) d0[q * L::YS] = v0[q];
        if (two) {
#pragma unroll
            for (int q = 0; q < 15; q++) d1[q * L::YS] = v1[q];
        }
    }
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // y-pass planes: gxx t^0,1,2 | gxy t^0,1,2 | gyy t^0,1,2 | gxc t^0,1 | gyc t^0,1 | gcc | w.
    // moments: mp[p][a][b] at 6 p + (0, 1, 2 | 3, 4 | 5) for (a, b) = (0..2, 0 | 0..1, 1 | 0, 2);
    // mr[p][a][b] at 18 + 3 p + (0, 1 | 2) for (a, b) = (0..1, 0 | 0, 1); gcc 24, w 25.
    // Every wavefront runs the same four jobs, of 3, 2, 1 and 1 x powers, on planes from these
    // tables (byte v for wavefront v): the third power of plane 1 (gxx t^1) is computed and
    // dropped, and wavefront 3 has no fourth job (it repeats plane 14 and stores nothing).
    constexpr unsigned P3 = 0x01060300u, B3 = 0x030c0600u;  // planes 0, 3, 6, 1 -> moments 0, 6, 12, 3
    constexpr unsigned P2 = 0x0b090704u, B2 = 0x15120f09u;  // planes 4, 7, 9, 11 -> 9, 15, 18, 21
    constexpr unsigned P1 = 0x0a080502u, B1 = 0x14110b05u;  // planes 2, 5, 8, 10 -> 5, 11, 17, 20
    constexpr unsigned Q1 = 0x0e0e0d0cu, C1 = 0xff191817u;  // planes 12, 13, 14 -> 23, 24, 25
    const int sh = 8 * wave;
    double res[7][XB_B];
    x_plane<R, 3, 0>(lds + ((P3 >> sh) & 255) * L::YS, lane, kt, res);
    x_plane<R, 2, 3>(lds + ((P2 >> sh) & 255) * L::YS, lane, kt, res);
    x_plane<R, 1, 5>(lds + ((P1 >> sh) & 255) * L::YS, lane, kt, res);
    x_plane<R, 1, 6>(lds + ((Q1 >> sh) & 255) * L::YS, lane, kt, res);
    // The 26 moments reach the solving threads in two rounds through the staging area: first the
    // 3- and 2-power jobs of every wavefront (19 moments), then the 1-power jobs (7). That takes
    // 38.9 KB of LDS instead of 53 for all 26 at once: four workgroups per CU instead of three.
    // Slots of the first round, by moment: 0-4 -> 0-4, 6-10 -> 5-9, 12-16 -> 10-14, 18-19 -> 15-16,
    // 21-22 -> 17-18; of the second: 5, 11, 17, 20, 23, 24, 25 -> 0-6.
    constexpr unsigned S3 = 0x030a0500u, S2 = 0x110f0d08u;  // first round: P3 and P2 jobs
    constexpr unsigned S1 = 0x03020100u, SQ = 0xff060504u;  // second round: P1 and Q1 jobs
    __syncthreads();
    double *mom = lds;
    const double *mm = mom + wave * 64 + lane;
    if (wave < 3)
        x_put<3, 0>(mom, (S3 >> sh) & 255, lane, res);
    else
        x_put<2, 0>(mom, (S3 >> sh) & 255, lane, res);
    x_put<2, 3>(mom, (S2 >> sh) & 255, lane, res);
    __syncthreads();
    double mp[3][3][3], mr[2][2][2];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        mp[p][0][0] = mm[(5 * p + 0) * XB_P];
        mp[p][1][0] = mm[(5 * p + 1) * XB_P];
        mp[p][2][0] = mm[(5 * p + 2) * XB_P];
        mp[p][0][1] = mm[(5 * p + 3) * XB_P];
        mp[p][1][1] = mm[(5 * p + 4) * XB_P];
        mp[p][1][2] = mp[p][2][1] = mp[p][2][2] = 0;
    }
#pragma unroll
    for (int p = 0; p < 2; p++) {
        mr[p][0][0] = mm[(15 + 2 * p) * XB_P];
        mr[p][1][0] = mm[(16 + 2 * p) * XB_P];
        mr[p][1][1] = 0;
    }
    __syncthreads();
    x_put<1, 5>(mom, (S1 >> sh) & 255, lane, res);
    if (wave < 3) x_put<1, 6>(mom, (SQ >> sh) & 255, lane, res);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 3; p++) mp[p][0][2] = mm[p * XB_P];
    mr[0][0][1] = mm[3 * XB_P];
    mr[1][0][1] = mm[4 * XB_P];
    const double mcc = mm[5 * XB_P], mw = mm[6 * XB_P];
    const int x = x0 + 4 * lane + wave;
    if (x >= w) return;
    solve_pixel(mp, mr, mcc, mw, (size_t)y * w + x, ux_in, uy_in, ux_out, uy_out, energy, resid, slope);
}

// Energy threshold (phaseflow.py:149): pixels with energy <= 1e-9 median + 1e-30 (or NaN) keep the
// input displacement and get resid = pi.
__global__ void k_energy_reset(const double *__restrict__ energy, const double *__restrict__ med,
                               const double *__restrict__ ux_in, const double *__restrict__ uy_in, double *ux,
                               double *uy, double *resid, double *slope, size_t n, const unsigned int *gate) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (gate && !*gate)) return;
    const double thr = 1e-9 * med[0] + 1e-30;
    if (!(energy[i] > thr)) {
        ux[i] = ux_in[i];
        uy[i] = uy_in[i];
        resid[i] = M_PI;
        slope[4 * i] = slope[4 * i + 1] = slope[4 * i + 2] = slope[4 * i + 3] = 0;
    }
}

// ------------------------------------------------------------------ quadratic shape function
// shape_order 2 (phaseflow.py:123-161): 12 unknowns (u, v and their first and second
// derivatives), moments up to total power 4. A y pass writes 23 planes (the 7 sums times the y
// powers each family needs), an x pass the 59 moments, and a solve pass the 12 x 12 Cholesky per
// pixel. kt holds the window taps times t^p for p = 0 .. 4 (row p at p * (2r + 1)).
constexpr int O2_YPLANES = 23, O2_MOMENTS = 59;

// y-plane of (source, y power): sources 0..2 (gxx, gxy, gyy) powers 0..4, 3..4 (gxc, gyc) 0..2,
// 5 (gcc), 6 (gw) power 0
__host__ __device__ constexpr int o2_yplane(int src, int py) {
    return src < 3 ? src * 5 + py : (src < 5 ? 15 + (src - 3) * 3 + py : 21 + (src - 5));
}
// moment of (source, x power, y power); the x powers of a y power run 0 .. (4 or 2) - py
__host__ __device__ constexpr int o2_moment(int src, int px, int py) {
    // sources 0..2: 15 moments each (px + py <= 4), ordered by py then px
    // offset of (py) within a source with max total T: sum_{q < py} (T + 1 - q)
    return src < 3 ? src * 15 + (py * 5 - py * (py - 1) / 2) + px
                   : (src < 5 ? 45 + (src - 3) * 6 + (py * 3 - py * (py - 1) / 2) + px : 57 + (src - 5));
}

__global__ void __launch_bounds__(256) k_moments2_y(const double *__restrict__ sums, double *__restrict__ out, int w,
                                                    int h, const double *__restrict__ kt, int r) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    const int klen = 2 * r + 1;
    double acc[O2_YPLANES];
#pragma unroll
    for (int q = 0; q < O2_YPLANES; q++) acc[q] = 0;
    for (int j = -r; j <= r; j++) {
        const int yy = min(max(y + j, 0), h - 1);
        const size_t i = (size_t)yy * w + x;
        double kp[5];
#pragma unroll
        for (int p = 0; p < 5; p++) kp[p] = kt[p * klen + j + r];
#pragma unroll
        for (int src = 0; src < 7; src++) {
            const double v = sums[src * n + i];
            const int np = src < 3 ? 5 : (src < 5 ? 3 : 1);
#pragma unroll
            for (int py = 0; py < 5; py++)
                if (py < np) acc[o2_yplane(src, py)] += kp[py] * v;
        }
    }
    const size_t i = (size_t)y * w + x;
#pragma unroll
    for (int q = 0; q < O2_YPLANES; q++) out[q * n + i] = acc[q];
}

__global__ void __launch_bounds__(128) k_moments2_x(const double *__restrict__ ym, double *__restrict__ mom, int w,
                                                    int h, const double *__restrict__ kt, int r) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    const int klen = 2 * r + 1;
    double acc[O2_MOMENTS];
#pragma unroll
    for (int q = 0; q < O2_MOMENTS; q++) acc[q] = 0;
    const double *row = ym + (size_t)y * w;
    for (int j = -r; j <= r; j++) {
        const int xx = min(max(x + j, 0), w - 1);
        double kp[5];
#pragma unroll
        for (int p = 0; p < 5; p++) kp[p] = kt[p * klen + j + r];
#pragma unroll
        for (int src = 0; src < 7; src++) {
            const int T = src < 3 ? 4 : (src < 5 ? 2 : 0);
#pragma unroll
            for (int py = 0; py <= 4; py++) {
                if (py > T) continue;
                const double v = row[o2_yplane(src, py) * n + xx];
#pragma unroll
                for (int px = 0; px <= 4; px++)
                    if (px + py <= T) acc[o2_moment(src, px, py)] += kp[px] * v;
            }
        }
    }
    const size_t i = (size_t)y * w + x;
#pragma unroll
    for (int q = 0; q < O2_MOMENTS; q++) mom[q * n + i] = acc[q];
}

// The 12 x 12 system per pixel (unknown order of phaseflow.py:123-129: u, v, then for each power
// (1,0), (0,1), (2,0), (1,1), (0,2) the u and v coefficients), as the 6 x 6 form: aperture test,
// ridge 1e-12 tr / n, Cholesky, residual, slope = the four first derivatives.
__global__ void __launch_bounds__(64) k_solve2(const double *__restrict__ mom, const double *__restrict__ ux_in,
                                               const double *__restrict__ uy_in, double *__restrict__ ux_out,
                                               double *__restrict__ uy_out, double *__restrict__ energy,
                                               double *__restrict__ resid, double *__restrict__ slope, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int N = 12;
    constexpr int uc[N] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1};
    constexpr int ua[N] = {0, 0, 1, 1, 0, 0, 2, 2, 1, 1, 0, 0};
    constexpr int ub[N] = {0, 0, 0, 0, 1, 1, 0, 0, 1, 1, 2, 2};
    double M[N][N], bv[N], b0[N];
#pragma unroll
    for (int a = 0; a < N; a++) {
#pragma unroll
        for (int b = 0; b <= a; b++) {
            const int src = (uc[a] == 0 && uc[b] == 0) ? 0 : ((uc[a] == 1 && uc[b] == 1) ? 2 : 1);
            M[a][b] = mom[(size_t)o2_moment(src, ua[a] + ua[b], ub[a] + ub[b]) * n + i];
        }
        bv[a] = mom[(size_t)o2_moment(3 + uc[a], ua[a], ub[a]) * n + i];
    }
    const double mcc = mom[(size_t)o2_moment(5, 0, 0) * n + i], mw = mom[(size_t)o2_moment(6, 0, 0) * n + i];
    const double e = M[0][0] + M[1][1];
    const double det2 = M[0][0] * M[1][1] - M[1][0] * M[1][0];
    energy[i] = e;
    bool ok = det2 > APERTURE_TOL * e * e;
    if (!ok) {
#pragma unroll
        for (int a = 0; a < N; a++) {
#pragma unroll
            for (int b = 0; b <= a; b++) M[a][b] = (a == b) ? 1.0 : 0.0;
            bv[a] = 0;
        }
    }
    double tr = 0;
#pragma unroll
    for (int a = 0; a < N; a++) tr += M[a][a];
#pragma unroll
    for (int a = 0; a < N; a++) {
        M[a][a] += RIDGE * tr / (double)N;
        b0[a] = bv[a];
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        double d = M[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) d -= M[j][k] * M[j][k];
        ok = ok && isfinite(d) && d > 0;
        const double l = sqrt(d);
        M[j][j] = l;
#pragma unroll
        for (int a = j + 1; a < N; a++) {
            double s = M[a][j];
#pragma unroll
            for (int k = 0; k < j; k++) s -= M[a][k] * M[j][k];
            M[a][j] = s / l;
        }
    }
#pragma unroll
    for (int a = 0; a < N; a++) {
        double s = bv[a];
#pragma unroll
        for (int k = 0; k < a; k++) s -= M[a][k] * bv[k];
        bv[a] = s / M[a][a];
    }
#pragma unroll
    for (int a = N - 1; a >= 0; a--) {
        double s = bv[a];
#pragma unroll
        for (int k = a + 1; k < N; k++) s -= M[k][a] * bv[k];
        bv[a] = s / M[a][a];
    }
#pragma unroll
    for (int a = 0; a < N; a++) ok = ok && isfinite(bv[a]);
    double *sl = slope + 4 * i;
    if (ok) {
        double tb = 0;
#pragma unroll
        for (int a = 0; a < N; a++) tb += bv[a] * b0[a];
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

// ------------------------------------------------------------------ exact median (np.median)
// Order-preserving 64-bit keys of the doubles (a sort by the keys is a sort by total order: -0.0
// before 0.0, NaN of either sign at the ends).
__device__ __forceinline__ unsigned long long order_key(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ double from_key(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    return __longlong_as_double((long long)b);
}

// Radix select over the keys, 8 bits per pass, with histograms privatised in LDS. Each pass
// derives the digit of the pass before from its histogram (select.rs:79-153 in one pass per digit).
struct SelectState {
    unsigned long long prefix[9];  // key digits fixed before pass p (prefix[8]: the selected key)
    unsigned long long rank[9];    // rank left within them
    unsigned int nan_count;
    unsigned int undecided;          // bounded reset: some pixel needs the exact median
    unsigned long long below_count;  // elements with key < the selected key
    unsigned long long below_max;    // largest key below it
    unsigned int hist[8][256];       // one 8-bit digit histogram per pass
};

__global__ void k_select_init(SelectState *st, unsigned long long rank) {
    if (threadIdx.x == 0) {
        st->prefix[0] = 0;
        st->rank[0] = rank;
        st->nan_count = 0;
        st->undecided = 0;
        st->below_count = 0;
        st->below_max = 0;
    }
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) st->hist[i >> 8][i & 255] = 0;
}

// The digit of pass p - 1 and the rank left after it, from that pass's histogram: the smallest
// digit d whose cumulative count exceeds the rank (255 if none does). Computed by one whole
// wavefront (every lane gets the result): 4 bins per lane, a shuffle scan over the lanes.
__device__ __forceinline__ void select_pick(const SelectState *st, int p, unsigned long long &prefix,
                                            unsigned long long &rank) {
    const int lane = threadIdx.x & 63;
    const unsigned long long r = st->rank[p - 1];
    const unsigned int *h = st->hist[p - 1];
    const unsigned long long c0 = h[4 * lane], c1 = h[4 * lane + 1], c2 = h[4 * lane + 2], c3 = h[4 * lane + 3];
    unsigned long long incl = c0 + c1 + c2 + c3;
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    const unsigned long long excl = incl - (c0 + c1 + c2 + c3);
    const unsigned long long hit = __ballot(incl > r);
    int d = 255;
    unsigned long long below = 0;
    if (hit) {
        const int l = __ffsll((long long)hit) - 1;
        // inside lane l's four bins
        unsigned long long cum = __shfl(excl, l);
        const unsigned long long b0 = __shfl(c0, l), b1 = __shfl(c1, l), b2 = __shfl(c2, l);
        int j = 0;
        if (cum + b0 > r) j = 0;
        else if (cum + b0 + b1 > r) { j = 1; cum += b0; }
        else if (cum + b0 + b1 + b2 > r) { j = 2; cum += b0 + b1; }
        else { j = 3; cum += b0 + b1 + b2; }
        d = 4 * l + j;
        below = cum;
    } else {
        below = __shfl(incl, 63);  // no bin reaches the rank: digit 255, every count below it
    }
    const int shift = 64 - 8 * p;
    prefix = st->prefix[p - 1] | ((unsigned long long)d << shift);
    rank = r - below;
}

// Pass p (digit bits 63 - 8p .. 56 - 8p): every workgroup first derives the digit of pass p - 1
// from its histogram (block 0 records it), then histograms the keys that share the prefix.
__global__ void __launch_bounds__(256) k_select_hist(const double *__restrict__ v, size_t n, int p,
                                                     SelectState *st, int gated) {
    if (gated && !st->undecided) return;  // the bounded reset decided every pixel
    __shared__ unsigned int lh[256];
    lh[threadIdx.x] = 0;
    unsigned long long prefix = 0;
    if (p > 0) {
        unsigned long long rank;
        select_pick(st, p, prefix, rank);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->prefix[p] = prefix;
            st->rank[p] = rank;
        }
    }
    __syncthreads();
    const int shift = 56 - 8 * p;
    const int hi = shift + 8;  // bits above the current digit must equal the prefix
    unsigned int nan_local = 0;
    // Counts one key. The first digits of the energies fall into a handful of bins, so a wavefront
    // whose counted keys share one bin adds them with a single LDS atomic.
    auto count = [&](double x, bool live) {
        if (p == 0 && live && isnan(x)) nan_local++;
        const unsigned long long k = order_key(x);
        const bool in = live && (hi >= 64 || (k >> hi) == (prefix >> hi));
        const unsigned int bin = (unsigned int)(k >> shift) & 255u;
        const unsigned long long act = __ballot(in);
        if (!act) return;
        const unsigned int lead = (unsigned int)__builtin_amdgcn_readlane((int)bin, __builtin_ctzll(act));
        const unsigned long long same = __ballot(in && bin == lead);
        if (same == act) {
            if ((threadIdx.x & 63) == (unsigned int)__builtin_ctzll(act)) atomicAdd(&lh[lead], (unsigned int)__popcll(act));
        } else if (in) {
            atomicAdd(&lh[bin], 1u);
        }
    };
    // eight loads in flight per thread, then the remainder one at a time; every lane of a wavefront
    // runs the same trip counts, so the ballots see the whole wavefront
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t first = (size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u), lane = threadIdx.x & 63;
    size_t i = first;
    for (; i + 7 * stride + 63 < n; i += 8 * stride) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = v[i + u * stride + lane];
#pragma unroll
        for (int u = 0; u < 8; u++) count(x[u], true);
    }
    for (; i < n; i += stride) {
        const bool live = i + lane < n;
        count(live ? v[i + lane] : 0.0, live);
    }
    if (nan_local) atomicAdd(&st->nan_count, nan_local);
    __syncthreads();
    if (lh[threadIdx.x]) atomicAdd(&st->hist[p][threadIdx.x], lh[threadIdx.x]);
}

// The selected key, from the last histogram (the first wavefront of workgroup 0 records it)
__device__ __forceinline__ unsigned long long select_key(SelectState *st) {
    unsigned long long key, rank;
    select_pick(st, 8, key, rank);
    return key;
}

__global__ void __launch_bounds__(256) k_select_below(const double *__restrict__ v, size_t n, SelectState *st,
                                                      int gated) {
    if (gated && !st->undecided) return;
    const unsigned long long key = select_key(st);
    unsigned long long cnt = 0, mx = 0;
    auto below = [&](double x) {
        const unsigned long long k = order_key(x);
        if (k < key) {
            cnt++;
            mx = k > mx ? k : mx;
        }
    };
    // eight loads in flight per thread, then the remainder one at a time
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 7 * stride < n; i += 8 * stride) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = v[i + u * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) below(x[u]);
    }
    for (; i < n; i += stride) below(v[i]);
    // wavefront reduction, then one atomic per workgroup
    __shared__ unsigned long long rc[4], rm[4];
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_down(cnt, off);
        unsigned long long o = __shfl_down(mx, off);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        rc[threadIdx.x >> 6] = cnt;
        rm[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = rc[0] + rc[1] + rc[2] + rc[3];
        unsigned long long m = max(max(rm[0], rm[1]), max(rm[2], rm[3]));
        if (c) {
            atomicAdd(&st->below_count, c);
            atomicMax(&st->below_max, m);
        }
    }
}

__global__ void k_select_finish(SelectState *st, unsigned long long k, int even, double *out, int gated) {
    if (gated && !st->undecided) return;
    const unsigned long long key = select_key(st);
    if (threadIdx.x != 0) return;
    st->prefix[8] = key;
    if (st->nan_count) {
        out[0] = NAN;
        return;
    }
    const double upper = from_key(key);
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

void prefilter_real(float *c, float *tmp, int w, int h, hipStream_t s) {
    if (h > 1) {
        const dim3 g((w + 255) / 256, (h + PF_SEG - 1) / PF_SEG);
        hipLaunchKernelGGL(k_prefilter_cols_causal, g, dim3(256), 0, s, (const float *)c, tmp, w, h);
        hipLaunchKernelGGL(k_prefilter_cols_anti, g, dim3(256), 0, s, (const float *)tmp, c, w, h);
    }
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_prefilter_rows, dim3(h), dim3(256), 2 * lp_len<4>(w) * sizeof(float), s, c, w);
    LAUNCHCHECK();
}

void prefilter_real2(const float *src0, const float *src1, float *dst0, float *dst1, float *tmp, int w, int h,
                     hipStream_t s) {
    const size_t n = (size_t)w * h;
    const PlanePair P{{src0, src1}, {tmp, tmp + n}, {dst0, dst1}};
    if (h > 1) {
        const dim3 g((w + 255) / 256, (h + PF_SEG - 1) / PF_SEG, 2);
        hipLaunchKernelGGL(k_prefilter_cols_causal2, g, dim3(256), 0, s, P, w, h);
        hipLaunchKernelGGL(k_prefilter_cols_anti2, g, dim3(256), 0, s, P, w, h);
    } else {
        dcopy(dst0, src0, n * sizeof(float), s);
        dcopy(dst1, src1, n * sizeof(float), s);
    }
    LAUNCHCHECK();
    // two rows of up to 4032 samples in two buffers: up to 68.5 KB, above the 64 KB default
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void *)k_prefilter_rows2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
    });
    hipLaunchKernelGGL(k_prefilter_rows2, dim3(h), dim3(256), 2 * lp_len<4>(2 * w) * sizeof(float), s, P, w);
    LAUNCHCHECK();
}

void blur(const double *in, double *tmp, float *out, int w, int h, const double *taps_host, int r, hipStream_t s) {
    if (r > 64) throw Error(OFX_EUNSUPPORTED, "sigma_s above 16 px is not supported by the GPU blur");
    BT bt;
    for (int j = 0; j <= 2 * r; j++) bt.k[j] = taps_host[j];
    if (r == 16)
        hipLaunchKernelGGL(k_blur_y_rb<16>, dim3((w + 255) / 256, (h + BLUR_ROWS - 1) / BLUR_ROWS), dim3(256), 0, s,
                           in, tmp, w, h, bt);
    else
        hipLaunchKernelGGL(k_blur_y, grid2(w, h), dim3(256), 0, s, in, tmp, w, h, bt, r);
    if (r == 16)
        hipLaunchKernelGGL(k_blur_x_rb<16>, dim3((w + 1023) / 1024, h), dim3(256), 0, s, tmp, out, w, bt);
    else
        hipLaunchKernelGGL(k_blur_x_lds, grid2(w, h), dim3(256), 0, s, tmp, out, w, bt, r);
    LAUNCHCHECK();
}

void warp(const float *c2, int w2, int h2, long shx, long shy, const float *sx, const float *sy, float *out, int w,
          int h, hipStream_t s) {
    hipLaunchKernelGGL(k_warp, grid2(w, h), dim3(256), 0, s, c2, w2, h2, (double)shx, (double)shy, sx, sy, out, w,
                       h);
    LAUNCHCHECK();
}

void refpack(const float2 *resp, int F, float4 *out, int w, int h, int wp, hipStream_t s) {
    const size_t np = (size_t)wp * (h + 2 * PAD);
    hipLaunchKernelGGL(k_refpack, dim3((w + 255) / 256, h, F), dim3(256), 0, s, resp, out, w, h, wp, F, np);
    LAUNCHCHECK();
}

void accumulate(const double *ux, const double *uy, const float *sx, const float *sy, const float *csx,
                const float *csy, const float2 *resp, const float4 *refp, double *sums, int w, int h, int F,
                hipStream_t s) {
    const int wp = w + 2 * PAD, hp = h + 2 * PAD;
    const int gx = (w + ACC_T - 1) / ACC_T, gy = (h + ACC_T - 1) / ACC_T;
    // an even padded width keeps every 16-byte pair of nodes aligned for the load-to-LDS staging
    if (wp % 2 == 0)
        hipLaunchKernelGGL(k_accumulate_tiled<true>, dim3(8 * ((gx + 7) / 8) * gy), dim3(256), 0, s,
                           ux, uy, sx, sy, csx, csy, resp, refp, sums, w, h, wp, hp, F);
    else
        hipLaunchKernelGGL(k_accumulate_tiled<false>, dim3(8 * ((gx + 7) / 8) * gy), dim3(256), 0, s,
                           ux, uy, sx, sy, csx, csy, resp, refp, sums, w, h, wp, hp, F);
    LAUNCHCHECK();
}

void moments_solve(const double *sums, double *ym, const double *ux_in, const double *uy_in, double *ux_out,
                   double *uy_out, double *energy, double *resid, double *slope, int w, int h, const double *kt,
                   const double *kt_host, int r, hipStream_t s) {
    if (r == 16 || r <= MT_MAXR) {
        KT kt_arg;
        const int klen = 2 * r + 1;
        for (int p = 0; p < 3; p++)
            for (int j = 0; j < klen; j++) kt_arg.k[p][j] = kt_host[p * klen + j];
        if (r == 16) {
            // one launch over the 7 sum planes: launches split by power count (164 and 100 VGPRs for
            // the two- and one-power planes) took 158 against 145 us at 2048^2
            hipLaunchKernelGGL(k_moments_y_rb<16>, dim3((w + 255) / 256, (h + RB_ROWS - 1) / RB_ROWS, 7), dim3(256), 0,
                               s, sums, ym, w, h, kt_arg);
            hipLaunchKernelGGL(k_moments_x_rb<16>, dim3((w + XB_P - 1) / XB_P, h), dim3(256), 0, s, ym, ux_in, uy_in,
                               ux_out, uy_out, energy, resid, slope, w, h, kt_arg);
            LAUNCHCHECK();
            return;
        }
        hipLaunchKernelGGL(k_moments_y_tiled, dim3((w + MT_TW - 1) / MT_TW, (h + MT_TH - 1) / MT_TH), dim3(256),
                           (size_t)(MT_TH + 2 * r) * MT_TW * sizeof(double), s, sums, ym, w, h, kt_arg, r);
        hipLaunchKernelGGL(k_moments_x_solve_tiled<0>, dim3((w + 255) / 256, h), dim3(256),
                           (size_t)15 * (256 + 2 * r) * sizeof(double), s, ym, ux_in, uy_in, ux_out, uy_out, energy,
                           resid, slope, w, h, kt_arg, r);
    } else {
        hipLaunchKernelGGL(k_moments_y, grid2(w, h), dim3(256), 0, s, sums, ym, w, h, kt, r);
        hipLaunchKernelGGL(k_moments_x_solve, grid2(w, h), dim3(256), 0, s, ym, ux_in, uy_in, ux_out, uy_out, energy,
                           resid, slope, w, h, kt, r);
    }
    LAUNCHCHECK();
}

size_t moments2_scratch_doubles(size_t n) { return (size_t)(O2_YPLANES + O2_MOMENTS) * n; }

void moments_solve2(const double *sums, double *scratch, const double *ux_in, const double *uy_in, double *ux_out,
                    double *uy_out, double *energy, double *resid, double *slope, int w, int h, const double *kt, int r,
                    hipStream_t s) {
    const size_t n = (size_t)w * h;
    double *ym = scratch, *mom = scratch + (size_t)O2_YPLANES * n;
    hipLaunchKernelGGL(k_moments2_y, grid2(w, h), dim3(256), 0, s, sums, ym, w, h, kt, r);
    hipLaunchKernelGGL(k_moments2_x, dim3((w + 127) / 128, h), dim3(128), 0, s, (const double *)ym, mom, w, h, kt, r);
    hipLaunchKernelGGL(k_solve2, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, (const double *)mom, ux_in, uy_in,
                       ux_out, uy_out, energy, resid, slope, n);
    LAUNCHCHECK();
}

void energy_reset(const double *energy, const double *med, const double *ux_in, const double *uy_in, double *ux,
                  double *uy, double *resid, double *slope, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_energy_reset, dim3((n + 255) / 256), dim3(256), 0, s, energy, med, ux_in, uy_in, ux, uy,
                       resid, slope, n, (const unsigned int *)nullptr);
    LAUNCHCHECK();
}

size_t median_scratch_bytes() { return sizeof(SelectState); }

// Workgroups of a histogram pass: each ends with up to 256 atomic adds into the one global
// histogram, so 512 longer workgroups (two per CU, eight loads in flight per thread) rather than
// 1024. OFX_SELECT_BLOCKS overrides the count, for timing.
int hist_blocks(size_t n) {
    static const int cap = [] {
        const char *v = std::getenv("OFX_SELECT_BLOCKS");
        const int b = v ? std::atoi(v) : 0;
        return b > 0 ? b : 512;
    }();
    return (int)std::max<size_t>(1, std::min<size_t>(cap, (n + 255) / 256));
}


// numpy.median of n values (NaN if any is NaN), for an even n the mean of the n/2 - 1 and n/2 order
// statistics (core/select.rs:19-42).
void median(const double *v, size_t n, void *scratch, double *out, hipStream_t s) {
    const unsigned long long k = n / 2;
    const int even = (n % 2 == 0) ? 1 : 0;
    SelectState *st = (SelectState *)scratch;
    hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, s, st, k);
    const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256), hblocks = hist_blocks(n);
    for (int p = 0; p < 8; p++) hipLaunchKernelGGL(k_select_hist, dim3(hblocks), dim3(256), 0, s, v, n, p, st, 0);
    if (even) hipLaunchKernelGGL(k_select_below, dim3(hblocks), dim3(256), 0, s, v, n, st, 0);
    hipLaunchKernelGGL(k_select_finish, dim3(1), dim3(64), 0, s, st, k, even, out, 0);
    LAUNCHCHECK();
}

// The energy threshold without the exact median where the pixels allow it. After two digit passes
// the median's key shares 16 bits (sign, exponent and 4 mantissa bits) with [lo, hi], the keys of
// its bucket, and thr(M) = 1e-9 M + 1e-30 is nondecreasing in M, so thr(lo) <= thr(M) <= thr(hi):
// energies above thr(hi) keep the solve, energies not above thr(lo) (NaN included) are reset, as
// with the exact median. Only an energy in (thr(lo), thr(hi)] needs M itself; such a pixel, a NaN
// in the values, a NaN bound or (n even) a lower middle element outside the bucket sets
// `undecided`, and the remaining passes, the exact median and the plain reset then run; otherwise
// they return at once and `out` is not written.
__global__ void __launch_bounds__(256) k_energy_reset_bounded(const double *__restrict__ energy, SelectState *st,
                                                              int even, const double *__restrict__ ux_in,
                                                              const double *__restrict__ uy_in, double *ux, double *uy,
                                                              double *resid, double *slope, size_t n) {
    __shared__ double thr_lo, thr_hi;
    __shared__ int ok;
    if (threadIdx.x < 64) {
        unsigned long long prefix, rank;
        select_pick(st, 2, prefix, rank);
        if (threadIdx.x == 0) {
            const double lo = from_key(prefix), hi = from_key(prefix | 0xFFFFFFFFFFFFull);
            // rank: elements of the bucket below the upper middle one; with one or more, the lower
            // middle element (n even) is in the bucket too
            ok = st->nan_count == 0 && (!even || rank >= 1) && !isnan(lo) && !isnan(hi);
            thr_lo = 1e-9 * lo + 1e-30;
            thr_hi = 1e-9 * hi + 1e-30;
        }
    }
    __syncthreads();
    if (!ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->undecided = 1;
        return;
    }
    const double tl = thr_lo, th = thr_hi;
    bool open = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double e = energy[i];
        if (e > th) continue;
        if (!(e > tl)) {
            ux[i] = ux_in[i];
            uy[i] = uy_in[i];
            resid[i] = M_PI;
            slope[4 * i] = slope[4 * i + 1] = slope[4 * i + 2] = slope[4 * i + 3] = 0;
        } else {
            open = true;
        }
    }
    if (open) st->undecided = 1;
}

void median_energy_reset(const double *energy, size_t n, void *scratch, double *med, const double *ux_in,
                         const double *uy_in, double *ux, double *uy, double *resid, double *slope, hipStream_t s) {
    const unsigned long long k = n / 2;
    const int even = (n % 2 == 0) ? 1 : 0;
    SelectState *st = (SelectState *)scratch;
    hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, s, st, k);
    const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256), hblocks = hist_blocks(n);
    for (int p = 0; p < 2; p++) hipLaunchKernelGGL(k_select_hist, dim3(hblocks), dim3(256), 0, s, energy, n, p, st, 0);
    hipLaunchKernelGGL(k_energy_reset_bounded, dim3(blocks), dim3(256), 0, s, energy, st, even, ux_in, uy_in, ux, uy,
                       resid, slope, n);
    for (int p = 2; p < 8; p++) hipLaunchKernelGGL(k_select_hist, dim3(hblocks), dim3(256), 0, s, energy, n, p, st, 1);
    if (even) hipLaunchKernelGGL(k_select_below, dim3(hblocks), dim3(256), 0, s, energy, n, st, 1);
    hipLaunchKernelGGL(k_select_finish, dim3(1), dim3(64), 0, s, st, k, even, med, 1);
    hipLaunchKernelGGL(k_energy_reset, dim3((n + 255) / 256), dim3(256), 0, s, energy, (const double *)med, ux_in,
                       uy_in, ux, uy, resid, slope, n, (const unsigned int *)&st->undecided);
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
