// The learned coarse proposal: a 6-layer 3x3 convolutional network in bfloat16 with float32
// accumulation on the matrix cores (v_mfma_f32_32x32x16_bf16), its training kernels (backward
// by data and by weight, loss, Adam) and the renderer of its training pairs. It has no counterpart
// in the reference; it replaces the coarse levels of the POC pyramid (poc.rs:694-766) by one
// network pass, and the level-0 POC grid checks what it proposes (engine.hip).
//
// Everything runs at 1/8 of the frame's resolution ("cells"): the two frames are box-averaged
// over 8 x 8 pixels and standardised, a fixed correlation volume compares them over
// +-CV_R cells (no weights), and the network maps that volume to the displacement in cells.
// Tensors are NHWC: element (b, y, x, c) of a batch of H x W maps with C channels sits at
// ((b H + y) W + x) C + c. Activations and weights are bfloat16 bit patterns (uint16_t),
// rounded to nearest even; sums are float32.
#include <cmath>
#include <cstring>
#include <random>

#include "cnn.hpp"

namespace ofx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite values)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
    const uint4 v = *(const uint4 *)p;
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// ------------------------------------------------------------------ frame to cells
// cell (i, j) = mean of the frame's pixels in rows 8i .. 8i + 7, columns 8j .. 8j + 7 that exist
__global__ void k_down8(const float *__restrict__ img, int w, int h, int w8, int h8, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= w8) return;
    const int x0 = 8 * j, y0 = 8 * i, x1 = min(x0 + 8, w), y1 = min(y0 + 8, h);
    float s = 0.f;
    for (int y = y0; y < y1; y++)
        for (int x = x0; x < x1; x++) s += img[(size_t)y * w + x];
    out[(size_t)i * w8 + j] = s / (float)((x1 - x0) * (y1 - y0));
}

// mean and 1 / standard deviation of each of the nb maps of n values, in two passes with a fixed
// reduction order: STAT_PARTS partial f64 sums per map, then one workgroup per map; a constant map
// gets 1 / sd = 0
constexpr int STAT_PARTS = 64;
__global__ void __launch_bounds__(256) k_stats_part(const float *__restrict__ a, int n, double *__restrict__ part) {
    __shared__ double s1[256], s2[256];
    const float *m = a + (size_t)blockIdx.y * n;
    double p = 0, q = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += STAT_PARTS * 256) {
        const double v = m[i];
        p += v;
        q += v * v;
    }
    s1[threadIdx.x] = p;
    s2[threadIdx.x] = q;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * ((size_t)blockIdx.y * STAT_PARTS + blockIdx.x)] = s1[0];
        part[2 * ((size_t)blockIdx.y * STAT_PARTS + blockIdx.x) + 1] = s2[0];
    }
}
__global__ void k_stats_final(const double *__restrict__ part, int n, float2 *__restrict__ out) {
    if (threadIdx.x) return;
    double p = 0, q = 0;
    for (int g = 0; g < STAT_PARTS; g++) {
        p += part[2 * ((size_t)blockIdx.x * STAT_PARTS + g)];
        q += part[2 * ((size_t)blockIdx.x * STAT_PARTS + g) + 1];
    }
    const double mean = p / n, var = fmax(q / n - mean * mean, 0.0);
    out[blockIdx.x] = make_float2((float)mean, var > 1e-20 ? (float)(1.0 / sqrt(var)) : 0.f);
}

// The correlation volume: channel (dy + R)(2R + 1) + (dx + R) of cell p is the mean over the 3 x 3
// neighbourhood q of n1(p + q) n2(p + q + d), n = (a - mean) / sd per map, zero outside the map;
// channels CV_CH .. CV_CPAD - 1 are 0. A 16 x 16 tile of cells per workgroup, its neighbourhoods
// staged in LDS.
constexpr int CT = 16;
__global__ void __launch_bounds__(256) k_costvol(const float *__restrict__ a1, const float *__restrict__ a2,
                                                 const float2 *__restrict__ st1, const float2 *__restrict__ st2,
                                                 int w, int h, uint16_t *__restrict__ out) {
    constexpr int S1 = CT + 2, S2 = CT + 2 + 2 * CV_R;
    __shared__ float t1[S1][S1], t2[S2][S2];
    const int b = blockIdx.z, tx0 = blockIdx.x * CT, ty0 = blockIdx.y * CT;
    const float *m1 = a1 + (size_t)b * w * h, *m2 = a2 + (size_t)b * w * h;
    const float2 q1 = st1[b], q2 = st2[b];
    for (int k = threadIdx.x; k < S1 * S1; k += 256) {
        const int yy = ty0 - 1 + k / S1, xx = tx0 - 1 + k % S1;
        t1[k / S1][k % S1] = (yy >= 0 && yy < h && xx >= 0 && xx < w) ? (m1[(size_t)yy * w + xx] - q1.x) * q1.y : 0.f;
    }
    for (int k = threadIdx.x; k < S2 * S2; k += 256) {
        const int yy = ty0 - 1 - CV_R + k / S2, xx = tx0 - 1 - CV_R + k % S2;
        t2[k / S2][k % S2] = (yy >= 0 && yy < h && xx >= 0 && xx < w) ? (m2[(size_t)yy * w + xx] - q2.x) * q2.y : 0.f;
    }
    __syncthreads();
    const int lx = threadIdx.x % CT, ly = threadIdx.x / CT, x = tx0 + lx, y = ty0 + ly;
    if (x >= w || y >= h) return;
    float c1[9];
    for (int q = 0; q < 9; q++) c1[q] = t1[ly + q / 3][lx + q % 3];
    uint16_t *o = out + (((size_t)b * h + y) * w + x) * CV_CPAD;
    alignas(16) uint16_t buf[8];
    int nb = 0;
    for (int c = 0; c < CV_CPAD; c++) {
        float v = 0.f;
        if (c < CV_CH) {
            const int dy = c / CV_D, dx = c % CV_D;  // offsets of the shifted window in t2
            for (int q = 0; q < 9; q++) v += c1[q] * t2[ly + q / 3 + dy][lx + q % 3 + dx];
            v *= (1.f / 9.f);
        }
        buf[nb++] = to_bf16(v);
        if (nb == 8) {
            *(uint4 *)(o + c - 7) = *(const uint4 *)buf;
            nb = 0;
        }
    }
}

// ------------------------------------------------------------------ 3 x 3 convolution on MFMA
// One wavefront computes 32 consecutive cells of a row (M) by 32 output channels (N) as an
// implicit GEMM over K = 9 taps x CIN input channels, 16 per v_mfma_f32_32x32x16_bf16:
//   A (32 cells x 16 channels): lane l holds cell l % 32, channels 8 (l / 32) .. + 7 of the tap;
//   B (16 channels x 32 outputs): lane l holds output l % 32, the same 8 channels;
//   D (32 x 32 f32): register r of lane l is cell 8 (r / 4) + 4 (l / 32) + r % 4, output l % 32.
// wt: [32][9][CIN] (rows past the layer's outputs are zero). With RELU the output is max(0, .);
// with a mask the output is multiplied by (mask > 0) (the derivative of the ReLU that made mask).
// Outputs n < NSTORE are stored with a stride of OSTRIDE channels, as bf16 or (OUTF32) float.
// The backward pass by data is the same kernel: the gradient w.r.t. the input of a layer is the
// convolution of the output gradient with the layer's weights flipped in space and transposed.
template <int CIN, int OSTRIDE, int NSTORE, bool RELU, bool MASK, bool OUTF32>
__global__ void __launch_bounds__(256) k_conv3(const uint16_t *__restrict__ in, const uint16_t *__restrict__ wt,
                                               const float *__restrict__ bias, const uint16_t *__restrict__ mask,
                                               void *__restrict__ outp, int nb, int h, int w) {
    static_assert(CIN % 16 == 0, "channels in blocks of 16");
    const int lane = threadIdx.x & 63;
    const int segs = (w + 31) / 32;
    const long task = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (task >= (long)nb * h * segs) return;
    const int seg = (int)(task % segs);
    const long row = task / segs;  // b h + y
    const int y = (int)(row % h);
    const int b = (int)(row / h);
    const int m = lane & 31, half = lane >> 5;
    const int x = seg * 32 + m;
    f32x16 acc;
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    const uint16_t *wl = wt + (size_t)m * 9 * CIN + 8 * half;  // output channel m of the B operand
    for (int t = 0; t < 9; t++) {
        const int ys = y + t / 3 - 1, xs = x + t % 3 - 1;
        const bool ok = ys >= 0 && ys < h && xs >= 0 && xs < w;
        const uint16_t *src = in + (((size_t)b * h + (ok ? ys : 0)) * w + (ok ? xs : 0)) * CIN + 8 * half;
#pragma unroll
        for (int cb = 0; cb < CIN / 16; cb++) {
            const bf16x8 av = ok ? load8(src + 16 * cb) : zero8();
            const bf16x8 bv = load8(wl + t * CIN + 16 * cb);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
        }
    }
    const int n = m;  // output channel of this lane's registers
    if (n >= NSTORE) return;
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int cell = 8 * (r >> 2) + 4 * half + (r & 3);
        const int xc = seg * 32 + cell;
        if (xc >= w) continue;
        const size_t p = ((size_t)b * h + y) * w + xc;
        float v = acc[r] + bn;
        if (RELU) v = fmaxf(v, 0.f);
        if (MASK) v = from_bf16(mask[p * NET_CH + n]) > 0.f ? v : 0.f;
        if (OUTF32)
            ((float *)outp)[p * OSTRIDE + n] = v;
        else
            ((uint16_t *)outp)[p * OSTRIDE + n] = to_bf16(v);
    }
}

// Weight gradient of a layer, one partial sum per group of RG rows (b h + y) of the batch:
// part[t][cb][g][co][ci] = sum over the group's cells p of dz[p][co] x[p + tap t][32 cb + ci], as a
// GEMM with M = co (32), N = ci (32), K = cells (16 per MFMA; a cell's shifted input outside the
// map is 0). DZC: channels of dz (co >= DZC read as 0); CIN: channels of x (ci >= CIN read as 0).
constexpr int WG_ROWS = 8;
template <int DZC, int CIN>
__global__ void __launch_bounds__(64) k_wgrad(const uint16_t *__restrict__ dz, const uint16_t *__restrict__ x,
                                              int nb, int h, int w, float *__restrict__ part) {
    const int lane = threadIdx.x, m = lane & 31, half = lane >> 5;
    const int g = blockIdx.x, t = blockIdx.y, cb = blockIdx.z, ng = gridDim.x;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int ci = 32 * cb + m;
    f32x16 acc;
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    const long rows = (long)nb * h;
    for (long row = (long)g * WG_ROWS; row < min(rows, (long)(g + 1) * WG_ROWS); row++) {
        const int y = (int)(row % h);
        const int ys = y + dy;
        const bool yok = ys >= 0 && ys < h;
        for (int x0 = 0; x0 < w; x0 += 16) {
            alignas(16) uint16_t ha[8], hb[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int xk = x0 + 8 * half + i;
                const bool in = xk < w;
                const size_t p = (size_t)row * w + xk;
                ha[i] = (in && m < DZC) ? dz[p * DZC + m] : (uint16_t)0;
                const int xs = xk + dx;
                const bool ok = in && yok && xs >= 0 && xs < w && ci < CIN;
                hb[i] = ok ? x[((size_t)(row + dy) * w + xs) * CIN + ci] : (uint16_t)0;
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, *(const uint4 *)ha),
                                                          __builtin_bit_cast(bf16x8, *(const uint4 *)hb), acc, 0, 0,
                                                          0);
        }
    }
    float *o = part + (((size_t)t * gridDim.z + cb) * ng + g) * 1024;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int co = 8 * (r >> 2) + 4 * half + (r & 3);
        o[co * 32 + m] = acc[r];
    }
}

// dW[co][t][ci] (padded layout [32][9][cin_pad]) = the partials summed over the groups in order
__global__ void k_wgrad_reduce(const float *__restrict__ part, int ng, int ncb, int cin_pad, float *__restrict__ dw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 32 * 9 * cin_pad) return;
    const int co = k / (9 * cin_pad), t = (k / cin_pad) % 9, ci = k % cin_pad;
    const int cb = ci / 32;
    if (cb >= ncb) {
        dw[k] = 0.f;
        return;
    }
    const float *p = part + ((size_t)t * ncb + cb) * ng * 1024 + co * 32 + (ci % 32);
    float s = 0.f;
    for (int g = 0; g < ng; g++) s += p[(size_t)g * 1024];
    dw[k] = s;
}

// bias gradient partials: bpart[g][c] = sum over the group's cells of dz[p][c] (c < C)
template <int C>
__global__ void k_bgrad(const uint16_t *__restrict__ dz, long rows, int w, float *__restrict__ bpart) {
    const int c = threadIdx.x, g = blockIdx.x;
    if (c >= 32) return;
    float s = 0.f;
    if (c < C)
        for (long row = (long)g * WG_ROWS; row < min(rows, (long)(g + 1) * WG_ROWS); row++)
            for (int x = 0; x < w; x++) s += from_bf16(dz[((size_t)row * w + x) * C + c]);
    bpart[(size_t)g * 32 + c] = s;
}
__global__ void k_bgrad_reduce(const float *__restrict__ bpart, int ng, float *__restrict__ db) {
    const int c = threadIdx.x;
    if (c >= 32) return;
    float s = 0.f;
    for (int g = 0; g < ng; g++) s += bpart[(size_t)g * 32 + c];
    db[c] = s;
}

// ------------------------------------------------------------------ loss
// Charbonnier end-point error in cells, sqrt(ex^2 + ey^2 + eps^2) - eps, averaged over the n cells;
// its gradient w.r.t. the prediction into dout (bf16, OUT_CPAD channels, the 2 used ones first),
// the per-block partial sums of the loss into lpart (f64).
constexpr float LOSS_EPS = 0.01f;
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ pred, const float *__restrict__ target, long n,
                                              uint16_t *__restrict__ dout, double *__restrict__ lpart) {
    __shared__ double red[256];
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0;
    if (p < n) {
        const float ex = pred[2 * p] - target[2 * p], ey = pred[2 * p + 1] - target[2 * p + 1];
        const float r = sqrtf(ex * ex + ey * ey + LOSS_EPS * LOSS_EPS);
        l = (double)r - LOSS_EPS;
        const float k = 1.f / (r * (float)n);
        uint16_t *o = dout + p * OUT_CPAD;
        o[0] = to_bf16(ex * k);
        o[1] = to_bf16(ey * k);
        for (int c = 2; c < OUT_CPAD; c++) o[c] = 0;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}
__global__ void k_loss_reduce(const double *__restrict__ lpart, int nblk, long n, double *__restrict__ out) {
    if (threadIdx.x || blockIdx.x) return;
    double s = 0;
    for (int i = 0; i < nblk; i++) s += lpart[i];
    *out = s / (double)n;
}

// ------------------------------------------------------------------ parameters
// The master parameters of a layer are float32 in the padded layout: weights [32][9][cin_pad]
// then bias [32]. The bf16 copies the kernels read: forward [32][9][cin_pad]; backward by data
// (layers after the first) [32 (= the layer's input channels)][9][bw_pad], wb[ci][8 - t][co] =
// w[co][t][ci] for co < bw_pad.
__global__ void k_cast_layer(const float *__restrict__ wm, int cin_pad, int bw_pad, uint16_t *__restrict__ wf,
                             uint16_t *__restrict__ wb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 32 * 9 * cin_pad) return;
    const int co = k / (9 * cin_pad), t = (k / cin_pad) % 9, ci = k % cin_pad;
    const uint16_t v = to_bf16(wm[k]);
    wf[k] = v;
    if (wb && co < bw_pad && ci < 32) wb[((size_t)ci * 9 + (8 - t)) * bw_pad + co] = v;
}

// Adam (Kingma & Ba 2015) on n parameters with the bias-corrected step lr_t; decoupled from
// nothing else (no weight decay)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, long n, float lr_t, float b1, float b2, float eps) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk, v = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = v;
    p[k] -= lr_t * a / (sqrtf(v) + eps);
}

// ------------------------------------------------------------------ training pairs
// The renderer of the training pairs on the GPU, the steps of synth.render: frame 1 samples the
// speckle pattern at the pixel, frame 2 at the inverse-mapped position X of x = X + u(X) (fixed
// point), bilinearly; both get Gaussian noise (a counter-based generator, Box-Muller) and are
// quantised to 8 bits. Each thread renders the 64 pixels of one cell and writes their means; the
// target is u at the cell's centre, in cells.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float gauss(uint32_t seed, uint32_t a, uint32_t b) {
    const uint32_t h1 = mix32(seed ^ mix32(a * 2654435761u ^ mix32(b + 0x9e3779b9u)));
    const uint32_t h2 = mix32(h1 + 0x632be5abu);
    const float u1 = ((h1 >> 8) + 1) * (1.f / 16777217.f), u2 = (h2 >> 8) * (1.f / 16777216.f);
    return sqrtf(-2.f * logf(u1)) * cosf(6.28318530718f * u2);
}
__device__ __forceinline__ void field_at(const TrainSample &s, float X, float Y, float &ux, float &uy) {
    ux = uy = 0.f;
    for (int k = 0; k < TRAIN_BUMPS; k++) {
        const float dx = X - s.cx[k], dy = Y - s.cy[k];
        const float g = __expf(-(dx * dx + dy * dy) * s.inv2s2[k]);
        ux += s.ax[k] * g;
        uy += s.ay[k] * g;
    }
}
__device__ __forceinline__ float pattern_at(const float *pat, int pw, int ph, float x, float y) {
    x = fminf(fmaxf(x, 0.f), pw - 1.000001f);
    y = fminf(fmaxf(y, 0.f), ph - 1.000001f);
    const int x0 = (int)x, y0 = (int)y;
    const float tx = x - x0, ty = y - y0;
    const float *r0 = pat + (size_t)y0 * pw + x0, *r1 = r0 + pw;
    return (1 - ty) * ((1 - tx) * r0[0] + tx * r0[1]) + ty * ((1 - tx) * r1[0] + tx * r1[1]);
}
__device__ __forceinline__ float quant8(float v) { return fminf(fmaxf(rintf(v * 255.f), 0.f), 255.f) * (1.f / 255.f); }

__global__ void k_train_pairs(const float *__restrict__ pats, int pw, int ph, const TrainSample *__restrict__ smp,
                              int w8, int h8, float noise, float *__restrict__ a1, float *__restrict__ a2,
                              float *__restrict__ target) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, b = blockIdx.z;
    if (j >= w8) return;
    const TrainSample s = smp[b];
    const float *pat = pats + (size_t)s.pattern * pw * ph;
    float s1 = 0.f, s2 = 0.f;
    for (int py = 0; py < 8; py++)
        for (int px = 0; px < 8; px++) {
            const int x = 8 * j + px, y = 8 * i + py;
            const uint32_t pix = (uint32_t)(y * 8 * w8 + x);
            const float v1 = pat[(size_t)(s.oy + y) * pw + (s.ox + x)] + noise * gauss(s.seed, pix, 1u);
            float X = (float)x, Y = (float)y, ux, uy;
            for (int it = 0; it < 20; it++) {
                field_at(s, X, Y, ux, uy);
                X = x - ux;
                Y = y - uy;
            }
            const float v2 = pattern_at(pat, pw, ph, s.ox + X, s.oy + Y) + noise * gauss(s.seed, pix, 2u);
            s1 += quant8(v1);
            s2 += quant8(v2);
        }
    const size_t c = ((size_t)b * h8 + i) * w8 + j;
    a1[c] = s1 * (1.f / 64.f);
    a2[c] = s2 * (1.f / 64.f);
    float ux, uy;
    field_at(s, 8.f * j + 3.5f, 8.f * i + 3.5f, ux, uy);
    target[2 * c] = ux * 0.125f;
    target[2 * c + 1] = uy * 0.125f;
}

// the proposal (cells, at cell centres 8 j + 3.5) interpolated bilinearly at the level-0 POC grid,
// in pixels, as the grid's starting guess
__global__ void k_prop_to_grid(const float *__restrict__ pred, int w8, int h8, int x0, int y0, int step, int nx,
                               int ny, double *__restrict__ ix, double *__restrict__ iy) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nx * ny) return;
    const int gi = k / nx, gj = k - gi * nx;
    const float fx = fminf(fmaxf((x0 + gj * step - 3.5f) * 0.125f, 0.f), (float)(w8 - 1));
    const float fy = fminf(fmaxf((y0 + gi * step - 3.5f) * 0.125f, 0.f), (float)(h8 - 1));
    const int j0 = min((int)fx, max(w8 - 2, 0)), i0 = min((int)fy, max(h8 - 2, 0));
    const int j1 = min(j0 + 1, w8 - 1), i1 = min(i0 + 1, h8 - 1);
    const float tx = fx - j0, ty = fy - i0;
    for (int c = 0; c < 2; c++) {
        const float v = (1 - ty) * ((1 - tx) * pred[2 * ((size_t)i0 * w8 + j0) + c] + tx * pred[2 * ((size_t)i0 * w8 + j1) + c]) +
                        ty * ((1 - tx) * pred[2 * ((size_t)i1 * w8 + j0) + c] + tx * pred[2 * ((size_t)i1 * w8 + j1) + c]);
        (c ? iy : ix)[k] = 8.0 * (double)v;
    }
}

}  // namespace

// ------------------------------------------------------------------ host side
const LayerShape NET_LAYERS[NET_NLAYERS] = {
    {CV_CH, CV_CPAD, NET_CH, true}, {NET_CH, NET_CH, NET_CH, true}, {NET_CH, NET_CH, NET_CH, true},
    {NET_CH, NET_CH, NET_CH, true}, {NET_CH, NET_CH, NET_CH, true}, {NET_CH, NET_CH, 2, false}};

size_t net_param_count() {
    size_t n = 0;
    for (const auto &L : NET_LAYERS) n += (size_t)L.cout * 9 * L.cin + L.cout;
    return n;
}
size_t layer_master_floats(int l) { return (size_t)32 * 9 * NET_LAYERS[l].cin_pad + 32; }
size_t net_master_floats() {
    size_t n = 0;
    for (int l = 0; l < NET_NLAYERS; l++) n += layer_master_floats(l);
    return n;
}

// flat (the weights [cout][9][cin] then the bias of each layer, unpadded) <-> padded master
void net_flat_to_master(const float *flat, std::vector<float> &master) {
    master.assign(net_master_floats(), 0.f);
    size_t f = 0, o = 0;
    for (int l = 0; l < NET_NLAYERS; l++) {
        const LayerShape &L = NET_LAYERS[l];
        for (int co = 0; co < L.cout; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < L.cin; ci++) master[o + ((size_t)co * 9 + t) * L.cin_pad + ci] = flat[f++];
        for (int co = 0; co < L.cout; co++) master[o + (size_t)32 * 9 * L.cin_pad + co] = flat[f++];
        o += layer_master_floats(l);
    }
}
void net_master_to_flat(const std::vector<float> &master, float *flat) {
    size_t f = 0, o = 0;
    for (int l = 0; l < NET_NLAYERS; l++) {
        const LayerShape &L = NET_LAYERS[l];
        for (int co = 0; co < L.cout; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < L.cin; ci++) flat[f++] = master[o + ((size_t)co * 9 + t) * L.cin_pad + ci];
        for (int co = 0; co < L.cout; co++) flat[f++] = master[o + (size_t)32 * 9 * L.cin_pad + co];
        o += layer_master_floats(l);
    }
}

CoarseNet::~CoarseNet() {
    for (void *p : {(void *)master, (void *)wf, (void *)wb}) (void)hipFree(p);
}

void CoarseNet::upload(const std::vector<float> &m, hipStream_t s) {
    if (!master) {
        HIPCHECK(hipMalloc(&master, net_master_floats() * sizeof(float)));
        size_t nf = 0, nb = 0;
        for (int l = 0; l < NET_NLAYERS; l++) {
            wf_off[l] = nf;
            nf += (size_t)32 * 9 * NET_LAYERS[l].cin_pad;
            wb_off[l] = nb;
            if (l > 0) nb += (size_t)32 * 9 * bw_pad(l);
        }
        HIPCHECK(hipMalloc(&wf, nf * sizeof(uint16_t)));
        HIPCHECK(hipMalloc(&wb, nb * sizeof(uint16_t)));
        size_t o = 0;
        for (int l = 0; l < NET_NLAYERS; l++) {
            m_off[l] = o;
            o += layer_master_floats(l);
        }
    }
    PinnedBuf pin;
    copy_h2d(pin, master, m.data(), m.size() * sizeof(float), s);
    cast(s);
    wait_stream(s);
}

int CoarseNet::bw_pad(int l) { return l == NET_NLAYERS - 1 ? OUT_CPAD : NET_CH; }

void CoarseNet::cast(hipStream_t s) {
    for (int l = 0; l < NET_NLAYERS; l++) {
        const int cp = NET_LAYERS[l].cin_pad, nw = 32 * 9 * cp;
        hipLaunchKernelGGL(k_cast_layer, dim3((nw + 255) / 256), dim3(256), 0, s, (const float *)(master + m_off[l]), cp,
                           bw_pad(l), wf + wf_off[l], l > 0 ? wb + wb_off[l] : (uint16_t *)nullptr);
    }
    LAUNCHCHECK();
}

const float *CoarseNet::bias(int l) const { return master + m_off[l] + (size_t)32 * 9 * NET_LAYERS[l].cin_pad; }

// activations of one forward pass over nb maps of h8 x w8 cells
void NetActs::ensure(int nb_, int h8_, int w8_) {
    nb = nb_;
    h8 = h8_;
    w8 = w8_;
    const size_t n = (size_t)nb * h8 * w8;
    a1 = a1b.get<float>(n);
    a2 = a2b.get<float>(n);
    st = stb.get<float2>(2 * (size_t)nb);
    stp = stpb.get<double>(2 * (size_t)STAT_PARTS * nb);
    cv = cvb.get<uint16_t>(n * CV_CPAD);
    for (int l = 0; l < NET_NLAYERS - 1; l++) hid[l] = hb[l].get<uint16_t>(n * NET_CH);
    pred = predb.get<float>(2 * n);
}

void net_features(NetActs &A, hipStream_t s) {
    const int n = A.h8 * A.w8;
    hipLaunchKernelGGL(k_stats_part, dim3(STAT_PARTS, A.nb), dim3(256), 0, s, (const float *)A.a1, n, A.stp);
    hipLaunchKernelGGL(k_stats_final, dim3(A.nb), dim3(64), 0, s, (const double *)A.stp, n, A.st);
    hipLaunchKernelGGL(k_stats_part, dim3(STAT_PARTS, A.nb), dim3(256), 0, s, (const float *)A.a2, n, A.stp);
    hipLaunchKernelGGL(k_stats_final, dim3(A.nb), dim3(64), 0, s, (const double *)A.stp, n, A.st + A.nb);
    hipLaunchKernelGGL(k_costvol, dim3((A.w8 + CT - 1) / CT, (A.h8 + CT - 1) / CT, A.nb), dim3(256), 0, s,
                       (const float *)A.a1, (const float *)A.a2, (const float2 *)A.st, (const float2 *)(A.st + A.nb),
                       A.w8, A.h8, A.cv);
    LAUNCHCHECK();
}

static unsigned conv_blocks(int nb, int h, int w) {
    const long tasks = (long)nb * h * ((w + 31) / 32);
    return (unsigned)((tasks + 3) / 4);
}

void net_forward(const CoarseNet &net, NetActs &A, hipStream_t s) {
    const dim3 g(conv_blocks(A.nb, A.h8, A.w8)), b(256);
    hipLaunchKernelGGL((k_conv3<CV_CPAD, NET_CH, NET_CH, true, false, false>), g, b, 0, s, (const uint16_t *)A.cv,
                       (const uint16_t *)(net.wf + net.wf_off[0]), net.bias(0), (const uint16_t *)nullptr,
                       (void *)A.hid[0], A.nb, A.h8, A.w8);
    for (int l = 1; l < NET_NLAYERS - 1; l++)
        hipLaunchKernelGGL((k_conv3<NET_CH, NET_CH, NET_CH, true, false, false>), g, b, 0, s,
                           (const uint16_t *)A.hid[l - 1], (const uint16_t *)(net.wf + net.wf_off[l]), net.bias(l),
                           (const uint16_t *)nullptr, (void *)A.hid[l], A.nb, A.h8, A.w8);
    hipLaunchKernelGGL((k_conv3<NET_CH, 2, 2, false, false, true>), g, b, 0, s,
                       (const uint16_t *)A.hid[NET_NLAYERS - 2],
                       (const uint16_t *)(net.wf + net.wf_off[NET_NLAYERS - 1]), net.bias(NET_NLAYERS - 1),
                       (const uint16_t *)nullptr, (void *)A.pred, A.nb, A.h8, A.w8);
    LAUNCHCHECK();
}

void net_propose(const CoarseNet &net, NetActs &A, const float *i1, const float *i2, int w, int h, hipStream_t s) {
    const int w8 = (w + 7) / 8, h8 = (h + 7) / 8;
    A.ensure(1, h8, w8);
    hipLaunchKernelGGL(k_down8, dim3((w8 + 255) / 256, h8), dim3(256), 0, s, i1, w, h, w8, h8, A.a1);
    hipLaunchKernelGGL(k_down8, dim3((w8 + 255) / 256, h8), dim3(256), 0, s, i2, w, h, w8, h8, A.a2);
    LAUNCHCHECK();
    net_features(A, s);
    net_forward(net, A, s);
}

void prop_to_grid(const NetActs &A, int x0, int y0, int step, int nx, int ny, double *ix, double *iy, hipStream_t s) {
    hipLaunchKernelGGL(k_prop_to_grid, dim3((nx * ny + 255) / 256), dim3(256), 0, s, (const float *)A.pred, A.w8, A.h8,
                       x0, y0, step, nx, ny, ix, iy);
    LAUNCHCHECK();
}

// ------------------------------------------------------------------ training state
void NetTrainer::ensure(int nb, int h8, int w8) {
    A.ensure(nb, h8, w8);
    const size_t n = (size_t)nb * h8 * w8;
    const size_t nm = net_master_floats();
    grad = gradb.get<float>(nm);
    m1 = m1b.get<float>(nm);
    m2 = m2b.get<float>(nm);
    target = targetb.get<float>(2 * n);
    dout = doutb.get<uint16_t>(n * OUT_CPAD);
    dz[0] = dzb[0].get<uint16_t>(n * NET_CH);
    dz[1] = dzb[1].get<uint16_t>(n * NET_CH);
    ng = (int)((nb * (long)h8 + WG_ROWS - 1) / WG_ROWS);
    part = partb.get<float>((size_t)9 * ((CV_CPAD + 31) / 32) * ng * 1024);
    bpart = bpartb.get<float>((size_t)ng * 32);
    lblk = (int)((n + 255) / 256);
    lpart = lpartb.get<double>(lblk);
}

void NetTrainer::zero_moments(hipStream_t s) {
    const size_t nm = net_master_floats();
    std::vector<float> z(nm, 0.f);
    copy_h2d(pin, m1, z.data(), nm * sizeof(float), s);
    copy_h2d(pin, m2, z.data(), nm * sizeof(float), s);
}

// loss of the current batch (A.a1, A.a2, target filled) into *loss_dev; the gradient of every master
// parameter into grad
void NetTrainer::forward_backward(const CoarseNet &net, double *loss_dev, hipStream_t s) {
    net_features(A, s);
    net_forward(net, A, s);
    const long n = (long)A.nb * A.h8 * A.w8;
    hipLaunchKernelGGL(k_loss, dim3(lblk), dim3(256), 0, s, (const float *)A.pred, (const float *)target, n, dout, lpart);
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(64), 0, s, (const double *)lpart, lblk, n, loss_dev);
    LAUNCHCHECK();
    const dim3 g(conv_blocks(A.nb, A.h8, A.w8)), b(256);
    const long rows = (long)A.nb * A.h8;
    // layer l's output gradient: dout for the last, dz[cur] for the others (already masked)
    const uint16_t *dy = dout;
    int cur = 0;
    for (int l = NET_NLAYERS - 1; l >= 0; l--) {
        const LayerShape &L = NET_LAYERS[l];
        const uint16_t *x = l == 0 ? A.cv : A.hid[l - 1];
        float *gw = grad + net.m_off[l], *gb = gw + (size_t)32 * 9 * L.cin_pad;
        const int ncb = (L.cin_pad + 31) / 32;
        const dim3 wg(ng, 9, ncb);
        if (l == NET_NLAYERS - 1) {
            hipLaunchKernelGGL((k_wgrad<OUT_CPAD, NET_CH>), wg, dim3(64), 0, s, dy, x, A.nb, A.h8, A.w8, part);
            hipLaunchKernelGGL(k_bgrad<OUT_CPAD>, dim3(ng), dim3(64), 0, s, dy, rows, A.w8, bpart);
        } else if (l == 0) {
            hipLaunchKernelGGL((k_wgrad<NET_CH, CV_CPAD>), wg, dim3(64), 0, s, dy, x, A.nb, A.h8, A.w8, part);
            hipLaunchKernelGGL(k_bgrad<NET_CH>, dim3(ng), dim3(64), 0, s, dy, rows, A.w8, bpart);
        } else {
            hipLaunchKernelGGL((k_wgrad<NET_CH, NET_CH>), wg, dim3(64), 0, s, dy, x, A.nb, A.h8, A.w8, part);
            hipLaunchKernelGGL(k_bgrad<NET_CH>, dim3(ng), dim3(64), 0, s, dy, rows, A.w8, bpart);
        }
        const int nw = 32 * 9 * L.cin_pad;
        hipLaunchKernelGGL(k_wgrad_reduce, dim3((nw + 255) / 256), dim3(256), 0, s, (const float *)part, ng, ncb,
                           L.cin_pad, gw);
        hipLaunchKernelGGL(k_bgrad_reduce, dim3(1), dim3(64), 0, s, (const float *)bpart, ng, gb);
        LAUNCHCHECK();
        if (l == 0) break;
        // the gradient w.r.t. this layer's input, times the derivative of the ReLU that made it
        uint16_t *dx = dz[cur];
        if (l == NET_NLAYERS - 1)
            hipLaunchKernelGGL((k_conv3<OUT_CPAD, NET_CH, NET_CH, false, true, false>), g, b, 0, s, dy,
                               (const uint16_t *)(net.wb + net.wb_off[l]), (const float *)nullptr,
                               (const uint16_t *)A.hid[l - 1], (void *)dx, A.nb, A.h8, A.w8);
        else
            hipLaunchKernelGGL((k_conv3<NET_CH, NET_CH, NET_CH, false, true, false>), g, b, 0, s, dy,
                               (const uint16_t *)(net.wb + net.wb_off[l]), (const float *)nullptr,
                               (const uint16_t *)A.hid[l - 1], (void *)dx, A.nb, A.h8, A.w8);
        LAUNCHCHECK();
        dy = dx;
        cur ^= 1;
    }
}

void NetTrainer::adam_step(CoarseNet &net, int step, double lr, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const double c1 = 1.0 - std::pow((double)b1, step), c2 = 1.0 - std::pow((double)b2, step);
    const float lr_t = (float)(lr * std::sqrt(c2) / c1);
    const long n = (long)net_master_floats();
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, net.master, (const float *)grad, m1,
                       m2, n, lr_t, b1, b2, 1e-8f);
    LAUNCHCHECK();
    net.cast(s);
}

void render_train_pairs(const float *pats, int pw, int ph, const TrainSample *smp, int nb, int h8, int w8, float noise,
                        float *a1, float *a2, float *target, hipStream_t s) {
    hipLaunchKernelGGL(k_train_pairs, dim3((w8 + 63) / 64, h8, nb), dim3(64), 0, s, pats, pw, ph, smp, w8, h8, noise,
                       a1, a2, target);
    LAUNCHCHECK();
}

// Training samples: a crop at a uniform place of a uniformly chosen pattern, inside a virtual frame
// of 4000 x 3000 px with TRAIN_BUMPS Gaussian bumps (centres uniform over that frame, scales
// uniform in [150, 600] px, directions uniform, magnitudes uniform in [0.3, 1]) rescaled so that
// the largest |u| over the frame (on a 64 px lattice) is uniform in [max_disp / 2, max_disp] (a
// frame's bumps then hold displacements from 0 up to that largest one, the large ones well
// represented).
void make_train_samples(uint64_t seed, int first, int count, int n_pat, int pw, int ph, int crop, double max_disp,
                        std::vector<TrainSample> &out) {
    out.resize(count);
    for (int k = 0; k < count; k++) {
        std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + (uint64_t)(first + k));
        std::uniform_real_distribution<double> U(0.0, 1.0);
        TrainSample &s = out[k];
        s.pattern = (int)(U(rng) * n_pat) % n_pat;
        const int lo = 64, hx = pw - crop - 64, hy = ph - crop - 64;
        s.ox = lo + (int)(U(rng) * (hx - lo + 1));
        s.oy = lo + (int)(U(rng) * (hy - lo + 1));
        const double VW = 4000, VH = 3000;
        const double fx = U(rng) * (VW - crop), fy = U(rng) * (VH - crop);  // the crop inside the frame
        double cx[TRAIN_BUMPS], cy[TRAIN_BUMPS], sg[TRAIN_BUMPS], ax[TRAIN_BUMPS], ay[TRAIN_BUMPS];
        for (int b = 0; b < TRAIN_BUMPS; b++) {
            cx[b] = U(rng) * VW;
            cy[b] = U(rng) * VH;
            sg[b] = 150.0 + 450.0 * U(rng);
            const double ang = 2 * M_PI * U(rng), mag = 0.3 + 0.7 * U(rng);
            ax[b] = mag * std::cos(ang);
            ay[b] = mag * std::sin(ang);
        }
        double mx = 0;
        for (double Y = 0; Y < VH; Y += 64)
            for (double X = 0; X < VW; X += 64) {
                double ux = 0, uy = 0;
                for (int b = 0; b < TRAIN_BUMPS; b++) {
                    const double g = std::exp(-((X - cx[b]) * (X - cx[b]) + (Y - cy[b]) * (Y - cy[b])) /
                                              (2 * sg[b] * sg[b]));
                    ux += ax[b] * g;
                    uy += ay[b] * g;
                }
                mx = std::max(mx, std::sqrt(ux * ux + uy * uy));
            }
        const double amp = (0.5 + 0.5 * U(rng)) * max_disp, scale = mx > 0 ? amp / mx : 0.0;
        for (int b = 0; b < TRAIN_BUMPS; b++) {
            s.cx[b] = (float)(cx[b] - fx);
            s.cy[b] = (float)(cy[b] - fy);
            s.inv2s2[b] = (float)(1.0 / (2 * sg[b] * sg[b]));
            s.ax[b] = (float)(ax[b] * scale);
            s.ay[b] = (float)(ay[b] * scale);
        }
        s.seed = (uint32_t)(rng() & 0xFFFFFFFFu);
    }
}

}  // namespace ofx
