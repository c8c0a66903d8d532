// Coarse stage on the GPU: windowed, band-limited phase-only correlation on a grid over a Gaussian
// pyramid, outlier cleaning and fill, and the dense start field (python/optflow_ref/poc.py:18-239,
// rust/crates/optflow-core/src/poc.rs:141-766, track.py:131-137). One workgroup per grid point
// holds the two windows' spectra in LDS (float32) and iterates the re-centred match; the cleaning
// passes run in one workgroup over the whole grid.
#include <cmath>
#include <cstring>
#include <mutex>

#include "fine.hpp"
#include "lines.hpp"

namespace ofx {

constexpr int POC_MAX = 4096;  // window elements a workgroup holds (64 x 64, 16 x 256, ...)

struct CoarseWs {
    int w = 0, h = 0, L = 0;
    std::vector<float *> pa, pb;   // pyramid levels 1..L (level 0 = the inputs)
    std::vector<size_t> pcap;      // capacity of pa[l], pb[l] in floats
    std::vector<int> lw, lh;
    double *tmp = nullptr;         // blur scratch
    double *gux = nullptr, *guy = nullptr, *gpk = nullptr, *ix = nullptr, *iy = nullptr;
    int *gflag = nullptr;          // per point: valid, out flags
    int *gwork = nullptr;          // clean: two lists of points to fill, then a counter
    float *hann = nullptr;         // Hann tables of the current level (wy + wx floats)
    size_t gcap = 0;
    size_t tmpcap = 0;
};

namespace {

int pow2_floor(int n) {
    if (n < 1) return 0;
    int p = 1;
    while (p * 2 <= n) p *= 2;
    return p;
}

// poc.py:172-187
bool level_window(int h, int w, int win, int &wy, int &wx) {
    const int lim_h = pow2_floor(2 * h / 3), lim_w = pow2_floor(2 * w / 3);
    if (win <= lim_h && win <= lim_w) {
        wy = wx = win;
        return true;
    }
    if (h <= w) {
        wy = std::min(win, lim_h);
        wx = std::min(lim_w, std::max(win, win * win / std::max(wy, 1)));
    } else {
        wx = std::min(win, lim_w);
        wy = std::min(lim_h, std::max(win, win * win / std::max(wx, 1)));
    }
    return std::min(wy, wx) >= 16;
}

int n_levels(double max_disp, int win, int h, int w) {
    const double v = std::ceil(std::log2(std::max(max_disp, 1e-9) / (win / 4.0)));
    int L = (int)std::max(0.0, v);
    int wy, wx;
    while (L > 0 && !level_window(h >> L, w >> L, win, wy, wx)) L--;
    return L;
}

}  // namespace

int pyramid_levels(double max_disp, int win, int h, int w) { return n_levels(max_disp, win, h, w); }
bool poc_window_at(int h, int w, int win, int &wy, int &wx) { return level_window(h, w, win, wy, wx); }

namespace {

// rust poc.rs:244-256: (start, count) of the grid positions along an axis of length n
void grid_axis(int n, int win, int step, int &start, int &count) {
    const int stop = n + 1 - (win - win / 2);
    start = win / 2;
    count = start >= stop ? 0 : (stop - start - 1) / step + 1;
}

// ------------------------------------------------------------------ pyramid (poc.py:164-165)
__global__ void k_down_y(const float *__restrict__ in, double *__restrict__ out, int w, int h, int oh) {
    // blur along y (sigma 1, 9 taps, replicated edges) at the even rows only
    const int x = blockIdx.x * blockDim.x + threadIdx.x, yo = blockIdx.y;
    if (x >= w || yo >= oh) return;
    const int y = 2 * yo;
    double k[9], s = 0;
    for (int j = -4; j <= 4; j++) {
        k[j + 4] = exp(-0.5 * j * j);
        s += k[j + 4];
    }
    double acc = 0;
    for (int j = -4; j <= 4; j++) acc += (k[j + 4] / s) * (double)in[(size_t)min(max(y + j, 0), h - 1) * w + x];
    out[(size_t)yo * w + x] = acc;
}

__global__ void k_down_x(const double *__restrict__ in, float *__restrict__ out, int w, int ow, int oh) {
    const int xo = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (xo >= ow || y >= oh) return;
    const int x = 2 * xo;
    double k[9], s = 0;
    for (int j = -4; j <= 4; j++) {
        k[j + 4] = exp(-0.5 * j * j);
        s += k[j + 4];
    }
    double acc = 0;
    const double *row = in + (size_t)y * w;
    for (int j = -4; j <= 4; j++) acc += (k[j + 4] / s) * row[min(max(x + j, 0), w - 1)];
    out[(size_t)y * ow + xo] = (float)acc;
}

// ------------------------------------------------------------------ POC per grid point
struct PocArgs {
    int W, H;        // level image size
    int wy, wx;      // window
    int x0, y0, step, nx, ny;  // grid positions x0 + j step, y0 + i step
    int nfx, nfy;    // FFT factorisations
    int facx[MAX_FACTORS], facy[MAX_FACTORS];
    int have_init;
};

// 2-D forward DFT of X (wy x wx, row-major) in LDS; result in column-major order ([kx][ky]) in
// the returned buffer, Y is scratch.
__device__ float2 *fft2_lds(float2 *X, float2 *Y, int wy, int wx, const PocArgs &P, const float2 *twx,
                            const float2 *twy) {
    float2 *r = fft_lines_lds(X, Y, wx, wy, P.facx, P.nfx, twx, threadIdx.x, blockDim.x);
    float2 *o = (r == X) ? Y : X;
    for (int i = threadIdx.x; i < wy * wx; i += blockDim.x) {
        const int y = i / wx, x = i - y * wx;
        o[x * wy + y] = r[i];
    }
    __syncthreads();
    return fft_lines_lds(o, r, wy, wx, P.facy, P.nfy, twy, threadIdx.x, blockDim.x);
}

// ---------------------------------------------------------------- 64 x 64 windows
// Element (y, x) of a 64 x 64 LDS tile lives at y * 65 + x: rows handed to consecutive lanes
// (row pass) and columns handed to consecutive lanes (column pass) both spread over the banks.
constexpr int T64 = 65;

// Radix-8 Stockham pass (p = P) of the 64 x 64 tile along its rows (COLS false) or columns, in
// place, for a workgroup of exactly 512 threads: butterfly item -> line item % 64, index item / 64.
// Every thread reads its butterfly's 8 inputs, the workgroup waits, every thread writes its 8
// outputs over the same tile, and the workgroup waits again.
template <int P, bool COLS>
__device__ __forceinline__ void pass64_inplace(float2 *buf, const float2 *tw) {
    const int item = threadIdx.x;
    const int line = item & 63, i = item >> 6;
    const int k = i % P;
    float2 x[8];
#pragma unroll
    for (int q = 0; q < 8; q++) x[q] = COLS ? buf[(i + 8 * q) * T64 + line] : buf[line * T64 + i + 8 * q];
    if constexpr (P > 1) {
#pragma unroll
        for (int q = 1; q < 8; q++) x[q] = cmul(x[q], tw[q * k]);
    }
    dft8(x);
    const int j = (i / P) * P * 8 + k;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (COLS) buf[(j + q * P) * T64 + line] = x[q];
        else buf[line * T64 + j + q * P] = x[q];
    }
    __syncthreads();
}

// Forward 2-D DFT of the 64 x 64 tile X in place (natural layout in and out), 512 threads.
__device__ __forceinline__ void fft64x64_inplace(float2 *X, const float2 *tw) {
    pass64_inplace<1, false>(X, tw);
    pass64_inplace<8, false>(X, tw);
    pass64_inplace<1, true>(X, tw);
    pass64_inplace<8, true>(X, tw);
}

__device__ double block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    __syncthreads();
    return s;
}

// Hann tables of a level: hann[0 .. wy) then hann[wy .. wy + wx), 0.5 - 0.5 cos(2 pi i / n) as float
// (poc.py:45-46).
__global__ void k_hann(float *hann, int wy, int wx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= wy + wx) return;
    const int n = i < wy ? wy : wx, e = i < wy ? i : i - wy;
    hann[i] = (float)(0.5 - 0.5 * cos(6.283185307179586 * (double)e / (double)n));
}

// Loads the mean-removed, Hann-windowed window at (y0, x0) of img into X (complex); element (y, x)
// goes to X[y * pitch + x].
__device__ void load_window(const float *img, int W, int y0, int x0, int wy, int wx, int pitch, const float *hy,
                            const float *hx, float2 *X, double *red) {
    double s = 0;
    for (int i = threadIdx.x; i < wy * wx; i += blockDim.x) {
        const int y = i / wx, x = i - y * wx;
        s += (double)img[(size_t)(y0 + y) * W + x0 + x];
    }
    const double mean = block_sum(s, red) / (double)(wy * wx);
    for (int i = threadIdx.x; i < wy * wx; i += blockDim.x) {
        const int y = i / wx, x = i - y * wx;
        const double v = ((double)img[(size_t)(y0 + y) * W + x0 + x] - mean) * (double)hy[y] * (double)hx[x];
        X[y * pitch + x] = make_float2((float)v, 0.f);
    }
    __syncthreads();
}

// |fftfreq(n)[i]| <= 0.25, i.e. 4 |k| <= n (exact: k / n rounds to 0.25 only when 4 |k| = n)
__device__ __forceinline__ bool lp_pass(int i, int n) {
    const int k = (i <= (n - 1) / 2) ? i : i - n;
    return 4 * abs(k) <= n;
}

__device__ double parabola(double m, double c, double p) {
    const double den = m - 2.0 * c + p;
    return den == 0 ? 0.0 : 0.5 * (m - p) / den;
}

constexpr int POC_NT = 512;               // threads of a POC workgroup
constexpr int POC_PER = POC_MAX / POC_NT;  // spectrum elements a thread keeps

// poc.py:67-92 with poc_shift (poc.py:39-54) inside. WIN = 64: 64 x 64 windows with the tiles in
// natural layout and every index a shift; WIN = 0: any window of at most POC_MAX samples, spectra
// transposed ([kx][ky]). Eight wavefronts per SIMD (64 VGPRs, a few spilled): with the 34 KB tile
// of WIN = 64, four workgroups per CU (445 -> 426 us per 2048^2 grid against three at 78 VGPRs).
template <int WIN>
__global__ void __launch_bounds__(POC_NT) __attribute__((amdgpu_waves_per_eu(8)))
k_poc_grid(const float *__restrict__ I1, const float *__restrict__ I2,
                                                     PocArgs P, const float2 *__restrict__ twx_g,
                                                     const float2 *__restrict__ twy_g, const float *__restrict__ hann_g,
                                                     const double *__restrict__ ix, const double *__restrict__ iy,
                                                     double *__restrict__ gux, double *__restrict__ guy,
                                                     double *__restrict__ gpk) {
    extern __shared__ float2 lds[];
    const int wy = WIN ? WIN : P.wy, wx = WIN ? WIN : P.wx;
    const int n = wy * wx;
    // WIN = 64: one tile, transformed in place; Y is a 64-element area for the block sums and the
    // argmax partials. Otherwise two buffers of n elements (the second the transforms' scratch).
    const int tile = WIN ? WIN * T64 : n;  // LDS elements of X
    const int ytile = WIN ? 64 : n;
    float2 *X = lds, *Y = lds + tile;
    float2 *twx = Y + ytile, *twy = WIN ? twx : twx + wx;
    float *hy = (float *)(twy + wy), *hx = hy + wy;
    for (int i = threadIdx.x; i < wx; i += blockDim.x) twx[i] = twx_g[i];
    if (!WIN)
        for (int i = threadIdx.x; i < wy; i += blockDim.x) twy[i] = twy_g[i];
    for (int i = threadIdx.x; i < wy + wx; i += blockDim.x) hy[i] = hann_g[i];
    double *red = (double *)Y;  // block sums run while Y is free
    float2 A[POC_PER];          // spectrum of the first window, element tid + POC_NT q
    const int pt = blockIdx.x;
    const int gi = pt / P.nx, gj = pt - gi * P.nx;
    const int y = P.y0 + gi * P.step, x = P.x0 + gj * P.step;
    const int hy2 = wy / 2, hx2 = wx / 2;
    double dx = P.have_init ? ix[pt] : 0.0, dy = P.have_init ? iy[pt] : 0.0;
    __syncthreads();
    // spectrum of the first window, kept in registers for every iteration
    load_window(I1, P.W, y - hy2, x - hx2, wy, wx, WIN ? T64 : wx, hy, hx, X, red);
    {
        float2 *r;
        if constexpr (WIN == 64) {
            fft64x64_inplace(X, twx);
            r = X;
        } else {
            r = fft2_lds(X, Y, wy, wx, P, twx, twy);
        }
#pragma unroll
        for (int q = 0; q < POC_PER; q++) {
            const int i = threadIdx.x + POC_NT * q;
            if constexpr (WIN == 64) A[q] = r[(i >> 6) * T64 + (i & 63)];
            else A[q] = i < n ? r[i] : make_float2(0.f, 0.f);
        }
        __syncthreads();
    }
    int nlp = 0;  // low-pass mask size (lp.mean() = nlp / n)
    {
        int cy = 0, cx = 0;
        for (int i = 0; i < wy; i++) cy += lp_pass(i, wy);
        for (int i = 0; i < wx; i++) cx += lp_pass(i, wx);
        nlp = cy * cx;
    }
    const double lpm = (double)nlp / (double)n;
    double pk = 0.0;
    for (int t = 0; t < 3; t++) {
        const int ox = (int)rint(dx), oy = (int)rint(dy);
        const int wy0 = y - hy2 + oy, wx0 = x - hx2 + ox;
        if (wy0 < 0 || wx0 < 0 || wy0 + wy > P.H || wx0 + wx > P.W) {
            pk = 0.0;
            break;
        }
        load_window(I2, P.W, wy0, wx0, wy, wx, WIN ? T64 : wx, hy, hx, X, red);
        float2 *B;
        if constexpr (WIN == 64) {
            fft64x64_inplace(X, twx);
            B = X;
        } else {
            B = fft2_lds(X, Y, wy, wx, P, twx, twy);
        }
        // normalised cross power spectrum, low-passed, conjugated for the inverse
#pragma unroll
        for (int q = 0; q < POC_PER; q++) {
            const int i = threadIdx.x + POC_NT * q;
            if (i >= n) break;
            int kx, ky, e;
            if constexpr (WIN == 64) {
                ky = i >> 6;
                kx = i & 63;
                e = ky * T64 + kx;
            } else {
                kx = i / wy;
                ky = i - kx * wy;
                e = i;
            }
            const float2 b = B[e], a = A[q];
            float2 c = make_float2(b.x * a.x + b.y * a.y, b.y * a.x - b.x * a.y);
            const float m = sqrtf(c.x * c.x + c.y * c.y) + 1e-12f;
            c = make_float2(c.x / m, c.y / m);
            if (!(lp_pass(kx, wx) && lp_pass(ky, wy))) c = make_float2(0.f, 0.f);
            B[e] = cconj(c);
        }
        __syncthreads();
        float2 *r2;
        int pitch;
        if constexpr (WIN == 64) {
            fft64x64_inplace(X, twx);  // forward transform of the conjugate = conjugate of the inverse
            r2 = X;
            pitch = T64;
        } else {
            // inverse: along ky (lines of length wy), transpose back to [ky][kx], along kx
            float2 *o = (B == X) ? Y : X;
            float2 *r1 = fft_lines_lds(B, o, wy, wx, P.facy, P.nfy, twy, threadIdx.x, blockDim.x);
            float2 *o2 = (r1 == B) ? o : B;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int kx = i / wy, yy = i - kx * wy;
                o2[yy * wx + kx] = r1[i];
            }
            __syncthreads();
            r2 = fft_lines_lds(o2, r1, wx, wy, P.facx, P.nfx, twx, threadIdx.x, blockDim.x);
            pitch = wx;
        }
        // the buffer r2 does not occupy holds the argmax partials
        float *bval = (float *)((r2 == X) ? Y : X);
        int *bidx = (int *)(bval + 64);
        // r = fftshift(real(ifft2)) / lpm; argmax in the shifted array (first maximum, NaN wins)
        const double sc = 1.0 / ((double)n * lpm);
        const int shy = (wy + 1) / 2, shx = (wx + 1) / 2;
        auto at = [&](int sy, int sx) {  // shifted (sy, sx) -> unshifted element
            int uy = sy + shy, ux = sx + shx;
            uy -= uy >= wy ? wy : 0;
            ux -= ux >= wx ? wx : 0;
            return r2[uy * pitch + ux].x;
        };
        float best = 0.f;
        int bi = -1;
        bool bnan = false;
        for (int s = threadIdx.x; s < n; s += blockDim.x) {
            int sy, sx;
            if constexpr (WIN == 64) {
                sy = s >> 6;
                sx = s & 63;
            } else {
                sy = s / wx;
                sx = s - sy * wx;
            }
            const float v = (float)((double)at(sy, sx) * sc);
            if (bnan) continue;
            if (isnan(v)) {
                bnan = true;
                best = v;
                bi = s;
            } else if (bi < 0 || v > best) {
                best = v;
                bi = s;
            }
        }
        // first maximum over the workgroup (NaN wins, then the larger value, then the smaller index):
        // wavefront shuffles, then the wavefront winners
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            const bool on = __shfl_xor((int)bnan, off) != 0;
            bool take;
            if (oi < 0) take = false;
            else if (bi < 0) take = true;
            else if (bnan) take = on && oi < bi;
            else if (on) take = true;
            else take = ov > best || (ov == best && oi < bi);
            if (take) {
                best = ov;
                bi = oi;
                bnan = on;
            }
        }
        if ((threadIdx.x & 63) == 0) {
            bval[threadIdx.x >> 6] = bnan ? NAN : best;
            bidx[threadIdx.x >> 6] = bi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float bv = 0.f;
            int bb = -1;
            bool bn = false;
            for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
                const int ki = bidx[k];
                if (ki < 0) continue;
                const float kv = bval[k];
                const bool kn = isnan(kv);
                bool take;
                if (bb < 0) take = true;
                else if (bn) take = kn && ki < bb;
                else if (kn) take = true;
                else take = kv > bv || (kv == bv && ki < bb);
                if (take) {
                    bv = kv;
                    bb = ki;
                    bn = kn;
                }
            }
            bidx[0] = bb;
        }
        __syncthreads();
        const int bsel = bidx[0];
        int py = bsel / wx, px = bsel - py * wx;
        py = min(max(py, 1), wy - 2);
        px = min(max(px, 1), wx - 2);
        auto R = [&](int sy, int sx) { return (double)(float)((double)at(sy, sx) * sc); };
        const double c0 = R(py, px);
        const double sx = parabola(R(py, px - 1), c0, R(py, px + 1));
        const double sy = parabola(R(py - 1, px), c0, R(py + 1, px));
        pk = c0;
        dx = ox + (px - wx / 2 + sx);
        dy = oy + (py - wy / 2 + sy);
        __syncthreads();
        if (fabs(px - wx / 2 + sx) < 0.5 && fabs(py - wy / 2 + sy) < 0.5) break;
    }
    if (threadIdx.x == 0) {
        gux[pt] = dx;
        guy[pt] = dy;
        gpk[pt] = pk;
    }
}

// ------------------------------------------------------------------ clean (poc.py:95-140)
__device__ double median_small(double *v, int n) {
    for (int i = 1; i < n; i++) {
        const double t = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > t) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = t;
    }
    return (n & 1) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// points of a raw grid that confirm their start: peak >= peak_min and |result - start| <= tol on
// both axes (one count per wavefront)
__global__ void k_raw_check(const double *__restrict__ gux, const double *__restrict__ guy,
                            const double *__restrict__ gpk, const double *__restrict__ ix, const double *__restrict__ iy,
                            int n, double peak_min, double tol, int *count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = k < n && gpk[k] >= peak_min && fabs(gux[k] - ix[k]) <= tol && fabs(guy[k] - iy[k]) <= tol;
    const unsigned long long m = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (int)__popcll(m));
}

__global__ void k_zero_int(int *p) {
    if (threadIdx.x == 0) p[0] = 0;
}

// flags: bit 0 valid (peak >= peak_min), bit 1 out (the point holds a usable value)
// Outlier test (poc.py:95-120) over the whole grid, one thread per point: a valid point with at
// least 3 valid 8-neighbours is an outlier if either component deviates from their median by more
// than thresh (MAD + 0.1). counters[0] counts the points left out.
__global__ void __launch_bounds__(256) k_clean_test(const double *__restrict__ ux, const double *__restrict__ uy,
                                                    const double *__restrict__ peak, int *__restrict__ flags,
                                                    int *__restrict__ counters, int ny, int nx, double peak_min,
                                                    double thresh) {
    const int n = ny * nx;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool out = false;
    if (k < n) {
        const bool valid = peak[k] >= peak_min;
        out = valid;
        if (valid) {
            const int i = k / nx, j = k - i * nx;
            int nb[8], c = 0;
            for (int di = -1; di <= 1; di++)
                for (int dj = -1; dj <= 1; dj++) {
                    if (!di && !dj) continue;
                    const int a = i + di, b = j + dj;
                    if (a >= 0 && a < ny && b >= 0 && b < nx && peak[a * nx + b] >= peak_min) nb[c++] = a * nx + b;
                }
            if (c >= 3) {
                for (int comp = 0; comp < 2; comp++) {
                    const double *u = comp ? uy : ux;
                    double vals[8], dev[8];
                    for (int q = 0; q < c; q++) vals[q] = u[nb[q]];
                    const double med = median_small(vals, c);
                    for (int q = 0; q < c; q++) dev[q] = fabs(u[nb[q]] - med);
                    const double res = median_small(dev, c);
                    if (fabs(u[k] - med) / (res + 0.1) > thresh) out = false;
                }
            }
        }
        flags[k] = (valid ? 1 : 0) | (out ? 2 : 0);
    }
    const unsigned long long b = __ballot(out);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&counters[0], (int)__popcll(b));
}

// Fill (poc.py:121-140) in one workgroup over the list of points still to fill: each pass gives
// every listed point with out neighbours the median of their values, then drops it from the list;
// at most 100 passes. Without any out point the grid falls back to the guess (or 0) with peak 0.
// Finally the peak of every invalid point is 0.
__global__ void __launch_bounds__(1024) k_clean_fill(double *ux, double *uy, double *peak, const double *fx,
                                                     const double *fy, int have_fb, int *flags, int *work,
                                                     const int *counters, int ny, int nx) {
    __shared__ int nwork, nnext;
    // work: two lists of up to ny nx points (this pass, next pass)
    const int n = ny * nx;
    if (counters[0] == 0) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            ux[k] = have_fb ? fx[k] : 0.0;
            uy[k] = have_fb ? fy[k] : 0.0;
            peak[k] = peak[k] * 0;
        }
        return;
    }
    int *cur = work, *nxt = work + n;
    if (threadIdx.x == 0) nwork = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x)
        if (!(flags[k] & 2)) cur[atomicAdd(&nwork, 1)] = k;
    __syncthreads();
    for (int pass = 0; pass < 100 && nwork > 0; pass++) {
        // values from the points out at the start of the pass; a filled point is marked with bit 2
        for (int t = threadIdx.x; t < nwork; t += blockDim.x) {
            const int k = cur[t];
            const int i = k / nx, j = k - i * nx;
            double vx[8], vy[8];
            int c = 0;
            for (int di = -1; di <= 1; di++)
                for (int dj = -1; dj <= 1; dj++) {
                    if (!di && !dj) continue;
                    const int a = i + di, b = j + dj;
                    if (a >= 0 && a < ny && b >= 0 && b < nx && (flags[a * nx + b] & 2)) {
                        vx[c] = ux[a * nx + b];
                        vy[c] = uy[a * nx + b];
                        c++;
                    }
                }
            if (c) {
                ux[k] = median_small(vx, c);
                uy[k] = median_small(vy, c);
                flags[k] |= 4;
            }
        }
        if (threadIdx.x == 0) nnext = 0;
        __syncthreads();
        // filled points become out; the others go to the next list (in any order: a point's fill
        // depends only on the state at the start of a pass)
        for (int t = threadIdx.x; t < nwork; t += blockDim.x) {
            const int k = cur[t];
            if (flags[k] & 4) flags[k] = (flags[k] & 1) | 2;
            else nxt[atomicAdd(&nnext, 1)] = k;
        }
        __syncthreads();
        int *tmp = cur;
        cur = nxt;
        nxt = tmp;
        if (threadIdx.x == 0) nwork = nnext;
        __syncthreads();
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x)
        if (!(flags[k] & 1)) peak[k] = 0.0;
}

// ------------------------------------------------------------------ grid interpolation (poc.py:143-161)
struct GridGeom {
    int x0, y0, step, nx, ny;
};

__device__ void grid_frac(double q, int q0, int step, int nn, double &t, int &i0) {
    if (nn == 1) {
        t = 0;
        i0 = 0;
        return;
    }
    double f = (q - q0) / (double)step;
    f = fmin(fmax(f, 0.0), (double)(nn - 1));
    i0 = min((int)floor(f), nn - 2);
    t = f - i0;
}

__device__ void grid_bilinear(const GridGeom &g, const double *ax, const double *ay, double xq, double yq, double &ox,
                              double &oy) {
    double tx, ty;
    int x0, y0;
    grid_frac(xq, g.x0, g.step, g.nx, tx, x0);
    grid_frac(yq, g.y0, g.step, g.ny, ty, y0);
    const int x1 = min(x0 + 1, g.nx - 1), y1 = min(y0 + 1, g.ny - 1);
    const double *A[2] = {ax, ay};
    double o[2];
    for (int c = 0; c < 2; c++) {
        const double *a = A[c];
        o[c] = (1 - ty) * ((1 - tx) * a[y0 * g.nx + x0] + tx * a[y0 * g.nx + x1]) +
               ty * ((1 - tx) * a[y1 * g.nx + x0] + tx * a[y1 * g.nx + x1]);
    }
    ox = o[0];
    oy = o[1];
}

// init of level lev from the cleaned grid of level lev + 1 (halve coordinates, double the field)
__global__ void k_level_init(GridGeom prev, const double *pux, const double *puy, GridGeom cur, double *ix,
                             double *iy) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cur.nx * cur.ny) return;
    const int i = k / cur.nx, j = k - i * cur.nx;
    double ox, oy;
    grid_bilinear(prev, pux, puy, (cur.x0 + j * cur.step) / 2.0, (cur.y0 + i * cur.step) / 2.0, ox, oy);
    ix[k] = 2.0 * ox;
    iy[k] = 2.0 * oy;
}

// coarsest level from a dense guess: init[Y sc, X sc] / sc
__global__ void k_dense_init(const double *gx, const double *gy, int W, GridGeom cur, int sc, double *ix, double *iy) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cur.nx * cur.ny) return;
    const int i = k / cur.nx, j = k - i * cur.nx;
    const size_t p = (size_t)((cur.y0 + i * cur.step) * sc) * W + (size_t)(cur.x0 + j * cur.step) * sc;
    ix[k] = (double)gx[p] / sc;
    iy[k] = (double)gy[p] / sc;
}

// correction relative to the dense guess at the grid nodes (track.py:133-137)
__global__ void k_grid_minus_init(const double *gux, const double *guy, const double *gx, const double *gy, int W,
                                  GridGeom g, double *cx, double *cy) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nx * g.ny) return;
    const int i = k / g.nx, j = k - i * g.nx;
    const size_t p = (size_t)(g.y0 + i * g.step) * W + (g.x0 + j * g.step);
    cx[k] = gux[k] - (double)gx[p];
    cy[k] = guy[k] - (double)gy[p];
}

__global__ void k_dense(GridGeom g, const double *gux, const double *guy, const double *init_x, const double *init_y,
                        int W, int H, double *ux, double *uy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    double ox, oy;
    grid_bilinear(g, gux, guy, x, y, ox, oy);
    const size_t i = (size_t)y * W + x;
    if (init_x) {
        ox += (double)init_x[i];
        oy += (double)init_y[i];
    }
    ux[i] = ox;
    uy[i] = oy;
}

template <class T>
void ensure(T *&p, size_t n, size_t &cap) {
    if (p && cap >= n) return;
    if (p) release_device(p);
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    cap = n;
}

// the grid buffers of the workspace, sized for npmax points
void ensure_grid(CoarseWs *ws, size_t npmax) {
    if (ws->gux && ws->gcap >= npmax) return;
    release_device(ws->gux); release_device(ws->guy); release_device(ws->gpk);
    release_device(ws->ix); release_device(ws->iy); release_device(ws->gflag); release_device(ws->gwork);
    HIPCHECK(hipMalloc(&ws->gux, npmax * sizeof(double) * 2));
    HIPCHECK(hipMalloc(&ws->guy, npmax * sizeof(double) * 2));
    HIPCHECK(hipMalloc(&ws->gpk, npmax * sizeof(double)));
    HIPCHECK(hipMalloc(&ws->ix, npmax * sizeof(double)));
    HIPCHECK(hipMalloc(&ws->iy, npmax * sizeof(double)));
    HIPCHECK(hipMalloc(&ws->gflag, npmax * sizeof(int)));
    HIPCHECK(hipMalloc(&ws->gwork, (2 * npmax + 64) * sizeof(int)));
    ws->gcap = npmax;
}

// One pyramid level (poc.rs:738-762): the POC grid of a, b (lw x lh) with windows wy x wx at the
// points of `cur`, started from ws->ix / ws->iy when have_init; the raw grid is in ws->gux, guy,
// gpk, and with raw_check (not NULL) the points whose peak is at least poc_peak_min and whose
// result lies within tol px of the start on both axes are counted into *raw_check; then `clean`
// with that start as the fallback.
void poc_level(const float *a, const float *b, int lw, int lh, int wy, int wx, const GridGeom &cur, int have_init,
               const ofx_config &c, CoarseWs *ws, int *raw_check, double tol, hipStream_t s);

}  // namespace

// The level-0 grid of the last coarse_dense call with this workspace: positions and (ux, uy,
// peak) per point, ny x nx, on the host.
void coarse_grid(const CoarseWs *ws, int w, int h, const ofx_config &c, std::vector<int> &xs, std::vector<int> &ys,
                 std::vector<double> &ux, std::vector<double> &uy, std::vector<double> &peak, PinnedBuf &pin,
                 hipStream_t s) {
    if (!ws || !ws->gux) throw Error(OFX_EINVAL, "no coarse grid computed");
    int x0, nx, y0, ny;
    grid_axis(w, c.poc_win, c.grid_step, x0, nx);
    grid_axis(h, c.poc_win, c.grid_step, y0, ny);
    xs.resize(nx);
    ys.resize(ny);
    for (int i = 0; i < nx; i++) xs[i] = x0 + i * c.grid_step;
    for (int j = 0; j < ny; j++) ys[j] = y0 + j * c.grid_step;
    const size_t np = (size_t)nx * ny;
    ux.resize(np);
    uy.resize(np);
    peak.resize(np);
    copy_d2h(pin, ux.data(), ws->gux, np * sizeof(double), s);
    copy_d2h(pin, uy.data(), ws->guy, np * sizeof(double), s);
    copy_d2h(pin, peak.data(), ws->gpk, np * sizeof(double), s);
}

void coarse_release(CoarseWs *&ws) {
    if (!ws) return;
    for (auto p : ws->pa) (void)hipFree(p);
    for (auto p : ws->pb) (void)hipFree(p);
    (void)hipFree(ws->tmp);
    (void)hipFree(ws->gux);
    (void)hipFree(ws->guy);
    (void)hipFree(ws->gpk);
    (void)hipFree(ws->ix);
    (void)hipFree(ws->iy);
    (void)hipFree(ws->gflag);
    (void)hipFree(ws->gwork);
    (void)hipFree(ws->hann);
    delete ws;
    ws = nullptr;
}

namespace {
void poc_level(const float *a, const float *b, int lw, int lh, int wy, int wx, const GridGeom &cur, int have_init,
               const ofx_config &c, CoarseWs *ws, int *raw_check, double tol, hipStream_t s) {
    const int x0 = cur.x0, y0 = cur.y0, nx = cur.nx, ny = cur.ny;
    const size_t np = (size_t)nx * ny;
    const FftPlan &PX = fft_plan(wx, s), &PY = fft_plan(wy, s);
    PocArgs P;
    std::memset(&P, 0, sizeof(P));
    P.W = lw;
    P.H = lh;
    P.wy = wy;
    P.wx = wx;
    P.x0 = x0;
    P.y0 = y0;
    P.step = c.grid_step;
    P.nx = nx;
    P.ny = ny;
    P.nfx = PX.nfac;
    P.nfy = PY.nfac;
    for (int i = 0; i < MAX_FACTORS; i++) {
        P.facx[i] = PX.fac[i];
        P.facy[i] = PY.fac[i];
    }
    P.have_init = have_init;
    hipLaunchKernelGGL(k_hann, dim3((wy + wx + 255) / 256), dim3(256), 0, s, ws->hann, wy, wx);
    LAUNCHCHECK();
    if (wy == 64 && wx == 64) {
        // one 64 x 64 tile (transformed in place), the 64-element scratch area, twiddles and Hann
        // tables: 34 KiB, so three or four workgroups share a CU
        const size_t lds = ((size_t)64 * T64 + 64 + 64) * sizeof(float2) + 128 * sizeof(float);
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute((const void *)k_poc_grid<64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024);
        });
        hipLaunchKernelGGL(k_poc_grid<64>, dim3((unsigned)np), dim3(POC_NT), lds, s, a, b, P, PX.tw, PY.tw,
                           ws->hann, ws->ix, ws->iy, ws->gux, ws->guy, ws->gpk);
    } else {
        const size_t lds = (2 * (size_t)wy * wx + wx + wy) * sizeof(float2) + (size_t)(wy + wx) * sizeof(float);
        if (lds > 65536) {
            static std::once_flag once;
            std::call_once(once, [] {
                (void)hipFuncSetAttribute((const void *)k_poc_grid<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          160 * 1024);
            });
        }
        hipLaunchKernelGGL(k_poc_grid<0>, dim3((unsigned)np), dim3(POC_NT), lds, s, a, b, P, PX.tw, PY.tw,
                           ws->hann, ws->ix, ws->iy, ws->gux, ws->guy, ws->gpk);
    }
    LAUNCHCHECK();
    if (raw_check) {
        hipLaunchKernelGGL(k_raw_check, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, (const double *)ws->gux,
                           (const double *)ws->guy, (const double *)ws->gpk, (const double *)ws->ix,
                           (const double *)ws->iy, (int)np, c.poc_peak_min, tol, raw_check);
        LAUNCHCHECK();
    }
    int *cnt = ws->gwork + 2 * ws->gcap;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, cnt);
    hipLaunchKernelGGL(k_clean_test, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, (const double *)ws->gux,
                       (const double *)ws->guy, (const double *)ws->gpk, ws->gflag, cnt, ny, nx, c.poc_peak_min,
                       c.median_thresh);
    hipLaunchKernelGGL(k_clean_fill, dim3(1), dim3(1024), 0, s, ws->gux, ws->guy, ws->gpk, ws->ix, ws->iy, have_init,
                       ws->gflag, ws->gwork, (const int *)cnt, ny, nx);
    LAUNCHCHECK();
}
}  // namespace

// The coarse stage from a proposal in place of the pyramid: the level-0 POC grid started from the
// proposal (guess fills the start at the grid points), cleaned with that start as the fallback,
// then the dense start. Returns the number of grid points and, in *confirmed (host), how many of
// them confirmed the proposal: a peak of at least poc_peak_min and a result within tol px of it.
int coarse_dense_from(const float *i1, const float *i2, int w, int h, const ofx_config &c, const GuessFn &guess,
                      double tol, double *ux, double *uy, CoarseWs *&ws, int *confirmed, hipStream_t s) {
    if (!ws) {
        ws = new CoarseWs;
        HIPCHECK(hipMalloc(&ws->hann, 2 * (size_t)POC_MAX * sizeof(float)));
    }
    ws->w = w;
    ws->h = h;
    ws->L = 0;
    const int wy = c.poc_win, wx = c.poc_win;
    if ((long)wy * wx > POC_MAX)
        throw Error(OFX_EUNSUPPORTED, "POC window " + std::to_string(wy) + "x" + std::to_string(wx) +
                                          " exceeds the 4096 samples a workgroup holds");
    int x0, nx, y0, ny;
    grid_axis(w, wx, c.grid_step, x0, nx);
    grid_axis(h, wy, c.grid_step, y0, ny);
    const size_t np = (size_t)nx * ny;
    if (np == 0) throw Error(OFX_EINVAL, "no POC grid point fits the image");
    ensure_grid(ws, np);
    const GridGeom cur{x0, y0, c.grid_step, nx, ny};
    guess(x0, y0, c.grid_step, nx, ny, ws->ix, ws->iy);
    int *dcount = ws->gwork + 2 * ws->gcap + 32;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, dcount);
    LAUNCHCHECK();
    poc_level(i1, i2, w, h, wy, wx, cur, 1, c, ws, dcount, tol, s);
    hipLaunchKernelGGL(k_dense, dim3((w + 255) / 256, h), dim3(256), 0, s, cur, (const double *)ws->gux,
                       (const double *)ws->guy, (const double *)nullptr, (const double *)nullptr, w, h, ux, uy);
    LAUNCHCHECK();
    PinnedBuf pin;
    copy_d2h(pin, confirmed, dcount, sizeof(int), s);
    return (int)np;
}

void coarse_dense(const float *i1, const float *i2, int w, int h, const ofx_config &c, const double *init_ux,
                  const double *init_uy, double *ux, double *uy, CoarseWs *&ws, hipStream_t s) {
    const int L = n_levels(c.max_disp, c.poc_win, h, w);
    // the workspace keeps every buffer while it is large enough (no frees when the size shrinks)
    if (!ws) {
        ws = new CoarseWs;
        HIPCHECK(hipMalloc(&ws->hann, 2 * (size_t)POC_MAX * sizeof(float)));
    }
    ws->w = w;
    ws->h = h;
    ws->L = L;
    ws->lw.assign(L, 0);
    ws->lh.assign(L, 0);
    {
        int lw = w, lh = h;
        for (int l = 0; l < L; l++) {
            const int nw = (lw + 1) / 2, nh = (lh + 1) / 2;
            if ((int)ws->pa.size() <= l) {
                ws->pa.push_back(nullptr);
                ws->pb.push_back(nullptr);
                ws->pcap.push_back(0);
            }
            size_t cap_b = ws->pcap[l];
            ensure(ws->pa[l], (size_t)nw * nh, ws->pcap[l]);
            ensure(ws->pb[l], (size_t)nw * nh, cap_b);
            ws->lw[l] = nw;
            ws->lh[l] = nh;
            lw = nw;
            lh = nh;
        }
        ensure(ws->tmp, (size_t)w * ((h + 1) / 2), ws->tmpcap);
    }
    // pyramid
    for (int l = 0; l < L; l++) {
        const int pw = l == 0 ? w : ws->lw[l - 1], ph = l == 0 ? h : ws->lh[l - 1];
        const float *sa = l == 0 ? i1 : ws->pa[l - 1], *sb = l == 0 ? i2 : ws->pb[l - 1];
        const int nw = ws->lw[l], nh = ws->lh[l];
        for (int img = 0; img < 2; img++) {
            hipLaunchKernelGGL(k_down_y, dim3((pw + 255) / 256, nh), dim3(256), 0, s, img ? sb : sa, ws->tmp, pw, ph, nh);
            hipLaunchKernelGGL(k_down_x, dim3((nw + 255) / 256, nh), dim3(256), 0, s, ws->tmp,
                               img ? ws->pb[l] : ws->pa[l], pw, nw, nh);
        }
        LAUNCHCHECK();
    }
    // grid buffers sized for the largest level; the previous level's cleaned grid is kept in the
    // upper halves of gux / guy
    size_t npmax = 0;
    for (int lev = L; lev >= 0; lev--) {
        const int lw = lev == 0 ? w : ws->lw[lev - 1], lh = lev == 0 ? h : ws->lh[lev - 1];
        int wy = c.poc_win, wx = c.poc_win;
        if (lev > 0) level_window(lh, lw, c.poc_win, wy, wx);
        int x0, nx, y0, ny;
        grid_axis(lw, wx, c.grid_step, x0, nx);
        grid_axis(lh, wy, c.grid_step, y0, ny);
        npmax = std::max(npmax, (size_t)nx * ny);
    }
    ensure_grid(ws, npmax);
    GridGeom prev{};
    bool have_prev = false;
    for (int lev = L; lev >= 0; lev--) {
        const int lw = lev == 0 ? w : ws->lw[lev - 1], lh = lev == 0 ? h : ws->lh[lev - 1];
        const float *a = lev == 0 ? i1 : ws->pa[lev - 1], *b = lev == 0 ? i2 : ws->pb[lev - 1];
        int wy = c.poc_win, wx = c.poc_win;
        if (lev > 0) level_window(lh, lw, c.poc_win, wy, wx);
        if ((long)wy * wx > POC_MAX)
            throw Error(OFX_EUNSUPPORTED, "POC window " + std::to_string(wy) + "x" + std::to_string(wx) +
                                              " exceeds the 4096 samples a workgroup holds");
        int x0, nx, y0, ny;
        grid_axis(lw, wx, c.grid_step, x0, nx);
        grid_axis(lh, wy, c.grid_step, y0, ny);
        const GridGeom cur{x0, y0, c.grid_step, nx, ny};
        const size_t np = (size_t)nx * ny;
        if (np == 0) throw Error(OFX_EINVAL, "no POC grid point fits the image");
        // the previous level's cleaned grid lives in the upper halves of gux / guy
        double *pux = ws->gux + ws->gcap, *puy = ws->guy + ws->gcap;
        int have_init = 0;
        if (have_prev) {
            dcopy(pux, ws->gux, (size_t)prev.nx * prev.ny * sizeof(double), s);
            dcopy(puy, ws->guy, (size_t)prev.nx * prev.ny * sizeof(double), s);
            hipLaunchKernelGGL(k_level_init, dim3((np + 255) / 256), dim3(256), 0, s, prev, pux, puy, cur, ws->ix, ws->iy);
            have_init = 1;
        } else if (init_ux) {
            hipLaunchKernelGGL(k_dense_init, dim3((np + 255) / 256), dim3(256), 0, s, init_ux, init_uy, w, cur, 1 << lev,
                               ws->ix, ws->iy);
            have_init = 1;
        }
        LAUNCHCHECK();
        poc_level(a, b, lw, lh, wy, wx, cur, have_init, c, ws, nullptr, 0.0, s);
        prev = cur;
        have_prev = true;
    }
    // dense start (track.py:131-137)
    double *gx = ws->gux, *gy = ws->guy;
    if (init_ux) {
        double *cx = ws->gux + ws->gcap, *cy = ws->guy + ws->gcap;
        hipLaunchKernelGGL(k_grid_minus_init, dim3((prev.nx * prev.ny + 255) / 256), dim3(256), 0, s, ws->gux, ws->guy,
                           init_ux, init_uy, w, prev, cx, cy);
        gx = cx;
        gy = cy;
    }
    hipLaunchKernelGGL(k_dense, dim3((w + 255) / 256, h), dim3(256), 0, s, prev, gx, gy, init_ux, init_uy, w, h, ux, uy);
    LAUNCHCHECK();
}

}  // namespace ofx
