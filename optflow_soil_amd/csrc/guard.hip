// Distance transforms, connected components, the trusted pre-warp fill, validity and strain.
//
// Distances follow scipy.ndimage.distance_transform_edt as the reference uses it: exact Euclidean
// distance to the nearest feature pixel, and for the pre-warp the nearest feature itself with
// scipy's tie rule (smallest column, then smallest row; rust/crates/optflow-core/src/edt.rs:175-183).
// The column pass is exact and unbounded (two sweeps per column). The row pass has two forms: a
// thresholded test d <= thr searches only |dx| <= floor(thr) (every column farther away is more
// than thr away, so the decision is exact), and the nearest-feature search widens |dx| until
// dx^2 exceeds the best squared distance found.
#include <cmath>

#include "fine.hpp"

namespace ofx {

namespace {

dim3 grid2(int w, int h) { return dim3((w + 255) / 256, h); }

// Column pass in chunks of CHUNK rows, one thread per (column, chunk): the first and last feature
// row of every chunk, then every chunk finds its nearest feature above and below from those
// summaries and sweeps its own rows (a tie between the two keeps the smaller row).
constexpr int CHUNK = 32;

// Both kernels hold a chunk's feature flags in registers: all CHUNK loads are issued before the
// first is used, and the sweeps run on registers (one store per output plane).
__global__ void k_col_summary(const uint8_t *__restrict__ feat, int w, int h, int *__restrict__ first,
                              int *__restrict__ last) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (x >= w) return;
    const int y0 = c * CHUNK, rows = min(CHUNK, h - y0);
    bool f8[CHUNK];
#pragma unroll
    for (int k = 0; k < CHUNK; k++) {
        // every row loaded (the chunk's last row again past the frame), so that the loads are not
        // split by a branch and are all in flight together
        const bool v = feat[(size_t)(y0 + min(k, rows - 1)) * w + x] != 0;
        f8[k] = k < rows && v;
    }
    int f = -1, l = -1;
#pragma unroll
    for (int k = CHUNK - 1; k >= 0; k--)
        if (f8[k]) f = y0 + k;
#pragma unroll
    for (int k = 0; k < CHUNK; k++)
        if (f8[k]) l = y0 + k;
    first[(size_t)c * w + x] = f;
    last[(size_t)c * w + x] = l;
}

__global__ void k_column_nearest(const uint8_t *__restrict__ feat, const int *__restrict__ first,
                                 const int *__restrict__ last, int w, int h, int *__restrict__ col_d2,
                                 int *__restrict__ col_row) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (x >= w) return;
    const int nch = (h + CHUNK - 1) / CHUNK;
    const int y0 = c * CHUNK, rows = min(CHUNK, h - y0);
    bool f8[CHUNK];
#pragma unroll
    for (int k = 0; k < CHUNK; k++) {
        // every row loaded (the chunk's last row again past the frame), so that the loads are not
        // split by a branch and are all in flight together
        const bool v = feat[(size_t)(y0 + min(k, rows - 1)) * w + x] != 0;
        f8[k] = k < rows && v;
    }
    int up = -1, down = -1;
    for (int cc = c - 1; cc >= 0 && up < 0; cc--) up = last[(size_t)cc * w + x];
    for (int cc = c + 1; cc < nch && down < 0; cc++) down = first[(size_t)cc * w + x];
    int ups[CHUNK];
#pragma unroll
    for (int k = 0; k < CHUNK; k++) {
        if (f8[k]) up = y0 + k;
        ups[k] = up;
    }
#pragma unroll
    for (int k = CHUNK - 1; k >= 0; k--) {
        if (k >= rows) continue;
        const int y = y0 + k;
        if (f8[k]) down = y;
        const int u = ups[k];
        const int du = u >= 0 ? y - u : 0x7fffffff, dd = down >= 0 ? down - y : 0x7fffffff;
        const size_t i = (size_t)y * w + x;
        // by selects, so that the stores of the chunk are not split by branches; a tie keeps the
        // smaller row, no feature in the column gives -1
        const bool none = u < 0 && down < 0, above = du <= dd;
        col_d2[i] = none ? -1 : above ? du * du : dd * dd;
        col_row[i] = none ? -1 : above ? u : down;
    }
}

__global__ void k_within(const int *__restrict__ col_d2, int w, int h, double thr, int T, uint8_t *__restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const int *row = col_d2 + (size_t)y * w;
    long long best = -1;
    for (int dx = -T; dx <= T; dx++) {
        const int xx = x + dx;
        if (xx < 0 || xx >= w) continue;
        const int g = row[xx];
        if (g < 0) continue;
        const long long d2 = (long long)dx * dx + g;
        if (best < 0 || d2 < best) best = d2;
    }
    out[(size_t)y * w + x] = best >= 0 && sqrt((double)best) <= thr;
}

// The same test with the row segment a workgroup reaches (its 256 pixels and T on each side)
// staged in LDS once, instead of 2T + 1 global loads per pixel; positions outside the row read as
// -1 (no column point), which the loop skips as k_within skips them. For T <= WITHIN_MAXT.
// The distance test on integers: sqrt(d2) <= thr holds for exactly the squared distances
// d2 <= B, B the largest integer whose correctly rounded square root is at most thr (B < 0 when
// thr < 0), found on the host. A squared distance here is at most 256^2 plus a squared column
// distance (below 7744^2), so 32 bits hold it; a column without a feature (-1) counts as none.
constexpr int WITHIN_MAXT = 256;
__global__ void __launch_bounds__(256) k_within_lds(const int *__restrict__ col_d2, int w, int B, int T,
                                                    uint8_t *__restrict__ out) {
    __shared__ int seg[256 + 2 * WITHIN_MAXT];
    const int x0 = blockIdx.x * 256, y = blockIdx.y;
    const int *row = col_d2 + (size_t)y * w;
    for (int c = threadIdx.x; c < 256 + 2 * T; c += 256) {
        const int xx = x0 - T + c;
        seg[c] = (xx >= 0 && xx < w) ? row[xx] : -1;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= w) return;
    const int *g0 = seg + threadIdx.x + T;  // g0[dx] is column x + dx
    int best = 0x7fffffff;
    for (int dx = -T; dx <= T; dx++) {
        const int g = g0[dx];
        best = min(best, g < 0 ? 0x7fffffff : dx * dx + g);
    }
    out[(size_t)y * w + x] = best <= B;
}

__global__ void k_zero_counters(Counters *cnt) {
    cnt->n_low = 0;
    cnt->n_ok = 0;
}

__global__ void __launch_bounds__(256) k_trusted(const uint8_t *__restrict__ low, const uint8_t *__restrict__ within_r,
                                                 size_t n, uint8_t *__restrict__ ok, Counters *cnt) {
    __shared__ unsigned int red[2][4];
    unsigned int nl = 0, no = 0;
    // four pixels per step as 32-bit words (the byte planes are allocations of their own, so their
    // starts are aligned), the last n mod 4 pixels one at a time. nz(v) has bit 7 of each byte set
    // where that byte of v is not zero (no carry crosses a byte).
    auto nz = [](unsigned int v) { return (((v & 0x7f7f7f7fu) + 0x7f7f7f7fu) | v) & 0x80808080u; };
    const size_t stride = (size_t)gridDim.x * blockDim.x, n4 = n / 4;
    const unsigned int *l4 = (const unsigned int *)low, *w4 = (const unsigned int *)within_r;
    unsigned int *o4 = (unsigned int *)ok;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
        const unsigned int o = (~nz(w4[q]) & 0x80808080u) >> 7;  // farther than R from every low pixel
        o4[q] = o;
        nl += __popc(nz(l4[q]));
        no += __popc(o);
    }
    for (size_t i = 4 * n4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint8_t o = !within_r[i];
        ok[i] = o;
        nl += low[i] != 0;
        no += o;
    }
    for (int off = 32; off > 0; off >>= 1) {
        nl += __shfl_down(nl, off);
        no += __shfl_down(no, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = nl;
        red[1][threadIdx.x >> 6] = no;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const unsigned int b = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (a) atomicAdd(&cnt->n_low, a);
        if (b) atomicAdd(&cnt->n_ok, b);
    }
}

// phaseflow.py:344-355: every untrusted pixel X takes u(N) + grad u(N) (X - N) from its nearest
// trusted pixel N, unless no pixel is low, none is trusted or all are.
__global__ void k_prewarp_fill(const double *__restrict__ ux, const double *__restrict__ uy,
                               const double *__restrict__ slope, const uint8_t *__restrict__ ok,
                               const int *__restrict__ col_d2, const int *__restrict__ col_row,
                               const Counters *__restrict__ cnt, int w, int h, double *__restrict__ fx,
                               double *__restrict__ fy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const unsigned int n_ok = cnt->n_ok;
    if (cnt->n_low == 0 || n_ok == 0 || n_ok == (unsigned int)((size_t)w * h) || ok[i]) {
        fx[i] = ux[i];
        fy[i] = uy[i];
        return;
    }
    const int *row = col_d2 + (size_t)y * w;
    long long best = -1;
    int bx = -1;
    const int maxdx = max(x, w - 1 - x);
    // the smallest squared distance, and among equal ones the smallest column (scipy's rule); a
    // tie can come from any |dx| up to the distance itself, so equal totals found later still
    // compete by column
    for (int dx = 0; dx <= maxdx; dx++) {
        const long long dd = (long long)dx * dx;
        if (best >= 0 && dd > best) break;
        const int c0 = x - dx, c1 = x + dx;
        if (c0 >= 0) {
            const int g = row[c0];
            const long long t = dd + g;
            if (g >= 0 && (best < 0 || t < best || (t == best && c0 < bx))) { best = t; bx = c0; }
        }
        if (dx > 0 && c1 < w) {
            const int g = row[c1];
            const long long t = dd + g;
            if (g >= 0 && (best < 0 || t < best || (t == best && c1 < bx))) { best = t; bx = c1; }
        }
    }
    if (bx < 0) {  // unreachable while n_ok > 0
        fx[i] = ux[i];
        fy[i] = uy[i];
        return;
    }
    const int by = col_row[(size_t)y * w + bx];
    const size_t j = (size_t)by * w + bx;
    const double ddx = (double)(x - bx), ddy = (double)(y - by);
    const double *g = slope + 4 * j;
    fx[i] = ux[j] + g[0] * ddx + g[2] * ddy;
    fy[i] = uy[j] + g[1] * ddx + g[3] * ddy;
}

// ---------------------------------------------------------------- components (union-find)
__device__ int find_root(int *parent, int x) {
    volatile int *p = parent;
    int r = x;
    while (true) {
        const int q = p[r];
        if (q == r) break;
        r = q;
    }
    // path compression: every node on the path above r gets r (a root when read) as parent. Every
    // write keeps parent[x] < x for a non-root, so no chain can close into a cycle. The walk stops
    // below r as well as at r: while it runs, r can be linked under a lower root and another
    // thread can compress part of the path straight to that root, so the walk can pass under r;
    // pointing a node there up at r would close a cycle (r -> ... -> node -> r) and every later
    // find_root on that component would never end. Roots are never written here, so the linking
    // CAS of unite stays the only write to a root.
    while (x > r) {
        const int q = p[x];
        if (q == x) break;
        if (q != r) p[x] = r;
        x = q;
    }
    return r;
}

__global__ void k_cc_init(const uint8_t *__restrict__ b, size_t n, int *__restrict__ parent,
                          unsigned int *__restrict__ sizes) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    parent[i] = b[i] ? (int)i : -1;
    sizes[i] = 0;
}

__device__ void unite(int *parent, int a, int c) {
    while (true) {
        a = find_root(parent, a);
        c = find_root(parent, c);
        if (a == c) return;
        if (a < c) {
            const int t = a;
            a = c;
            c = t;
        }
        // link the larger root below the smaller one; retry if it stopped being a root
        if (atomicCAS(&parent[a], a, c) == a) return;
    }
}

__global__ void k_cc_union(const uint8_t *__restrict__ b, int w, int h, int *parent) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const int i = y * w + x;
    if (!b[i]) return;
    if (x + 1 < w && b[i + 1]) unite(parent, i, i + 1);
    if (y + 1 < h && b[i + w]) unite(parent, i, i + w);
}

__global__ void k_cc_flatten(const uint8_t *__restrict__ b, size_t n, int *parent) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !b[i]) return;
    find_root(parent, (int)i);
}

// Component sizes; lanes of a wavefront that share a root add their count with one atomic (the
// border strips and large holes are single components of many thousand pixels).
__global__ void k_cc_count(const uint8_t *__restrict__ b, size_t n, int *parent, unsigned int *sizes) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (i < n && b[i]) ? find_root(parent, (int)i) : -1;
    const int lane = threadIdx.x & 63;
    unsigned long long active = __ballot(r >= 0);
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int rl = __shfl(r, leader);
        const unsigned long long same = __ballot(r == rl);
        if (lane == leader) atomicAdd(&sizes[rl], (unsigned int)__popcll(same));
        active &= ~same;
    }
}

__global__ void k_cc_keep(const uint8_t *__restrict__ b, size_t n, int *parent,
                          const unsigned int *__restrict__ sizes, double min_area, uint8_t *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!b[i]) {
        out[i] = 0;
        return;
    }
    const int r = find_root(parent, (int)i);
    out[i] = (double)sizes[r] >= min_area;
}

__global__ void k_guard_sources(const uint8_t *__restrict__ large, const uint8_t *__restrict__ mask, size_t n,
                                uint8_t *__restrict__ src) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    src[i] = large[i] || (mask && !mask[i]);
}

__global__ void k_low(const float *__restrict__ conf, double conf_min, size_t n, uint8_t *__restrict__ low) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    low[i] = (double)conf[i] < conf_min;
}

// track.py:66-98: invalid = low conf, outside the ROI, within the guard of a source, or within the
// guard of the image border
__global__ void k_validity(const uint8_t *__restrict__ low, const uint8_t *__restrict__ mask,
                           const uint8_t *__restrict__ near_src, double guard, int guard_on,
                           const double *__restrict__ ux, const double *__restrict__ uy, int w, int h,
                           double *__restrict__ mx, double *__restrict__ my) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    bool bad = low[i] || (mask && !mask[i]);
    if (guard_on) {
        bad = bad || near_src[i];
        const int bd = min(min(x, y), min(w - 1 - x, h - 1 - y));
        bad = bad || (double)bd < guard;
    }
    mx[i] = bad ? NAN : ux[i];
    my[i] = bad ? NAN : uy[i];
}

// ---------------------------------------------------------------- strain (strain.py:61-112)
// y pass with zero padding: 11 planes, m t^0..4 | u t^0..2 | v t^0..2 (m = valid, u, v zero-filled)
// R > 0: the radius at compile time (the loop unrolled, every load issued before the sums use
// them); R = 0: the radius r at run time.
template <int R>
__global__ void k_strain_y(const double *__restrict__ mx, const double *__restrict__ my, int w, int h, int r_rt,
                           double *__restrict__ out) {
    const int r = R > 0 ? R : r_rt;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t n = (size_t)w * h;
    double a[11];
#pragma unroll
    for (int q = 0; q < 11; q++) a[q] = 0;
    // row y + j of the window; rows outside the frame count as invalid pixels (zero terms, which
    // leave every sum as it is)
    auto add = [&](int j, double u, double v) {
        const int yy = y + j;
        const bool ok = yy >= 0 && yy < h && isfinite(u) && isfinite(v);
        const double m = ok ? 1.0 : 0.0, u0 = ok ? u : 0.0, v0 = ok ? v : 0.0;
        const double t = j, t2 = t * t;
        const double tp[5] = {1.0, t, t2, t2 * t, t2 * t2};
#pragma unroll
        for (int q = 0; q < 5; q++) a[q] += tp[q] * m;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            a[5 + q] += tp[q] * u0;
            a[8 + q] += tp[q] * v0;
        }
    };
    if constexpr (R > 0) {
        // every row of the window loaded (clamped to the frame) before the sums use them
        double us[2 * R + 1], vs[2 * R + 1];
#pragma unroll
        for (int k = 0; k <= 2 * R; k++) {
            const size_t i = (size_t)min(max(y - R + k, 0), h - 1) * w + x;
            us[k] = mx[i];
            vs[k] = my[i];
        }
#pragma unroll
        for (int k = 0; k <= 2 * R; k++) add(k - R, us[k], vs[k]);
    } else {
        for (int j = -r; j <= r; j++) {
            const size_t i = (size_t)min(max(y + j, 0), h - 1) * w + x;
            add(j, mx[i], my[i]);
        }
    }
    const size_t i = (size_t)y * w + x;
#pragma unroll
    for (int q = 0; q < 11; q++) out[q * n + i] = a[q];
}

// strain.rs:64-108: Cholesky with a relative pivot test, two right-hand sides. Each column's
// diagonal is kept as its reciprocal (one reciprocal square root, within an ulp or two of
// 1 / sqrt), so every division of the factorisation and of the two substitutions is a multiply.
__device__ bool solve6(double m[6][6], double rhs[6][2], double c[6][2]) {
    double l[6][6], il[6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) l[i][j] = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = m[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) d -= l[j][k] * l[j][k];
        if (isnan(d) || d <= 1e-12 * m[j][j]) return false;
        il[j] = rsqrt(d);
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double s = m[i][j];
#pragma unroll
            for (int k = 0; k < j; k++) s -= l[i][k] * l[j][k];
            l[i][j] = s * il[j];
        }
    }
#pragma unroll
    for (int col = 0; col < 2; col++) {
        double yv[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double s = rhs[i][col];
#pragma unroll
            for (int k = 0; k < i; k++) s -= l[i][k] * yv[k];
            yv[i] = s * il[i];
        }
#pragma unroll
        for (int i = 5; i >= 0; i--) {
            double s = yv[i];
#pragma unroll
            for (int k = i + 1; k < 6; k++) s -= l[k][i] * c[k][col];
            c[i][col] = s * il[i];
        }
    }
    return true;
}

// LDS = true: the 11 planes' row segments this workgroup reaches (its 256 columns and r on each
// side) are staged in LDS once, and the tap loop reads them there instead of making 11 (2r + 1)
// global loads per pixel (same values, same order of sums).
template <bool LDS>
__global__ void __launch_bounds__(256) k_strain_x(const double *__restrict__ ym, const double *__restrict__ mx,
                                                  const double *__restrict__ my, int w, int h, int r, int win,
                                                  float *ux_out, float *uy_out, uint8_t *valid, float *exx, float *eyy,
                                                  float *exy, float *gmax, float *vol) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const size_t n = (size_t)w * h;
    extern __shared__ double seg[];  // LDS: 11 planes x (256 + 2r)
    const int cols = 256 + 2 * r;
    if constexpr (LDS) {
        const int x0 = blockIdx.x * blockDim.x;
        for (int idx = threadIdx.x; idx < 11 * cols; idx += blockDim.x) {
            const int q = idx / cols, c = idx - q * cols;
            const int xx = x0 - r + c;
            seg[idx] = (xx >= 0 && xx < w) ? ym[(size_t)q * n + (size_t)y * w + xx] : 0.0;
        }
        __syncthreads();
    }
    if (x >= w) return;
    const double *s0 = seg + threadIdx.x + r;  // LDS: s0[q * cols + j] is plane q at column x + j
    double mm[5][5], mu[3][3], mv[3][3];
#pragma unroll
    for (int a = 0; a < 5; a++)
#pragma unroll
        for (int b = 0; b < 5; b++) mm[a][b] = 0;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) mu[a][b] = mv[a][b] = 0;
    const double *row = ym + (size_t)y * w;
    for (int j = -r; j <= r; j++) {
        const int xx = x + j;
        if (xx < 0 || xx >= w) continue;
        const double t = j, t2 = t * t;
        const double tp[5] = {1.0, t, t2, t2 * t, t2 * t2};
#pragma unroll
        for (int b = 0; b < 5; b++) {
            const double v = LDS ? s0[b * cols + j] : row[b * n + xx];
#pragma unroll
            for (int a = 0; a + b < 5; a++) mm[a][b] += tp[a] * v;
        }
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const double vu = LDS ? s0[(5 + b) * cols + j] : row[(5 + b) * n + xx];
            const double vv = LDS ? s0[(8 + b) * cols + j] : row[(8 + b) * n + xx];
#pragma unroll
            for (int a = 0; a + b < 3; a++) {
                mu[a][b] += tp[a] * vu;
                mv[a][b] += tp[a] * vv;
            }
        }
    }
    const int BA[6] = {0, 1, 0, 2, 1, 0}, BB[6] = {0, 0, 1, 0, 1, 2};
    double M[6][6], R[6][2], C[6][2];
#pragma unroll
    for (int i = 0; i < 6; i++) {
#pragma unroll
        for (int j = 0; j < 6; j++) M[i][j] = mm[BA[i] + BA[j]][BB[i] + BB[j]];
        R[i][0] = mu[BA[i]][BB[i]];
        R[i][1] = mv[BA[i]][BB[i]];
    }
    const double count = mm[0][0];
    // strain.rs:37-46: count >= win^2 / 2 decided on integers
    bool ok = 2.0 * rint(count) >= (double)win * (double)win;
    if (ok) ok = solve6(M, R, C);
    double a = NAN, b = NAN, c = NAN, d = NAN;
    if (ok) {
        a = C[1][0];
        b = C[2][0];
        c = C[1][1];
        d = C[2][1];
    }
    const size_t i = (size_t)y * w + x;
    const double Exx = a + 0.5 * (a * a + c * c);
    const double Eyy = d + 0.5 * (b * b + d * d);
    const double Exy = 0.5 * (b + c) + 0.5 * (a * b + c * d);
    const double hd = (Exx - Eyy) / 2;
    if (exx) exx[i] = (float)Exx;
    if (eyy) eyy[i] = (float)Eyy;
    if (exy) exy[i] = (float)Exy;
    if (gmax) gmax[i] = (float)(2.0 * sqrt(hd * hd + Exy * Exy));
    if (vol) vol[i] = (float)((1 + a) * (1 + d) - b * c - 1.0);
    const double u = mx[i], v = my[i];
    if (ux_out) ux_out[i] = (float)u;
    if (uy_out) uy_out[i] = (float)v;
    if (valid) valid[i] = isfinite(u) && isfinite(v);
}

size_t nblk(size_t n) { return (n + 255) / 256; }

}  // namespace

void column_nearest(const uint8_t *feature, int w, int h, int *col_d2, int *col_row, int *summary, hipStream_t s) {
    const int nch = (h + CHUNK - 1) / CHUNK;
    int *first = summary, *last = summary + (size_t)nch * w;
    hipLaunchKernelGGL(k_col_summary, dim3((w + 255) / 256, nch), dim3(256), 0, s, feature, w, h, first, last);
    hipLaunchKernelGGL(k_column_nearest, dim3((w + 255) / 256, nch), dim3(256), 0, s, feature, (const int *)first,
                       (const int *)last, w, h, col_d2, col_row);
    LAUNCHCHECK();
}

size_t column_summary_ints(int w, int h) { return 2 * (size_t)((h + CHUNK - 1) / CHUNK) * w; }

void within_distance(const int *col_d2, int w, int h, double thr, uint8_t *within, hipStream_t s) {
    const int T = thr >= 0 ? (int)std::floor(thr) : -1;
    if (T <= WITHIN_MAXT) {
        // the largest B with sqrt(B) <= thr, from floor(thr^2) corrected by the same correctly
        // rounded square root the device takes
        long long B = -1;
        if (thr >= 0) {
            B = (long long)std::floor(thr * thr);
            while (std::sqrt((double)(B + 1)) <= thr) B++;
            while (B >= 0 && std::sqrt((double)B) > thr) B--;
        }
        hipLaunchKernelGGL(k_within_lds, dim3((w + 255) / 256, h), dim3(256), 0, s, col_d2, w, (int)B, T, within);
    } else
        hipLaunchKernelGGL(k_within, grid2(w, h), dim3(256), 0, s, col_d2, w, h, thr, T, within);
    LAUNCHCHECK();
}

void prewarp_trusted(const uint8_t *low, const uint8_t *within_r, int w, int h, uint8_t *ok, Counters *cnt,
                     hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(1), 0, s, cnt);
    hipLaunchKernelGGL(k_trusted, dim3((unsigned)std::min<size_t>(1024, nblk(n))), dim3(256), 0, s, low, within_r, n, ok,
                       cnt);
    LAUNCHCHECK();
}

void prewarp_fill(const double *ux, const double *uy, const double *slope, const uint8_t *ok, const int *col_d2,
                  const int *col_row, const Counters *cnt, int w, int h, double *fx, double *fy, hipStream_t s) {
    hipLaunchKernelGGL(k_prewarp_fill, grid2(w, h), dim3(256), 0, s, ux, uy, slope, ok, col_d2, col_row, cnt, w, h, fx,
                       fy);
    LAUNCHCHECK();
}

void large_components(const uint8_t *b, int w, int h, double min_area, int *parent, unsigned int *sizes,
                      uint8_t *out, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_cc_init, dim3(nblk(n)), dim3(256), 0, s, b, n, parent, sizes);
    hipLaunchKernelGGL(k_cc_union, grid2(w, h), dim3(256), 0, s, b, w, h, parent);
    hipLaunchKernelGGL(k_cc_flatten, dim3(nblk(n)), dim3(256), 0, s, b, n, parent);
    hipLaunchKernelGGL(k_cc_count, dim3(nblk(n)), dim3(256), 0, s, b, n, parent, sizes);
    hipLaunchKernelGGL(k_cc_keep, dim3(nblk(n)), dim3(256), 0, s, b, n, parent, sizes, min_area, out);
    LAUNCHCHECK();
}

void guard_sources(const uint8_t *large, const uint8_t *mask, int w, int h, uint8_t *src, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_guard_sources, dim3(nblk(n)), dim3(256), 0, s, large, mask, n, src);
    LAUNCHCHECK();
}

void low_mask(const float *conf, double conf_min, size_t n, uint8_t *low, hipStream_t s) {
    hipLaunchKernelGGL(k_low, dim3(nblk(n)), dim3(256), 0, s, conf, conf_min, n, low);
    LAUNCHCHECK();
}

void validity(const uint8_t *low, const uint8_t *mask, const uint8_t *near_src, double guard, int guard_on,
              const double *ux, const double *uy, int w, int h, double *mx, double *my, hipStream_t s) {
    hipLaunchKernelGGL(k_validity, grid2(w, h), dim3(256), 0, s, low, mask, near_src, guard, guard_on, ux, uy, w, h, mx,
                       my);
    LAUNCHCHECK();
}

void strain(const double *mx, const double *my, int w, int h, int win, double *ytmp, float *ux_out, float *uy_out,
            uint8_t *valid, float *exx, float *eyy, float *exy, float *gmax, float *vol, hipStream_t s) {
    const int r = win / 2;
    if (r == 7)  // the default 15 px window
        hipLaunchKernelGGL(k_strain_y<7>, grid2(w, h), dim3(256), 0, s, mx, my, w, h, r, ytmp);
    else
        hipLaunchKernelGGL(k_strain_y<0>, grid2(w, h), dim3(256), 0, s, mx, my, w, h, r, ytmp);
    const size_t lds = (size_t)11 * (256 + 2 * r) * sizeof(double);
    if (lds <= 65536)
        hipLaunchKernelGGL(k_strain_x<true>, grid2(w, h), dim3(256), lds, s, ytmp, mx, my, w, h, r, win, ux_out, uy_out,
                           valid, exx, eyy, exy, gmax, vol);
    else
        hipLaunchKernelGGL(k_strain_x<false>, grid2(w, h), dim3(256), 0, s, ytmp, mx, my, w, h, r, win, ux_out, uy_out,
                           valid, exx, eyy, exy, gmax, vol);
    LAUNCHCHECK();
}

}  // namespace ofx
