// Temporal averaging (rust/crates/optflow-cli/src/tavg.rs:23-207, python/optflow_ref/tavg.py) and
// frame conversion (rust/crates/optflow-cli/src/io.rs:44-52, :259-281, python/optflow_ref/io.py:15-54).
//
// The mean is accumulated in f64 one frame per pass in the reference's order (the centre, then the
// others in window order), so it is bit-identical to the sequential sum. The affine motion field
// and the grey levels are computed with explicitly rounded products and sums (__dmul_rn,
// __dadd_rn), so no fused multiply-add changes a bit against the reference's left-to-right f64
// arithmetic. Each kernel is elementwise over the frame: HBM-bound, one read and one write per
// plane.
#include <algorithm>
#include <cmath>

#include "fine.hpp"

namespace ofx {

namespace {

__global__ void k_mean_add(double *__restrict__ acc, const float *__restrict__ f, size_t n, int first) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = (double)f[i];
    acc[i] = first ? v : __dadd_rn(acc[i], v);
}

__global__ void k_mean_finish(const double *__restrict__ acc, double count, float *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)__ddiv_rn(acc[i], count);
}

__global__ void k_affine_field(double cx0, double cx1, double cx2, double cy0, double cy1, double cy2, int w, int h,
                               float *__restrict__ ux, float *__restrict__ uy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const double fx = x, fy = y;
    const size_t i = (size_t)y * w + x;
    ux[i] = (float)__dadd_rn(__dadd_rn(cx0, __dmul_rn(cx1, fx)), __dmul_rn(cx2, fy));
    uy[i] = (float)__dadd_rn(__dadd_rn(cy0, __dmul_rn(cy1, fx)), __dmul_rn(cy2, fy));
}

__global__ void k_f64_to_f32(const double *__restrict__ in, float *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// One output pixel per thread: the factor x factor box of grey levels, each row summed left to
// right and the row sums added in order, then divided by factor^2 (io.rs:259-281). Grey level:
// v / 255 or v / 65535, colour 0.299 R + 0.587 G + 0.114 B left to right (io.rs:44-52).
template <typename T>
__device__ double grey(const T *__restrict__ px, size_t pix, int ch, double scale) {
    const T *p = px + pix * ch;
    if (ch < 3) return __ddiv_rn((double)p[0], scale);
    const double r = __ddiv_rn((double)p[0], scale), g = __ddiv_rn((double)p[1], scale),
                 b = __ddiv_rn((double)p[2], scale);
    return __dadd_rn(__dadd_rn(__dmul_rn(0.299, r), __dmul_rn(0.587, g)), __dmul_rn(0.114, b));
}

template <typename T>
__global__ void k_gray_down(const T *__restrict__ px, int ch, double scale, int w, int factor, int ow, int oh,
                            float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= ow || j >= oh) return;
    double v;
    if (factor == 1) {
        v = grey(px, (size_t)j * w + i, ch, scale);
    } else {
        double sum = 0.0;
        for (int dy = 0; dy < factor; dy++) {
            const size_t base = (size_t)(j * factor + dy) * w + (size_t)i * factor;
            double row = 0.0;
            for (int dx = 0; dx < factor; dx++) row = __dadd_rn(row, grey(px, base + dx, ch, scale));
            sum = __dadd_rn(sum, row);
        }
        v = __ddiv_rn(sum, (double)factor * (double)factor);
    }
    out[(size_t)j * ow + i] = (float)v;
}

dim3 blocks1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// np.median of a non-empty vector
double median_of(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return (n % 2) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// tavg.rs:35-57: elimination with partial pivoting on the augmented 3 x 4 system; the pivot is the
// last row of the largest magnitude; no solution when it is below 1e-12
bool solve3(double a[3][4], double x[3]) {
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int i = c + 1; i < 3; i++)
            if (std::fabs(a[i][c]) >= std::fabs(a[p][c])) p = i;
        if (std::fabs(a[p][c]) < 1e-12) return false;
        if (p != c) std::swap(a[c], a[p]);
        for (int i = c + 1; i < 3; i++) {
            const double f = a[i][c] / a[c][c];
            for (int j = c; j < 4; j++) a[i][j] -= f * a[c][j];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double s = a[i][3];
        for (int j = i + 1; j < 3; j++) s -= a[i][j] * x[j];
        x[i] = s / a[i][i];
    }
    return true;
}

// tavg.rs:62-82: v = c0 + c1 x + c2 y over the selected points, solved in coordinates centred on
// their mean
bool lstsq(const std::vector<double> &xs, const std::vector<double> &ys, const std::vector<double> &v,
           const std::vector<char> &sel, double c[3]) {
    double mx = 0, my = 0, cnt = 0;
    for (size_t i = 0; i < xs.size(); i++)
        if (sel[i]) {
            mx += xs[i];
            my += ys[i];
            cnt += 1;
        }
    mx /= cnt;
    my /= cnt;
    double m[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    for (size_t i = 0; i < xs.size(); i++) {
        if (!sel[i]) continue;
        const double row[3] = {1.0, xs[i] - mx, ys[i] - my};
        for (int r = 0; r < 3; r++) {
            for (int k = 0; k < 3; k++) m[r][k] += row[r] * row[k];
            m[r][3] += row[r] * v[i];
        }
    }
    double s[3];
    if (!solve3(m, s)) return false;
    c[0] = s[0] - s[1] * mx - s[2] * my;
    c[1] = s[1];
    c[2] = s[2];
    return true;
}

}  // namespace

bool fit_affine(const std::vector<double> &xs, const std::vector<double> &ys, const std::vector<double> &ux,
                const std::vector<double> &uy, double cx[3], double cy[3]) {
    const size_t n = xs.size();
    std::vector<char> sel(n, 1);
    bool have = false;
    for (int pass = 0; pass < 4; pass++) {
        if ((size_t)std::count(sel.begin(), sel.end(), 1) < 6) return false;
        double ax[3], ay[3];
        if (!lstsq(xs, ys, ux, sel, ax) || !lstsq(xs, ys, uy, sel, ay)) return false;
        std::copy(ax, ax + 3, cx);
        std::copy(ay, ay + 3, cy);
        have = true;
        std::vector<double> r(n), kept;
        for (size_t i = 0; i < n; i++) {
            r[i] = std::hypot(ax[0] + ax[1] * xs[i] + ax[2] * ys[i] - ux[i], ay[0] + ay[1] * xs[i] + ay[2] * ys[i] - uy[i]);
            if (sel[i]) kept.push_back(r[i]);
        }
        const double lim = std::max(3.0 * (1.4826 * median_of(kept)), 0.05);
        for (size_t i = 0; i < n; i++) sel[i] = r[i] <= lim;
    }
    return have;
}

void mean_add(double *acc, const float *frame, size_t n, bool first, hipStream_t s) {
    hipLaunchKernelGGL(k_mean_add, blocks1(n), dim3(256), 0, s, acc, frame, n, first ? 1 : 0);
    LAUNCHCHECK();
}

void mean_finish(const double *acc, double count, float *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_mean_finish, blocks1(n), dim3(256), 0, s, acc, count, out, n);
    LAUNCHCHECK();
}

void affine_field(const double cx[3], const double cy[3], int w, int h, float *ux, float *uy, hipStream_t s) {
    hipLaunchKernelGGL(k_affine_field, dim3((w + 255) / 256, h), dim3(256), 0, s, cx[0], cx[1], cx[2], cy[0], cy[1],
                       cy[2], w, h, ux, uy);
    LAUNCHCHECK();
}

void f64_to_f32(const double *in, float *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_f64_to_f32, blocks1(n), dim3(256), 0, s, in, out, n);
    LAUNCHCHECK();
}

void frame_to_gray(const void *pixels, int bits, int channels, int w, int h, int factor, float *out, hipStream_t s) {
    const int ow = w / factor, oh = h / factor;
    if (ow < 1 || oh < 1) return;
    const dim3 grid((ow + 255) / 256, oh);
    if (bits == 16)
        hipLaunchKernelGGL(k_gray_down<uint16_t>, grid, dim3(256), 0, s, (const uint16_t *)pixels, channels, 65535.0, w,
                           factor, ow, oh, out);
    else
        hipLaunchKernelGGL(k_gray_down<uint8_t>, grid, dim3(256), 0, s, (const uint8_t *)pixels, channels, 255.0, w,
                           factor, ow, oh, out);
    LAUNCHCHECK();
}

}  // namespace ofx
