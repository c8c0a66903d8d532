// Sequence helpers on the GPU (python/optflow_ref/track.py:231-328;
// rust/crates/optflow-core/src/track.rs:866-1023): bilinear sampling with NaN outside, composition
// of displacement fields, link confidence and the mean confidence of a leg.
#include <cmath>

#include "fine.hpp"

namespace ofx {

namespace {

// track.py:231-242
__device__ double sample_bilinear(const double *a, int w, int h, double x, double y) {
    if (!(isfinite(x) && isfinite(y)) || x < 0 || x > w - 1 || y < 0 || y > h - 1) return NAN;
    const int x0 = min((int)floor(x), w - 2), y0 = min((int)floor(y), h - 2);
    const double tx = x - x0, ty = y - y0;
    const double *r0 = a + (size_t)y0 * w, *r1 = r0 + w;
    return (1 - ty) * ((1 - tx) * r0[x0] + tx * r0[x0 + 1]) + ty * ((1 - tx) * r1[x0] + tx * r1[x0 + 1]);
}

__global__ void k_compose(const double *ax, const double *ay, const double *bx, const double *by, int w, int h,
                          double *ox, double *oy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const double px = x + ax[i], py = y + ay[i];
    const double vx = sample_bilinear(bx, w, h, px, py), vy = sample_bilinear(by, w, h, px, py);
    ox[i] = ax[i] + vx;
    oy[i] = ay[i] + vy;
}

__global__ void k_conf_at(const double *conf, const double *ux, const double *uy, int w, int h, double *out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const double v = sample_bilinear(conf, w, h, x + ux[i], y + uy[i]);
    out[i] = isfinite(v) ? v : 0.0;
}

// reportable (phaseflow.py:297-317) as 0/1 doubles
__global__ void k_reportable(const double *raw_ux, const double *raw_uy, double B, int w, int h, double *rep) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const double X = x + raw_ux[i], Y = y + raw_uy[i];
    const double b1 = fmin(fmin((double)x, (double)y), fmin((double)(w - 1 - x), (double)(h - 1 - y)));
    const double b2 = fmin(fmin(X, Y), fmin((double)(w - 1) - X, (double)(h - 1) - Y));
    rep[i] = (b1 >= B && b2 >= B) ? 1.0 : 0.0;
}

// track.py:317-328: sum and count of the leg's confidence at X + u0r over the reportable ROI
// pixels (the per-row partial sums are added in row order on the host side of the reduction)
__global__ void __launch_bounds__(256) k_mean_conf_rows(const double *conf, const double *rep, const double *ux,
                                                        const double *uy, const uint8_t *mask, int w, int h,
                                                        double *row_sum, double *row_cnt) {
    __shared__ double ss[256], sc[256];
    const int y = blockIdx.x;
    double s = 0, c = 0;
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
        const size_t i = (size_t)y * w + x;
        const double X = x + ux[i], Y = y + uy[i];
        const double cv = sample_bilinear(conf, w, h, X, Y);
        const double rv = sample_bilinear(rep, w, h, X, Y);
        if (isfinite(cv) && rv > 0.5 && (!mask || mask[i])) {
            s += cv;
            c += 1;
        }
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            ss[threadIdx.x] += ss[threadIdx.x + off];
            sc[threadIdx.x] += sc[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        row_sum[y] = ss[0];
        row_cnt[y] = sc[0];
    }
}

__global__ void k_mean_conf_total(const double *row_sum, const double *row_cnt, int h, double *acc) {
    if (threadIdx.x != 0) return;
    double s = 0, c = 0;
    for (int y = 0; y < h; y++) {
        s += row_sum[y];
        c += row_cnt[y];
    }
    acc[0] = s;
    acc[1] = c;
}

}  // namespace

void compose(const double *ax, const double *ay, const double *bx, const double *by, int w, int h, double *ox,
             double *oy, hipStream_t s) {
    hipLaunchKernelGGL(k_compose, dim3((w + 255) / 256, h), dim3(256), 0, s, ax, ay, bx, by, w, h, ox, oy);
    LAUNCHCHECK();
}

void conf_at(const double *conf, const double *ux, const double *uy, int w, int h, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_conf_at, dim3((w + 255) / 256, h), dim3(256), 0, s, conf, ux, uy, w, h, out);
    LAUNCHCHECK();
}

// acc must hold 2 + 3 h + w h doubles of scratch: [sum, count], row sums, row counts, reportable map
void mean_conf(const double *conf, const double *raw_ux, const double *raw_uy, const double *ux, const double *uy,
               const uint8_t *mask, double B, int w, int h, double *acc, hipStream_t s) {
    double *row_sum = acc + 2, *row_cnt = acc + 2 + h, *rep = acc + 2 + 2 * h;
    hipLaunchKernelGGL(k_reportable, dim3((w + 255) / 256, h), dim3(256), 0, s, raw_ux, raw_uy, B, w, h, rep);
    hipLaunchKernelGGL(k_mean_conf_rows, dim3(h), dim3(256), 0, s, conf, rep, ux, uy, mask, w, h, row_sum, row_cnt);
    hipLaunchKernelGGL(k_mean_conf_total, dim3(1), dim3(64), 0, s, row_sum, row_cnt, h, acc);
    LAUNCHCHECK();
}

}  // namespace ofx
