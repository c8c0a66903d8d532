// Host-side launchers of the path's kernels (one HIP stream per call).
#pragma once
#include "internal.hpp"

namespace ofx {

// fine.hip
// cubic B-spline coefficients of a real plane in place (tmp: w x h scratch)
void prefilter_real(float *c, float *tmp, int w, int h, hipStream_t s);
// taps_host: the 2r + 1 weights on the host (passed as a kernel argument)
void blur(const double *in, double *tmp, float *out, int w, int h, const double *taps_host, int r, hipStream_t s);
// out = I2((x + shx) + sx, (y + shy) + sy) from the spline c2 (w2 x h2) of image 2.
void warp(const float *c2, int w2, int h2, long shx, long shy, const float *sx, const float *sy, float *out, int w,
          int h, hipStream_t s);
void refpack(const float2 *r, const float2 *rx, const float2 *ry, float4 *out, int w, int h, int wp, hipStream_t s);
void accumulate(const double *ux, const double *uy, const float *sx, const float *sy, const float *csx,
                const float *csy, const float2 *resp, const float4 *refp, double *sums, int w, int h, int F,
                hipStream_t s);
void moments_solve(const double *sums, double *ym, const double *ux_in, const double *uy_in, double *ux_out,
                   double *uy_out, double *energy, double *resid, double *slope, int w, int h, const double *kt,
                   const double *kt_host, int r, hipStream_t s);
void energy_reset(const double *energy, const double *med, const double *ux_in, const double *uy_in, double *ux,
                  double *uy, double *resid, double *slope, size_t n, hipStream_t s);
size_t median_scratch_bytes();
void median(const double *v, size_t n, void *scratch, double *out, hipStream_t s);
void confidence(const double *energy, const double *resid, const double *med, const double *ux, const double *uy,
                double phase_tol, double conf_min, double B, int w, int h, uint8_t *low, float *conf_out,
                double *conf64, hipStream_t s);

// guard.hip: distance transforms, components, pre-warp fill, validity, strain
struct Counters {
    unsigned int n_low, n_ok;
};
// Exact 1-D nearest feature per column (squared distance, row), -1 where the column has none.
void column_nearest(const uint8_t *feature, int w, int h, int *col_d2, int *col_row, int *summary, hipStream_t s);
size_t column_summary_ints(int w, int h);
// within[i] = (Euclidean distance to the nearest feature) <= thr, from column_nearest's output.
void within_distance(const int *col_d2, int w, int h, double thr, uint8_t *within, hipStream_t s);
// pre-warp (phaseflow.py:326-355): counts, trusted mask, fill of the untrusted pixels
void prewarp_trusted(const uint8_t *low, const uint8_t *within_r, int w, int h, uint8_t *ok, Counters *cnt,
                     hipStream_t s);
void prewarp_fill(const double *ux, const double *uy, const double *slope, const uint8_t *ok, const int *col_d2,
                  const int *col_row, const Counters *cnt, int w, int h, double *fx, double *fy, hipStream_t s);
// 4-connected components of b with at least min_area pixels (scipy.ndimage.label default)
void large_components(const uint8_t *b, int w, int h, double min_area, int *parent, unsigned int *sizes,
                      uint8_t *out, hipStream_t s);
// guard sources: large low components or outside the ROI
void guard_sources(const uint8_t *large, const uint8_t *mask, int w, int h, uint8_t *src, hipStream_t s);
// validity (track.py:56-108) and NaN-masked displacement as double planes for the strain
void validity(const float *conf, double conf_min, const uint8_t *mask, const uint8_t *near_src, double guard,
              int guard_on, const double *ux, const double *uy, int w, int h, double *mx, double *my, hipStream_t s);
// strain (strain.py:61-112) from the NaN-masked field; writes the outputs of a result
void strain(const double *mx, const double *my, int w, int h, int win, double *ytmp, float *ux_out, float *uy_out,
            uint8_t *valid, float *exx, float *eyy, float *exy, float *gmax, float *vol, hipStream_t s);
void low_mask(const float *conf, double conf_min, size_t n, uint8_t *low, hipStream_t s);

// poc.hip: coarse stage (poc.py) producing the dense start field
struct CoarseWs;
void coarse_dense(const float *i1, const float *i2, int w, int h, const ofx_config &c, const double *init_ux,
                  const double *init_uy, double *ux, double *uy, CoarseWs *&ws, hipStream_t s);
void coarse_release(CoarseWs *&ws);

// seq.hip: sequence helpers (track.py:231-328)
void compose(const double *ax, const double *ay, const double *bx, const double *by, int w, int h, double *ox,
             double *oy, hipStream_t s);
void conf_at(const double *conf, const double *ux, const double *uy, int w, int h, double *out, hipStream_t s);
void mean_conf(const double *conf, const double *raw_ux, const double *raw_uy, const double *ux, const double *uy,
               const uint8_t *mask, double B, int w, int h, double *acc /* [sum, count] */, hipStream_t s);

}  // namespace ofx
