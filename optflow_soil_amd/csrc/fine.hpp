// Host-side launchers of the path's kernels (one HIP stream per call).
#pragma once
#include "internal.hpp"

namespace ofx {

// fine.hip
// cubic B-spline coefficients of a real plane in place (tmp: w x h scratch)
void prefilter_real(float *c, float *tmp, int w, int h, hipStream_t s);
// The splines of two planes, src0 -> dst0 and src1 -> dst1 (sources kept), in the same launches;
// tmp: 2 w h floats.
void prefilter_real2(const float *src0, const float *src1, float *dst0, float *dst1, float *tmp, int w, int h,
                     hipStream_t s);
// taps_host: the 2r + 1 weights on the host (passed as a kernel argument)
void blur(const double *in, double *tmp, float *out, int w, int h, const double *taps_host, int r, hipStream_t s);
// out = I2((x + shx) + sx, (y + shy) + sy) from the spline c2 (w2 x h2) of image 2.
void warp(const float *c2, int w2, int h2, long shx, long shy, const float *sx, const float *sy, float *out, int w,
          int h, hipStream_t s);
// the reference pack of all F filters from the 3F inverses (r, rx, ry planes of filter f at f,
// F + f, 2F + f), one launch
void refpack(const float2 *resp, int F, float4 *out, int w, int h, int wp, hipStream_t s);
void accumulate(const double *ux, const double *uy, const float *sx, const float *sy, const float *csx,
                const float *csy, const float2 *resp, const float4 *refp, double *sums, int w, int h, int F,
                hipStream_t s);
void moments_solve(const double *sums, double *ym, const double *ux_in, const double *uy_in, double *ux_out,
                   double *uy_out, double *energy, double *resid, double *slope, int w, int h, const double *kt,
                   const double *kt_host, int r, hipStream_t s);
// shape_order 2: the same outputs from the 12 x 12 system; scratch of moments2_scratch_doubles(n)
// doubles; kt holds the taps times t^0 .. t^4
size_t moments2_scratch_doubles(size_t n);
void moments_solve2(const double *sums, double *scratch, const double *ux_in, const double *uy_in, double *ux_out,
                    double *uy_out, double *energy, double *resid, double *slope, int w, int h, const double *kt, int r,
                    hipStream_t s);
void energy_reset(const double *energy, const double *med, const double *ux_in, const double *uy_in, double *ux,
                  double *uy, double *resid, double *slope, size_t n, hipStream_t s);
size_t median_scratch_bytes();
void median(const double *v, size_t n, void *scratch, double *out, hipStream_t s);
// energy_reset with the threshold from two digit passes where every pixel is decided by them; the
// exact median (written to med) and the plain reset run only when some pixel is not. The results
// equal median + energy_reset; med is left as it was when the exact median was not needed.
void median_energy_reset(const double *energy, size_t n, void *scratch, double *med, const double *ux_in,
                         const double *uy_in, double *ux, double *uy, double *resid, double *slope, hipStream_t s);
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
void validity(const uint8_t *low, const uint8_t *mask, const uint8_t *near_src, double guard, int guard_on,
              const double *ux, const double *uy, int w, int h, double *mx, double *my, hipStream_t s);
// strain (strain.py:61-112) from the NaN-masked field; writes the outputs of a result
void strain(const double *mx, const double *my, int w, int h, int win, double *ytmp, float *ux_out, float *uy_out,
            uint8_t *valid, float *exx, float *eyy, float *exy, float *gmax, float *vol, hipStream_t s);
void low_mask(const float *conf, double conf_min, size_t n, uint8_t *low, hipStream_t s);

// poc.hip: coarse stage (poc.py) producing the dense start field
struct CoarseWs;
void coarse_dense(const float *i1, const float *i2, int w, int h, const ofx_config &c, const double *init_ux,
                  const double *init_uy, double *ux, double *uy, CoarseWs *&ws, hipStream_t s);
void coarse_release(CoarseWs *&ws);
// the start of a grid (x0 + j step, y0 + i step; nx x ny) into ix, iy (device), in pixels
using GuessFn = std::function<void(int x0, int y0, int step, int nx, int ny, double *ix, double *iy)>;
int coarse_dense_from(const float *i1, const float *i2, int w, int h, const ofx_config &c, const GuessFn &guess,
                      double tol, double *ux, double *uy, CoarseWs *&ws, int *confirmed, hipStream_t s);
void coarse_grid(const CoarseWs *ws, int w, int h, const ofx_config &c, std::vector<int> &xs, std::vector<int> &ys,
                 std::vector<double> &ux, std::vector<double> &uy, std::vector<double> &peak, PinnedBuf &pin,
                 hipStream_t s);

// tavg.hip: temporal averaging and frame conversion (tavg.rs, io.rs)
// acc (f64, n) = frame (first) or acc += frame (in the order of the calls)
void mean_add(double *acc, const float *frame, size_t n, bool first, hipStream_t s);
void mean_finish(const double *acc, double count, float *out, size_t n, hipStream_t s);
// ux = (float)(cx0 + cx1 x + cx2 y), likewise uy, every product and sum rounded as written
void affine_field(const double cx[3], const double cy[3], int w, int h, float *ux, float *uy, hipStream_t s);
void f64_to_f32(const double *in, float *out, size_t n, hipStream_t s);
// pixels (h x w x channels, 8 or 16 bit) -> grey levels, box downscaled by factor
void frame_to_gray(const void *pixels, int bits, int channels, int w, int h, int factor, float *out, hipStream_t s);
// tavg.rs:87-110: the robust affine fit (host)
bool fit_affine(const std::vector<double> &xs, const std::vector<double> &ys, const std::vector<double> &ux,
                const std::vector<double> &uy, double cx[3], double cy[3]);

// seq.hip: sequence helpers (track.py:231-328)
void compose(const double *ax, const double *ay, const double *bx, const double *by, int w, int h, double *ox,
             double *oy, hipStream_t s);
void conf_at(const double *conf, const double *ux, const double *uy, int w, int h, double *out, hipStream_t s);
void mean_conf(const double *conf, const double *raw_ux, const double *raw_uy, const double *ux, const double *uy,
               const uint8_t *mask, double B, int w, int h, double *acc /* [sum, count] */, hipStream_t s);

// hybrid.hip: seq_mode "hybrid" (track/hybrid.rs) -- per-pixel flags and counters
enum : uint8_t {
    HYB_DIRECT_OK = 1, HYB_GOOD = 2, HYB_CAND_RE = 4, HYB_NEED = 8,
    HYB_HAS_FB = 16, HYB_RE_NOFB = 32, HYB_RE = 64, HYB_FB_CAND = 128
};
enum {
    HYB_ANY_NEED = 0, HYB_ANY_RE_NOFB, HYB_ANY_CAND, HYB_NFB_MAX, HYB_N_FB_OK, HYB_N_REJECTED, HYB_N_RE, HYB_N_ROI,
    HYB_NCOUNTERS
};
void hyb_zero_counters(int *counters, hipStream_t s);
void hyb_good_of(const uint8_t *flags, uint8_t *good, size_t n, hipStream_t s);
void hyb_refmove(const uint8_t *good_prev, const double *cp, double *c0r, uint8_t *anch, size_t n, hipStream_t s);
void hyb_classify(const double *cx, const double *cy, const double *leg_c, const double *c0r, const uint8_t *anch,
                  const uint8_t *mask, const double *upx, const double *upy, const int *nfb, double conf_min,
                  int fb_max, int reacq_ok, double *link, uint8_t *flags, int *counters, size_t n, hipStream_t s);
void hyb_carried(const double *cax, const double *cay, const double *upx, const double *upy, double *ax, double *ay,
                 size_t n, hipStream_t s);
void hyb_has_fb(uint8_t *flags, const double *ca, double conf_min, const uint8_t *mask, int *counters, size_t n,
                hipStream_t s);
void hyb_reacquire(uint8_t *flags, const uint8_t *gate0, const double *cx, const double *cy, const double *ax,
                   const double *ay, const double *leg_c, const double *ca, double fb_tol, double conf_min, int fb_max,
                   double *c0r, uint8_t *anch, int *nfb, const uint8_t *mask, int *counters, size_t n, hipStream_t s);
void hyb_output(const uint8_t *flags, const uint8_t *gate, const double *cx, const double *cy, const double *ax,
                const double *ay, const double *link, const double *ca, const int *nfb, int fb_max,
                const uint8_t *mask, double *ux, double *uy, float *conf, uint8_t *low, int32_t *fb, int *counters,
                size_t n, hipStream_t s);
// the local affine gate (hybrid.rs:147-260) of the direct-good ROI pixels (flags & HYB_GOOD,
// mask) at targets (tx, ty); scratch of gate_scratch_bytes
size_t gate_scratch_bytes(int w, int h);
void hyb_gate(const uint8_t *flags, const uint8_t *mask, const double *cx, const double *cy, const double *tx,
              const double *ty, int w, int h, double radius, long min_support, double tol, void *scratch,
              uint8_t *ok, hipStream_t s);

}  // namespace ofx
