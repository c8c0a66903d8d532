/* FP64 CPU oracle of the optflow-soil dense flow path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in plain double precision, the algorithm of
 * the reference's Python specification (python/optflow_ref/{interp,phaseflow,poc,strain,track}.py)
 * with the Rust port's deviations where they are documented (Cholesky solve, odd-window grid end,
 * exact integer strain window test). It is the checker for the HIP path and the timed CPU
 * baseline; nothing in the product links, loads or calls it.
 *
 * Conventions: images are row-major h x w (index y * w + x), coordinates are (x = column,
 * y = row), complex arrays are interleaved (re, im) doubles.
 */
#ifndef OFX_ORACLE_H
#define OFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SCALES 16

/* Settings the path consumes (python/optflow_ref/config.py:24-120). */
typedef struct orc_cfg {
    int grid_step;
    int poc_win;
    double max_disp;
    double poc_peak_min;
    double median_thresh;
    double lam;
    int n_scales;
    double scales[ORC_MAX_SCALES];
    int n_orient;
    double bandwidth;
    double dc_cut;
    double sigma_w;
    int shape_order;
    double sigma_s;
    int outer_iters;
    int inner_iters;
    double conf_min;
    double edge_guard;
    double edge_guard_min_area;
    double phase_tol;
    int strain_win;
    double ref_update_conf;
    int seq_hybrid; /* 0 chain, 1 hybrid */
    int fb_max;
    double fb_radius;
    int fb_min_support;
    double fb_tol;
    int prewarp_freeze_at;
    int tile; /* 0 whole frame; N: tiles of N px for frames with a side over N (tile.rs) */
} orc_cfg;

/* ---- FFT (any length; mixed radix with direct DFT for prime factors) ---- */
void orc_fft2(const double *in, int h, int w, int inverse, double *out);
void orc_fft1(const double *in, int n, int inverse, double *out);

/* ---- interpolation (interp.py) ---- */
void orc_bspline_coeffs(const double *in, int h, int w, double *out);
void orc_bspline_coeffs_c(const double *in, int h, int w, double *out); /* complex */
void orc_sample(const double *coeffs, int h, int w, const double *x, const double *y, long n,
                double *out);
void orc_warp(const double *img, int h, int w, const double *ux, const double *uy, double *out);
void orc_gauss_blur(const double *in, int h, int w, double sigma, double *out);

/* ---- Gabor bank and reference phase (phaseflow.py:73-111) ---- */
int orc_n_filters(const orc_cfg *cfg);
void orc_gabor_bank(int hp, int wp, const orc_cfg *cfg, double *bank /* F x hp x wp */);
void orc_reference_phase(const double *i1, int h, int w, const orc_cfg *cfg, double *r_re,
                         double *r_im, double *gx, double *gy, double *amp);

/* ---- fine stage (phaseflow.py:114-511) ---- */
double orc_median(const double *v, long n);
void orc_local_solve(const double *sums /* 7 x n */, const double *ux, const double *uy, int h,
                     int w, double sigma, int order, double *oux, double *ouy, double *energy,
                     double *resid, double *slope /* n x 4 */);
int orc_refine_shifted(const double *i1, int h, int w, const double *i2, int h2, int w2, long shx,
                       long shy, const double *ux0, const double *uy0, const orc_cfg *cfg, double *ux,
                       double *uy, double *conf, double *resid, double *energy);
int orc_refine(const double *i1, const double *i2, int h, int w, const double *ux0,
               const double *uy0, const orc_cfg *cfg, double *ux, double *uy, double *conf,
               double *resid, double *energy);

/* ---- coarse stage (poc.py) ---- */
void orc_poc_shift(const double *a, const double *b, int h, int w, double *dx, double *dy,
                   double *peak);
int orc_level_window(int h, int w, int win, int *wy, int *wx);
int orc_n_levels(double max_disp, int win, int h, int w);
void orc_grid_dims(int h, int w, const orc_cfg *cfg, int *nx, int *ny);
int orc_coarse_flow(const double *i1, const double *i2, int h, int w, const orc_cfg *cfg,
                    const double *init_ux, const double *init_uy, double *gux, double *guy,
                    double *peak);
/* the dense start field of track.py:131-137 (coarse_flow's grid at every pixel), h x w each */
void orc_coarse_dense(const double *i1, const double *i2, int h, int w, const orc_cfg *cfg,
                      const double *init_ux, const double *init_uy, double *ux, double *uy);

/* ---- temporal averaging and frame conversion (tavg.rs, io.rs) ---- */
int orc_fit_affine(const double *xs, const double *ys, const double *ux, const double *uy, long n,
                   double *cx, double *cy);
int orc_align_to(const double *centre, const double *other, int h, int w, int mode, const uint8_t *mask,
                 const orc_cfg *cfg, double *out);
int orc_tavg_window(const double *frames, int nf, int centre, int h, int w, int mode, const uint8_t *mask,
                    int keep_unaligned, const orc_cfg *cfg, double *out);
void orc_to_gray(const void *pixels, int bits, int channels, int h, int w, int factor, double *out);

/* ---- distance transform and components (scipy.ndimage semantics) ---- */
void orc_edt(const uint8_t *input, int h, int w, double *dist, int32_t *nearest);
void orc_large_components(const uint8_t *b, int h, int w, double min_area, uint8_t *out);

/* ---- strain (strain.py:61-112) ---- */
void orc_displacement_gradient(const double *ux, const double *uy, int h, int w, int win,
                               double *dudx, double *dudy, double *dvdx, double *dvdy);

/* ---- pair and sequence (track.py) ---- */
typedef struct orc_result {
    double *ux, *uy, *conf, *raw_ux, *raw_uy;
    double *exx, *eyy, *exy, *gmax, *vol;
    int32_t *fb; /* hybrid only (may be NULL) */
    double t_coarse, t_fine, t_strain;
    int reference;
    double mean_conf;
    /* hybrid only: timing shares of track/hybrid.rs:520-531 */
    int n_fb_max;
    double fallback_frac, fb_gate_rejected, reacquired;
} orc_result;

/* hybrid helpers: zero-padded (2r+1)^2 window sums and the local affine gate (track/hybrid.rs) */
void orc_box_sum(const double *a, int h, int w, int r, double *out);
void orc_local_affine_agreement(const double *ux, const double *uy, const uint8_t *support, const double *tx,
                                const double *ty, int h, int w, double radius, long min_support, double tol,
                                uint8_t *ok);
int orc_track_pair(const double *i1, const double *i2, int h, int w, const uint8_t *mask,
                   const double *init_ux, const double *init_uy, const orc_cfg *cfg,
                   orc_result *res);
/* frames: n x h x w; results: n - 1 entries, frame k = 1..n-1 */
int orc_track_sequence(const double *frames, int n, int h, int w, const uint8_t *mask,
                       const orc_cfg *cfg, orc_result *res);

const char *orc_last_error(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
