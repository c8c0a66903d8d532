/* ofx.h -- C ABI of the MI355X dense flow engine (libofx.so).
 *
 * The drop-in boundary of the reference's fine-stage plugin: the `FineEngine` trait
 * (rust/crates/optflow-core/src/engine.rs:31-125) with its error type `EngineError(String)`
 * (engine.rs:22-24), extended by whole-pair and sequence entry points that run the coarse stage,
 * confidence, guard and strain on the GPU as well (the reference runs those on the CPU,
 * rust/README.md:137-138).
 *
 * Conventions
 *  - Images are row-major, width w (columns, x) by height h (rows, y), index y * w + x; the
 *    reference's Image<f32> (rust/crates/optflow-core/src/image.rs:1-26).
 *  - Every pointer argument is caller-owned. Functions named *_device take device (HBM) pointers
 *    valid on the engine's device; all others take host pointers. No pointer is kept after a call
 *    returns.
 *  - Every function returns OFX_OK (0) or an error status; the message of the last failure on
 *    the calling thread is ofx_last_error(). There is no CPU fallback: a failure is reported,
 *    never hidden (rust/README.md:118-122).
 *  - One engine runs one call at a time per lane; a call that finds every lane busy fails with
 *    OFX_EBUSY instead of waiting (rust/crates/optflow-gpu/src/engine.rs:664-681).
 */
#ifndef OFX_H
#define OFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OFX_OK 0
#define OFX_EINVAL 1    /* bad argument or configuration (the reference's validate messages) */
#define OFX_ETOOLARGE 2 /* window exceeds device limits (GpuError::TooLarge, gpu/src/lib.rs:14-31) */
#define OFX_EDEVICE 3   /* HIP runtime error (GpuError::Command) */
#define OFX_ENODEVICE 4 /* no usable HIP device (GpuError::NoDevice) */
#define OFX_EBUSY 5     /* every lane of the engine is in use */
#define OFX_EUNSUPPORTED 6

#define OFX_MAX_SCALES 16

/* Settings of the path: the fields and defaults of python/optflow_ref/config.py:24-90 and
 * rust/crates/optflow-core/src/config.rs:487-557 that the engine consumes. */
typedef struct ofx_config {
    int grid_step;            /* 16 */
    int poc_win;              /* 64 */
    double max_disp;          /* 32 */
    double poc_peak_min;      /* 0.1 */
    double median_thresh;     /* 2.0 */
    double lam;               /* 8 */
    int n_scales;             /* 3 */
    double scales[OFX_MAX_SCALES]; /* 0.6, 1.0, 1.6 */
    int n_orient;             /* 4 */
    double bandwidth;         /* 0.5 */
    double dc_cut;            /* 2.0 */
    double sigma_w;           /* 4 */
    int shape_order;          /* 1: affine, 2: quadratic shape function */
    double sigma_s;           /* 4 */
    int outer_iters;          /* 8 */
    int inner_iters;          /* 2 */
    double conf_min;          /* 0.2 */
    double edge_guard;        /* -1: automatic */
    double edge_guard_min_area; /* -1: automatic */
    double phase_tol;         /* 0.4 */
    int strain_win;           /* 15 */
    double ref_update_conf;   /* 0.5 */
    int seq_hybrid;           /* 0: seq_mode "chain", 1: "hybrid" */
    int fb_max;               /* 10 */
    double fb_radius;         /* -1 */
    int fb_min_support;       /* 20 */
    double fb_tol;            /* 0.5 */
    int prewarp_freeze_at;    /* 0: off */
    int tile;                 /* 0: whole frame; N >= 64: N x N tiles for frames with a side over N
                                 (the reference's tiled fine stage with N = 1024, tile.rs:56-98);
                                 the tiles run in three groups side by side, on the lane's stream
                                 and two more streams of the engine */
    int coarse_learned;       /* 0: the POC pyramid (the reference's coarse stage); 1: the level-0
                                 POC grid started from the learned coarse proposal
                                 (ofx_coarse_net_load), with the pyramid as the fallback */
    double coarse_gate;       /* 0.9: share of the grid points that must confirm the proposal */
} ofx_config;

typedef struct ofx_engine ofx_engine;

/* Fills *cfg with the reference defaults. */
void ofx_config_default(ofx_config *cfg);

/* Engine on HIP device `device` with `lanes` concurrent calls (one HIP stream each; the
 * reference's OPTFLOW_GPU_LANES, gpu/src/engine.rs:72-95). Replaces GpuEngine::new. */
int ofx_create(int device, int lanes, ofx_engine **out);
void ofx_destroy(ofx_engine *e);

/* Message of the last failed call on this thread (EngineError's text). */
const char *ofx_last_error(void);

/* FineEngine::lanes (engine.rs:68-75). */
int ofx_lanes(const ofx_engine *e);

/* Per-stage GPU time, the GPU counterpart of the reference's OPTFLOW_GPU_PROFILE=1 stage timer
 * (rust/crates/optflow-gpu/src/engine.rs:111; stage names of rust/crates/optflow-core/src/prof.rs:17-44).
 * ofx_profile(e, 1) starts recording (and zeroes the totals) with one HIP event per stage boundary
 * on each lane's stream; ofx_profile_read sums the milliseconds and the number of times each
 * stage ran over all calls since; ofx_profile_stage(i) names stage i (NULL past the last). */
int ofx_profile(ofx_engine *e, int enable);
int ofx_profile_read(ofx_engine *e, double *ms, int64_t *calls, int n);
const char *ofx_profile_stage(int i);

/* FineEngine::returns_slope (engine.rs:82-86): 0, the slope stays on the device. */
int ofx_returns_slope(const ofx_engine *e);

/* FineEngine::check_window (engine.rs:42-46; gpu/src/engine.rs:1014-1071): OFX_ETOOLARGE before
 * any work if a w x h window does not fit the device, OFX_EINVAL for a bad configuration. */
int ofx_check_window(ofx_engine *e, int w, int h, const ofx_config *cfg);

/* FineEngine::release_caches (engine.rs:114-117): frees the per-size device workspaces. */
int ofx_release_caches(ofx_engine *e);

/* FineEngine::refine_raw (engine.rs:31-41, contract of phaseflow.rs:1441-1456, output RawFlow of
 * phaseflow.rs:1408-1414): the fine stage of i1 (w x h) from the start field (ux, uy, w x h)
 * against i2, any w2 x h2 window of the second frame in which pixel (x, y) of i1 sits at
 * (x + shift_x, y + shift_y) before the field displaces it (positions outside i2 are clamped to
 * its border, as the reference's spline does). w2, h2 >= 1 and <= 7680 (OFX_ETOOLARGE above).
 * Outputs are f64, w x h each; energy and resid are genuine f64 (planes_exact_in_f32 = false). */
int ofx_refine_raw(ofx_engine *e, const float *i1, int w, int h, const float *i2, int w2, int h2,
                   int64_t shift_x, int64_t shift_y, const float *ux, const float *uy,
                   const ofx_config *cfg, double *out_ux, double *out_uy, double *out_energy,
                   double *out_resid);

/* The coarse stage alone: coarse_flow (rust/crates/optflow-core/src/poc.rs:694-766,
 * python/optflow_ref/poc.py:216-239) on the pyramid of i1, i2 (w x h each), then the dense start
 * field of track_pair (track.rs:496-554, track.py:131-137): the grid interpolated to every
 * pixel. init_ux/init_uy: dense guess or NULL (both or neither). Outputs f64, w x h. Host
 * pointers. */
int ofx_coarse_flow(ofx_engine *e, const float *i1, const float *i2, int w, int h, const float *init_ux,
                    const float *init_uy, const ofx_config *cfg, double *out_ux, double *out_uy);

/* Outputs of one pair or sequence frame: the reference's TrackResult fields and .npy outputs
 * (rust/crates/optflow-core/src/track.rs:157-209, cli/io.rs:558-600). Any pointer may be NULL
 * (not written). ux, uy: NaN where invalid; valid: 1 where ux is finite; raw_*: unmasked
 * field; fb: hybrid sequences only. */
typedef struct ofx_result {
    float *ux, *uy, *conf, *raw_ux, *raw_uy;
    float *exx, *eyy, *exy, *gmax, *vol;
    uint8_t *valid;
    int32_t *fb;
    double t_coarse, t_fine, t_strain; /* seconds (Timing, track.rs:148-161) */
    int reference;                     /* sequences: reference frame index */
    double mean_conf;                  /* sequences: mean confidence of the leg */
    /* hybrid sequences: the timing shares of track/hybrid.rs:520-531 (over ROI pixels) */
    int n_fb_max;                      /* most consecutive frames any pixel has been carried */
    double fallback_frac;              /* valid by an accepted fallback */
    double fb_gate_rejected;           /* fallback candidates the gate rejected */
    double reacquired;                 /* re-acquired by the direct leg */
    /* pairs with coarse_learned: the share of the level-0 grid points that confirmed the learned
     * proposal, and 1 when that share was below coarse_gate and the POC pyramid ran instead;
     * -1 and 0 when the pyramid ran by configuration */
    double coarse_confirmed;
    int coarse_fallback;
} ofx_result;

/* track_pair (rust/crates/optflow-core/src/track.rs:688-861; python/optflow_ref/track.py:115-146):
 * coarse POC, fine stage, confidence, guard and strain for I1 -> I2 (both w x h). mask: ROI
 * (nonzero = inside) or NULL; init_ux/init_uy: dense start guess or NULL. Host pointers. */
int ofx_track_pair(ofx_engine *e, const float *i1, const float *i2, int w, int h,
                   const uint8_t *mask, const float *init_ux, const float *init_uy,
                   const ofx_config *cfg, ofx_result *out);

/* One pair of a batch: the host arguments of ofx_track_pair. */
typedef struct ofx_pair {
    const float *i1, *i2;
    int w, h;
    const uint8_t *mask;             /* ROI or NULL */
    const float *init_ux, *init_uy;  /* dense start guess or NULL */
} ofx_pair;

/* Independent pairs of any sizes (many uneven pairs at once; no reference counterpart, the
 * reference runs them one by one through track_pair, track.rs:688-861): the pairs are spread over
 * the engine's free lanes, each lane running one pair at a time on its own stream, largest
 * first. results[i] is what ofx_track_pair writes for pairs[i], bit for bit. Every pair is
 * checked before any work; a pair that fails at run time does not stop the others, and the call
 * returns the first failure with the pair's index in the message. */
int ofx_track_pairs(ofx_engine *e, int n, const ofx_pair *pairs, const ofx_config *cfg, ofx_result *results);

/* The same with every image, mask, guess and output array in device memory (HBM). */
int ofx_track_pair_device(ofx_engine *e, const float *d_i1, const float *d_i2, int w, int h,
                          const uint8_t *d_mask, const float *d_init_ux, const float *d_init_uy,
                          const ofx_config *cfg, ofx_result *d_out);

/* track_sequence (track.rs:1112-1225; track.py:260-315) in seq_mode "chain", or "hybrid" when
 * cfg->seq_hybrid (direct-first with the gated adjacent-frame fallback, track/hybrid.rs:262-540;
 * track.py:383-505; results carry fb and the hybrid timing shares): frames is n x h x w (host),
 * results has n - 1 entries for frames 1 .. n - 1. */
int ofx_track_sequence(ofx_engine *e, const float *frames, int n, int w, int h,
                       const uint8_t *mask, const ofx_config *cfg, ofx_result *results);

/* Temporal averaging (rust/crates/optflow-cli/src/tavg.rs:118-207; python/optflow_ref/tavg.py:30-96):
 * out (w x h) is the mean of the nf frames (nf x h x w, host) around frames[centre], summed in f64
 * in the order centre, then the others. align 0: the plain mean; 1: every other frame is first
 * aligned to the centre by the robust affine fit to the coarse POC grid (tavg_model "affine";
 * only grid points inside mask, if given); 2: by the dense coarse field (tavg_model "dense").
 * A frame whose motion cannot be estimated is left out, or enters unaligned with
 * keep_unaligned. *n_used (may be NULL) is the number of frames in the mean. */
int ofx_tavg_window(ofx_engine *e, const float *frames, int nf, int centre, int w, int h, int align,
                    const uint8_t *mask, int keep_unaligned, const ofx_config *cfg, float *out, int *n_used);

/* numpy.median of n doubles (host array), as the engine takes it of the per-pixel energy every
 * inner iteration (rust/crates/optflow-core/src/select.rs:19-42): NaN if any value is NaN or n is
 * 0; for an even n the mean of the two middle order statistics, the sign of zero kept as a sort by
 * total order keeps it. */
int ofx_median(ofx_engine *e, const double *values, size_t n, double *out);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): ofx_median of n
 * doubles already in device memory (d_values), the result written to d_out (one double, device
 * memory). The selection's histograms live in a fixed array of the library (16 KiB per GPU, held
 * until the process ends), so the call allocates nothing. At most 4294967295 values (the
 * histogram bins count in 32 bits; OFX_ETOOLARGE above). Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU race, so run them one at a time per GPU. */
int ofx_median_device_noalloc(ofx_engine *e, const double *d_values, size_t n, double *d_out);

/* Frame conversion (rust/crates/optflow-cli/src/io.rs:44-52, :259-281; python/optflow_ref/io.py:15-54):
 * pixels (h x w x channels, channels 1, 3 or 4, 8 or 16 bit, host) -> grey levels in [0, 1]
 * (v / 255 or v / 65535; colour 0.299 R + 0.587 G + 0.114 B), box downscaled by factor (the
 * remainder cropped), f32 (w / factor x h / factor), bit-identical to the reference's f64 values
 * rounded to f32. */
int ofx_frame_to_gray(ofx_engine *e, const void *pixels, int bits, int channels, int w, int h, int factor,
                      float *out);

/* The streaming form of track_sequence, one frame at a time (the Sequence iterator of
 * track.rs:1112-1225; track.py:260-315 and track/hybrid.rs:262-540): ofx_seq_begin takes frame 0
 * (w x h, host) and the optional ROI, and holds one lane of the engine until ofx_seq_end;
 * ofx_seq_step takes frame k = 1, 2, ... and writes its result on the frame-0 grid, as
 * ofx_track_sequence does for results[k - 1]. ofx_seq_step_device takes a device frame and device
 * output arrays (as ofx_track_pair_device). Errors as ofx_track_sequence; a failed step leaves the
 * sequence unusable, and ofx_seq_end must still be called. */
typedef struct ofx_sequence ofx_sequence;
int ofx_seq_begin(ofx_engine *e, const float *frame0, int w, int h, const uint8_t *mask, const ofx_config *cfg,
                  ofx_sequence **out);
int ofx_seq_step(ofx_sequence *seq, const float *frame, ofx_result *out);
int ofx_seq_step_device(ofx_sequence *seq, const float *d_frame, ofx_result *d_out);
void ofx_seq_end(ofx_sequence *seq);

/* No counterpart in the reference: the engine's schedule, the knobs its autotuner
 * (optflow_soil_amd/autotune.py) chooses among. tile_groups: tiles refined side by side in a tiled
 * fine stage (1 .. 8, default 3; each group beyond the first has its own stream and workspace);
 * graphs: 1 (default) replays the whole-frame fine stage from HIP graphs, 0 runs it eagerly. Takes
 * effect for calls that start after it; results do not depend on it. */
typedef struct ofx_schedule {
    int tile_groups;
    int graphs;
} ofx_schedule;
int ofx_set_schedule(ofx_engine *e, const ofx_schedule *sc);
int ofx_get_schedule(const ofx_engine *e, ofx_schedule *sc);

/* No counterpart in the reference. The whole-frame fine stage runs from a HIP graph once it has
 * run eagerly with the same sizes, settings, inputs and buffers (captured on the second such call,
 * replayed from the third; results bit-identical; OFX_GRAPHS=0 in the environment turns this off,
 * and it is off while ofx_profile records). The number of captures and replays so far, summed over
 * the engine's lanes. */
int ofx_graph_stats(ofx_engine *e, int64_t *captures, int64_t *launches);

/* ---------------------------------------------------------------- learned coarse proposal
 * No counterpart in the reference: a 6-layer 3 x 3 convolutional network in bfloat16 (32 hidden
 * channels, float32 accumulation on the matrix cores) that maps the correlation volume of the two
 * frames at 1/8 resolution (displacements of -6 .. 6 cells, 169 channels) to the displacement of
 * every 8 x 8 cell. It stands in for the coarse levels of the POC pyramid (poc.rs:694-766): the
 * level-0 POC grid starts from its proposal and checks it, and the pyramid runs instead when the
 * check fails (ofx_config.coarse_learned, coarse_gate). Parameters are float32 in a flat layout:
 * per layer the weights [cout][3 x 3][cin] then the biases [cout], layers 169 -> 32 -> 32 -> 32 ->
 * 32 -> 32 -> 2 (ReLU after all but the last). */
size_t ofx_coarse_net_params(void);

/* Loads the network every lane of the engine uses for coarse_learned (n = ofx_coarse_net_params()).
 * OFX_EBUSY while a call runs on the engine. */
int ofx_coarse_net_load(ofx_engine *e, const float *params, size_t n);

/* The proposal for frames i1, i2 (w x h, host): out is ((h + 7) / 8) x ((w + 7) / 8) x 2, the
 * displacement (ux, uy) in pixels of each 8 x 8 cell, with the loaded network. */
int ofx_coarse_net_propose(ofx_engine *e, const float *i1, const float *i2, int w, int h, float *out);

/* One forward and backward pass of the network with parameters params (n floats, flat layout) on a
 * batch of nb pairs of cell maps a1, a2 (nb x h8 x w8 each, host; the 8 x 8 means of the frames)
 * against target (nb x h8 x w8 x 2, cells): the mean Charbonnier end-point error
 * sqrt(ex^2 + ey^2 + 0.01^2) - 0.01 into *loss, the gradient of every parameter into grad (n, flat
 * layout), and the prediction (nb x h8 x w8 x 2, cells) into pred if not NULL. */
int ofx_coarse_net_gradient(ofx_engine *e, const float *params, size_t n, const float *a1, const float *a2, int nb,
                            int w8, int h8, const float *target, double *loss, float *grad, float *pred);

/* Training on pairs from the renderer, on the GPU: each step renders a batch of pairs (crops of
 * crop x crop px of one of the n_pat speckle patterns (pw x ph each, host, values in [0, 1]) under
 * a field of 6 Gaussian bumps over a 4000 x 3000 px frame whose largest |u| is uniform in
 * [max_disp / 2, max_disp] px, 8-bit frames with noise of sd `noise`),
 * runs forward and backward and takes an Adam step (learning rate from lr down to lr_final by a
 * cosine). Sample k of the n_pairs distinct pairs is step k / batch of an epoch; the samples are a
 * function of seed. params: initial parameters in, trained parameters out (flat layout); loss:
 * the loss of every step (steps doubles), may be NULL. */
typedef struct ofx_net_train {
    int steps, batch, crop, n_pairs;
    double lr, lr_final, max_disp, noise;
    uint64_t seed;
} ofx_net_train;
int ofx_coarse_net_train(ofx_engine *e, const ofx_net_train *t, const float *patterns, int n_pat, int pw, int ph,
                         float *params, double *loss);

/* ---------------------------------------------------------------- output folder (host only)
 * The reference CLI's output side (rust/crates/optflow-cli/src/crop.rs, run.rs:74-110,
 * :185-207, io.rs:557-621), so that a folder written through this library has the bytes the
 * reference writes. None of these functions touches the device. */

/* The automatic ROI crop (crop.rs:43-130): the bounding box of the nonzero pixels of mask
 * (w x h), grown by *margin px, or with margin NULL by the default margin (the largest
 * displacement + 4 sigma_w + the guard band + 32 + (poc_win / 2 + grid_step) 2^L, doubled when
 * seq), its origin moved down to a multiple of grid_step 2^L (the frame's coarse grid), clipped to
 * the frame. Errors as the reference: "ROI mask is empty", "--crop-margin must be >= 0, got m",
 * "image smaller than poc_win", "the crop WxH is smaller than poc_win". */
typedef struct ofx_crop {
    int x0, y0, x1, y1; /* half-open rectangle of the frame */
    int64_t margin;
} ofx_crop;
int ofx_crop_plan(const uint8_t *mask, int w, int h, const ofx_config *cfg, const int64_t *margin, int seq,
                  ofx_crop *out);

/* crop.json (run.rs:192-207, Crop::json crop.rs:139-144): the rectangle and margin with "mode":
 * "auto", or {"mode": "off"} for crop NULL, written into dir (created with its parents). */
int ofx_write_crop_json(const char *dir, const ofx_crop *crop);

/* One result folder (run.rs:74-110 save, io.rs:557-621 save_result_placed): ux, uy, conf, Exx,
 * Eyy, Exy, gmax, vol as <f4 .npy, valid (|u1, 1 where the written ux is finite), fb (<i4) when
 * res->fb, raw_ux and raw_uy with raw, and timing.json (Timing of track.rs:157-184 with "device":
 * "gpu" after "coarse"; reference and mean_conf when seq; the hybrid shares when seq and res->fb).
 * res holds w x h host maps: the whole frame (crop NULL), or the crop's, each then placed in a
 * frame_w x frame_h map that is NaN outside the crop (conf and fb 0). The .npy bytes are those of
 * numpy.save (format 1.0); dir is created with its parents. */
int ofx_write_result(const char *dir, const ofx_result *res, int w, int h, const ofx_crop *crop, int frame_w,
                     int frame_h, int raw, int seq);

/* The coarse stage's pyramid (poc.rs:610-666, python/optflow_ref/poc.py:172-195): the number of
 * levels below the full image for a w x h frame, and the POC window (wy, wx) at a level of w x h
 * (0, 0 when none fits). */
int ofx_pyramid_levels(const ofx_config *cfg, int w, int h, int *levels);
int ofx_poc_level_window(int w, int h, int win, int *wy, int *wx);

/* v as the reference CLI's JSON writer (serde_json) spells it: the shortest digits that read back
 * to v, "1.0" for integers, 1e-7 style exponents outside [1e-5, 1e16); null when not finite. */
int ofx_format_json_number(double v, char *buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
