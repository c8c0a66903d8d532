//! The C ABI of libofx.so (include/ofx.h), declared for Rust. Every struct mirrors its C
//! counterpart field for field (tests/test_rust_binding.py checks the order and types against the
//! header); every function keeps the header's argument order.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_double, c_int};

pub const OFX_OK: c_int = 0;
pub const OFX_EINVAL: c_int = 1;
pub const OFX_ETOOLARGE: c_int = 2;
pub const OFX_EDEVICE: c_int = 3;
pub const OFX_ENODEVICE: c_int = 4;
pub const OFX_EBUSY: c_int = 5;
pub const OFX_EUNSUPPORTED: c_int = 6;
pub const OFX_MAX_SCALES: usize = 16;

/// `ofx_config`: the settings the engine reads (python/optflow_ref/config.py:24-90).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct OfxConfig {
    pub grid_step: c_int,
    pub poc_win: c_int,
    pub max_disp: c_double,
    pub poc_peak_min: c_double,
    pub median_thresh: c_double,
    pub lam: c_double,
    pub n_scales: c_int,
    pub scales: [c_double; OFX_MAX_SCALES],
    pub n_orient: c_int,
    pub bandwidth: c_double,
    pub dc_cut: c_double,
    pub sigma_w: c_double,
    pub shape_order: c_int,
    pub sigma_s: c_double,
    pub outer_iters: c_int,
    pub inner_iters: c_int,
    pub conf_min: c_double,
    pub edge_guard: c_double,
    pub edge_guard_min_area: c_double,
    pub phase_tol: c_double,
    pub strain_win: c_int,
    pub ref_update_conf: c_double,
    pub seq_hybrid: c_int,
    pub fb_max: c_int,
    pub fb_radius: c_double,
    pub fb_min_support: c_int,
    pub fb_tol: c_double,
    pub prewarp_freeze_at: c_int,
    pub tile: c_int,
    pub coarse_learned: c_int,
    pub coarse_gate: c_double,
}

/// `ofx_result`: the planes of one pair or sequence frame (caller-owned; NULL = not written) and
/// the timing record (TrackResult, optflow-core/src/track.rs:157-209).
#[repr(C)]
pub struct OfxResult {
    pub ux: *mut f32,
    pub uy: *mut f32,
    pub conf: *mut f32,
    pub raw_ux: *mut f32,
    pub raw_uy: *mut f32,
    pub exx: *mut f32,
    pub eyy: *mut f32,
    pub exy: *mut f32,
    pub gmax: *mut f32,
    pub vol: *mut f32,
    pub valid: *mut u8,
    pub fb: *mut i32,
    pub t_coarse: c_double,
    pub t_fine: c_double,
    pub t_strain: c_double,
    pub reference: c_int,
    pub mean_conf: c_double,
    pub n_fb_max: c_int,
    pub fallback_frac: c_double,
    pub fb_gate_rejected: c_double,
    pub reacquired: c_double,
    pub coarse_confirmed: c_double,
    pub coarse_fallback: c_int,
}

/// `ofx_crop`: the automatic ROI crop (optflow-cli/src/crop.rs:80-88).
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct OfxCrop {
    pub x0: c_int,
    pub y0: c_int,
    pub x1: c_int,
    pub y1: c_int,
    pub margin: i64,
}

/// `ofx_schedule`: tiles refined side by side and the graph switch (no counterpart in the
/// reference; results do not depend on it).
#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct OfxSchedule {
    pub tile_groups: c_int,
    pub graphs: c_int,
}

#[repr(C)]
pub struct OfxEngine {
    _private: [u8; 0],
}

#[repr(C)]
pub struct OfxSequence {
    _private: [u8; 0],
}

extern "C" {
    pub fn ofx_config_default(cfg: *mut OfxConfig);
    pub fn ofx_create(device: c_int, lanes: c_int, out: *mut *mut OfxEngine) -> c_int;
    pub fn ofx_destroy(e: *mut OfxEngine);
    pub fn ofx_last_error() -> *const c_char;
    pub fn ofx_lanes(e: *const OfxEngine) -> c_int;
    pub fn ofx_returns_slope(e: *const OfxEngine) -> c_int;
    pub fn ofx_check_window(e: *mut OfxEngine, w: c_int, h: c_int, cfg: *const OfxConfig) -> c_int;
    pub fn ofx_release_caches(e: *mut OfxEngine) -> c_int;
    pub fn ofx_refine_raw(
        e: *mut OfxEngine,
        i1: *const f32,
        w: c_int,
        h: c_int,
        i2: *const f32,
        w2: c_int,
        h2: c_int,
        shift_x: i64,
        shift_y: i64,
        ux: *const f32,
        uy: *const f32,
        cfg: *const OfxConfig,
        out_ux: *mut f64,
        out_uy: *mut f64,
        out_energy: *mut f64,
        out_resid: *mut f64,
    ) -> c_int;
    pub fn ofx_track_pair(
        e: *mut OfxEngine,
        i1: *const f32,
        i2: *const f32,
        w: c_int,
        h: c_int,
        mask: *const u8,
        init_ux: *const f32,
        init_uy: *const f32,
        cfg: *const OfxConfig,
        out: *mut OfxResult,
    ) -> c_int;
    pub fn ofx_seq_begin(
        e: *mut OfxEngine,
        frame0: *const f32,
        w: c_int,
        h: c_int,
        mask: *const u8,
        cfg: *const OfxConfig,
        out: *mut *mut OfxSequence,
    ) -> c_int;
    pub fn ofx_seq_step(q: *mut OfxSequence, frame: *const f32, out: *mut OfxResult) -> c_int;
    pub fn ofx_seq_end(q: *mut OfxSequence);
    pub fn ofx_crop_plan(
        mask: *const u8,
        w: c_int,
        h: c_int,
        cfg: *const OfxConfig,
        margin: *const i64,
        seq: c_int,
        out: *mut OfxCrop,
    ) -> c_int;
    pub fn ofx_write_crop_json(dir: *const c_char, crop: *const OfxCrop) -> c_int;
    pub fn ofx_write_result(
        dir: *const c_char,
        res: *const OfxResult,
        w: c_int,
        h: c_int,
        crop: *const OfxCrop,
        frame_w: c_int,
        frame_h: c_int,
        raw: c_int,
        seq: c_int,
    ) -> c_int;
    pub fn ofx_set_schedule(e: *mut OfxEngine, sc: *const OfxSchedule) -> c_int;
    pub fn ofx_get_schedule(e: *const OfxEngine, sc: *mut OfxSchedule) -> c_int;
    pub fn ofx_graph_stats(e: *mut OfxEngine, captures: *mut i64, launches: *mut i64) -> c_int;
}
