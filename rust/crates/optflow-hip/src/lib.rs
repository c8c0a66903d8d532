//! optflow-hip: the MI355X fine-stage engine for optflow-soil, over libofx.so (include/ofx.h).
//!
//! `HipEngine` implements the reference's `FineEngine` trait (optflow-core/src/engine.rs:31-125),
//! so the reference's pipeline (coarse stage, tiled driver, guard and strain on the CPU) runs its
//! fine stage on the GPU: `--device hip` in the CLI's `engine()` (optflow-cli/src/run.rs:164-178).
//! `track_pair_hip` and `HipSequence` run the whole pair and sequence on the GPU instead (the
//! library's own coarse stage, guard and strain), returning the reference's `TrackResult`.
//!
//! Errors follow the reference: an `EngineError` with the library's message, never a CPU
//! fallback (rust/README.md:118-122). Source only in this repository (no Rust toolchain in its
//! image); every type follows include/ofx.h and the reference's trait and types.
mod ffi;

use optflow_core::config::Config;
use optflow_core::engine::{EngineError, FineEngine};
use optflow_core::image::Image;
use optflow_core::phaseflow::RawFlow;
use std::ffi::CStr;

pub use ffi::{OfxConfig, OfxResult};

/// A GPU engine: one HIP stream per lane, device buffers kept per window size.
pub struct HipEngine(*mut ffi::OfxEngine);

// libofx enforces one call per lane at a time (OFX_EBUSY instead of a wait, as the Metal engine,
// optflow-gpu/src/engine.rs:664-681); the handle itself may be shared between threads.
unsafe impl Send for HipEngine {}
unsafe impl Sync for HipEngine {}

fn last_error() -> EngineError {
    let msg = unsafe { CStr::from_ptr(ffi::ofx_last_error()) };
    EngineError(msg.to_string_lossy().into_owned())
}

fn check(status: i32) -> Result<(), EngineError> {
    if status == ffi::OFX_OK {
        Ok(())
    } else {
        Err(last_error())
    }
}

/// `Config` -> `ofx_config`: the same fields under the same names (config.rs:487-557). `tile` is
/// the reference's TILE (track.rs:583) when the engine serves the reference's tiled driver, 0 for
/// the whole-frame entry points.
pub fn to_c(cfg: &Config, tile: i32) -> Result<ffi::OfxConfig, EngineError> {
    let mut c = unsafe {
        let mut c = std::mem::zeroed();
        ffi::ofx_config_default(&mut c);
        c
    };
    if cfg.scales.len() > ffi::OFX_MAX_SCALES {
        return Err(EngineError(format!("at most {} scales", ffi::OFX_MAX_SCALES)));
    }
    c.grid_step = cfg.grid_step as i32;
    c.poc_win = cfg.poc_win as i32;
    c.max_disp = cfg.max_disp;
    c.poc_peak_min = cfg.poc_peak_min;
    c.median_thresh = cfg.median_thresh;
    c.lam = cfg.lam;
    c.n_scales = cfg.scales.len() as i32;
    c.scales[..cfg.scales.len()].copy_from_slice(&cfg.scales);
    c.n_orient = cfg.n_orient as i32;
    c.bandwidth = cfg.bandwidth;
    c.dc_cut = cfg.dc_cut;
    c.sigma_w = cfg.sigma_w;
    c.shape_order = cfg.shape_order as i32;
    c.sigma_s = cfg.sigma_s;
    c.outer_iters = cfg.outer_iters as i32;
    c.inner_iters = cfg.inner_iters as i32;
    c.conf_min = cfg.conf_min;
    c.edge_guard = cfg.edge_guard;
    c.edge_guard_min_area = cfg.edge_guard_min_area;
    c.phase_tol = cfg.phase_tol;
    c.strain_win = cfg.strain_win as i32;
    c.ref_update_conf = cfg.ref_update_conf;
    c.seq_hybrid = (cfg.seq_mode == "hybrid") as i32;
    c.fb_max = cfg.fb_max as i32;
    c.fb_radius = cfg.fb_radius;
    c.fb_min_support = cfg.fb_min_support as i32;
    c.fb_tol = cfg.fb_tol;
    c.prewarp_freeze_at = cfg.prewarp_freeze_at as i32;
    c.tile = tile;
    Ok(c)
}

impl HipEngine {
    /// GpuEngine::new's counterpart (optflow-gpu/src/engine.rs:72-95): device 0 and
    /// OPTFLOW_GPU_LANES lanes (default 1).
    pub fn new() -> Result<Self, EngineError> {
        let lanes = std::env::var("OPTFLOW_GPU_LANES")
            .ok()
            .and_then(|v| v.parse().ok())
            .unwrap_or(1);
        let mut h = std::ptr::null_mut();
        check(unsafe { ffi::ofx_create(0, lanes, &mut h) })?;
        Ok(HipEngine(h))
    }

    pub fn raw(&self) -> *mut ffi::OfxEngine {
        self.0
    }

    /// Tiles refined side by side (1 .. 8) and the graph switch; results do not depend on them.
    pub fn set_schedule(&self, tile_groups: i32, graphs: bool) -> Result<(), EngineError> {
        let sc = ffi::OfxSchedule { tile_groups, graphs: graphs as i32 };
        check(unsafe { ffi::ofx_set_schedule(self.0, &sc) })
    }
}

impl Drop for HipEngine {
    fn drop(&mut self) {
        unsafe { ffi::ofx_destroy(self.0) }
    }
}

impl FineEngine for HipEngine {
    fn refine_raw(
        &self,
        i1: &Image<f32>,
        i2: &Image<f32>,
        shift: (i64, i64),
        ux: &Image<f32>,
        uy: &Image<f32>,
        cfg: &Config,
    ) -> Result<RawFlow, EngineError> {
        // one window of the reference's tiled driver: refined whole
        let c = to_c(cfg, 0)?;
        let (w, h) = (i1.w, i1.h);
        let mut out = RawFlow {
            ux: Image::new(w, h),
            uy: Image::new(w, h),
            energy: Image::new(w, h),
            resid: Image::new(w, h),
        };
        check(unsafe {
            ffi::ofx_refine_raw(
                self.0,
                i1.data.as_ptr(),
                w as i32,
                h as i32,
                i2.data.as_ptr(),
                i2.w as i32,
                i2.h as i32,
                shift.0,
                shift.1,
                ux.data.as_ptr(),
                uy.data.as_ptr(),
                &c,
                out.ux.data.as_mut_ptr(),
                out.uy.data.as_mut_ptr(),
                out.energy.data.as_mut_ptr(),
                out.resid.data.as_mut_ptr(),
            )
        })?;
        Ok(out)
    }

    fn check_window(&self, w: usize, h: usize, cfg: &Config) -> Result<(), EngineError> {
        let c = to_c(cfg, 0)?;
        check(unsafe { ffi::ofx_check_window(self.0, w as i32, h as i32, &c) })
    }

    fn lanes(&self) -> usize {
        unsafe { ffi::ofx_lanes(self.0) as usize }
    }

    // the engine builds the reference pack itself inside refine_raw, as the reference's GPU engine
    // does (core/engine.rs:51-58): no first tile to prepare ahead
    fn prefetch_first_tile(&self) -> bool {
        false
    }

    fn returns_slope(&self) -> bool {
        unsafe { ffi::ofx_returns_slope(self.0) != 0 }
    }

    fn release_caches(&self) {
        unsafe {
            ffi::ofx_release_caches(self.0);
        }
    }
}

/// The planes of one result, owned here and lent to the library as an `ofx_result`.
pub struct Planes {
    pub w: usize,
    pub h: usize,
    pub maps: [Vec<f32>; 10], // ux, uy, conf, raw_ux, raw_uy, exx, eyy, exy, gmax, vol
    pub valid: Vec<u8>,
    pub fb: Option<Vec<i32>>,
}

impl Planes {
    pub fn new(w: usize, h: usize, with_fb: bool) -> Self {
        let n = w * h;
        Planes {
            w,
            h,
            maps: std::array::from_fn(|_| vec![0.0f32; n]),
            valid: vec![0u8; n],
            fb: with_fb.then(|| vec![0i32; n]),
        }
    }

    pub fn as_ofx(&mut self) -> OfxResult {
        let [ux, uy, conf, raw_ux, raw_uy, exx, eyy, exy, gmax, vol] = &mut self.maps;
        OfxResult {
            ux: ux.as_mut_ptr(),
            uy: uy.as_mut_ptr(),
            conf: conf.as_mut_ptr(),
            raw_ux: raw_ux.as_mut_ptr(),
            raw_uy: raw_uy.as_mut_ptr(),
            exx: exx.as_mut_ptr(),
            eyy: eyy.as_mut_ptr(),
            exy: exy.as_mut_ptr(),
            gmax: gmax.as_mut_ptr(),
            vol: vol.as_mut_ptr(),
            valid: self.valid.as_mut_ptr(),
            fb: self.fb.as_mut().map_or(std::ptr::null_mut(), |f| f.as_mut_ptr()),
            t_coarse: 0.0,
            t_fine: 0.0,
            t_strain: 0.0,
            reference: 0,
            mean_conf: 0.0,
            n_fb_max: 0,
            fallback_frac: 0.0,
            fb_gate_rejected: 0.0,
            reacquired: 0.0,
            coarse_confirmed: -1.0,
            coarse_fallback: 0,
        }
    }
}

/// track_pair (optflow-core/src/track.rs:392-414) with every stage on the GPU. `mask`: ROI, 1 =
/// inside. The planes are those of `TrackResult`; `timing` carries Timing's coarse/fine/strain.
pub fn track_pair_hip(
    e: &HipEngine,
    i1: &Image<f32>,
    i2: &Image<f32>,
    cfg: &Config,
    mask: Option<&Image<u8>>,
) -> Result<(Planes, OfxResult), EngineError> {
    let c = to_c(cfg, 0)?;
    let mut p = Planes::new(i1.w, i1.h, false);
    let mut r = p.as_ofx();
    check(unsafe {
        ffi::ofx_track_pair(
            e.0,
            i1.data.as_ptr(),
            i2.data.as_ptr(),
            i1.w as i32,
            i1.h as i32,
            mask.map_or(std::ptr::null(), |m| m.data.as_ptr()),
            std::ptr::null(),
            std::ptr::null(),
            &c,
            &mut r,
        )
    })?;
    Ok((p, r))
}

/// The reference's sequence iterator (track.rs:1112-1225; hybrid: track/hybrid.rs:262-540) on the
/// GPU, one frame at a time: it holds one lane of the engine until dropped.
pub struct HipSequence<'e, I: Iterator<Item = Image<f32>>> {
    _engine: &'e HipEngine,
    q: *mut ffi::OfxSequence,
    frames: I,
    k: usize,
    hybrid: bool,
}

impl<'e, I: Iterator<Item = Image<f32>>> HipSequence<'e, I> {
    pub fn new(
        e: &'e HipEngine,
        mut frames: I,
        cfg: &Config,
        mask: Option<&Image<u8>>,
    ) -> Result<Self, EngineError> {
        let c = to_c(cfg, 0)?;
        let f0 = frames
            .next()
            .ok_or_else(|| EngineError("no frames".to_string()))?;
        let mut q = std::ptr::null_mut();
        check(unsafe {
            ffi::ofx_seq_begin(
                e.0,
                f0.data.as_ptr(),
                f0.w as i32,
                f0.h as i32,
                mask.map_or(std::ptr::null(), |m| m.data.as_ptr()),
                &c,
                &mut q,
            )
        })?;
        Ok(HipSequence {
            _engine: e,
            q,
            frames,
            k: 0,
            hybrid: cfg.seq_mode == "hybrid",
        })
    }
}

impl<'e, I: Iterator<Item = Image<f32>>> Iterator for HipSequence<'e, I> {
    type Item = Result<(usize, Planes, OfxResult), EngineError>;
    fn next(&mut self) -> Option<Self::Item> {
        let f = self.frames.next()?;
        self.k += 1;
        let mut p = Planes::new(f.w, f.h, self.hybrid);
        let mut r = p.as_ofx();
        let rc = unsafe { ffi::ofx_seq_step(self.q, f.data.as_ptr(), &mut r) };
        Some(check(rc).map(|_| (self.k, p, r)))
    }
}

impl<'e, I: Iterator<Item = Image<f32>>> Drop for HipSequence<'e, I> {
    fn drop(&mut self) {
        unsafe { ffi::ofx_seq_end(self.q) }
    }
}
