"""Processing settings of the dense flow path.

Same keys, defaults, JSON form and validation messages as the reference's settings
(python/optflow_ref/config.py:22-272, rust/crates/optflow-core/src/config.rs:487-619) for every key
the pair / sequence path consumes. The experimental crack keys of the Python reference
(config.py:18-19) are not part of this path: a JSON that sets one to a non-default value is
refused with an error rather than silently ignored.
"""
import json
import math
from dataclasses import asdict, dataclass, fields


def _is_number(v) -> bool:
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _is_finite(v) -> bool:
    return _is_number(v) and math.isfinite(v)


CRACK_DEFAULTS = {"crack_split": False, "split_tol": 0.25, "split_gain": 0.7, "split_from_outer": 0,
                  "crack_out": False, "crack_r": 8.0, "crack_a": 12.0, "crack_sides": False, "side_from_outer": 2,
                  "crack_guard_frac": 0.5, "side_band": "post", "crack_model": "off", "xdic_RB": -1.0,
                  "xdic_iters": 2}


@dataclass
class Config:
    grid_step: int = 16
    poc_win: int = 64
    max_disp: float = 32.0
    poc_peak_min: float = 0.1
    median_thresh: float = 2.0
    lam: float = 8.0
    scales: tuple = (0.6, 1.0, 1.6)
    n_orient: int = 4
    bandwidth: float = 0.5
    dc_cut: float = 2.0
    sigma_w: float = 4.0
    shape_order: int = 1
    sigma_s: float = 4.0
    outer_iters: int = 8
    inner_iters: int = 2
    conf_min: float = 0.2
    edge_guard: float = -1.0
    edge_guard_min_area: float = -1.0
    phase_tol: float = 0.4
    strain_win: int = 15
    ref_update_conf: float = 0.5
    seq_mode: str = "chain"
    fb_max: int = 10
    fb_radius: float = -1.0
    fb_min_support: int = 20
    fb_tol: float = 0.5
    tavg: int = 0
    tavg_align: bool = False
    tavg_model: str = "affine"
    prewarp_freeze_at: int = 0
    # Not a key of the reference's config: the fine stage's tiling. 0 refines the whole frame at
    # once; N > 0 refines frames with a side over N px in N x N cores with the reference's margin,
    # as its tiled driver does with N = 1024 (rust/crates/optflow-core/src/tile.rs:56-98, track.rs:583).
    tile: int = 0
    # Not keys of the reference's config either: the coarse stage. "poc" is the reference's POC
    # pyramid; "learned" starts the level-0 POC grid from the proposal of the engine's learned
    # coarse network instead of the coarser pyramid levels, and falls back to the pyramid when fewer
    # than coarse_gate of the grid points confirm the proposal (a POC peak of at least poc_peak_min
    # and a result within poc_win / 8 px of the proposal).
    coarse_model: str = "poc"
    coarse_gate: float = 0.9

    def to_json(self) -> str:
        return json.dumps(asdict(self), indent=2)

    @classmethod
    def from_json(cls, text: str) -> "Config":
        """Parse a JSON object with the reference's type rules (config.py:133-164)."""
        data = json.loads(text)
        if not isinstance(data, dict):
            raise ValueError("config must be a JSON object")
        known = {f.name for f in fields(cls)}
        crack = {k: v for k, v in data.items() if k in CRACK_DEFAULTS}
        for k, v in crack.items():
            if v != CRACK_DEFAULTS[k]:
                raise ValueError(f"config key {k!r} (crack handling) is not supported by this engine")
        data = {k: v for k, v in data.items() if k not in CRACK_DEFAULTS}
        unknown = set(data) - known
        if unknown:
            raise ValueError(f"unknown config keys: {sorted(unknown)}")
        defaults = cls()
        for key, v in data.items():
            d = getattr(defaults, key)
            if isinstance(d, tuple):
                if not isinstance(v, list) or not all(_is_number(x) for x in v):
                    raise ValueError(f"config key {key!r} must be a list of numbers, got {v!r}")
                data[key] = tuple(float(x) for x in v)
            elif isinstance(d, bool):
                if not isinstance(v, bool):
                    raise ValueError(f"config key {key!r} must be true or false, got {v!r}")
            elif isinstance(d, str):
                if not isinstance(v, str):
                    raise ValueError(f"config key {key!r} must be a string, got {v!r}")
            elif isinstance(d, int):
                if isinstance(v, bool) or not isinstance(v, int):
                    raise ValueError(f"config key {key!r} must be an integer, got {v!r}")
            elif not _is_number(v):
                raise ValueError(f"config key {key!r} must be a number, got {v!r}")
            else:
                data[key] = float(v)
        cfg = cls(**data)
        cfg.validate()
        return cfg

    def validate(self) -> None:
        """The reference's checks and messages, in its order (config.py:166-245)."""
        if self.seq_mode not in ("chain", "hybrid"):
            raise ValueError(f"seq_mode must be 'chain' or 'hybrid', got {self.seq_mode!r}")
        if not self.scales:
            raise ValueError("scales must not be empty")
        if self.poc_win < 16 or self.poc_win & (self.poc_win - 1):
            raise ValueError("poc_win must be a power of two >= 16")
        if self.grid_step < 1 or self.strain_win < 5 or self.strain_win % 2 == 0:
            raise ValueError("grid_step >= 1 and strain_win odd >= 5 required")
        if self.dc_cut <= 0:
            raise ValueError("dc_cut must be positive")
        if self.phase_tol <= 0:
            raise ValueError("phase_tol must be positive")
        if self.shape_order not in (1, 2):
            raise ValueError("shape_order must be 1 or 2")
        if self.lam * min(self.scales) < 3 or self.sigma_w <= 0 or self.sigma_s <= 0:
            raise ValueError("shortest wavelength lam*min(scales) must be >= 3 px; sigmas must be positive")
        for name in ("max_disp", "poc_peak_min", "median_thresh", "lam", "bandwidth", "dc_cut", "sigma_w",
                     "sigma_s", "conf_min", "edge_guard", "edge_guard_min_area", "phase_tol", "ref_update_conf"):
            if not _is_finite(getattr(self, name)):
                raise ValueError(f"{name} must be a finite number")
        if not all(_is_finite(x) for x in self.scales):
            raise ValueError("scales must be finite numbers")
        if self.bandwidth <= 0 or self.max_disp <= 0 or self.median_thresh <= 0:
            raise ValueError("bandwidth, max_disp and median_thresh must be positive")
        if self.n_orient < 1 or self.outer_iters < 1 or self.inner_iters < 1:
            raise ValueError("n_orient, outer_iters and inner_iters must be >= 1")
        if not 0 <= self.conf_min <= 1 or not 0 <= self.poc_peak_min <= 1:
            raise ValueError("conf_min and poc_peak_min must lie in [0, 1]")
        if self.fb_max < 0:
            raise ValueError("fb_max must be >= 0")
        if not _is_finite(self.fb_radius) or not _is_finite(self.fb_tol) or self.fb_tol <= 0 or self.fb_radius == 0:
            raise ValueError("fb_radius must be finite and nonzero (negative: 4 sigma_w), fb_tol finite and positive")
        if self.fb_min_support < 0 or self.tavg < 0 or self.tavg > 64:
            raise ValueError("fb_min_support must be >= 0 and tavg in [0, 64]")
        if self.tavg_model not in ("affine", "dense"):
            raise ValueError(f"tavg_model must be 'affine' or 'dense', got {self.tavg_model!r}")
        if not isinstance(self.tavg_align, bool):
            raise ValueError("tavg_align must be true or false")
        if self.prewarp_freeze_at < 0:
            raise ValueError("prewarp_freeze_at must be >= 0")
        if not isinstance(self.tile, int) or isinstance(self.tile, bool) or (self.tile != 0 and self.tile < 64):
            raise ValueError("tile must be 0 (whole frame) or an integer >= 64")
        if self.coarse_model not in ("poc", "learned"):
            raise ValueError(f"coarse_model must be 'poc' or 'learned', got {self.coarse_model!r}")
        if not _is_finite(self.coarse_gate) or not 0 <= self.coarse_gate <= 1:
            raise ValueError("coarse_gate must lie in [0, 1]")
        if (self.n_orient * len(self.scales) > 64 or self.sigma_w > 256 or self.sigma_s > 256
                or self.strain_win > 255 or self.poc_win > 1024):
            raise ValueError("at most 64 filters (n_orient * len(scales)); sigma_w, sigma_s <= 256, "
                             "strain_win <= 255 and poc_win <= 1024 required")
        if self.grid_step > 4096 or self.edge_guard_px() > 256:
            raise ValueError("grid_step <= 4096 and an edge guard (edge_guard, or the automatic one from lam, "
                             "scales and bandwidth) <= 256 px required")

    def fb_radius_px(self) -> float:
        return float(self.fb_radius) if self.fb_radius > 0 else 4.0 * self.sigma_w

    def edge_guard_min_area_px(self) -> float:
        if self.edge_guard_min_area >= 0:
            return float(self.edge_guard_min_area)
        return 2.0 * (self.lam * max(self.scales)) ** 2

    def auto_edge_guard_px(self) -> float:
        return 3.0 * self.lam * max(self.scales) / (2 * math.pi * self.bandwidth)

    def edge_guard_px(self) -> float:
        if self.edge_guard >= 0:
            return float(self.edge_guard)
        return self.auto_edge_guard_px()

    def n_filters(self) -> int:
        return self.n_orient * len(self.scales)
