"""The C ABI library loads and exports every symbol include/ofx.h declares; the Python binding's
structure layouts agree with the header (no GPU needed, no compute call)."""
import ctypes
import os
import re

import numpy as np

import optflow_soil_amd as ofx
from conftest import ROOT
from optflow_soil_amd.config import Config


def test_header_symbols_are_exported():
    header = open(os.path.join(ROOT, "include", "ofx.h")).read()
    declared = set(re.findall(r"\b(ofx_[a-z_]+)\s*\(", header))
    assert declared == set(ofx.SYMBOLS)
    lib = ofx.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_config_defaults_match_header_defaults():
    c = ofx.OfxConfig()
    ofx.lib().ofx_config_default(ctypes.byref(c))
    mine = ofx.to_c(Config())
    for name, _ in ofx.OfxConfig._fields_:
        a, b = getattr(c, name), getattr(mine, name)
        if name == "scales":
            assert list(a) == list(b)
        else:
            assert a == b, name


def test_config_validation_messages():
    import pytest
    with pytest.raises(ValueError, match="poc_win must be a power of two"):
        Config(poc_win=48).validate()
    with pytest.raises(ValueError, match="strain_win odd"):
        Config(strain_win=14).validate()
    with pytest.raises(ValueError, match="unknown config keys"):
        Config.from_json('{"nonsense": 1}')
    assert Config.from_json(Config().to_json()) == Config()


def test_no_device_is_an_error_not_a_fallback():
    try:
        import torch
        if torch.cuda.is_available():
            return
    except Exception:
        pass
    import pytest
    with pytest.raises(RuntimeError, match="no HIP device"):
        ofx.Engine(0)


def test_synthetic_workloads_are_seeded():
    from optflow_soil_amd import synth
    f1, truth = synth.workload_c1()
    f2, _ = synth.workload_c1()
    assert len(f1) == 5 and f1[0].shape == (256, 256) and f1[0].dtype == np.uint8
    assert all((a == b).all() for a, b in zip(f1, f2))
    ux, uy = truth(4)
    assert abs(ux[128, 128] - 3.2) < 1e-12 and abs(ux[128, 228] - (3.2 + 0.02 * 100)) < 1e-9
    m = synth.convex_polygon_mask(200, 200, seed=3)
    assert 0.55 < m.mean() < 0.65


def test_output_folder_has_the_reference_files_and_dtypes(tmp_path):
    """io.py:203-223: the .npy names and dtypes of a result folder, byte-identical to numpy.save."""
    from optflow_soil_amd import Result, io
    h, w = 5, 7
    rng = np.random.default_rng(0)
    f = lambda: rng.standard_normal((h, w)).astype(np.float32)
    ux = f()
    ux[0, 0] = np.nan
    strain = {k: f() for k in ("Exx", "Eyy", "Exy", "gmax", "vol")}
    res = Result(ux, f(), f(), strain, f(), f(), np.isfinite(ux), {"coarse": 0.1}, np.zeros((h, w), np.int32))
    io.save_result(str(tmp_path), res, raw=True)
    want = {"ux": "<f4", "uy": "<f4", "conf": "<f4", "Exx": "<f4", "Eyy": "<f4", "Exy": "<f4", "gmax": "<f4",
            "vol": "<f4", "raw_ux": "<f4", "raw_uy": "<f4", "valid": "|u1", "fb": "<i4"}
    for name, dt in want.items():
        a = np.load(tmp_path / f"{name}.npy")
        assert a.dtype.str == dt and a.shape == (h, w), name
    assert np.load(tmp_path / "valid.npy")[0, 0] == 0
    # byte-identical to numpy.save of the same array
    import io as pyio
    buf = pyio.BytesIO()
    np.save(buf, np.ascontiguousarray(res.strain["gmax"], dtype="<f4"))
    assert (tmp_path / "gmax.npy").read_bytes() == buf.getvalue()
    import json
    assert json.loads((tmp_path / "timing.json").read_text()) == {"coarse": 0.1}
    assert io.sequence_dir("out", 3).endswith("frame_00003")


def test_crop_plan_and_embed():
    """crop.py:81-103: bounding box + margin, origin on the coarse grid, clipped; embed puts the
    crop back with NaN / 0 outside."""
    from optflow_soil_amd import Result, io
    cfg = Config()
    mask = np.zeros((600, 800), bool)
    mask[200:300, 250:400] = True
    c = io.Crop.plan(mask, cfg, margin=50)
    A = cfg.grid_step * 2 ** io.n_levels(cfg.max_disp, cfg.poc_win, mask.shape)
    assert (c.x0, c.y0) == ((250 - 50) // A * A, (200 - 50) // A * A)
    assert (c.x1, c.y1) == (400 + 50, 300 + 50)
    assert io.n_levels(32, 64, (2048, 2048)) == 1 and io.level_window((56, 56), 64) == (32, 32)
    h, w = c.shape
    z = np.ones((h, w), np.float32)
    res = Result(z, z, z, {k: z for k in ("Exx", "Eyy", "Exy", "gmax", "vol")}, z, z, np.ones((h, w), bool), {})
    full = c.embed(res, mask.shape)
    assert np.isnan(full.ux[0, 0]) and full.conf[0, 0] == 0 and not full.valid[0, 0]
    assert full.ux[c.y0, c.x0] == 1 and np.isfinite(full.ux).sum() == h * w
    assert '"mode": "auto"' in c.json()
    import pytest
    with pytest.raises(ValueError, match="ROI mask is empty"):
        io.Crop.plan(np.zeros((100, 100), bool), cfg)


def test_sequence_handle_errors_need_no_device():
    """A NULL or ended sequence is an OFX_EINVAL with a message, not a crash; ofx_seq_end(NULL) is
    a no-op (these paths touch no device)."""
    lib = ofx.lib()
    assert lib.ofx_seq_step(None, None, None) == ofx.OFX_EINVAL
    assert b"sequence is NULL" in lib.ofx_last_error()
    assert lib.ofx_seq_step_device(None, None, None) == ofx.OFX_EINVAL
    lib.ofx_seq_end(None)


def test_batched_pair_checks_name_the_pair():
    """Argument checks of ofx_track_pairs run before any device work and name the failing pair."""
    import pytest
    ofx_pairs = (ofx.OfxPair * 1)()
    ofx_pairs[0] = ofx.OfxPair(None, None, 64, 64, None, None, None)
    res = (ofx.OfxResult * 1)()
    c = ofx.to_c(Config())
    assert ofx.lib().ofx_track_pairs(None, 1, ofx_pairs, ctypes.byref(c), res) == ofx.OFX_EINVAL
    assert b"pair 0" in ofx.lib().ofx_last_error()
