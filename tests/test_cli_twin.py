"""The command-line twin (optflow_soil_amd/ofx_cli, cli/ofx_cli.cpp) of the reference's
`optflow pair` / `optflow seq` (rust/crates/optflow-cli/src/args.rs, run.rs). Without a GPU: the
argument errors and their exit status 2, and config.json byte-identical to the configuration's own
JSON (Python's float spelling). With a GPU: a pair with an ROI (auto crop, placed full-frame maps,
crop.json, raw maps, timing.json) and a 5-frame chain sequence, every file against the Python
binding's result, and the maps against the FP64 oracle to the parity bar."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load
from optflow_soil_amd.config import Config

CLI = os.path.join(ROOT, "optflow_soil_amd", "ofx_cli")


def run(*args):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout, p.stderr


def test_arguments_are_checked_as_the_reference_cli_checks_them(tmp_path):
    assert run()[0] == 2
    rc, _, err = run("pair", "--ref", "a.npy", "--out", str(tmp_path))
    assert rc == 2 and "error: the following arguments are required: --ref, --def" in err
    rc, _, err = run("pair", "--ref", "a.npy", "--def", "b.npy", "--out", str(tmp_path), "--scale", "0")
    assert rc == 2 and "error: scale must be >= 1, got 0" in err
    rc, _, err = run("seq", "--frames", "x", "--out", str(tmp_path), "--step", "abc")
    assert rc == 2 and "invalid int value: 'abc'" in err
    rc, _, err = run("pair", "--ref", "a.npy", "--def", "b.npy", "--out", str(tmp_path), "--set", "nonsense=1")
    assert rc == 2 and "unknown config keys: ['nonsense']" in err


def test_config_json_is_the_configurations_own_json(tmp_path):
    """config.json is written before any frame or device is touched; its bytes are those of
    Config.to_json() (json.dumps with indent 2) for the same settings."""
    out = tmp_path / "o"
    run("pair", "--ref", "missing_a.npy", "--def", "missing_b.npy", "--out", str(out), "--set", "sigma_w=6",
        "--set", "seq_mode=hybrid", "--set", "scales=[0.5, 1.25, 2]", "--set", "edge_guard=0.0001",
        "--set", "fb_tol=1e-05", "--set", "max_disp=48")
    cfg = Config(sigma_w=6.0, seq_mode="hybrid", scales=(0.5, 1.25, 2.0), edge_guard=0.0001, fb_tol=1e-05,
                 max_disp=48.0)
    assert (out / "config.json").read_text() == cfg.to_json()
    cj = tmp_path / "c.json"
    cj.write_text(json.dumps({"outer_iters": 4, "lam": 7.5, "coarse_model": "learned"}))
    out2 = tmp_path / "o2"
    run("pair", "--ref", "a.npy", "--def", "b.npy", "--out", str(out2), "--config", str(cj))
    assert (out2 / "config.json").read_text() == Config(outer_iters=4, lam=7.5, coarse_model="learned").to_json()


@pytest.mark.gpu
def test_pair_folder_matches_the_binding_and_the_oracle(tmp_path):
    import oracle as O
    import optflow_soil_amd as ofx
    from optflow_soil_amd import io
    a, b = load("pair_a").astype(np.float32), load("pair_b").astype(np.float32)
    roi = load("pair_roi") > 0.5
    np.save(tmp_path / "a.npy", a)
    np.save(tmp_path / "b.npy", b)
    np.save(tmp_path / "roi.npy", roi)
    out = tmp_path / "out"
    print("ofx_cli pair on the pair fixture, ROI, crop-margin 24, raw maps", flush=True)
    rc, so, err = run("pair", "--ref", str(tmp_path / "a.npy"), "--def", str(tmp_path / "b.npy"), "--mask",
                      str(tmp_path / "roi.npy"), "--crop-margin", "24", "--raw", "--out", str(out))
    assert rc == 0, err
    cfg = Config()
    crop = io.CropWindow.plan(roi, cfg, margin=24)
    assert (out / "crop.json").read_text() == crop.to_json()
    eng = ofx.Engine(0)
    r = eng.track_pair(crop.cut(a), crop.cut(b), cfg, mask=crop.cut(roi))
    eng.close()
    want = tmp_path / "want"
    io.write_result(str(want), r, raw=True, crop=crop, frame_shape=a.shape)
    for name in sorted(os.listdir(want)):
        if name.endswith(".npy"):
            assert (out / name).read_bytes() == (want / name).read_bytes(), name
    t = json.loads((out / "timing.json").read_text())
    assert list(t) == ["coarse", "device", "fine", "strain"] and t["device"] == "gpu"
    # the maps against the oracle's pair of the same crop, placed the same way
    o = O.track_pair(crop.cut(a).astype(np.float64), crop.cut(b).astype(np.float64), cfg, mask=crop.cut(roi))
    ux = np.load(out / "ux.npy")
    sl = np.s_[crop.y0:crop.y1, crop.x0:crop.x1]
    assert np.isnan(ux[:crop.y0]).all() and np.isnan(ux[:, :crop.x0]).all()
    both = np.isfinite(ux[sl]) & np.isfinite(o["ux"])
    assert (np.isfinite(ux[sl]) == np.isfinite(o["ux"]))[crop.cut(roi)].mean() >= 0.999
    assert np.sqrt(np.mean((ux[sl][both] - o["ux"][both]) ** 2)) <= 1e-3
    exx = np.load(out / "Exx.npy")[sl]
    fin = np.isfinite(exx) & np.isfinite(o["Exx"])
    assert np.sqrt(np.mean((exx[fin] - o["Exx"][fin]) ** 2)) <= 1e-5


@pytest.mark.gpu
def test_sequence_folders_match_the_binding_and_the_oracle(tmp_path):
    import oracle as O
    import optflow_soil_amd as ofx
    from optflow_soil_amd import io, synth
    frames, _ = synth.workload_c1()
    d = tmp_path / "frames"
    d.mkdir()
    for k, f in enumerate(frames):
        np.save(d / f"f{k:03d}.npy", f)
    out = tmp_path / "out"
    print("ofx_cli seq on the 5-frame chain workload", flush=True)
    rc, so, err = run("seq", "--frames", str(d), "--out", str(out))
    assert rc == 0, err
    assert (out / "crop.json").read_text() == '{\n  "mode": "off"\n}'
    ff = [synth.as_float(f) for f in frames]
    eng = ofx.Engine(0)
    got = eng.track_sequence(ff, Config())
    eng.close()
    ref = O.track_sequence([f.astype(np.float64) for f in ff], Config())
    for (k, r), o in zip(got, ref):
        folder = out / f"frame_{k:05d}"
        want = tmp_path / f"want_{k}"
        io.write_result(str(want), r, sequence=True)
        for name in sorted(os.listdir(want)):
            if name.endswith(".npy"):
                assert (folder / name).read_bytes() == (want / name).read_bytes(), (k, name)
        t = json.loads((folder / "timing.json").read_text())
        assert list(t) == ["coarse", "device", "fine", "strain", "reference", "mean_conf"]
        assert t["reference"] == o["timing"]["reference"]
        ux = np.load(folder / "ux.npy")
        both = np.isfinite(ux) & np.isfinite(o["ux"])
        assert (np.isfinite(ux) == np.isfinite(o["ux"])).mean() >= 0.999
        assert np.sqrt(np.mean((ux[both] - o["ux"][both]) ** 2)) <= 1e-3


@pytest.mark.gpu
def test_sequence_pipeline_ends_on_a_bad_frame_and_on_a_failed_write(tmp_path):
    """The sequence reads frames on one thread, tracks on another and writes on a third. A frame
    file that cannot be read ends the run with status 2 after the frames before it are written; a
    frame folder that cannot be made does the same. Neither leaves a thread waiting."""
    from optflow_soil_amd import synth
    frames, _ = synth.workload_c1()
    d = tmp_path / "frames"
    d.mkdir()
    for k, f in enumerate(frames):
        np.save(d / f"f{k:03d}.npy", f)
    (d / "f003.npy").write_bytes(b"not an array")
    out = tmp_path / "out"
    print("ofx_cli seq with a bad frame 3", flush=True)
    rc, so, err = run("seq", "--frames", str(d), "--out", str(out))
    assert rc == 2 and "f003.npy: not a .npy file" in err, err
    assert (out / "frame_00001" / "ux.npy").exists() and (out / "frame_00002" / "ux.npy").exists()
    assert not (out / "frame_00003").exists()
    np.save(d / "f003.npy", frames[3])
    out2 = tmp_path / "out2"
    out2.mkdir()
    (out2 / "frame_00002").write_text("a file where a folder goes")
    print("ofx_cli seq with an unwritable frame 2", flush=True)
    rc, so, err = run("seq", "--frames", str(d), "--out", str(out2))
    assert rc == 2 and "frame_00002" in err, err
    assert (out2 / "frame_00001" / "ux.npy").exists()
    assert not (out2 / "frame_00004").exists()
