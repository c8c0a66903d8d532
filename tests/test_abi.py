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
