"""Copies the reference's own golden fixtures for this path into tests/golden/reference/.

The files are data written by the reference's fixture generator (python/bench/make_parity_fixtures.py,
run with Python 3.12.14, numpy 2.5.4, scipy 1.18.1 per meta.json) and checked in at
rust/crates/optflow-core/tests/fixtures/. Only the sets that pin the dense flow path are copied
(interp, gabor, POC, refine, strain, EDT, pair, chain and hybrid sequences, meta.json); the crack
model sets (xdic/) are out of scope. Run once in a container where /root/reference exists.
"""
import glob
import os
import shutil

SRC = "/root/reference/rust/crates/optflow-core/tests/fixtures"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference")
PREFIXES = ("interp_", "gabor_", "poc_", "refine_", "strain_", "edt_", "pair_", "seq_", "seqc1_", "hyb")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    n = 0
    for f in sorted(glob.glob(os.path.join(SRC, "*.npy"))) + [os.path.join(SRC, "meta.json")]:
        name = os.path.basename(f)
        if name == "meta.json" or name.startswith(PREFIXES):
            shutil.copy(f, os.path.join(DST, name))
            n += 1
    print(f"copied {n} files into {DST}")
