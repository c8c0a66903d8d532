"""Builds libofx.so (the HIP engine behind include/ofx.h) in-tree with hipcc for gfx950."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libofx.so")
SOURCES = ["copy.hip", "fft.hip", "fine.hip", "guard.hip", "poc.hip", "seq.hip", "hybrid.hip", "tavg.hip", "cnn.hip", "engine.hip", "output.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         # packed FP32 instructions stay off: they have given wrong results on this hardware
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(jobs=8, verbose=False):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        objs.append(obj)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs, verbose)
            procs = []
    _wait(procs, verbose)
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    build_cli()
    return LIB


CLI = os.path.join(HERE, "ofx_cli")


def build_cli():
    """The command-line twin (cli/ofx_cli.cpp): host C++ over libofx.so, found beside it at run time."""
    tmp = CLI + ".tmp"
    subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-pthread", os.path.join(HERE, "cli", "ofx_cli.cpp"),
                    "-L" + HERE, "-lofx", "-Wl,-rpath,$ORIGIN", "-Wl,--allow-shlib-undefined", "-o", tmp], check=True)
    os.replace(tmp, CLI)
    return CLI


def _wait(procs, verbose):
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write(f"--- {src}\n{text}\n")
        raise RuntimeError("hipcc failed on " + ", ".join(s for s, _ in failed))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
