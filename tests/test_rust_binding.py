"""The reference-side Rust binding (rust/crates/optflow-hip) against the C header it binds. There
is no Rust toolchain here, so the check is textual: every #[repr(C)] struct of src/ffi.rs has the
fields of its C struct in include/ofx.h, in order, with matching types, and every function the
FFI block declares exists in the header with the same number of arguments and is exported by
libofx.so. No GPU."""
import os
import re

import optflow_soil_amd as ofx
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ofx.h")
FFI = os.path.join(ROOT, "rust", "crates", "optflow-hip", "src", "ffi.rs")

C_TO_RUST = {"int": "c_int", "double": "c_double", "float *": "*mut f32", "uint8_t *": "*mut u8",
             "int32_t *": "*mut i32", "int64_t": "i64", "double[OFX_MAX_SCALES]": "[c_double; OFX_MAX_SCALES]"}


def c_struct(text, name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        m = re.match(r"([A-Za-z_0-9]+)\s*(\*?)\s*(.*)", decl)
        base, star, names = m.groups()
        for nm in names.split(","):
            nm = nm.strip()
            ptr = star or ("*" if nm.startswith("*") else "")
            nm = nm.lstrip("*").strip()
            arr = re.match(r"(\w+)\[(\w+)\]", nm)
            if arr:
                fields.append((arr.group(1), f"{base}[{arr.group(2)}]"))
            else:
                fields.append((nm, f"{base} *" if ptr else base))
    return fields


def rust_struct(text, name):
    body = re.search(r"pub struct %s \{(.*?)\n\}" % name, text, re.S).group(1)
    return re.findall(r"pub (\w+): ([^,\n]+),", body)


def test_ffi_structs_match_the_header():
    h, r = open(HEADER).read(), open(FFI).read()
    for cname, rname in (("ofx_config", "OfxConfig"), ("ofx_result", "OfxResult"), ("ofx_crop", "OfxCrop"),
                         ("ofx_schedule", "OfxSchedule")):
        cf, rf = c_struct(h, cname), rust_struct(r, rname)
        assert [n for n, _ in cf] == [n for n, _ in rf], cname
        for (n, ct), (_, rt) in zip(cf, rf):
            assert C_TO_RUST[ct] == rt.strip(), (cname, n, ct, rt)


def test_ffi_functions_exist_in_the_header_and_the_library():
    h, r = open(HEADER).read(), open(FFI).read()
    block = r[r.index('extern "C" {'):]
    for name, args in re.findall(r"pub fn (ofx_\w+)\((.*?)\)", block, re.S):
        decl = re.search(r"^(?:const )?\w+\s*\*?\s*%s\s*\(([^;]*)\);" % name, h, re.M)
        assert decl, name
        n_c = 0 if decl.group(1).strip() in ("", "void") else decl.group(1).count(",") + 1
        n_r = len([a for a in args.split(",") if a.strip()])
        assert n_c == n_r, (name, n_c, n_r)
        assert hasattr(ofx.lib(), name), name
