// Links libofx.so from OFX_LIB_DIR (the directory `python -m optflow_soil_amd.build` writes it to).
fn main() {
    let dir = std::env::var("OFX_LIB_DIR").unwrap_or_else(|_| "/opt/optflow-mi355x/lib".into());
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=ofx");
    println!("cargo:rerun-if-env-changed=OFX_LIB_DIR");
}
