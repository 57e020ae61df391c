// build.rs for the reference crate's `mi355x` feature (compile-unverified:
// this build has no cargo/rustc).  With the feature on, it builds the MI355X
// library with hipcc through its own in-tree build script and links it;
// without the feature it does nothing, so the crate's CPU build is unchanged.
//
// SPECTRAL_MI355X_ROOT points at this repository (the directory that holds
// spectral-init_amd/ and include/).
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    if env::var_os("CARGO_FEATURE_MI355X").is_none() {
        return;
    }
    let root = PathBuf::from(
        env::var("SPECTRAL_MI355X_ROOT").expect("set SPECTRAL_MI355X_ROOT to the MI355X build tree"),
    );
    let build = root.join("spectral-init_amd").join("build.py");
    let status = Command::new("python3")
        .arg(&build)
        .env("SI_OFFLOAD_ARCH", env::var("SI_OFFLOAD_ARCH").unwrap_or_else(|_| "gfx950".into()))
        .status()
        .expect("running the hipcc build of libspectral_mi355x");
    assert!(status.success(), "hipcc build of libspectral_mi355x failed");
    let lib = root.join("spectral-init_amd").join("lib");
    println!("cargo:rustc-link-search=native={}", lib.display());
    println!("cargo:rustc-link-lib=dylib=spectral_mi355x");
    // the library's own run path to rocBLAS / rocSOLVER, and ours to it
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib.display());
    println!("cargo:rerun-if-changed={}", root.join("spectral-init_amd").join("csrc").display());
    println!("cargo:rerun-if-changed={}", root.join("include").display());
    println!("cargo:rerun-if-env-changed=SPECTRAL_MI355X_ROOT");
}
