// Links libboojum_amd.so, built by `make -C zksync-crypto-gpu_amd/csrc` (hipcc, gfx950).
// BOOJUM_AMD_DIR names the directory holding it (default: the in-tree build).
fn main() {
    let dir = std::env::var("BOOJUM_AMD_DIR").unwrap_or_else(|_| {
        let manifest = std::env::var("CARGO_MANIFEST_DIR").unwrap();
        format!("{manifest}/../..")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=boojum_amd");
    println!("cargo:rerun-if-env-changed=BOOJUM_AMD_DIR");
}
