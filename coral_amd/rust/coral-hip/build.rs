// Links libcoralhip.so.  CORAL_HIP_LIB_DIR names the directory that holds it
// (coral_amd/ of the repository after `make -C coral_amd/csrc`); without it the
// crate looks two levels up from its own directory, where the build puts it.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("CORAL_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("..").join(".."),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=coralhip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    println!("cargo:rerun-if-env-changed=CORAL_HIP_LIB_DIR");
}
