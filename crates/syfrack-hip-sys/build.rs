// Link search path of libsyfrack_hip.so: SYFRACK_HIP_LIB_DIR (the syfrack_amd/ directory of a built tree).
fn main() {
    println!("cargo:rerun-if-env-changed=SYFRACK_HIP_LIB_DIR");
    if let Ok(dir) = std::env::var("SYFRACK_HIP_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
    println!("cargo:rustc-link-lib=dylib=syfrack_hip");
}
