// Links libbvht_hip.so.  BVHT_LIB_DIR names the directory that holds a built
// library; without it the library is built in-tree by the repository's
// Python build script (hipcc --offload-arch=gfx950), which puts it in
// <repository>/bvhtracer_amd/.
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    println!("cargo:rerun-if-env-changed=BVHT_LIB_DIR");
    let dir = match env::var("BVHT_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
            let repo = manifest.parent().and_then(|p| p.parent()).expect("crate sits at <repository>/rust/<crate>");
            let status = Command::new("python3")
                .args(["-m", "bvhtracer_amd.build"])
                .current_dir(repo)
                .status()
                .expect("running python3 -m bvhtracer_amd.build");
            assert!(status.success(), "building libbvht_hip.so failed");
            println!("cargo:rerun-if-changed={}", repo.join("bvhtracer_amd/csrc").display());
            repo.join("bvhtracer_amd")
        }
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=bvht_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
