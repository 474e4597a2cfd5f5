//! The `Backend::Hip` arm of the reference crate, NOT COMPILED HERE (no Rust
//! toolchain in this environment).  A maintainer adds, in a checkout of the
//! reference:
//!
//!   src/rust/src/hip_sys.rs      this directory's hip_sys.rs
//!   src/rust/src/hip_backend.rs  this file
//!   src/rust/src/lib.rs          `#[cfg(feature = "hip")] mod hip_sys;`
//!                                `#[cfg(feature = "hip")] mod hip_backend;`
//!   src/rust/src/gpu.rs:7-11     a `Hip` variant in `enum Backend`
//!   src/rust/src/gpu.rs:15-21    `"hip" | "rocm" => Backend::Hip` in from_name
//!   src/rust/src/chain_kernel.rs:528 (before the unavailable-backend panic)
//!       #[cfg(feature = "hip")]
//!       if backend == crate::gpu::Backend::Hip {
//!           return crate::hip_backend::run_chains(code, consts, prior_code,
//!               prior_consts, n_params, data, n_cols, n_obs, init,
//!               proposal_sd, temperatures, n_iter, seed, proposal_mode,
//!               gamma, de_noise, proposal_l);
//!       }
//!   src/rust/src/lib.rs:51-61    in rust_available_backends:
//!       if cfg!(feature = "hip") && crate::hip_backend::available() {
//!           v.push("hip".to_string());
//!       }
//!   src/rust/Cargo.toml [features]  hip = []   (hip_cargo_features.toml)
//!
//! Argument meaning, draw layout and the ChainResult are the reference's;
//! errors become the panic the R guard (R/api.R:12-34) turns into hints.
use crate::chain_kernel::ChainResult;
use crate::hip_sys::*;
use std::ffi::CStr;

pub fn available() -> bool {
    unsafe { gpum_hip_available() == 1 }
}

fn last_error() -> String {
    unsafe { CStr::from_ptr(gpum_hip_last_error()) }
        .to_string_lossy()
        .into_owned()
}

fn opt_ptr(s: &[f64]) -> *const f64 {
    if s.is_empty() {
        std::ptr::null()
    } else {
        s.as_ptr()
    }
}

/// Copy a library-owned result into the reference's ChainResult and free it.
fn take(mut r: GpumChainResult) -> ChainResult {
    let n = (r.n_iter * r.n_chains * r.n_params) as usize;
    let c = r.n_chains as usize;
    let out = unsafe {
        ChainResult {
            n_iter: r.n_iter as usize,
            n_chains: c,
            n_params: r.n_params as usize,
            draws: std::slice::from_raw_parts(r.draws, n).to_vec(),
            accept_rate: std::slice::from_raw_parts(r.accept_rate, c).to_vec(),
            last_log_post: std::slice::from_raw_parts(r.last_log_post, c).to_vec(),
        }
    };
    unsafe { gpum_hip_result_free(&mut r) };
    out
}

/// gpu_metropolis_chain (chain_kernel.rs:468-488) on the HIP backend.
#[allow(clippy::too_many_arguments)]
pub fn run_chains(
    code: &[u32], consts: &[f64], prior_code: &[u32], prior_consts: &[f64],
    n_params: usize, data: &[f64], n_cols: usize, n_obs: usize, init: &[f64],
    proposal_sd: &[f64], temperatures: &[f64], n_iter: usize, seed: u32,
    proposal_mode: u32, gamma: f64, de_noise: f64, proposal_l: &[f64],
) -> ChainResult {
    let n_chains = init.len() / n_params;
    let mut r = GpumChainResult::empty();
    let rc = unsafe {
        gpum_hip_run_chains(
            code.as_ptr(), code.len(), opt_ptr(consts), consts.len(),
            prior_code.as_ptr(), prior_code.len(), opt_ptr(prior_consts),
            prior_consts.len(), n_params as u32, data.as_ptr(), n_cols as u32,
            n_obs as u32, init.as_ptr(), n_chains as u32, proposal_sd.as_ptr(),
            temperatures.as_ptr(), n_iter as u32, seed, proposal_mode, gamma,
            de_noise, opt_ptr(proposal_l), &mut r,
        )
    };
    if rc != 0 {
        panic!("{}", last_error());
    }
    take(r)
}

/// run_de_sync (cpu_native.rs:400-506) on the HIP backend, behind
/// rust_gpu_metropolis_de_sync (lib.rs:215-260).
#[allow(clippy::too_many_arguments)]
pub fn run_de_sync(
    code: &[u32], consts: &[f64], prior_code: &[u32], prior_consts: &[f64],
    n_params: usize, data: &[f64], n_cols: usize, n_obs: usize, init: &[f64],
    proposal_sd: &[f64], n_iter: usize, seed: u32, gamma: f64, de_noise: f64,
) -> ChainResult {
    let n_chains = init.len() / n_params;
    let mut r = GpumChainResult::empty();
    let rc = unsafe {
        gpum_hip_run_de_sync(
            code.as_ptr(), code.len(), opt_ptr(consts), consts.len(),
            prior_code.as_ptr(), prior_code.len(), opt_ptr(prior_consts),
            prior_consts.len(), n_params as u32, data.as_ptr(), n_cols as u32,
            n_obs as u32, init.as_ptr(), n_chains as u32, proposal_sd.as_ptr(),
            n_iter as u32, seed, gamma, de_noise, &mut r,
        )
    };
    if rc != 0 {
        panic!("{}", last_error());
    }
    take(r)
}

/// loglik_pointwise (cpu_native.rs:371-391) and grad_batch (:306-329) on the
/// device in FP64, behind rust_loglik_pointwise / rust_grad_batch
/// (lib.rs:158-207).  `Err` leaves the CPU path as the caller's fallback
/// (more than 16 parameters or 1024 instructions).
#[allow(clippy::too_many_arguments)]
pub fn eval_points(
    gradient: bool, code: &[u32], consts: &[f64], n_params: usize,
    data: &[f64], n_cols: usize, n_obs: usize, points: &[f64],
) -> Result<Vec<f64>, String> {
    let n_points = if n_params == 0 { 0 } else { points.len() / n_params };
    let len = n_points * if gradient { n_params } else { n_obs };
    let mut out = vec![0.0f64; len];
    let f = if gradient { gpum_hip_grad_batch } else { gpum_hip_loglik_pointwise };
    let rc = unsafe {
        f(code.as_ptr(), code.len(), opt_ptr(consts), consts.len(),
          n_params as u32, data.as_ptr(), n_cols as u32, n_obs as u32,
          opt_ptr(points), n_points as u32, out.as_mut_ptr())
    };
    if rc != 0 {
        return Err(last_error());
    }
    Ok(out)
}
