//! `extern "C"` binding of the HIP backend (`include/gpum_hip.h`) for the
//! reference crate (`src/rust/src/hip_sys.rs` in a checkout of it).
//!
//! NOT COMPILED HERE: this environment has no Rust toolchain.  The structs
//! mirror the header field for field; `tests/test_abi.py` checks the same
//! layout (sizeof and every offsetof, as gcc lays the header out) against
//! this repository's ctypes mirror.  Build with the `hip` Cargo feature
//! (`hip_cargo_features.toml`) and link `-lgpum_hip`.
#![allow(non_camel_case_types, dead_code)]

use std::os::raw::c_char;

/// gpum_chain_result (header) == ChainResult (chain_kernel.rs:360-369) with
/// library-owned buffers, released by `gpum_hip_result_free`.
#[repr(C)]
pub struct GpumChainResult {
    pub n_iter: u32,
    pub n_chains: u32,
    pub n_params: u32,
    pub draws: *mut f32,
    pub accept_rate: *mut f32,
    pub last_log_post: *mut f64,
}

impl GpumChainResult {
    pub fn empty() -> Self {
        GpumChainResult {
            n_iter: 0,
            n_chains: 0,
            n_params: 0,
            draws: std::ptr::null_mut(),
            accept_rate: std::ptr::null_mut(),
            last_log_post: std::ptr::null_mut(),
        }
    }
}

/// gpum_run_ext: zero-initialise and set `struct_size`.
#[repr(C)]
pub struct GpumRunExt {
    pub struct_size: u32,
    pub stream_id: u32,
    pub trace_plp: *mut f64,
    pub trace_prop: *mut f64,
    pub trace_accept: *mut u8,
    pub final_state: *mut f64,
    pub init_log_post: *mut f64,
    pub chains_per_block: i32,
    pub waves_per_chain: i32,
    pub precompute_data: i32,
    pub fuse: i32,
    pub kernel_ms: f64,
    pub iter_offset: u32,
    pub swap_every: u32,
    pub swap_group: u32,
    pub adapt_batches: u32,
    pub batch0: u32,
    pub init_scale: *const f64,
    pub swap_hist: *mut u8,
    pub acc_hist: *mut f32,
    pub final_sd: *mut f64,
    pub final_scale: *mut f64,
    pub trace_ll: *mut f64,
    pub chains_per_wave: i32,
}

#[repr(C)]
pub struct GpumModelInfo {
    pub n_vm_instr: u32,
    pub vm_depth: u32,
    pub ref_depth: u32,
    pub n_uniform: u32,
    pub n_cols_total: u32,
    pub n_hoist_words: u32,
    pub w_exec: f64,
    pub w_bytecode: f64,
    pub sum_mode: u32,
    pub n_sum_tail: u32,
}

#[repr(C)]
pub struct GpumTrainOpts {
    pub struct_size: u32,
    pub steps: u32,
    pub batch: u32,
    pub seed: u32,
    pub lr: f64,
    pub lr_final: f64,
    pub beta1: f64,
    pub beta2: f64,
    pub eps: f64,
}

#[repr(C)]
pub struct GpumDaExt {
    pub struct_size: u32,
    pub stream_id: u32,
    pub iter_offset: u32,
    pub work_items: u32,
    pub work_groups: u32,
    pub final_state: *mut f64,
    pub stage1_rate: *mut f64,
    pub trace_s: *mut f64,
    pub trace_plp: *mut f64,
    pub trace_stage: *mut u8,
    pub kernel_ms: f64,
    pub exact_evals: f64,
    pub trace_prop: *mut f64,
    pub trace_ll: *mut f64,
    pub chains_per_wave: u32,
    pub gamma: f64,
    pub de_noise: f64,
}

#[repr(C)]
pub struct GpumBatchModel {
    pub model: *mut GpumHipModel,
    pub n_chains: u32,
    pub init: *const f64,
    pub proposal_sd: *const f64,
    pub seed: u32,
    pub stream_id: u32,
    pub gamma: f64,
    pub de_noise: f64,
    pub draws: *mut f32,
    pub accept_hist: *mut f32,
    pub final_sd: *mut f64,
    pub last_log_post: *mut f64,
    pub final_state: *mut f64,
}

#[repr(C)]
pub struct GpumBatchOpts {
    pub struct_size: u32,
    pub proposal_mode: u32,
    pub n_iter: u32,
    pub batch_len: u32,
    pub warmup_batches: u32,
    pub adapt: i32,
    pub chains_per_block: i32,
    pub waves_per_chain: i32,
    pub use_graph: i32,
    pub kernel_ms: f64,
    pub n_workgroups: u32,
    pub chains_per_wave: i32,
}

#[repr(C)]
pub struct GpumWarmupOpts {
    pub struct_size: u32,
    pub warmup: u32,
    pub seed: u32,
    pub auto_stop: i32,
    pub use_graph: i32,
    pub final_init: *mut f64,
    pub final_sd: *mut f64,
    pub final_scale: *mut f64,
    pub final_l: *mut f64,
    pub accept_hist: *mut f32,
    pub n_batches_planned: u32,
    pub n_batches: u32,
    pub warmup_used: u32,
    pub has_l: u32,
    pub kernel_ms: f64,
}

/// Opaque handles.
#[repr(C)]
pub struct GpumHipModel {
    _p: [u8; 0],
}
#[repr(C)]
pub struct GpumHipSurrogate {
    _p: [u8; 0],
}

#[link(name = "gpum_hip")]
extern "C" {
    pub fn gpum_hip_available() -> i32;
    pub fn gpum_hip_last_error() -> *const c_char;
    pub fn gpum_hip_run_chains(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        prior_code: *const u32, prior_len: usize, prior_consts: *const f64,
        n_prior_consts: usize, n_params: u32, data: *const f64, n_cols: u32,
        n_obs: u32, init: *const f64, n_chains: u32, proposal_sd: *const f64,
        temperatures: *const f64, n_iter: u32, seed: u32, proposal_mode: u32,
        gamma: f64, de_noise: f64, proposal_l: *const f64,
        out: *mut GpumChainResult,
    ) -> i32;
    pub fn gpum_hip_run_chains_ex(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        prior_code: *const u32, prior_len: usize, prior_consts: *const f64,
        n_prior_consts: usize, n_params: u32, data: *const f64, n_cols: u32,
        n_obs: u32, init: *const f64, n_chains: u32, proposal_sd: *const f64,
        temperatures: *const f64, n_iter: u32, seed: u32, proposal_mode: u32,
        gamma: f64, de_noise: f64, proposal_l: *const f64,
        ext: *mut GpumRunExt, out: *mut GpumChainResult,
    ) -> i32;
    pub fn gpum_hip_result_free(r: *mut GpumChainResult);
    pub fn gpum_hip_loglik_batch(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        n_params: u32, data: *const f64, n_cols: u32, n_obs: u32,
        points: *const f64, n_points: u32, out: *mut f64,
    ) -> i32;
    pub fn gpum_hip_loglik_pointwise(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        n_params: u32, data: *const f64, n_cols: u32, n_obs: u32,
        points: *const f64, n_points: u32, out: *mut f64,
    ) -> i32;
    pub fn gpum_hip_grad_batch(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        n_params: u32, data: *const f64, n_cols: u32, n_obs: u32,
        points: *const f64, n_points: u32, out: *mut f64,
    ) -> i32;
    pub fn gpum_hip_run_de_sync(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        prior_code: *const u32, prior_len: usize, prior_consts: *const f64,
        n_prior_consts: usize, n_params: u32, data: *const f64, n_cols: u32,
        n_obs: u32, init: *const f64, n_chains: u32, proposal_sd: *const f64,
        n_iter: u32, seed: u32, gamma: f64, de_noise: f64,
        out: *mut GpumChainResult,
    ) -> i32;
    pub fn gpum_hip_model_create(
        code: *const u32, code_len: usize, consts: *const f64, n_consts: usize,
        prior_code: *const u32, prior_len: usize, prior_consts: *const f64,
        n_prior_consts: usize, n_params: u32, data: *const f64, n_cols: u32,
        n_obs: u32, precompute_data: i32, fuse: i32,
    ) -> *mut GpumHipModel;
    pub fn gpum_hip_model_free(m: *mut GpumHipModel);
    pub fn gpum_hip_model_info(m: *const GpumHipModel, out: *mut GpumModelInfo) -> i32;
    pub fn gpum_hip_model_set_jit(m: *mut GpumHipModel, enable: i32,
                                  compile_ms: *mut f64) -> i32;
    pub fn gpum_hip_model_run(
        m: *mut GpumHipModel, init: *const f64, n_chains: u32,
        proposal_sd: *const f64, temperatures: *const f64, n_iter: u32,
        seed: u32, proposal_mode: u32, gamma: f64, de_noise: f64,
        proposal_l: *const f64, ext: *mut GpumRunExt, out: *mut GpumChainResult,
    ) -> i32;
    pub fn gpum_hip_batch_run(models: *mut GpumBatchModel, n_models: u32,
                              opts: *mut GpumBatchOpts) -> i32;
    pub fn gpum_hip_model_warmup(
        m: *mut GpumHipModel, init: *const f64, n_chains: u32,
        proposal_sd: *const f64, opts: *mut GpumWarmupOpts,
    ) -> i32;
    pub fn gpum_hip_model_run_da(
        m: *mut GpumHipModel, s: *mut GpumHipSurrogate, init: *const f64,
        n_chains: u32, proposal_sd: *const f64, temperatures: *const f64,
        n_iter: u32, seed: u32, proposal_mode: u32, proposal_l: *const f64,
        ext: *mut GpumDaExt, out: *mut GpumChainResult,
    ) -> i32;
}
