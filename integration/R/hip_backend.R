# R-side wiring of backend = "hip" for the reference package (NOT RUN HERE:
# this environment has no R).  Each block names the reference line it
# changes; the code is the replacement or addition.

## R/api.R:254 -- the backend argument accepts "hip":
##   backend = c("auto", "cpu", "cuda", "vulkan", "hip")

## R/api.R:351-371 -- "auto" prefers the HIP backend when it was compiled in
## and a gfx950 device answers (rust_available_backends() lists "hip" only
## then, see hip_backend.rs), then CUDA, then Vulkan, else the CPU.
.gpum_auto_backend <- function(avail) {
  for (b in c("hip", "cuda", "vulkan")) {
    if (b %in% avail) return(b)
  }
  "cpu"
}
## ... and in gpu_metropolis(): if (identical(backend, "auto"))
##   backend <- .gpum_auto_backend(avail)   (keeping the CPU message)

## R/api.R:373-377 -- MALA runs on "hip" (the Langevin step and its FP64
## gradient run on the device); the refusal stays for cuda / vulkan:
.gpum_mala_backend_ok <- function(backend) backend %in% c("cpu", "hip")
##   if (identical(method, "mala") && !.gpum_mala_backend_ok(backend)) stop(...)

## R/api.R:464-472 -- synchronous DE (de_sync = TRUE) runs on "hip"; the
## Rust side dispatches rust_gpu_metropolis_de_sync to gpum_hip_run_de_sync
## under the hip feature:
.gpum_de_sync_backend_ok <- function(backend) backend %in% c("cpu", "hip")

## R/api.R:12-34 (.gpum_backend_guard) needs no change: the library's error
## strings contain "out of memory", "kernel", "device" or "driver" in the
## cases the guard's patterns match.

## tools/config.R:73-129 -- detect a HIP toolchain as nvcc / vulkaninfo
## are detected, with the same override variables:
.gpum_detect_hip <- function(selected_backends) {
  env_hip <- Sys.getenv("GPUMETROPOLIS_HIP", unset = "")
  if (nzchar(Sys.which("hipcc"))) selected_backends <- c(selected_backends, "hip")
  if (nzchar(env_hip)) {
    v <- tolower(env_hip)
    if (v %in% c("0", "false", "no", "off")) selected_backends <- setdiff(selected_backends, "hip")
    if (v %in% c("1", "true", "yes", "on")) selected_backends <- union(selected_backends, "hip")
  }
  selected_backends
}
## ... called after the nvcc / vulkaninfo checks; the final filter becomes
##   intersect(selected_backends, c("cuda", "vulkan", "hip"))
## and, when "hip" is selected, @GPUM_HIP_LIBDIR@ in src/Makevars.in is set
## to the directory holding libgpum_hip.so (GPUMETROPOLIS_HIP_LIBDIR, else
## the in-tree gpumetropolis_amd/lib of this backend's checkout).
.gpum_hip_libdir <- function(default) {
  d <- Sys.getenv("GPUMETROPOLIS_HIP_LIBDIR", unset = "")
  if (nzchar(d)) d else default
}
