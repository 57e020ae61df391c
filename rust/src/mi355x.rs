//! `mi355x` feature of the reference crate: its public entry point and its
//! operator seams on one MI355X through the C ABI of include/spectral_mi355x.h
//! (compile-unverified: this build has no cargo/rustc; the C ABI itself is
//! exercised by the repository's tests through ctypes).
//!
//! Wiring (see rust/Cargo-feature.toml):
//!   * `src/lib.rs`: `#[cfg(feature = "mi355x")] pub mod mi355x;` and, in
//!     `spectral_init` (src/lib.rs:198-204), `#[cfg(feature = "mi355x")] return
//!     mi355x::spectral_init(graph, n_components, seed, data, config);` as its
//!     first statement -- same signature, same errors;
//!   * `src/operator.rs`: `DeviceOperator` below implements the crate's
//!     `LinearOperator` (src/operator.rs:6-12) and `spmv_csr` is the device
//!     form of the raw-slice replacement point (src/operator.rs:16-57).
use std::cell::RefCell;
use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

use ndarray::{Array2, ArrayView2};
use sprs::CsMatI;

use crate::operator::LinearOperator;
use crate::{ComputeMode, SpectralError, SpectralInitConfig};

// ---------------------------------------------------------------- C ABI
#[repr(C)]
pub struct SiContext {
    _private: [u8; 0],
}
#[repr(C)]
pub struct SiLaplacian {
    _private: [u8; 0],
}

/// si_config (include/spectral_mi355x.h); `struct_size` first.
#[repr(C)]
pub struct SiConfig {
    pub struct_size: u32,
    pub compute_mode: i32,
    pub solver: i32,
    pub block_size: i32,
    pub cheb_degree: i32,
    pub max_outer: i32,
    pub tol: f64,
    pub dense_threshold: i64,
    pub warm_start: *const f64,
    pub batch_components: i32,
}

/// si_report; fields past `laplacian_ms` are written when `struct_size` covers them.
#[repr(C)]
pub struct SiReport {
    pub struct_size: u32,
    pub level: i32,
    pub restarts: i32,
    pub block_size: i32,
    pub iterations: i64,
    pub spmm_calls: i64,
    pub max_residual: f64,
    pub eigenvalues: [f64; 8],
    pub solve_ms: f64,
    pub laplacian_ms: f64,
    pub passes: i32,
    pub reserved: i32,
    pub level_ms: [f64; 6],
}

pub const SI_OK: c_int = 0;
pub const SI_ERR_CONVERGENCE: c_int = 1;
pub const SI_ERR_INVALID_GRAPH: c_int = 2;
pub const SI_ERR_TOO_FEW_NODES: c_int = 3;
pub const SI_F64: i32 = 1;

extern "C" {
    pub fn si_config_default(cfg: *mut SiConfig);
    pub fn si_context_create(device: c_int, out: *mut *mut SiContext) -> c_int;
    pub fn si_context_destroy(ctx: *mut SiContext);
    pub fn si_last_error(ctx: *const SiContext) -> *const c_char;
    pub fn si_spectral_init(
        ctx: *mut SiContext, n: i64, indptr: *const i64, indices: *const i32, data: *const f32,
        n_components: i32, seed: u64, raw_data: *const f32, n_features: i64,
        cfg: *const SiConfig, out: *mut f32, rep: *mut SiReport,
    ) -> c_int;
    pub fn si_laplacian_build(
        ctx: *mut SiContext, n: i64, indptr: *const i64, indices: *const i32, data: *const f32,
        input_on_device: c_int, compute_mode: i32, out: *mut *mut SiLaplacian,
    ) -> c_int;
    pub fn si_laplacian_destroy(lap: *mut SiLaplacian);
    pub fn si_operator_apply_host(
        ctx: *mut SiContext, lap: *const SiLaplacian, x: *const f64, y: *mut f64, k: i32,
    ) -> c_int;
    pub fn si_spmv_csr(
        ctx: *mut SiContext, n_rows: i64, indptr: *const u64, indices: *const u64,
        data: *const f64, data_len: i64, x: *const f64, x_len: i64, y: *mut f64,
    ) -> c_int;
}

// ---------------------------------------------------------------- context
/// One si_context per host thread (it owns a HIP stream and the solver
/// workspaces); device 0, or SPECTRAL_MI355X_DEVICE.
struct Ctx(*mut SiContext);

impl Drop for Ctx {
    fn drop(&mut self) {
        unsafe { si_context_destroy(self.0) }
    }
}

thread_local! {
    static CTX: RefCell<Option<Ctx>> = const { RefCell::new(None) };
}

fn with_ctx<R>(f: impl FnOnce(*mut SiContext) -> R) -> R {
    CTX.with(|cell| {
        let mut slot = cell.borrow_mut();
        if slot.is_none() {
            let device = std::env::var("SPECTRAL_MI355X_DEVICE")
                .ok()
                .and_then(|s| s.parse::<c_int>().ok())
                .unwrap_or(0);
            let mut raw = std::ptr::null_mut();
            let st = unsafe { si_context_create(device, &mut raw) };
            assert_eq!(st, SI_OK, "spectral-mi355x: no HIP device {device} (status {st})");
            *slot = Some(Ctx(raw));
        }
        f(slot.as_ref().unwrap().0)
    })
}

fn last_error(ctx: *mut SiContext) -> String {
    unsafe {
        let p = si_last_error(ctx);
        if p.is_null() { String::new() } else { CStr::from_ptr(p).to_string_lossy().into_owned() }
    }
}

/// si_status -> SpectralError (src/lib.rs:167-182); device and internal
/// failures stay panics, as the reference's internal errors are.
fn check(ctx: *mut SiContext, st: c_int, n: usize, dims: usize) -> Result<(), SpectralError> {
    match st {
        SI_OK => Ok(()),
        SI_ERR_CONVERGENCE => Err(SpectralError::ConvergenceFailure),
        SI_ERR_INVALID_GRAPH => {
            let msg = last_error(ctx);
            let msg = msg.strip_prefix("invalid graph: ").unwrap_or(&msg).to_string();
            Err(SpectralError::InvalidGraph(msg))
        }
        SI_ERR_TOO_FEW_NODES => Err(SpectralError::TooFewNodes { n, dims }),
        _ => panic!("spectral-mi355x: {} (status {st})", last_error(ctx)),
    }
}

// ---------------------------------------------------------------- spectral_init
/// `spectral_init` (src/lib.rs:198-204) on the device: the same arguments,
/// result and errors.
pub fn spectral_init(
    graph: &CsMatI<f32, u32, usize>,
    n_components: usize,
    seed: u64,
    data: Option<ArrayView2<'_, f32>>,
    config: SpectralInitConfig,
) -> Result<Array2<f32>, SpectralError> {
    let n = graph.rows();
    let indptr: Vec<i64> = graph.indptr().raw_storage().iter().map(|&p| p as i64).collect();
    let indices: Vec<i32> = graph.indices().iter().map(|&j| j as i32).collect();
    let mut cfg = std::mem::MaybeUninit::<SiConfig>::uninit();
    let mut cfg = unsafe {
        si_config_default(cfg.as_mut_ptr());
        cfg.assume_init()
    };
    cfg.compute_mode = match config.compute_mode {
        ComputeMode::PythonCompat => 0,
        _ => 1,
    };
    let owned;
    let (raw, nf) = match data {
        Some(d) => {
            owned = d.as_standard_layout().into_owned();
            (owned.as_ptr(), owned.ncols() as i64)
        }
        None => (std::ptr::null(), 0),
    };
    let mut out = Array2::<f32>::zeros((n, n_components));
    with_ctx(|ctx| {
        let st = unsafe {
            si_spectral_init(
                ctx, n as i64, indptr.as_ptr(), indices.as_ptr(), graph.data().as_ptr(),
                n_components as i32, seed, raw, nf, &cfg, out.as_mut_ptr(), std::ptr::null_mut(),
            )
        };
        check(ctx, st, n, n_components)
    })?;
    Ok(out)
}

// ---------------------------------------------------------------- operator seams
/// The normalized Laplacian of a graph resident on the device, as the crate's
/// `LinearOperator` (src/operator.rs:6-12): `apply` computes y = L x with the
/// exact FP64 Laplacian on the device for host blocks [n x k].
pub struct DeviceOperator {
    lap: *mut SiLaplacian,
    n: usize,
}

impl DeviceOperator {
    pub fn new(graph: &CsMatI<f32, u32, usize>, mode: ComputeMode) -> Result<Self, SpectralError> {
        let n = graph.rows();
        let indptr: Vec<i64> = graph.indptr().raw_storage().iter().map(|&p| p as i64).collect();
        let indices: Vec<i32> = graph.indices().iter().map(|&j| j as i32).collect();
        let mode = match mode {
            ComputeMode::PythonCompat => 0,
            _ => 1,
        };
        let mut lap = std::ptr::null_mut();
        with_ctx(|ctx| {
            let st = unsafe {
                si_laplacian_build(ctx, n as i64, indptr.as_ptr(), indices.as_ptr(),
                                   graph.data().as_ptr(), 0, mode, &mut lap)
            };
            check(ctx, st, n, 0)
        })?;
        Ok(DeviceOperator { lap, n })
    }
}

impl Drop for DeviceOperator {
    fn drop(&mut self) {
        unsafe { si_laplacian_destroy(self.lap) }
    }
}

impl LinearOperator for DeviceOperator {
    fn apply(&self, x: ArrayView2<f64>, y: &mut Array2<f64>) {
        let k = x.ncols();
        let xs = x.as_standard_layout();
        let mut ys = Array2::<f64>::zeros((self.n, k));
        with_ctx(|ctx| {
            let st = unsafe {
                si_operator_apply_host(ctx, self.lap, xs.as_ptr(), ys.as_mut_ptr(), k as i32)
            };
            assert_eq!(st, SI_OK, "spectral-mi355x: {}", last_error(ctx));
        });
        y.assign(&ys);
    }

    fn size(&self) -> usize {
        self.n
    }
}

/// `spmv_csr` (src/operator.rs:16-57) on the device: the same raw-slice
/// signature, y bit-identical to the scalar loop; the reference's precondition
/// panics stay panics.
pub fn spmv_csr(indptr: &[usize], indices: &[usize], data: &[f64], x: &[f64], y: &mut [f64]) {
    assert!(indptr.len() == y.len() + 1, "CSR invariant violated: indptr.len() != y.len()+1");
    let ip: Vec<u64> = indptr.iter().map(|&p| p as u64).collect();
    let ix: Vec<u64> = indices.iter().map(|&j| j as u64).collect();
    with_ctx(|ctx| {
        let st = unsafe {
            si_spmv_csr(ctx, y.len() as i64, ip.as_ptr(), ix.as_ptr(), data.as_ptr(),
                        data.len() as i64, x.as_ptr(), x.len() as i64, y.as_mut_ptr())
        };
        assert_eq!(st, SI_OK, "{}", last_error(ctx));
    });
}
