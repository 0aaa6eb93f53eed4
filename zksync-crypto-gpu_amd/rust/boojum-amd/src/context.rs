//! `boojum_cuda::context::Context` (crates/boojum-cuda/src/context.rs:157-296): power tables of
//! omega_{2^30}, 7 and 7^-1 resident in HBM.  The most recently created live context of a
//! device serves the context-free calls (ntt, ops_complex, barycentric), as the reference's
//! process-global tables do.
use crate::cudart::result::CudaResult;
use crate::cudart::stream::CudaStream;
use crate::{check, sys};

pub const OMEGA_LOG_ORDER: u32 = sys::BA_OMEGA_LOG_ORDER;

pub struct Context {
    handle: sys::ba_context_t,
}

unsafe impl Send for Context {}
unsafe impl Sync for Context {}

impl Context {
    pub fn create(powers_of_w_coarse_log_count: u32, powers_of_g_coarse_log_count: u32) -> CudaResult<Self> {
        let mut h = core::ptr::null_mut();
        check(unsafe { sys::ba_context_create(powers_of_w_coarse_log_count, powers_of_g_coarse_log_count, &mut h) },
              "Context::create")?;
        Ok(Self { handle: h })
    }

    pub fn destroy(mut self) -> CudaResult<()> {
        let h = core::mem::replace(&mut self.handle, core::ptr::null_mut());
        check(unsafe { sys::ba_context_destroy(h) }, "Context::destroy")
    }

    pub fn raw(&self) -> sys::ba_context_t {
        self.handle
    }

    /// Times every pass split and per-pass columns-per-workgroup for log_n in the range and
    /// keeps the fastest on this context (no reference counterpart).
    pub fn autotune(&self, log_n_min: u32, log_n_max: u32, elems_per_trial: u64, stream: &CudaStream)
                    -> CudaResult<()> {
        check(unsafe { sys::ba_autotune(self.handle, log_n_min, log_n_max, elems_per_trial, stream.raw()) },
              "autotune")
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        if !self.handle.is_null() {
            unsafe { sys::ba_context_destroy(self.handle) };
        }
    }
}
