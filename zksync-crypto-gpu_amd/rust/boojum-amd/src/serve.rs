//! Serve mode (no reference counterpart): coset-LDE requests admitted by memory budget and
//! batched continuously on the device; p50/p99 latency in `stats`.
use crate::context::Context;
use crate::cudart::result::CudaResult;
use crate::{check, sys};

pub use sys::{ba_serve_request as Request, ba_serve_result as Outcome, ba_serve_stats as Stats};

pub struct Server {
    handle: sys::ba_server_t,
}

// the library's server is internally synchronised: submit / wait / release from any thread
unsafe impl Send for Server {}
unsafe impl Sync for Server {}

impl Server {
    pub fn create(ctx: &Context, memory_budget_bytes: usize, max_inflight_batches: u32) -> CudaResult<Self> {
        let mut h = core::ptr::null_mut();
        check(unsafe { sys::ba_server_create(ctx.raw(), memory_budget_bytes, max_inflight_batches, &mut h) },
              "server_create")?;
        Ok(Self { handle: h })
    }
    /// # Safety
    /// Host or device evaluations named by the request stay valid until the request completes.
    pub unsafe fn submit(&self, request: &Request) -> CudaResult<u64> {
        let mut t = 0u64;
        check(sys::ba_server_submit(self.handle, request, &mut t), "server_submit")?;
        Ok(t)
    }
    pub fn wait(&self, ticket: u64) -> CudaResult<Outcome> {
        let mut r = core::mem::MaybeUninit::<Outcome>::zeroed();
        check(unsafe { sys::ba_server_wait(self.handle, ticket, r.as_mut_ptr()) }, "server_wait")?;
        Ok(unsafe { r.assume_init() })
    }
    pub fn release(&self, ticket: u64) -> CudaResult<()> {
        check(unsafe { sys::ba_server_release(self.handle, ticket) }, "server_release")
    }
    pub fn stats(&self) -> CudaResult<Stats> {
        let mut s = Stats::default();
        check(unsafe { sys::ba_server_stats(self.handle, &mut s) }, "server_stats")?;
        Ok(s)
    }
}

impl Drop for Server {
    fn drop(&mut self) {
        unsafe { sys::ba_server_destroy(self.handle) };
    }
}
