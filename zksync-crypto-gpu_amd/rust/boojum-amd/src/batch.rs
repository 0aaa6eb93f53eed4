//! Uneven batches (no reference counterpart): many NTT / LDE jobs planned once into grouped
//! launches, launched any number of times (capturable in a graph).
use crate::context::Context;
use crate::cudart::result::CudaResult;
use crate::cudart::stream::CudaStream;
use crate::{check, sys};

pub use sys::ba_job as Job;

pub struct BatchPlan {
    handle: sys::ba_batch_plan_t,
}

impl BatchPlan {
    /// # Safety
    /// The jobs' pointers are device memory that outlives every launch of the plan.
    pub unsafe fn create(ctx: &Context, jobs: &[Job]) -> CudaResult<Self> {
        let mut h = core::ptr::null_mut();
        check(sys::ba_batch_plan_create(ctx.raw(), jobs.as_ptr(), jobs.len() as u32, &mut h), "batch_plan_create")?;
        Ok(Self { handle: h })
    }
    pub fn launch(&self, stream: &CudaStream) -> CudaResult<()> {
        check(unsafe { sys::ba_batch_plan_launch(self.handle, stream.raw()) }, "batch_plan_launch")
    }
    /// (grouped launches, dependency levels, workgroups)
    pub fn stats(&self) -> CudaResult<(u32, u32, u64)> {
        let (mut l, mut v, mut w) = (0u32, 0u32, 0u64);
        check(unsafe { sys::ba_batch_plan_stats(self.handle, &mut l, &mut v, &mut w) }, "batch_plan_stats")?;
        Ok((l, v, w))
    }
}

impl Drop for BatchPlan {
    fn drop(&mut self) {
        unsafe { sys::ba_batch_plan_destroy(self.handle) };
    }
}
