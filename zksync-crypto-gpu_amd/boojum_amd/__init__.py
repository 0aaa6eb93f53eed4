"""Python view of libboojum_amd.so, named after the Rust API it mirrors.

The product is the C-ABI library (include/boojum_amd.h); this module is the thin host-side
binding that tests and the benchmark drive it through, with the names of the reference's
crates:
  runtime   -- era_cudart (crates/cudart): DeviceAllocation, CudaStream, CudaEvent
  context   -- boojum_cuda::context::Context (crates/boojum-cuda/src/context.rs:157-296)
  ntt       -- boojum_cuda::ntt (crates/boojum-cuda/src/ntt.rs:179-369)
  ops       -- boojum_cuda::ops_complex::bit_reverse* (crates/boojum-cuda/src/ops_complex.rs:262-331)
There is no CPU fallback: every call goes to the HIP library, and loading fails loudly when
the library is missing.
"""
from ._lib import BoojumAmdError, check, lib, library_path  # noqa: F401
from .runtime import DeviceAllocation, DeviceSlice, CudaStream, CudaEvent  # noqa: F401
from .runtime import device_count, synchronize, mem_get_info, set_device  # noqa: F401
from .context import Context, OMEGA_LOG_ORDER  # noqa: F401
from . import barycentric, batch, lde, ntt, ops, poseidon2, profile, serve, tune  # noqa: F401
