"""boojum_cuda::ntt mirror (crates/boojum-cuda/src/ntt.rs:179-369) over the C-ABI.

Same argument order and meaning as the Rust functions; DeviceSlices carry (pointer, length).
Invalid arguments (the reference's assert! preconditions) raise BoojumAmdError with status 1.
"""
from ._lib import check, lib


def _h(stream):
    return stream.handle if stream is not None else None


def batch_ntt_internal(inputs_ptr, outputs_ptr, log_n, num_ntts, stride_between_input_arrays,
                       stride_between_output_arrays, bitrev_inputs, inverse, log_extension_degree,
                       coset_index, stream=None):
    check(lib().ba_batch_ntt_internal(inputs_ptr, outputs_ptr, log_n, num_ntts,
                                      stride_between_input_arrays, stride_between_output_arrays,
                                      int(bitrev_inputs), int(inverse), log_extension_degree,
                                      coset_index, _h(stream)), "batch_ntt_internal")


def batch_ntt_out_of_place(inputs_matrix, outputs_matrix, log_n, num_ntts, inputs_offset,
                           outputs_offset, stride_between_input_arrays,
                           stride_between_output_arrays, bitrev_inputs, inverse,
                           log_extension_degree, coset_index, stream=None):
    check(lib().ba_batch_ntt_out_of_place(inputs_matrix.ptr, inputs_matrix.len, outputs_matrix.ptr,
                                          outputs_matrix.len, log_n, num_ntts, inputs_offset,
                                          outputs_offset, stride_between_input_arrays,
                                          stride_between_output_arrays, int(bitrev_inputs),
                                          int(inverse), log_extension_degree, coset_index,
                                          _h(stream)), "batch_ntt_out_of_place")


def batch_ntt_in_place(inputs_matrix, log_n, num_ntts, inputs_offset, stride_between_input_arrays,
                       bitrev_inputs, inverse, log_extension_degree, coset_index, stream=None):
    check(lib().ba_batch_ntt_in_place(inputs_matrix.ptr, inputs_matrix.len, log_n, num_ntts,
                                      inputs_offset, stride_between_input_arrays, int(bitrev_inputs),
                                      int(inverse), log_extension_degree, coset_index, _h(stream)),
          "batch_ntt_in_place")


def batch_ntt_ctx(ctx, inputs_ptr, outputs_ptr, log_n, num_ntts, stride_in, stride_out,
                  bitrev_inputs, inverse, log_extension_degree=0, coset_index=0, stream=None):
    check(lib().ba_batch_ntt_ctx(ctx.handle, inputs_ptr, outputs_ptr, log_n, num_ntts, stride_in,
                                 stride_out, int(bitrev_inputs), int(inverse), log_extension_degree,
                                 coset_index, _h(stream)), "batch_ntt_ctx")
