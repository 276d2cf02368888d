// dense_kernels.h -- small dense linear algebra on the device, shared by the
// Broyden engine (J0 blocks) and the solver's block preconditioner.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace tfs {

// In-place inverse of `nblocks` row-major blocks (block b: sizes[b]^2 entries
// at A + off[b]) by Gauss-Jordan elimination with partial pivoting, one
// workgroup per block.  piv (piv_off[b] + sizes[b] entries per block) is
// scratch; status[b] = 1 if a pivot of block b is zero or not finite, else 0.
// max_size bounds sizes[] (it sizes the dynamic LDS: 2 * max_size doubles).
hipError_t k_batched_inverse(double* A, const int64_t* off, const int32_t* sizes, int nblocks, int max_size,
                             int32_t* piv, const int32_t* piv_off, int32_t* status, hipStream_t stream);

// The same for symmetric positive definite blocks of at most 64 rows, without
// pivoting, one wave per block (max_size > 64 is an invalid value).  status[b]
// = 1 if a pivot of block b is not positive: the block is then not inverted
// and the caller falls back to k_batched_inverse on a fresh copy.
hipError_t k_batched_inverse_spd(double* A, const int64_t* off, const int32_t* sizes, int nblocks, int max_size,
                                 int32_t* status, hipStream_t stream);

}  // namespace tfs
