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

// ---- dense SPD inverse of a large matrix (one body's blob mobility) ----------
// Work space of k_spd_inverse_blocked for an n_pad x n_pad matrix (n_pad a
// multiple of kDenseTile): 64 x 64 pivot block, its pivots, status flags.
constexpr int kDenseTile = 64;
struct DenseWork {
    double* pivot = nullptr;    // kDenseTile^2
    int64_t* off = nullptr;     // 1 (zero)
    int32_t* size = nullptr;    // 1 (kDenseTile)
    int32_t* piv = nullptr;     // kDenseTile
    int32_t* piv_off = nullptr; // 1 (zero)
    int32_t* status = nullptr;  // n_pad / kDenseTile, one per pivot step
};
// A (row-major n_pad x n_pad, symmetric positive definite) := A^{-1} by
// blocked Gauss-Jordan elimination in 64 x 64 tiles: per step k the pivot
// block is inverted (partial pivoting inside it), then the row panel, the
// trailing update and the column panel are 64^3-FMA tile products, one
// workgroup per tile.  work.status[k] = 1 if step k met a singular pivot.
hipError_t k_spd_inverse_blocked(double* A, int n_pad, const DenseWork& work, hipStream_t stream);
// y = A x for a symmetric n_pad x n_pad A (n_pad a multiple of 64; x and y
// of n_pad entries) from its lower triangle: per 64 x 64 tile the row sums
// and, off the diagonal, the column sums into partial (chunks x n_pad,
// chunks = n_pad / 64), then summed in a fixed order.
hipError_t k_symv_dense(const double* A, int n_pad, const double* x, double* y, double* partial, int chunks,
                        hipStream_t stream);

}  // namespace tfs
