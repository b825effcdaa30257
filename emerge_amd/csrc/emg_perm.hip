// Numbering maps between this library's DOF order and a caller's.
//
// The library numbers edges and faces lexicographically by their vertex ids;
// the reference numbers them in the iteration order of Python sets
// (fem/mesh3d.py:252-271), so a drop-in that hands matrices and vectors back
// to the reference's callers must renumber them.  All of it stays on the GPU:
//
//   * emg_csr_permute: the caller-order CSR of a matrix on this library's
//     pattern (rows permuted, columns relabelled and re-sorted) and the map
//     from each caller entry to the library's entry, so every later value
//     array is one gather (tet_mass_stiffness_matrices,
//     fem/physics/edm/optimized_assembly.py:43-64, and assemble_freq_matrix,
//     fem/physics/edm/assembler.py:312-388, return caller-order CSR);
//   * emg_csr_map: the library position of every entry of a caller-order CSR
//     (SolveRoutine.solve, fem/solver.py:405-469, receives one);
//   * emg_gather_c128 / emg_scatter_c128: vectors and value arrays through
//     such maps.
//
// A permutation row of the pattern has a few dozen entries, so the column
// sort is a rank count per entry (each entry counts the smaller relabelled
// columns of its row): no global sort, deterministic.

#include <cstring>
#include <string>

#include <rocprim/rocprim.hpp>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

namespace {

// len[r] = length of library row p[r]
__global__ void __launch_bounds__(256) k_perm_lengths(int64_t n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ p, int32_t *__restrict__ len) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t i = p[r];
    len[r] = (i >= 0 && i < n) ? rowptr[i + 1] - rowptr[i] : 0;
}

// One wavefront per caller row r (library row i = p[r]): lane j takes entries
// j, j + 64, ... and writes each at its rank among the row's relabelled columns.
__global__ void __launch_bounds__(256) k_perm_rows(int64_t n, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ cols, const int32_t *__restrict__ p,
                                                   const int32_t *__restrict__ pinv,
                                                   const int32_t *__restrict__ rowptr_out, int32_t *__restrict__ cols_out,
                                                   int32_t *__restrict__ map_out) {
    const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x % 64;
    if (r >= n) return;
    const int32_t i = p[r];
    if (i < 0 || i >= n) return;
    const int32_t b = rowptr[i], len = rowptr[i + 1] - b;
    const int32_t o = rowptr_out[r];
    for (int32_t j = lane; j < len; j += 64) {
        const int32_t c = pinv[cols[b + j]];
        int32_t rank = 0;
        for (int32_t k = 0; k < len; k++) rank += pinv[cols[b + k]] < c;
        cols_out[o + rank] = c;
        map_out[o + rank] = b + j;
    }
}

// pos[j] = library position of caller entry j (row r, column cols_c[j]) or -1
__global__ void __launch_bounds__(256) k_csr_map(int64_t n, const int32_t *__restrict__ rowptr_c,
                                                 const int32_t *__restrict__ cols_c, const int32_t *__restrict__ p,
                                                 const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                                                 int32_t *__restrict__ pos) {
    const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x % 64;
    if (r >= n) return;
    const int32_t i = p[r];
    const bool ok = i >= 0 && i < n;
    const int32_t lo0 = ok ? rowptr[i] : 0, hi0 = ok ? rowptr[i + 1] : 0;
    for (int32_t j = rowptr_c[r] + lane; j < rowptr_c[r + 1]; j += 64) {
        const int32_t cc = cols_c[j];
        if (!ok || cc < 0 || cc >= n) {
            pos[j] = -1;
            continue;
        }
        const int32_t c = p[cc];
        int32_t lo = lo0, hi = hi0;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (cols[mid] < c) lo = mid + 1; else hi = mid;
        }
        pos[j] = (lo < hi0 && cols[lo] == c) ? lo : -1;
    }
}

__global__ void __launch_bounds__(256) k_gather_c128(int64_t n, const int32_t *__restrict__ idx, int64_t n_src,
                                                     const double2 *__restrict__ src, double2 *__restrict__ dst) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t s = idx[k];
    dst[k] = (s >= 0 && s < n_src) ? src[s] : make_double2(0.0, 0.0);
}

__global__ void __launch_bounds__(256) k_scatter_c128(int64_t n, const int32_t *__restrict__ idx, int64_t n_dst,
                                                      const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                      unsigned long long *dropped) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t d = idx[k];
    const double2 v = src[k];
    if (d >= 0 && d < n_dst) dst[d] = v;
    else if (dropped != nullptr && (v.x != 0.0 || v.y != 0.0)) atomicAdd(dropped, 1ull);
}

}  // namespace

extern "C" int emg_csr_permute(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const int32_t *d_p,
                               const int32_t *d_pinv, int32_t *d_rowptr_out, int32_t *d_cols_out, int32_t *d_map,
                               void *stream) {
    EMG_CHECK(n > 0 && n < (1ll << 31), "emg_csr_permute: bad n");
    EMG_CHECK(d_rowptr && d_cols && d_p && d_pinv && d_rowptr_out && d_cols_out && d_map,
              "emg_csr_permute: null argument");
    hipStream_t st = (hipStream_t)stream;
    int32_t *len = nullptr;
    void *tmp = nullptr;
    size_t need = 0;
    EMG_HIP(hipMalloc(&len, (size_t)(n + 1) * sizeof(int32_t)));
    k_perm_lengths<<<emg::grid_for(n, 256), 256, 0, st>>>(n, d_rowptr, d_p, len);
    hipError_t e = hipMemsetAsync(len + n, 0, sizeof(int32_t), st);
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(nullptr, need, len, d_rowptr_out, (int32_t)0, (size_t)(n + 1),
                                    rocprim::plus<int32_t>(), st);
    if (e == hipSuccess) e = hipMalloc(&tmp, need);
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(tmp, need, len, d_rowptr_out, (int32_t)0, (size_t)(n + 1),
                                    rocprim::plus<int32_t>(), st);
    if (e == hipSuccess) {
        k_perm_rows<<<emg::grid_for(n * 64, 256), 256, 0, st>>>(n, d_rowptr, d_cols, d_p, d_pinv, d_rowptr_out,
                                                                d_cols_out, d_map);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    (void)hipFree(len);
    if (e != hipSuccess) {
        emg::set_error(std::string("emg_csr_permute: ") + hipGetErrorString(e));
        return -2;
    }
    return 0;
}

extern "C" int emg_csr_map(int64_t n, const int32_t *d_rowptr_c, const int32_t *d_cols_c, const int32_t *d_p,
                           const int32_t *d_rowptr, const int32_t *d_cols, int32_t *d_pos, void *stream) {
    EMG_CHECK(n > 0 && d_rowptr_c && d_cols_c && d_p && d_rowptr && d_cols && d_pos, "emg_csr_map: bad argument");
    k_csr_map<<<emg::grid_for(n * 64, 256), 256, 0, (hipStream_t)stream>>>(n, d_rowptr_c, d_cols_c, d_p, d_rowptr,
                                                                            d_cols, d_pos);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gather_c128(int64_t n, const int32_t *d_idx, int64_t n_src, const double *d_src, double *d_dst,
                               void *stream) {
    if (n == 0) return 0;
    EMG_CHECK(n > 0 && d_idx && d_src && d_dst, "emg_gather_c128: bad argument");
    k_gather_c128<<<emg::grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(
        n, d_idx, n_src, reinterpret_cast<const double2 *>(d_src), reinterpret_cast<double2 *>(d_dst));
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_scatter_c128(int64_t n, const int32_t *d_idx, int64_t n_dst, const double *d_src, double *d_dst,
                                unsigned long long *d_dropped, void *stream) {
    if (n == 0) return 0;
    EMG_CHECK(n > 0 && d_idx && d_src && d_dst, "emg_scatter_c128: bad argument");
    k_scatter_c128<<<emg::grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(
        n, d_idx, n_dst, reinterpret_cast<const double2 *>(d_src), reinterpret_cast<double2 *>(d_dst), d_dropped);
    EMG_LAUNCH_CHECK();
    return 0;
}
