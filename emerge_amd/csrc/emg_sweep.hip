// Batched frequency-sweep kernels on one shared sparsity pattern.
//
// The reference sweep (fem/physics/edm/emfreq3d.py:607-732) assembles and
// factorises K(f) = E - k0^2 B + sum_t gamma_t(f) P_t once per frequency and
// solves one port at a time.  Every K(f) shares the pattern of E and B, so a
// block of (frequency, port) pairs is treated here as one multi-vector:
//
//   * emg_sweep_residual: r_j = mask.(c_j q_{p_j} - (E - s_j B) x_j - z_j) and
//     its FP64 norms for up to 16 vectors in ONE pass over E, B and the CSR
//     indices (an affine SpMM: the matrix bytes are read once per block, not
//     once per frequency, and K(f) is never materialised);
//   * emg_port_apply: z_j = sum_t g_{t,j} P_t x_j on the few rows that carry
//     port / absorbing-boundary blocks (assembler.py:100-144, 382-383);
//   * emg_basis_combine: X = V Y, the trial solutions of a block of pairs from
//     a basis of earlier solutions, written in the interleaved layout the
//     residual kernel gathers from.
//
// Multi-vectors are interleaved: X[row * nvec + j] (complex128), so the
// gather of one column index fetches the nvec values of that DOF as one
// contiguous run.

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

namespace {

struct c2 {
    double re, im;
};

__device__ __forceinline__ c2 ldc(const double *p, int64_t i) {
    const double2 v = reinterpret_cast<const double2 *>(p)[i];
    return c2{v.x, v.y};
}
__device__ __forceinline__ void stc(double *p, int64_t i, c2 v) {
    reinterpret_cast<double2 *>(p)[i] = make_double2(v.re, v.im);
}

constexpr int RES_BLOCK = 256;
constexpr int ROW_LANES = 64;        // one wavefront per entity (two CSR rows)
constexpr int MAX_VEC = 16;
constexpr int RES_GRID_MAX = 4096;   // persistent grid: partial sums per block

// One wavefront per entity e of the pattern (emg_pattern_csr: DOF = mode * ne +
// entity; rows (0, e) and (1, e) list the same column entities e'_k, a row's
// (1, e'_k) entries deg after its (0, e'_k) ones, row (1, e) half = rowptr[ne]
// entries after row (0, e)).  Lane l handles vector j = l % NV over the
// column entities l / NV, l / NV + 64 / NV, ...: the two NV-vector runs of x
// gathered for e'_k serve four matrix entries, and E, B and the column index
// of an entry are loaded once per group of NV lanes and shared by the NV vectors.
template <int NV>
__global__ void __launch_bounds__(RES_BLOCK)
    k_sweep_residual(int64_t ne, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                     const double *__restrict__ E, const double *__restrict__ B, const uint8_t *__restrict__ mask,
                     const double *__restrict__ s, const double *__restrict__ X, const double *__restrict__ Q,
                     const int32_t *__restrict__ qidx, const double *__restrict__ c,
                     const int32_t *__restrict__ prow_of_row, const double *__restrict__ Zp,
                     double *__restrict__ R, double *__restrict__ partial) {
    constexpr int SPLIT = ROW_LANES / NV;   // entry strides per vector
    const int lane = threadIdx.x % ROW_LANES;
    const int j = lane % NV;
    const int part = lane / NV;
    const double sj = s[j];
    const int qj = qidx[j];
    const c2 cj = ldc(c, j);
    const int64_t n = 2 * ne;
    const int64_t half = rowptr[ne];
    double rr = 0.0, bb = 0.0;
    const int64_t groups = (int64_t)gridDim.x * (RES_BLOCK / ROW_LANES);
    for (int64_t e = blockIdx.x * (int64_t)(RES_BLOCK / ROW_LANES) + threadIdx.x / ROW_LANES; e < ne; e += groups) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};   // row (0, e) re, im; row (1, e) re, im
        const int32_t r0 = rowptr[e];
        const int32_t deg = (rowptr[e + 1] - r0) >> 1;
        for (int32_t k = part; k < deg; k += SPLIT) {
            const int32_t col = cols[r0 + k];
            if (col < 0 || col >= ne) continue;   // not an entity pattern: no out-of-range reads
            const c2 x0 = ldc(X, (int64_t)col * NV + j), x1 = ldc(X, (int64_t)(ne + col) * NV + j);
            const int64_t pos[4] = {r0 + k, r0 + k + deg, half + r0 + k, half + r0 + k + deg};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const c2 ev = ldc(E, pos[q]), bv = ldc(B, pos[q]);
                const c2 x = (q & 1) ? x1 : x0;
                const double kre = ev.re - sj * bv.re, kim = ev.im - sj * bv.im;
                acc[(q >> 1) * 2] += kre * x.re - kim * x.im;
                acc[(q >> 1) * 2 + 1] += kre * x.im + kim * x.re;
            }
        }
#pragma unroll
        for (int off = ROW_LANES / 2; off >= NV; off >>= 1)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] += __shfl_xor(acc[q], off, ROW_LANES);
        if (part == 0) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int64_t row = h * ne + e;
                double are = acc[2 * h], aim = acc[2 * h + 1];
                const int32_t pr = prow_of_row ? prow_of_row[row] : -1;
                if (pr >= 0) {
                    const c2 z = ldc(Zp, (int64_t)pr * NV + j);
                    are += z.re;
                    aim += z.im;
                }
                double rre = 0.0, rim = 0.0;
                if (mask == nullptr || mask[row]) {
                    const c2 q = ldc(Q, (int64_t)qj * n + row);
                    const double bre = cj.re * q.re - cj.im * q.im, bim = cj.re * q.im + cj.im * q.re;
                    rre = bre - are;
                    rim = bim - aim;
                    rr += rre * rre + rim * rim;
                    bb += bre * bre + bim * bim;
                }
                if (R != nullptr) stc(R, row * NV + j, c2{rre, rim});
            }
        }
    }
    // block reduction of (rr, bb) per vector
    __shared__ double sr[RES_BLOCK], sb[RES_BLOCK];
    sr[threadIdx.x] = (part == 0) ? rr : 0.0;
    sb[threadIdx.x] = (part == 0) ? bb : 0.0;
    __syncthreads();
    if (threadIdx.x < NV) {
        double a = 0.0, d = 0.0;
        for (int t = threadIdx.x; t < RES_BLOCK; t += ROW_LANES) {
            a += sr[t];
            d += sb[t];
        }
        partial[((int64_t)blockIdx.x * NV + threadIdx.x) * 2 + 0] = a;
        partial[((int64_t)blockIdx.x * NV + threadIdx.x) * 2 + 1] = d;
    }
}

// norms[2j] = sum over blocks of |r_j|^2, norms[2j+1] = |b_j|^2
__global__ void __launch_bounds__(256) k_sweep_norms(int nblk, int nv, const double *__restrict__ partial,
                                                     double *__restrict__ norms) {
    const int j = blockIdx.x / 2, w = blockIdx.x % 2;
    double a = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) a += partial[((int64_t)b * nv + j) * 2 + w];
    __shared__ double sh[256];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) norms[2 * j + w] = sh[0];
}

// Zp[i, j] = sum over entries e of port row i of g[term_e, j] * val_e * X[col_e, j]
__global__ void __launch_bounds__(256) k_port_apply(int64_t nprow, int nv, const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ pcol, const int32_t *__restrict__ pterm,
                                                    const double *__restrict__ pval, const double *__restrict__ g,
                                                    const double *__restrict__ X, double *__restrict__ Zp) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nprow * nv) return;
    const int64_t i = t / nv;
    const int j = (int)(t % nv);
    double are = 0.0, aim = 0.0;
    for (int32_t e = ptr[i]; e < ptr[i + 1]; e++) {
        const c2 v = ldc(pval, e);
        const c2 gg = ldc(g, (int64_t)pterm[e] * nv + j);
        const c2 x = ldc(X, (int64_t)pcol[e] * nv + j);
        const double kre = gg.re * v.re - gg.im * v.im, kim = gg.re * v.im + gg.im * v.re;
        are += kre * x.re - kim * x.im;
        aim += kre * x.im + kim * x.re;
    }
    stc(Zp, i * nv + j, c2{are, aim});
}

// X[row, j] = sum_i V[i, row] * Y[j, i]   (Y staged in LDS)
constexpr int COMB_MAX = 2048;   // nvec * m complex entries held in LDS (32 KB)
template <int NV>
__global__ void __launch_bounds__(256) k_basis_combine(int64_t n, int m, const double *__restrict__ V,
                                                       const double *__restrict__ Y, double *__restrict__ X) {
    __shared__ double2 sy[COMB_MAX];
    for (int t = threadIdx.x; t < NV * m; t += blockDim.x) sy[t] = reinterpret_cast<const double2 *>(Y)[t];
    __syncthreads();
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= n) return;
    double are[NV], aim[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) { are[j] = 0.0; aim[j] = 0.0; }
    for (int i = 0; i < m; i++) {
        const c2 v = ldc(V, (int64_t)i * n + row);
#pragma unroll
        for (int j = 0; j < NV; j++) {
            const double2 y = sy[j * m + i];
            are[j] += v.re * y.x - v.im * y.y;
            aim[j] += v.re * y.y + v.im * y.x;
        }
    }
#pragma unroll
    for (int j = 0; j < NV; j++) stc(X, row * NV + j, c2{are[j], aim[j]});
}

template <int NV>
void launch_residual(hipStream_t st, int grid, int64_t n, const int32_t *rowptr, const int32_t *cols, const double *E,
                     const double *B, const uint8_t *mask, const double *s, const double *X, const double *Q,
                     const int32_t *qidx, const double *c, const int32_t *prow, const double *Zp, double *R,
                     double *partial) {
    k_sweep_residual<NV><<<grid, RES_BLOCK, 0, st>>>(n / 2, rowptr, cols, E, B, mask, s, X, Q, qidx, c, prow, Zp, R,
                                                     partial);
}

template <int NV>
void launch_combine(hipStream_t st, int64_t n, int m, const double *V, const double *Y, double *X) {
    k_basis_combine<NV><<<emg::grid_for(n, 256), 256, 0, st>>>(n, m, V, Y, X);
}

bool valid_nvec(int nv) { return nv == 1 || nv == 2 || nv == 4 || nv == 8 || nv == 16; }

}  // namespace

extern "C" int emg_sweep_residual(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_E,
                                  const double *d_B, const uint8_t *d_mask, int nvec, const double *d_s,
                                  const double *d_X, const double *d_Q, const int32_t *d_qidx, const double *d_c,
                                  const int32_t *d_prow_of_row, const double *d_Zp, double *d_R, double *d_norms,
                                  double *d_work, void *stream) {
    EMG_CHECK(n > 0 && d_rowptr && d_cols && d_E && d_B && d_s && d_X && d_Q && d_qidx && d_c && d_norms && d_work,
              "emg_sweep_residual: null argument");
    EMG_CHECK(valid_nvec(nvec), "emg_sweep_residual: nvec must be 1, 2, 4, 8 or 16");
    EMG_CHECK((d_prow_of_row == nullptr) == (d_Zp == nullptr), "emg_sweep_residual: port rows need Zp");
    hipStream_t st = (hipStream_t)stream;
    EMG_CHECK(n % 2 == 0, "emg_sweep_residual: n must be 2 * entities (an entity pattern)");
    constexpr int ents_per_block = RES_BLOCK / ROW_LANES;
    const int grid = (int)std::min<int64_t>(RES_GRID_MAX, (n / 2 + ents_per_block - 1) / ents_per_block);
    switch (nvec) {
        case 1: launch_residual<1>(st, grid, n, d_rowptr, d_cols, d_E, d_B, d_mask, d_s, d_X, d_Q, d_qidx, d_c, d_prow_of_row, d_Zp, d_R, d_work); break;
        case 2: launch_residual<2>(st, grid, n, d_rowptr, d_cols, d_E, d_B, d_mask, d_s, d_X, d_Q, d_qidx, d_c, d_prow_of_row, d_Zp, d_R, d_work); break;
        case 4: launch_residual<4>(st, grid, n, d_rowptr, d_cols, d_E, d_B, d_mask, d_s, d_X, d_Q, d_qidx, d_c, d_prow_of_row, d_Zp, d_R, d_work); break;
        case 8: launch_residual<8>(st, grid, n, d_rowptr, d_cols, d_E, d_B, d_mask, d_s, d_X, d_Q, d_qidx, d_c, d_prow_of_row, d_Zp, d_R, d_work); break;
        default: launch_residual<16>(st, grid, n, d_rowptr, d_cols, d_E, d_B, d_mask, d_s, d_X, d_Q, d_qidx, d_c, d_prow_of_row, d_Zp, d_R, d_work); break;
    }
    k_sweep_norms<<<2 * nvec, 256, 0, st>>>(grid, nvec, d_work, d_norms);
    EMG_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------ no-allocation variant
// The residual's per-block partial sums live in a file-scope device array
// instead of a caller workspace, for callers that cannot allocate (code under
// stream capture, serving loops).  The host side never uses the array's name
// as an address: its device address is looked up once per device with
// hipGetSymbolAddress and cached.  Every block of the residual kernel writes
// its own partial entries before k_sweep_norms reads them, so the array needs
// no clearing between calls; calls that overlap on one device race on it.
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = (int64_t)RES_GRID_MAX * MAX_VEC * 2;   // 1 MiB
__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

int scratch_for_current_device(double **out) {
    int dev = 0;
    EMG_HIP(hipGetDevice(&dev));
    EMG_CHECK(dev >= 0 && dev < MAX_DEVICES, "emg_sweep_residual_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        EMG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        EMG_CHECK(p != nullptr, "emg_sweep_residual_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" int emg_sweep_residual_noalloc(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols,
                                          const double *d_E, const double *d_B, const uint8_t *d_mask, int nvec,
                                          const double *d_s, const double *d_X, const double *d_Q,
                                          const int32_t *d_qidx, const double *d_c, const int32_t *d_prow_of_row,
                                          const double *d_Zp, double *d_R, double *d_norms, void *stream) {
    EMG_CHECK(valid_nvec(nvec), "emg_sweep_residual_noalloc: nvec must be 1, 2, 4, 8 or 16");
    EMG_CHECK(n > 0 && n % 2 == 0, "emg_sweep_residual_noalloc: n must be 2 * entities (an entity pattern)");
    constexpr int ents_per_block = RES_BLOCK / ROW_LANES;
    const int64_t grid = std::min<int64_t>(RES_GRID_MAX, (n / 2 + ents_per_block - 1) / ents_per_block);
    EMG_CHECK(grid * nvec * 2 <= NOALLOC_SCRATCH_DOUBLES,
              "emg_sweep_residual_noalloc: partial sums exceed the 1 MiB scratch array");
    double *work = nullptr;
    const int rc = scratch_for_current_device(&work);
    if (rc != 0) return rc;
    return emg_sweep_residual(n, d_rowptr, d_cols, d_E, d_B, d_mask, nvec, d_s, d_X, d_Q, d_qidx, d_c,
                              d_prow_of_row, d_Zp, d_R, d_norms, work, stream);
}

extern "C" int emg_port_apply(int64_t nprow, int nvec, const int32_t *d_ptr, const int32_t *d_pcol,
                              const int32_t *d_pterm, const double *d_pval, const double *d_g, const double *d_X,
                              double *d_Zp, void *stream) {
    EMG_CHECK(nvec >= 1 && nvec <= MAX_VEC, "emg_port_apply: bad nvec");
    if (nprow == 0) return 0;
    EMG_CHECK(d_ptr && d_pcol && d_pterm && d_pval && d_g && d_X && d_Zp, "emg_port_apply: null argument");
    k_port_apply<<<emg::grid_for(nprow * nvec, 256), 256, 0, (hipStream_t)stream>>>(nprow, nvec, d_ptr, d_pcol, d_pterm,
                                                                                  d_pval, d_g, d_X, d_Zp);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_basis_combine(int64_t n, int m, const double *d_V, int nvec, const double *d_Y, double *d_X,
                                 void *stream) {
    EMG_CHECK(n > 0 && m > 0 && d_V && d_Y && d_X, "emg_basis_combine: null argument");
    EMG_CHECK(valid_nvec(nvec), "emg_basis_combine: nvec must be 1, 2, 4, 8 or 16");
    EMG_CHECK((int64_t)nvec * m <= COMB_MAX, "emg_basis_combine: nvec * m exceeds 2048");
    hipStream_t st = (hipStream_t)stream;
    switch (nvec) {
        case 1: launch_combine<1>(st, n, m, d_V, d_Y, d_X); break;
        case 2: launch_combine<2>(st, n, m, d_V, d_Y, d_X); break;
        case 4: launch_combine<4>(st, n, m, d_V, d_Y, d_X); break;
        case 8: launch_combine<8>(st, n, m, d_V, d_Y, d_X); break;
        default: launch_combine<16>(st, n, m, d_V, d_Y, d_X); break;
    }
    EMG_LAUNCH_CHECK();
    return 0;
}
