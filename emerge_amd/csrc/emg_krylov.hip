// Complex sparse Krylov solver for the frequency-domain systems on MI355X.
//
// Replaces the per-frequency direct solve of the reference
// (fem/solver.py:405-469: PEC elimination, PARDISO / SuperLU / GMRES+ILU) with
// right-preconditioned restarted GMRES on the GPU:
//   * PEC DOFs are eliminated implicitly: the operator is
//     y = mask.(K (mask.x)) + (1 - mask).x, so x_PEC = 0 exactly and no
//     reduced matrix is ever formed (solver.py:430-435 builds one per call);
//   * SpMV is HBM-bound: 16-lane sub-waves per CSR row, coalesced loads of
//     columns and values, complex128 or complex64 values;
//   * Gram-Schmidt is classical, as a tiled multi-dot and a fused multi-AXPY
//     + norm, so each Arnoldi step streams the basis twice instead of j
//     times; a second pass runs only when the first cancelled 99% of the
//     vector's squared norm (loss-of-orthogonality test on the norms);
//   * the preconditioner is the inverse of the 2x2 blocks coupling the two
//     modes of each edge / face;
//   * precision 32 runs GMRES in complex64 inside FP64 iterative refinement:
//     the residual b - K x is always formed in complex128.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

using emg::cplx;

struct cplxf {
    float re, im;
    __host__ __device__ cplxf() : re(0), im(0) {}
    __host__ __device__ cplxf(float r, float i = 0) : re(r), im(i) {}
};

template <typename T>
struct Cx;
template <>
struct Cx<double> {
    using C = cplx;
};
template <>
struct Cx<float> {
    using C = cplxf;
};

template <typename R>
struct cx {
    R re, im;
    __device__ __forceinline__ cx() : re(0), im(0) {}
    __device__ __forceinline__ cx(R r, R i) : re(r), im(i) {}
};
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) { return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return cx<R>(a.re + b.re, a.im + b.im); }

template <typename R>
__device__ __forceinline__ cx<R> ld(const R *p, int64_t i) {
    if constexpr (sizeof(R) == 8) {
        const double2 v = reinterpret_cast<const double2 *>(p)[i];
        return cx<R>(v.x, v.y);
    } else {
        const float2 v = reinterpret_cast<const float2 *>(p)[i];
        return cx<R>(v.x, v.y);
    }
}
template <typename R>
__device__ __forceinline__ void st(R *p, int64_t i, cx<R> v) {
    if constexpr (sizeof(R) == 8) {
        reinterpret_cast<double2 *>(p)[i] = make_double2(v.re, v.im);
    } else {
        reinterpret_cast<float2 *>(p)[i] = make_float2(v.re, v.im);
    }
}

static constexpr int SUB = 16;  // lanes per CSR row
// (SpMV rows stay in dispatch order: giving each XCD a contiguous share of the
// rows fetched a third fewer bytes but ran 30 % slower.)

// y = mask.(A (mask.x)) + (1-mask).x   (mask == nullptr: plain y = A x)
// If `b` is given, y = b - (that).  L lanes share one CSR row.  MC: apply the
// column mask per entry (false when the masked columns are already zero in
// vals, as in the solver's complex64 copy).
template <typename R, int L, bool MC>
__global__ void __launch_bounds__(256) k_spmv(int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                                              const R *__restrict__ vals, const uint8_t *__restrict__ mask,
                                              const R *__restrict__ x, const R *__restrict__ b, R *__restrict__ y) {
    const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / L;
    const int lane = threadIdx.x % L;
    if (row >= n) return;
    const bool keep = mask == nullptr || mask[row];
    cx<R> acc;
    if (keep) {
        const int32_t p0 = rowptr[row], p1 = rowptr[row + 1];
        for (int32_t p = p0 + lane; p < p1; p += L) {
            const int32_t c = cols[p];
            if (MC && mask != nullptr && !mask[c]) continue;
            acc = cadd(acc, cmul(ld(vals, p), ld(x, c)));
        }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        acc.re += __shfl_xor(acc.re, off, L);
        acc.im += __shfl_xor(acc.im, off, L);
    }
    if (lane == 0) {
        cx<R> v = keep ? acc : ld(x, row);
        if (b != nullptr) {
            const cx<R> bb = ld(b, row);
            v = cx<R>(bb.re - v.re, bb.im - v.im);
        }
        st(y, row, v);
    }
}

// Lanes per CSR row for the SpMV: EMG_SPMV_LANES (4, 8, 16 or 32) if set, else 16.
static int spmv_lanes() {
    static int lanes = 0;
    if (lanes == 0) {
        const char *e = getenv("EMG_SPMV_LANES");
        const int v = e ? atoi(e) : 16;
        lanes = (v == 4 || v == 8 || v == 16 || v == 32) ? v : 16;
    }
    return lanes;
}

template <typename R, bool MC>
static void launch_spmv_l(hipStream_t st, int64_t n, const int32_t *rowptr, const int32_t *cols, const R *vals,
                          const uint8_t *mask, const R *x, const R *b, R *y) {
    const int L = spmv_lanes();
    const unsigned grid = (unsigned)((n * L + 255) / 256);
    switch (L) {
        case 4: k_spmv<R, 4, MC><<<grid, 256, 0, st>>>(n, rowptr, cols, vals, mask, x, b, y); break;
        case 8: k_spmv<R, 8, MC><<<grid, 256, 0, st>>>(n, rowptr, cols, vals, mask, x, b, y); break;
        case 32: k_spmv<R, 32, MC><<<grid, 256, 0, st>>>(n, rowptr, cols, vals, mask, x, b, y); break;
        default: k_spmv<R, 16, MC><<<grid, 256, 0, st>>>(n, rowptr, cols, vals, mask, x, b, y); break;
    }
}

// mask_cols false: vals already holds zeros in the masked columns.
template <typename R>
static void launch_spmv(hipStream_t st, int64_t n, const int32_t *rowptr, const int32_t *cols, const R *vals,
                        const uint8_t *mask, const R *x, const R *b, R *y, bool mask_cols = true) {
    if (mask_cols)
        launch_spmv_l<R, true>(st, n, rowptr, cols, vals, mask, x, b, y);
    else
        launch_spmv_l<R, false>(st, n, rowptr, cols, vals, mask, x, b, y);
}

// SpMV of a matrix on an entity pattern (emg_pattern_csr): DOF = mode * ne +
// entity, and the rows (0, e) and (1, e) hold the same entities e'_j, row
// (m, e) storing its (0, e'_j) entries at r_m + j and its (1, e'_j) entries at
// r_m + deg + j, r_0 = rowptr[e], r_1 = r_0 + half (half = nnz / 2).  One
// group of L lanes computes both rows of entity e: each column index and each
// pair of x gathers serves four matrix entries.  Same masking as k_spmv.
template <typename R, int L, bool MC>
__global__ void __launch_bounds__(256) k_spmv_blk(int64_t ne, int64_t half, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ cols, const R *__restrict__ vals,
                                                  const uint8_t *__restrict__ mask, const R *__restrict__ x,
                                                  const R *__restrict__ b, R *__restrict__ y) {
    const int64_t e = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / L;
    const int lane = threadIdx.x % L;
    if (e >= ne) return;
    const bool k0 = mask == nullptr || mask[e], k1 = mask == nullptr || mask[ne + e];
    cx<R> a0, a1;
    if (k0 || k1) {
        const int32_t r0 = rowptr[e];
        // a row beyond the first half is not an entity pattern: read nothing past nnz
        const int32_t deg = (rowptr[e + 1] <= half && rowptr[e + 1] >= r0) ? (rowptr[e + 1] - r0) >> 1 : 0;
        for (int32_t j = lane; j < deg; j += L) {
            const int32_t c = cols[r0 + j];
            if (c < 0 || c >= ne) continue;   // not an entity pattern: no out-of-range reads
            cx<R> x0 = ld(x, c), x1 = ld(x, ne + c);
            if (MC && mask != nullptr) {
                if (!mask[c]) x0 = cx<R>();
                if (!mask[ne + c]) x1 = cx<R>();
            }
            const int64_t p = r0 + j;
            a0 = cadd(a0, cadd(cmul(ld(vals, p), x0), cmul(ld(vals, p + deg), x1)));
            a1 = cadd(a1, cadd(cmul(ld(vals, half + p), x0), cmul(ld(vals, half + p + deg), x1)));
        }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        a0.re += __shfl_xor(a0.re, off, L);
        a0.im += __shfl_xor(a0.im, off, L);
        a1.re += __shfl_xor(a1.re, off, L);
        a1.im += __shfl_xor(a1.im, off, L);
    }
    if (lane == 0) {
        cx<R> v0 = k0 ? a0 : ld(x, e), v1 = k1 ? a1 : ld(x, ne + e);
        if (b != nullptr) {
            const cx<R> b0 = ld(b, e), b1 = ld(b, ne + e);
            v0 = cx<R>(b0.re - v0.re, b0.im - v0.im);
            v1 = cx<R>(b1.re - v1.re, b1.im - v1.im);
        }
        st(y, e, v0);
        st(y, ne + e, v1);
    }
}

// Lanes per entity row of the blocked SpMV when nothing was tuned (the public
// emg_spmv_blocked, or EMG_SPMV_BLK_LANES set): EMG_SPMV_BLK_LANES (4, 8, 16 or
// 32) if set, else 16.  Solvers tune their own (tune_spmv).
static int spmv_blk_lanes() {
    static int lanes = 0;
    if (lanes == 0) {
        const char *e = getenv("EMG_SPMV_BLK_LANES");
        const int v = e ? atoi(e) : 16;
        lanes = (v == 4 || v == 8 || v == 16 || v == 32) ? v : 16;
    }
    return lanes;
}

template <typename R>
static void launch_spmv_blk(hipStream_t st, int64_t n, int64_t nnz, const int32_t *rowptr, const int32_t *cols,
                            const R *vals, const uint8_t *mask, const R *x, const R *b, R *y, int L,
                            bool mask_cols = true) {
    const int64_t ne = n / 2, half = nnz / 2;
    if (L != 4 && L != 8 && L != 32) L = 16;
    const unsigned grid = (unsigned)((ne * L + 255) / 256);
#define EMG_BLK(LL, MCV) k_spmv_blk<R, LL, MCV><<<grid, 256, 0, st>>>(ne, half, rowptr, cols, vals, mask, x, b, y)
    if (mask_cols) {
        if (L == 4) EMG_BLK(4, true); else if (L == 8) EMG_BLK(8, true); else if (L == 32) EMG_BLK(32, true);
        else EMG_BLK(16, true);
    } else {
        if (L == 4) EMG_BLK(4, false); else if (L == 8) EMG_BLK(8, false); else if (L == 32) EMG_BLK(32, false);
        else EMG_BLK(16, false);
    }
#undef EMG_BLK
}

// ---------------------------------------------------------------- BLAS-1
static constexpr int RB = 256;   // reduction block
static constexpr int DT = 8;     // dot products per tile

// partial[tile][blk][k] = sum over the block's elements of conj(V_{t*DT+k}) . w
template <typename R>
__global__ void __launch_bounds__(RB) k_mdot(int64_t n, int nv, const R *__restrict__ V, int64_t ldv,
                                             const R *__restrict__ w, double *__restrict__ partial,
                                             const double *__restrict__ run) {
    if (run != nullptr && run[0] == 0.0) return;   // device-side skip (second Gram-Schmidt pass not needed)
    const int tile = blockIdx.y;
    const int i0 = tile * DT;
    double ar[DT], ai[DT];
#pragma unroll
    for (int k = 0; k < DT; k++) { ar[k] = 0; ai[k] = 0; }
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const cx<R> wv = ld(w, e);
#pragma unroll
        for (int k = 0; k < DT; k++) {
            if (i0 + k < nv) {
                const cx<R> v = ld(V + (int64_t)(i0 + k) * ldv * 2, e);
                ar[k] += (double)v.re * wv.re + (double)v.im * wv.im;
                ai[k] += (double)v.re * wv.im - (double)v.im * wv.re;
            }
        }
    }
    __shared__ double sr[DT][RB / 64], si[DT][RB / 64];
#pragma unroll
    for (int k = 0; k < DT; k++) {
        double r = ar[k], i = ai[k];
        for (int off = 32; off > 0; off >>= 1) {
            r += __shfl_xor(r, off, 64);
            i += __shfl_xor(i, off, 64);
        }
        if ((threadIdx.x & 63) == 0) { sr[k][threadIdx.x >> 6] = r; si[k][threadIdx.x >> 6] = i; }
    }
    __syncthreads();
    if (threadIdx.x < DT) {
        double r = 0, i = 0;
        for (int q = 0; q < RB / 64; q++) { r += sr[threadIdx.x][q]; i += si[threadIdx.x][q]; }
        const int64_t o = ((int64_t)tile * gridDim.x + blockIdx.x) * DT + threadIdx.x;
        partial[2 * o] = r;
        partial[2 * o + 1] = i;
    }
}

// h[i] (+)= sum over blocks of the partials (complex, i < nv); one block per value
__global__ void k_mdot_finish(int nv, int nblk, const double *__restrict__ partial, double *__restrict__ h, int accumulate,
                              const double *__restrict__ run) {
    const int i = blockIdx.x;
    if (i >= nv) return;
    if (run != nullptr && run[0] == 0.0) {
        if (threadIdx.x == 0 && !accumulate) { h[2 * i] = 0.0; h[2 * i + 1] = 0.0; }
        return;
    }
    const int tile = i / DT, k = i % DT;
    double r = 0, im = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const int64_t o = ((int64_t)tile * nblk + b) * DT + k;
        r += partial[2 * o];
        im += partial[2 * o + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    __shared__ double s[2][4];
    if ((threadIdx.x & 63) == 0) { s[0][threadIdx.x >> 6] = r; s[1][threadIdx.x >> 6] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        r = 0; im = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) { r += s[0][q]; im += s[1][q]; }
        if (accumulate) { h[2 * i] += r; h[2 * i + 1] += im; }
        else { h[2 * i] = r; h[2 * i + 1] = im; }
    }
}

// w -= sum_i h_i V_i ; partial norm^2 of the result per block
template <typename R>
__global__ void __launch_bounds__(RB) k_maxpy_norm(int64_t n, int nv, const R *__restrict__ V, int64_t ldv,
                                                   const double *__restrict__ h, R *__restrict__ w,
                                                   double *__restrict__ partial, const double *__restrict__ run) {
    if (run != nullptr && run[0] == 0.0) return;
    __shared__ double hs[2 * 512];
    for (int i = threadIdx.x; i < 2 * nv; i += RB) hs[i] = h[i];
    __syncthreads();
    double nrm = 0;
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        cx<R> wv = ld(w, e);
        double wr = wv.re, wi = wv.im;
        for (int i = 0; i < nv; i++) {
            const cx<R> v = ld(V + (int64_t)i * ldv * 2, e);
            wr -= hs[2 * i] * v.re - hs[2 * i + 1] * v.im;
            wi -= hs[2 * i] * v.im + hs[2 * i + 1] * v.re;
        }
        st(w, e, cx<R>((R)wr, (R)wi));
        nrm += (double)(R)wr * (double)(R)wr + (double)(R)wi * (double)(R)wi;
    }
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off, 64);
    __shared__ double s[RB / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < RB / 64; q++) t += s[q];
        partial[blockIdx.x] = t;
    }
}

// x += sum_i y_i V_i   (y on device, complex)
template <typename R>
__global__ void __launch_bounds__(RB) k_mcomb(int64_t n, int nv, const R *__restrict__ V, int64_t ldv,
                                              const double *__restrict__ y, R *__restrict__ x, int overwrite) {
    __shared__ double ys[2 * 512];
    for (int i = threadIdx.x; i < 2 * nv; i += RB) ys[i] = y[i];
    __syncthreads();
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        double xr = 0, xi = 0;
        if (!overwrite) {
            const cx<R> xv = ld(x, e);
            xr = xv.re; xi = xv.im;
        }
        for (int i = 0; i < nv; i++) {
            const cx<R> v = ld(V + (int64_t)i * ldv * 2, e);
            xr += ys[2 * i] * v.re - ys[2 * i + 1] * v.im;
            xi += ys[2 * i] * v.im + ys[2 * i + 1] * v.re;
        }
        st(x, e, cx<R>((R)xr, (R)xi));
    }
}

template <typename R>
__global__ void __launch_bounds__(RB) k_norm2(int64_t n, const R *__restrict__ x, double *__restrict__ partial) {
    double s = 0;
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const cx<R> v = ld(x, e);
        s += (double)v.re * v.re + (double)v.im * v.im;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double sh[RB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < RB / 64; q++) t += sh[q];
        partial[blockIdx.x] = t;
    }
}

// y = alpha * x (alpha real, on device as 1/sqrt(norm2 sum) when from_norm)
template <typename R>
__global__ void __launch_bounds__(RB) k_scale_by_norm(int64_t n, const double *__restrict__ partial, int nblk,
                                                      const R *__restrict__ x, R *__restrict__ y, double *__restrict__ out_norm) {
    __shared__ double inv;
    if (threadIdx.x < 64) {
        double s = 0;
        for (int b = threadIdx.x; b < nblk; b += 64) s += partial[b];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (threadIdx.x == 0) {
            const double nv = sqrt(s);
            inv = nv > 0 ? 1.0 / nv : 0.0;
            if (out_norm && blockIdx.x == 0) *out_norm = nv;
        }
    }
    __syncthreads();
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const cx<R> v = ld(x, e);
        st(y, e, cx<R>((R)(v.re * inv), (R)(v.im * inv)));
    }
}

// 2x2 mode-block Jacobi: z = blk^{-1} r on DOF pairs (e, e + n_ent).
// inv holds per entity the 4 complex entries of the inverse block (row-major).
template <typename R>
__global__ void __launch_bounds__(RB) k_bj_apply(int64_t n_ent, const R *__restrict__ inv, const R *__restrict__ r,
                                                 R *__restrict__ z) {
    const int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (e >= n_ent) return;
    const cx<R> a = ld(inv, 4 * e), b = ld(inv, 4 * e + 1), c = ld(inv, 4 * e + 2), d = ld(inv, 4 * e + 3);
    const cx<R> r0 = ld(r, e), r1 = ld(r, e + n_ent);
    st(z, e, cadd(cmul(a, r0), cmul(b, r1)));
    st(z, e + n_ent, cadd(cmul(c, r0), cmul(d, r1)));
}

__device__ __forceinline__ int32_t find_col(const int32_t *cols, int32_t p0, int32_t p1, int32_t c) {
    int32_t lo = p0, hi = p1 - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (cols[mid] < c) lo = mid + 1; else hi = mid;
    }
    return (lo < p1 && cols[lo] == c) ? lo : -1;
}

// Build the inverse 2x2 blocks from K (complex128); PEC DOFs get identity.
// ---------------------------------------------------------------- gradient correction
static constexpr int G_ROW_MAX = 8;   // entries per row of G (edge rows 3, face rows 6)

// dS[p] += g_ip * sum_j K_ij g_jp over the unmasked entries of row i, for the
// P2 nodes p of row i of G: diag(G^T K G) with one atomic per entry of G.
__global__ void __launch_bounds__(RB) k_gdiag(int64_t n, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ cols, const double *__restrict__ K,
                                              const uint8_t *__restrict__ mask, const int32_t *__restrict__ grp,
                                              const int32_t *__restrict__ gcol, const double *__restrict__ gval,
                                              double *__restrict__ dS) {
    const int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (i >= n || (mask && !mask[i])) return;
    const int32_t q0 = grp[i], ni = min(grp[i + 1] - q0, G_ROW_MAX);
    if (ni == 0) return;
    int32_t pk[G_ROW_MAX];
    double ar[G_ROW_MAX], ai[G_ROW_MAX];
    for (int k = 0; k < G_ROW_MAX; k++) { pk[k] = k < ni ? gcol[q0 + k] : -1; ar[k] = 0.0; ai[k] = 0.0; }
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
        const int32_t j = cols[e];
        if (mask && !mask[j]) continue;
        const int32_t r0 = grp[j], r1 = grp[j + 1];
        if (r0 == r1) continue;
        const double kr = K[2 * e], ki = K[2 * e + 1];
        for (int32_t q = r0; q < r1; q++) {
            const int32_t pj = gcol[q];
            const double gj = gval[q];
#pragma unroll
            for (int k = 0; k < G_ROW_MAX; k++)
                if (pk[k] == pj) { ar[k] += kr * gj; ai[k] += ki * gj; }
        }
    }
    for (int k = 0; k < ni; k++) {
        const double gi = gval[q0 + k];
        atomicAdd(&dS[2 * pk[k]], gi * ar[k]);
        atomicAdd(&dS[2 * pk[k] + 1], gi * ai[k]);
    }
}

__global__ void __launch_bounds__(RB) k_dsinv(int64_t np2, const double *__restrict__ dS, double *__restrict__ inv64,
                                              float *__restrict__ inv32) {
    const int64_t p = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (p >= np2) return;
    const double a = dS[2 * p], b = dS[2 * p + 1], d = a * a + b * b;
    const double r = d > 0.0 ? a / d : 0.0, im = d > 0.0 ? -b / d : 0.0;
    inv64[2 * p] = r;
    inv64[2 * p + 1] = im;
    if (inv32) { inv32[2 * p] = (float)r; inv32[2 * p + 1] = (float)im; }
}

// t[p] = dsinv[p] * sum over row p of G^T of g * r, GL lanes per row (vertex rows of G^T
// hold tens of entries, edge rows about ten; the solver's tuner picks GL)
template <typename R, int GL>
__global__ void __launch_bounds__(256) k_gt_apply(int64_t np2, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const R *__restrict__ gv, const R *__restrict__ dsinv,
                                                  const R *__restrict__ r, R *__restrict__ t) {
    const int64_t p = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / GL;
    const int lane = threadIdx.x % GL;
    if (p >= np2) return;
    double ar = 0.0, ai = 0.0;
    for (int32_t e = rp[p] + lane; e < rp[p + 1]; e += GL) {
        const cx<R> v = ld(r, ci[e]);
        ar += (double)gv[e] * v.re;
        ai += (double)gv[e] * v.im;
    }
#pragma unroll
    for (int off = GL / 2; off > 0; off >>= 1) {
        ar += __shfl_xor(ar, off, GL);
        ai += __shfl_xor(ai, off, GL);
    }
    if (lane == 0) {
        const cx<R> d = ld(dsinv, p);
        st(t, p, cx<R>((R)(d.re * ar - d.im * ai), (R)(d.re * ai + d.im * ar)));
    }
}

template <typename R>
static void launch_gt(hipStream_t st, int L, int64_t np2, const int32_t *rp, const int32_t *ci, const R *gv,
                      const R *dsinv, const R *r, R *t) {
    if (L != 4 && L != 16) L = 8;
    const unsigned grid = (unsigned)((np2 * L + 255) / 256);
    if (L == 4) k_gt_apply<R, 4><<<grid, 256, 0, st>>>(np2, rp, ci, gv, dsinv, r, t);
    else if (L == 16) k_gt_apply<R, 16><<<grid, 256, 0, st>>>(np2, rp, ci, gv, dsinv, r, t);
    else k_gt_apply<R, 8><<<grid, 256, 0, st>>>(np2, rp, ci, gv, dsinv, r, t);
}

// z = BJ(r) + G t, both modes of entity e per thread
template <typename R>
__global__ void __launch_bounds__(RB) k_bj_grad_apply(int64_t n_ent, const R *__restrict__ inv, const R *__restrict__ r,
                                                      const int32_t *__restrict__ grp, const int32_t *__restrict__ gcol,
                                                      const R *__restrict__ gv, const R *__restrict__ t,
                                                      R *__restrict__ z) {
    const int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (e >= n_ent) return;
    const cx<R> a = ld(inv, 4 * e), b = ld(inv, 4 * e + 1), c = ld(inv, 4 * e + 2), d = ld(inv, 4 * e + 3);
    const cx<R> r0 = ld(r, e), r1 = ld(r, e + n_ent);
    cx<R> z0 = cadd(cmul(a, r0), cmul(b, r1));
    cx<R> z1 = cadd(cmul(c, r0), cmul(d, r1));
    const int32_t q0 = grp[e], q1 = grp[e + 1], q2 = grp[e + n_ent], q3 = grp[e + n_ent + 1];
    if (q1 - q0 == q3 - q2) {
        // both modes' rows list the same P2 nodes in the same order (the entity's nodes less the
        // dropped ones): one column index and one gather of t serve both rows
        for (int32_t k = 0; k < q1 - q0; k++) {
            const cx<R> tv = ld(t, gcol[q0 + k]);
            const R g0 = gv[q0 + k], g1 = gv[q2 + k];
            z0 = cx<R>(z0.re + g0 * tv.re, z0.im + g0 * tv.im);
            z1 = cx<R>(z1.re + g1 * tv.re, z1.im + g1 * tv.im);
        }
    } else {
        for (int32_t q = q0; q < q1; q++) {
            const cx<R> tv = ld(t, gcol[q]);
            z0 = cx<R>(z0.re + gv[q] * tv.re, z0.im + gv[q] * tv.im);
        }
        for (int32_t q = q2; q < q3; q++) {
            const cx<R> tv = ld(t, gcol[q]);
            z1 = cx<R>(z1.re + gv[q] * tv.re, z1.im + gv[q] * tv.im);
        }
    }
    st(z, e, z0);
    st(z, e + n_ent, z1);
}

__global__ void __launch_bounds__(RB) k_bj_setup(int64_t n_ent, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ cols, const double *__restrict__ vals,
                                                 const uint8_t *__restrict__ mask, double *__restrict__ inv64,
                                                 float *__restrict__ inv32) {
    const int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (e >= n_ent) return;
    const int64_t r0 = e, r1 = e + n_ent;
    const bool m0 = mask == nullptr || mask[r0], m1 = mask == nullptr || mask[r1];
    auto get = [&](int64_t r, int64_t c) -> cx<double> {
        const int32_t p = find_col(cols, rowptr[r], rowptr[r + 1], (int32_t)c);
        return p < 0 ? cx<double>(0, 0) : ld(vals, p);
    };
    cx<double> a = m0 ? get(r0, r0) : cx<double>(1, 0);
    cx<double> d = m1 ? get(r1, r1) : cx<double>(1, 0);
    cx<double> b = (m0 && m1) ? get(r0, r1) : cx<double>(0, 0);
    cx<double> c = (m0 && m1) ? get(r1, r0) : cx<double>(0, 0);
    const cx<double> ad = cmul(a, d), bc = cmul(b, c);
    cx<double> det(ad.re - bc.re, ad.im - bc.im);
    const double dn = det.re * det.re + det.im * det.im;
    cx<double> idet(det.re / dn, -det.im / dn);
    if (!(dn > 0) || !std::isfinite(dn)) {  // singular block: fall back to the diagonal
        idet = cx<double>(1, 0);
        b = c = cx<double>(0, 0);
        const double na = a.re * a.re + a.im * a.im, nd = d.re * d.re + d.im * d.im;
        a = na > 0 ? cx<double>(a.re / na, -a.im / na) : cx<double>(1, 0);
        d = nd > 0 ? cx<double>(d.re / nd, -d.im / nd) : cx<double>(1, 0);
        const cx<double> tmp = a;
        a = d;
        d = tmp;  // undone by the swap below
    }
    const cx<double> ia = cmul(d, idet), ib = cmul(cx<double>(-b.re, -b.im), idet);
    const cx<double> ic = cmul(cx<double>(-c.re, -c.im), idet), id = cmul(a, idet);
    st(inv64, 4 * e, ia); st(inv64, 4 * e + 1, ib); st(inv64, 4 * e + 2, ic); st(inv64, 4 * e + 3, id);
    if (inv32) {
        st(inv32, 4 * e, cx<float>((float)ia.re, (float)ia.im));
        st(inv32, 4 * e + 1, cx<float>((float)ib.re, (float)ib.im));
        st(inv32, 4 * e + 2, cx<float>((float)ic.re, (float)ic.im));
        st(inv32, 4 * e + 3, cx<float>((float)id.re, (float)id.im));
    }
}

__global__ void __launch_bounds__(RB) k_to_c64(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * n; e += (int64_t)gridDim.x * RB) b[e] = (float)a[e];
}
// complex64 copy of the matrix values with the masked (PEC) columns set to zero
__global__ void __launch_bounds__(RB) k_to_c64_masked(int64_t nnz, const int32_t *__restrict__ cols,
                                                      const uint8_t *__restrict__ mask, const double *__restrict__ a,
                                                      float *__restrict__ b) {
    for (int64_t p = blockIdx.x * (int64_t)RB + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * RB) {
        const bool keep = mask == nullptr || mask[cols[p]];
        b[2 * p] = keep ? (float)a[2 * p] : 0.0f;
        b[2 * p + 1] = keep ? (float)a[2 * p + 1] : 0.0f;
    }
}
__global__ void __launch_bounds__(RB) k_add_c64_to_c128(int64_t n, const float *__restrict__ d, double *__restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * n; e += (int64_t)gridDim.x * RB) x[e] += (double)d[e];
}
// y = x (scaled to c64) * s  (s real)
__global__ void __launch_bounds__(RB) k_c128_to_c64_scaled(int64_t n, const double *__restrict__ a, double s, float *__restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * n; e += (int64_t)gridDim.x * RB) b[e] = (float)(a[e] * s);
}

// x (complex128) += scale * d (element type R)
template <typename R>
__global__ void __launch_bounds__(RB) k_add_scaled(int64_t n, const R *__restrict__ d, double scale, double *__restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * n; e += (int64_t)gridDim.x * RB) x[e] += scale * (double)d[e];
}
// u += z (same element type)
template <typename R>
__global__ void __launch_bounds__(RB) k_add(int64_t n, const R *__restrict__ z, R *__restrict__ u) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * n; e += (int64_t)gridDim.x * RB) u[e] += z[e];
}

// K = E - s B (complex128 values on the shared pattern)
__global__ void __launch_bounds__(RB) k_affine(int64_t nnz, const double *__restrict__ E, const double *__restrict__ B,
                                               double s, double *__restrict__ K) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * nnz; e += (int64_t)gridDim.x * RB) K[e] = E[e] - s * B[e];
}
// K[pos[i]] += g * v[i]   (complex; positions may repeat -> atomics)
__global__ void __launch_bounds__(RB) k_scatter_add(int64_t m, const int32_t *__restrict__ pos, const double *__restrict__ v,
                                                    double gre, double gim, double *__restrict__ K) {
    const int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (i >= m) return;
    const int32_t p = pos[i];
    if (p < 0) return;
    const double vr = v[2 * i], vi = v[2 * i + 1];
    atomicAdd(&K[2 * (int64_t)p], gre * vr - gim * vi);
    atomicAdd(&K[2 * (int64_t)p + 1], gre * vi + gim * vr);
}
// pos[i] = CSR position of (rows[i], cols[i]) or -1
__global__ void __launch_bounds__(RB) k_csr_find(int64_t m, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ cols, const int32_t *__restrict__ qr,
                                                 const int32_t *__restrict__ qc, int32_t *__restrict__ pos) {
    const int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (i >= m) return;
    const int32_t r = qr[i];
    pos[i] = find_col(cols, rowptr[r], rowptr[r + 1], qc[i]);
}

// ---------------------------------------------------------------- solver object
struct emg_solver {
    int64_t n = 0, nnz = 0, n_ent = 0;
    int restart = 0, precision = 64;
    const int32_t *rowptr = nullptr, *cols = nullptr;
    const double *vals = nullptr;   // complex128 K (caller-owned)
    const uint8_t *mask = nullptr;  // caller-owned, may be null
    float *vals32 = nullptr;        // complex64 copy for precision 32
    double *inv64 = nullptr;
    float *inv32 = nullptr;
    double *V = nullptr;            // (restart+1) basis vectors, element type per precision
    double *w = nullptr, *z = nullptr, *r = nullptr, *d = nullptr;
    double *partial = nullptr, *h = nullptr, *nrm = nullptr;
    int nblk = 0;
    // Arnoldi steps run from captured HIP graphs (chunks of CHUNK_IT steps)
    // with the Hessenberg column, Givens rotations and residual estimate
    // updated on the device; the host reads the estimates once per chunk.
    double *arn = nullptr;        // device Arnoldi state (see arn_off)
    double *pn = nullptr;         // per-block partial norms of the new basis vector
    double *flag = nullptr;       // [0]: 1 = run the second Gram-Schmidt pass
    double *res_host = nullptr;   // pinned copy of the residual estimates
    hipStream_t cap = nullptr;    // capture stream (graphs are launched on the caller's stream)
    // optional gradient correction of the smoother: G (n x np2), G^T (np2 x n)
    int64_t np2 = 0;
    const int32_t *g_rowptr = nullptr, *g_cols = nullptr, *gt_rowptr = nullptr, *gt_cols = nullptr;
    const double *g_vals = nullptr, *gt_vals = nullptr;   // caller-owned
    float *g_vals32 = nullptr, *gt_vals32 = nullptr;      // owned copies for complex64 steps
    double *dS = nullptr, *dsinv64 = nullptr, *tg = nullptr;
    float *dsinv32 = nullptr;
    // optional learned weights of the smoother (caller-owned complex128):
    // w_e per entity scales its 2x2 block inverse, w_p per P2 node its diagonal
    const double *w_e = nullptr, *w_p = nullptr;
    // optional polynomial acceleration of the smoother (emg_solver_set_poly):
    // M p(K M) with 1 - z p(z) = prod_i (1 - z c_i), c_i = 1/theta_i on the device
    int poly_deg = 1;
    double *poly_c = nullptr;                               // POLY_MAX complex coefficients
    double *pp = nullptr, *pt = nullptr, *py = nullptr, *pu = nullptr;   // work vectors (n complex each)
    // blocked-SpMV lanes per entity, [0] complex128, [1] complex64: 0 until tuned on
    // the first matrix (tune_spmv), with the measured milliseconds per variant
    int lanes[2] = {0, 0};
    float tune_ms[2][4] = {};
    // lanes per row of the smoother's G^T product (k_gt_apply): 0 until tuned with the
    // SpMV (gradient set), then 4, 8 or 16, with the measured milliseconds per launch
    int gt_lanes = 0;
    float gt_ms[3] = {};
    struct graphs {
        std::vector<hipGraphExec_t> exec;
        const void *vals = nullptr, *inv = nullptr, *mask = nullptr;
        int64_t np2 = 0;
        int poly = 1;
    } gs[2];                      // [0] complex128 steps, [1] complex64 steps
};
static constexpr int POLY_MAX = 16;

static void free_solver(emg_solver *s) {
    if (!s) return;
    for (auto &g : s->gs)
        for (hipGraphExec_t e : g.exec) (void)hipGraphExecDestroy(e);
    void *ps[] = {s->vals32, s->inv64, s->inv32, s->V, s->w, s->z, s->r, s->d, s->partial, s->h, s->nrm,
                  s->arn, s->pn, s->flag, s->g_vals32, s->gt_vals32, s->dS, s->dsinv64, s->tg, s->dsinv32,
                  s->poly_c, s->pp, s->pt, s->py, s->pu};
    for (void *p : ps) (void)hipFree(p);
    if (s->res_host) (void)hipHostFree(s->res_host);
    if (s->cap) (void)hipStreamDestroy(s->cap);
    delete s;
}

// The smoother: 2x2 block Jacobi, plus the gradient correction when set.
template <typename R>
static void apply_prec(emg_solver *s, const R *inv, const R *x, R *y, hipStream_t st) {
    if (s->np2 > 0) {
        const R *gv = sizeof(R) == 8 ? reinterpret_cast<const R *>(s->g_vals) : reinterpret_cast<const R *>(s->g_vals32);
        const R *gtv = sizeof(R) == 8 ? reinterpret_cast<const R *>(s->gt_vals) : reinterpret_cast<const R *>(s->gt_vals32);
        const R *dsi = sizeof(R) == 8 ? reinterpret_cast<const R *>(s->dsinv64) : reinterpret_cast<const R *>(s->dsinv32);
        R *t = reinterpret_cast<R *>(s->tg);
        launch_gt<R>(st, s->gt_lanes, s->np2, s->gt_rowptr, s->gt_cols, gtv, dsi, x, t);
        k_bj_grad_apply<R><<<emg::grid_for(s->n_ent, RB), RB, 0, st>>>(s->n_ent, inv, x, s->g_rowptr, s->g_cols, gv, t, y);
    } else {
        k_bj_apply<R><<<emg::grid_for(s->n_ent, RB), RB, 0, st>>>(s->n_ent, inv, x, y);
    }
}


// y = c[0] x   (c: complex128 coefficients on the device)
template <typename R>
__global__ void __launch_bounds__(RB) k_poly_first(int64_t n, const double *__restrict__ c, const R *__restrict__ x,
                                                   R *__restrict__ y) {
    const double cr = c[0], ci = c[1];
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const cx<R> v = ld(x, e);
        st(y, e, cx<R>((R)(cr * v.re - ci * v.im), (R)(cr * v.im + ci * v.re)));
    }
}
// p = q - c[i] t ;  y += c[i+1] p   (q may be p)
template <typename R>
__global__ void __launch_bounds__(RB) k_poly_step(int64_t n, const double *__restrict__ c, int i, const R *q,
                                                  const R *__restrict__ t, R *p, R *__restrict__ y) {
    const double ar = c[2 * i], ai = c[2 * i + 1], br = c[2 * i + 2], bi = c[2 * i + 3];
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const cx<R> qv = ld(q, e), tv = ld(t, e), yv = ld(y, e);
        const double pr = qv.re - (ar * tv.re - ai * tv.im), pi = qv.im - (ar * tv.im + ai * tv.re);
        st(p, e, cx<R>((R)pr, (R)pi));
        st(y, e, cx<R>((R)(yv.re + br * pr - bi * pi), (R)(yv.im + br * pi + bi * pr)));
    }
}

// The preconditioner: the smoother M, or with a polynomial of degree d > 1,
// M p(K M) x evaluated as  y = c_0 x, q_0 = x;  q_{i+1} = q_i - c_i K M q_i,
// y += c_{i+1} q_{i+1}  (i < d-1);  out = M y.  d - 1 extra SpMVs and d - 1
// extra smoother applications per call.
template <typename R>
static void apply_prec_poly(emg_solver *s, const R *vals, const R *inv, const R *x, R *out, hipStream_t st) {
    if (s->poly_deg <= 1) {
        apply_prec<R>(s, inv, x, out, st);
        return;
    }
    const int64_t n = s->n;
    const unsigned g = (unsigned)std::min<int64_t>((n + RB - 1) / RB, 2048);
    R *pp = reinterpret_cast<R *>(s->pp), *pt = reinterpret_cast<R *>(s->pt);
    R *py = reinterpret_cast<R *>(s->py), *pu = reinterpret_cast<R *>(s->pu);
    k_poly_first<R><<<g, RB, 0, st>>>(n, s->poly_c, x, py);
    const R *q = x;
    for (int i = 0; i < s->poly_deg - 1; i++) {
        apply_prec<R>(s, inv, q, pu, st);
        launch_spmv_blk<R>(st, n, s->nnz, s->rowptr, s->cols, vals, s->mask, pu, nullptr, pt, s->lanes[sizeof(R) == 4],
                           sizeof(R) == 8);
        k_poly_step<R><<<g, RB, 0, st>>>(n, s->poly_c, i, q, pt, pp, py);
        q = pp;
    }
    apply_prec<R>(s, inv, py, out, st);
}

// Blocks per tile of the multi-dot: about 2048 blocks in total, so each block
// streams many rows and the per-block reduction is amortised.
static int mdot_blocks(int64_t n, int nv) {
    const int ntile = (nv + DT - 1) / DT;
    const int64_t cap = std::max<int64_t>(64, 2048 / ntile);
    return (int)std::max<int64_t>(1, std::min<int64_t>({(n + RB - 1) / RB, 1024, cap}));
}

// Offsets (in doubles) into the device Arnoldi state for restart length m.
struct arn_off {
    int64_t H, cs, sn, g, res, inv, total;
    explicit arn_off(int m) {
        H = 0;                                  // column j at 2 * j * (m + 1), rows 0..m
        cs = 2 * (int64_t)m * (m + 1);
        sn = cs + 2 * m;
        g = sn + 2 * m;                         // m + 1 complex
        res = g + 2 * (m + 1);                  // |g[j+1]| after step j
        inv = res + m;                          // 1 / ||new vector|| of the last step
        total = inv + 8;
    }
};
static constexpr int CHUNK_IT = 10;   // Arnoldi steps per captured graph

// Sets *bad when the CSR is not on an entity pattern (the layout k_spmv_blk and
// the smoother rely on): rows (0, e) and (1, e) list the same column entities
// c_j < ne, the (1, c_j) entries of a row follow its (0, c_j) entries, and row
// (1, e) starts half = nnz / 2 entries after row (0, e).
__global__ void __launch_bounds__(256) k_check_entity_pattern(int64_t ne, int64_t nnz,
                                                              const int32_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ cols, int *bad) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e > ne) return;
    const int64_t half = nnz / 2;
    if (e == ne) {
        if (rowptr[0] != 0 || rowptr[ne] != half || rowptr[2 * ne] != nnz || (nnz & 1)) atomicOr(bad, 1);
        return;
    }
    const int64_t r0 = rowptr[e], r1 = rowptr[e + 1];
    if (r1 < r0 || r1 > half || ((r1 - r0) & 1) || rowptr[ne + e] != r0 + half || rowptr[ne + e + 1] != r1 + half) {
        atomicOr(bad, 1);
        return;
    }
    const int64_t deg = (r1 - r0) / 2;
    for (int64_t j = 0; j < deg; j++) {
        const int32_t c = cols[r0 + j];
        if (c < 0 || c >= ne || cols[r0 + deg + j] != c + ne || cols[half + r0 + j] != c ||
            cols[half + r0 + deg + j] != c + ne) {
            atomicOr(bad, 1);
            return;
        }
    }
}

extern "C" int emg_solver_create(int64_t n, int64_t nnz, const int32_t *d_rowptr, const int32_t *d_cols, int restart,
                                 int precision, emg_solver **out) {
    EMG_CHECK(out && d_rowptr && d_cols, "emg_solver_create: null argument");
    EMG_CHECK(n > 0 && (n % 2) == 0, "emg_solver_create: n must be 2 * n_entities");
    EMG_CHECK(restart >= 2 && restart <= 500, "emg_solver_create: restart must be in [2, 500]");
    EMG_CHECK(precision == 64 || precision == 32, "emg_solver_create: precision must be 64 or 32");
    auto s = new emg_solver();
    s->n = n; s->nnz = nnz; s->n_ent = n / 2; s->restart = restart; s->precision = precision;
    s->rowptr = d_rowptr; s->cols = d_cols;
    const size_t es = precision == 64 ? 16 : 8;   // bytes per complex element of the basis
    s->nblk = (int)std::min<int64_t>((n + RB - 1) / RB, 1024);
    const int ntile = (restart + 1 + DT - 1) / DT;
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, bytes); };
    A((void **)&s->V, (size_t)(restart + 1) * n * es);
    A((void **)&s->w, n * 16); A((void **)&s->z, n * 16); A((void **)&s->r, n * 16); A((void **)&s->d, n * 16);
    A((void **)&s->partial, (size_t)ntile * s->nblk * DT * 2 * sizeof(double) + 1024 * sizeof(double));
    A((void **)&s->h, 3 * 2 * 512 * sizeof(double));   // h pass 1 | h pass 2 | y
    A((void **)&s->nrm, 4 * sizeof(double));
    A((void **)&s->inv64, s->n_ent * 4 * 16);
    A((void **)&s->arn, arn_off(restart).total * sizeof(double));
    A((void **)&s->pn, 1024 * sizeof(double));
    A((void **)&s->flag, 8 * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->res_host, (restart + 8) * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking);
    if (precision == 32) { A((void **)&s->inv32, s->n_ent * 4 * 8); A((void **)&s->vals32, nnz * 8); }
    A((void **)&s->poly_c, 2 * POLY_MAX * sizeof(double));
    A((void **)&s->pp, n * 16); A((void **)&s->pt, n * 16); A((void **)&s->py, n * 16); A((void **)&s->pu, n * 16);
    if (e != hipSuccess) {
        emg::set_error(std::string("emg_solver_create: ") + hipGetErrorString(e));
        free_solver(s);
        return -2;
    }
    // the blocked SpMV and the 2x2 block smoother need an entity pattern: check it once
    int bad = 0;
    if (nnz <= 0 || nnz >= ((int64_t)1 << 31)) bad = 1;
    if (!bad) {
        int *d_bad = reinterpret_cast<int *>(s->flag);
        e = hipMemset(d_bad, 0, sizeof(int));
        if (e == hipSuccess) {
            k_check_entity_pattern<<<emg::grid_for(s->n_ent + 1, 256), 256>>>(s->n_ent, nnz, d_rowptr, d_cols, d_bad);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            emg::set_error(std::string("emg_solver_create: pattern check: ") + hipGetErrorString(e));
            free_solver(s);
            return -2;
        }
    }
    if (bad) {
        emg::set_error("emg_solver_create: the CSR is not on an entity pattern (emg_pattern_csr layout: rows e and "
                       "e + n/2 share their column entities, mode-1 columns after mode-0 ones, row e + n/2 starting "
                       "nnz/2 after row e); pass the full unreduced matrix and the PEC mask");
        free_solver(s);
        return -1;
    }
    *out = s;
    return 0;
}

extern "C" void emg_solver_destroy(emg_solver *s) { free_solver(s); }

// inv blocks of entity e *= w_e[e] (complex64 copy too); dsinv[p] *= w_p[p]
__global__ void __launch_bounds__(RB) k_scale_tables(int64_t n_ent, const double *__restrict__ we, double *__restrict__ inv64,
                                                     float *__restrict__ inv32, int64_t np2, const double *__restrict__ wp,
                                                     double *__restrict__ ds64, float *__restrict__ ds32) {
    const int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (we != nullptr && i < n_ent) {
        const double wr = we[2 * i], wi = we[2 * i + 1];
        for (int q = 0; q < 4; q++) {
            double *a = inv64 + 2 * (4 * i + q);
            const double ar = a[0], ai = a[1];
            a[0] = wr * ar - wi * ai;
            a[1] = wr * ai + wi * ar;
            if (inv32) { inv32[2 * (4 * i + q)] = (float)a[0]; inv32[2 * (4 * i + q) + 1] = (float)a[1]; }
        }
    }
    if (wp != nullptr && i < np2) {
        const double wr = wp[2 * i], wi = wp[2 * i + 1];
        double *a = ds64 + 2 * i;
        const double ar = a[0], ai = a[1];
        a[0] = wr * ar - wi * ai;
        a[1] = wr * ai + wi * ar;
        if (ds32) { ds32[2 * i] = (float)a[0]; ds32[2 * i + 1] = (float)a[1]; }
    }
}

// Autotuner of the blocked SpMV: each lanes-per-entity variant (4, 8, 16, 32) runs
// once untimed and five times between HIP events on the solver's first matrix, in
// the precision the solver's steps use; the fastest is kept for the solver's life
// (the sparsity pattern fixes the row lengths that decide it).  Vectors: zeroed
// solver work buffers.  EMG_SPMV_BLK_LANES, when set, overrides the search.
static const int TUNE_LANES[4] = {4, 8, 16, 32};

template <typename R>
static int tune_spmv_one(emg_solver *s, const R *vals, int slot, hipStream_t st) {
    hipEvent_t ev[2];
    EMG_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        emg::set_error("tune_spmv: hipEventCreate failed");
        return -2;
    }
    hipError_t e = hipSuccess;
    float best = 3.4e38f;
    int bl = 16;
    for (int i = 0; i < 4 && e == hipSuccess; i++) {
        const int L = TUNE_LANES[i];
        auto run = [&]() {
            launch_spmv_blk<R>(st, s->n, s->nnz, s->rowptr, s->cols, vals, s->mask, reinterpret_cast<const R *>(s->z),
                               nullptr, reinterpret_cast<R *>(s->pp), L, sizeof(R) == 8);
        };
        run();
        e = hipEventRecord(ev[0], st);
        for (int r = 0; r < 5; r++) run();
        if (e == hipSuccess) e = hipEventRecord(ev[1], st);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        ms /= 5.f;
        s->tune_ms[slot][i] = ms;
        if (e == hipSuccess && ms < best) { best = ms; bl = L; }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    EMG_HIP(e);
    s->lanes[slot] = bl;
    return 0;
}

// The smoother's G^T product with 4, 8 and 16 lanes per row, timed like the SpMV (one
// untimed and five timed launches, in the precision of the solver's steps, on zeroed
// vectors); EMG_GT_LANES (4, 8 or 16) overrides.  Needs the gradient correction set.
static const int GT_LANES[3] = {4, 8, 16};

static int tune_gt(emg_solver *s, hipStream_t st) {
    if (s->np2 <= 0) return 0;
    if (const char *env = getenv("EMG_GT_LANES")) {
        const int v = atoi(env);
        s->gt_lanes = (v == 4 || v == 16) ? v : 8;
        return 0;
    }
    EMG_HIP(hipMemsetAsync(s->z, 0, s->n * 16, st));
    hipEvent_t ev[2];
    EMG_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        emg::set_error("tune_gt: hipEventCreate failed");
        return -2;
    }
    hipError_t e = hipSuccess;
    float best = 3.4e38f;
    int bl = 8;
    for (int i = 0; i < 3 && e == hipSuccess; i++) {
        const int L = GT_LANES[i];
        auto run = [&]() {
            if (s->precision == 32)
                launch_gt<float>(st, L, s->np2, s->gt_rowptr, s->gt_cols, s->gt_vals32, s->dsinv32,
                                 reinterpret_cast<const float *>(s->z), reinterpret_cast<float *>(s->tg));
            else
                launch_gt<double>(st, L, s->np2, s->gt_rowptr, s->gt_cols, s->gt_vals, s->dsinv64, s->z, s->tg);
        };
        run();
        e = hipEventRecord(ev[0], st);
        for (int r = 0; r < 5; r++) run();
        if (e == hipSuccess) e = hipEventRecord(ev[1], st);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (e == hipSuccess) e = hipGetLastError();
        ms /= 5.f;
        s->gt_ms[i] = ms;
        if (e == hipSuccess && ms < best) { best = ms; bl = L; }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    EMG_HIP(e);
    s->gt_lanes = bl;
    return 0;
}

static int tune_spmv(emg_solver *s, hipStream_t st) {
    if (getenv("EMG_SPMV_BLK_LANES") != nullptr) {
        s->lanes[0] = s->lanes[1] = spmv_blk_lanes();
        return tune_gt(s, st);
    }
    EMG_HIP(hipMemsetAsync(s->z, 0, s->n * 16, st));
    int rc = tune_spmv_one<double>(s, s->vals, 0, st);
    if (rc) return rc;
    if (s->precision == 32) {
        rc = tune_spmv_one<float>(s, s->vals32, 1, st);
        if (rc) return rc;
    } else {
        s->lanes[1] = s->lanes[0];
    }
    return tune_gt(s, st);
}

extern "C" int emg_solver_smoother_tuning(emg_solver *s, int *h_lanes, float *h_ms) {
    EMG_CHECK(s && h_lanes && h_ms, "emg_solver_smoother_tuning: null argument");
    *h_lanes = s->gt_lanes;
    for (int i = 0; i < 3; i++) h_ms[i] = s->gt_ms[i];
    return 0;
}

extern "C" int emg_solver_spmv_tuning(emg_solver *s, int *h_lanes, float *h_ms) {
    EMG_CHECK(s && h_lanes && h_ms, "emg_solver_spmv_tuning: null argument");
    for (int p = 0; p < 2; p++) {
        h_lanes[p] = s->lanes[p];
        for (int i = 0; i < 4; i++) h_ms[4 * p + i] = s->tune_ms[p][i];
    }
    return 0;
}

extern "C" int emg_solver_set_matrix(emg_solver *s, const double *d_vals, const uint8_t *d_mask, void *stream) {
    EMG_CHECK(s && d_vals, "emg_solver_set_matrix: null argument");
    hipStream_t st = (hipStream_t)stream;
    s->vals = d_vals;
    s->mask = d_mask;
    k_bj_setup<<<emg::grid_for(s->n_ent, RB), RB, 0, st>>>(s->n_ent, s->rowptr, s->cols, d_vals, d_mask, s->inv64, s->inv32);
    if (s->precision == 32) k_to_c64_masked<<<2048, RB, 0, st>>>(s->nnz, s->cols, d_mask, d_vals, s->vals32);
    if (s->np2 > 0) {
        EMG_HIP(hipMemsetAsync(s->dS, 0, s->np2 * 16, st));
        k_gdiag<<<emg::grid_for(s->n, RB), RB, 0, st>>>(s->n, s->rowptr, s->cols, d_vals, d_mask, s->g_rowptr,
                                                         s->g_cols, s->g_vals, s->dS);
        k_dsinv<<<emg::grid_for(s->np2, RB), RB, 0, st>>>(s->np2, s->dS, s->dsinv64, s->dsinv32);
    }
    if (s->w_e != nullptr || (s->w_p != nullptr && s->np2 > 0)) {
        const int64_t nmax = std::max<int64_t>(s->n_ent, s->np2);
        k_scale_tables<<<emg::grid_for(nmax, RB), RB, 0, st>>>(s->n_ent, s->w_e, s->inv64, s->inv32, s->np2,
                                                               s->np2 > 0 ? s->w_p : nullptr, s->dsinv64, s->dsinv32);
    }
    EMG_LAUNCH_CHECK();
    if (s->lanes[0] == 0) {
        const int rc = tune_spmv(s, st);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int emg_solver_set_weights(emg_solver *s, const double *d_we, const double *d_wp) {
    EMG_CHECK(s != nullptr, "emg_solver_set_weights: null solver");
    s->w_e = d_we;
    s->w_p = d_wp;
    return 0;
}

// The two parts of the smoother for a vector r (complex128): u = BJ(r) (n) and,
// with the gradient correction set, t = diag(G^T K G)^{-1} G^T r (np2).
extern "C" int emg_solver_smoother_parts(emg_solver *s, const double *d_r, double *d_u, double *d_t, void *stream) {
    EMG_CHECK(s && d_r && d_u, "emg_solver_smoother_parts: null argument");
    hipStream_t st = (hipStream_t)stream;
    k_bj_apply<double><<<emg::grid_for(s->n_ent, RB), RB, 0, st>>>(s->n_ent, s->inv64, d_r, d_u);
    if (s->np2 > 0 && d_t != nullptr)
        launch_gt<double>(st, s->gt_lanes, s->np2, s->gt_rowptr, s->gt_cols, s->gt_vals, s->dsinv64, d_r, d_t);
    EMG_LAUNCH_CHECK();
    return 0;
}

// Polynomial of the preconditioner: degree d (1 = none) and its coefficients
// c_i = 1/theta_i (complex128, host array of 2 d doubles) in application order.
extern "C" int emg_solver_set_poly(emg_solver *s, int degree, const double *h_coef) {
    EMG_CHECK(s != nullptr, "emg_solver_set_poly: null solver");
    EMG_CHECK(degree >= 1 && degree <= POLY_MAX, "emg_solver_set_poly: degree must be in [1, 16]");
    EMG_CHECK(degree == 1 || h_coef != nullptr, "emg_solver_set_poly: null coefficients");
    if (degree > 1) {
        for (int i = 0; i < 2 * degree; i++)
            EMG_CHECK(std::isfinite(h_coef[i]), "emg_solver_set_poly: non-finite coefficient");
        EMG_HIP(hipMemcpy(s->poly_c, h_coef, 2 * degree * sizeof(double), hipMemcpyHostToDevice));
    }
    s->poly_deg = degree;
    return 0;
}

__global__ void __launch_bounds__(RB) k_scale_c128(int64_t n, double a, double *__restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < 2 * n; e += (int64_t)gridDim.x * RB) x[e] *= a;
}

// d steps of Arnoldi (complex128, two classical Gram-Schmidt passes) on K M
// (M = the smoother without polynomial, K masked) from d_v0; writes the
// (d + 1) x d Hessenberg matrix column-major as complex128 to h_out (host).
// Uses the solver's Krylov basis storage, so it must not run during a solve.
extern "C" int emg_solver_arnoldi_probe(emg_solver *s, int d, const double *d_v0, double *h_out, void *stream) {
    EMG_CHECK(s && d_v0 && h_out, "emg_solver_arnoldi_probe: null argument");
    EMG_CHECK(s->vals != nullptr, "emg_solver_arnoldi_probe: emg_solver_set_matrix was not called");
    const int64_t n = s->n;
    const size_t es = s->precision == 64 ? 16 : 8;
    EMG_CHECK(d >= 1 && d <= POLY_MAX && (size_t)(d + 1) * 16 <= (size_t)(s->restart + 1) * es,
              "emg_solver_arnoldi_probe: d out of range for this solver");
    hipStream_t st = (hipStream_t)stream;
    const int nb = s->nblk;
    double *V = s->V;   // d + 1 complex128 vectors
    double *z = s->z;
    std::vector<double> H(2 * (size_t)(d + 1) * d, 0.0), pn(nb), hh(2 * (d + 1));
    auto host_norm = [&](const double *partial) -> double {
        EMG_HIP(hipMemcpyAsync(pn.data(), partial, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipStreamSynchronize(st));
        double t = 0;
        for (double v : pn) t += v;
        return std::sqrt(t);
    };
    EMG_HIP(hipMemcpyAsync(V, d_v0, n * 16, hipMemcpyDeviceToDevice, st));
    k_norm2<double><<<nb, RB, 0, st>>>(n, V, s->partial);
    const double b0 = host_norm(s->partial);
    EMG_CHECK(b0 > 0 && std::isfinite(b0), "emg_solver_arnoldi_probe: zero or non-finite start vector");
    k_scale_c128<<<nb, RB, 0, st>>>(n, 1.0 / b0, V);
    for (int j = 0; j < d; j++) {
        double *vj = V + (size_t)j * n * 2, *w = V + (size_t)(j + 1) * n * 2;
        const int nv = j + 1, nbm = mdot_blocks(n, nv);
        const dim3 grid(nbm, (nv + DT - 1) / DT);
        apply_prec<double>(s, s->inv64, vj, z, st);
        launch_spmv_blk<double>(st, n, s->nnz, s->rowptr, s->cols, s->vals, s->mask, z, nullptr, w, s->lanes[0]);
        double wn = 0;
        for (int pass = 0; pass < 2; pass++) {
            k_mdot<double><<<grid, RB, 0, st>>>(n, nv, V, n, w, s->partial, nullptr);
            k_mdot_finish<<<nv, 256, 0, st>>>(nv, nbm, s->partial, s->h, 0, nullptr);
            k_maxpy_norm<double><<<nb, RB, 0, st>>>(n, nv, V, n, s->h, w, s->pn, nullptr);
            EMG_HIP(hipMemcpyAsync(hh.data(), s->h, 2 * nv * sizeof(double), hipMemcpyDeviceToHost, st));
            wn = host_norm(s->pn);
            for (int i = 0; i < nv; i++) {
                H[2 * ((size_t)j * (d + 1) + i)] += hh[2 * i];
                H[2 * ((size_t)j * (d + 1) + i) + 1] += hh[2 * i + 1];
            }
        }
        H[2 * ((size_t)j * (d + 1) + j + 1)] = wn;
        EMG_CHECK(wn > 0 && std::isfinite(wn), "emg_solver_arnoldi_probe: breakdown");
        k_scale_c128<<<nb, RB, 0, st>>>(n, 1.0 / wn, w);
    }
    EMG_HIP(hipStreamSynchronize(st));
    memcpy(h_out, H.data(), H.size() * sizeof(double));
    return 0;
}

__global__ void __launch_bounds__(RB) k_f64_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) b[e] = (float)a[e];
}

extern "C" int emg_solver_set_gradient(emg_solver *s, int64_t np2, const int32_t *d_g_rowptr, const int32_t *d_g_cols,
                                       const double *d_g_vals, const int32_t *d_gt_rowptr, const int32_t *d_gt_cols,
                                       const double *d_gt_vals, void *stream) {
    EMG_CHECK(s != nullptr && np2 >= 0, "emg_solver_set_gradient: bad argument");
    hipStream_t st = (hipStream_t)stream;
    if (np2 == 0) {
        s->np2 = 0;
        return 0;
    }
    EMG_CHECK(d_g_rowptr && d_g_cols && d_g_vals && d_gt_rowptr && d_gt_cols && d_gt_vals,
              "emg_solver_set_gradient: null argument");
    int32_t nnz_g = 0, nnz_gt = 0;
    EMG_HIP(hipMemcpyAsync(&nnz_g, d_g_rowptr + s->n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EMG_HIP(hipMemcpyAsync(&nnz_gt, d_gt_rowptr + np2, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EMG_HIP(hipStreamSynchronize(st));
    EMG_CHECK(nnz_g == nnz_gt && nnz_g > 0, "emg_solver_set_gradient: G and G^T entry counts differ");
    EMG_CHECK(nnz_g <= G_ROW_MAX * s->n, "emg_solver_set_gradient: more than 8 entries per row of G");
    // captured step graphs hold the old gradient buffers: drop them before the buffers go
    EMG_HIP(hipStreamSynchronize(st));
    for (auto &g : s->gs) {
        for (hipGraphExec_t e : g.exec) (void)hipGraphExecDestroy(e);
        g.exec.clear();
    }
    void *old[] = {s->g_vals32, s->gt_vals32, s->dS, s->dsinv64, s->tg, s->dsinv32};
    for (void *p : old) (void)hipFree(p);
    s->g_vals32 = s->gt_vals32 = nullptr;
    s->dS = s->dsinv64 = s->tg = nullptr;
    s->dsinv32 = nullptr;
    s->np2 = 0;
    EMG_HIP(hipMalloc((void **)&s->g_vals32, (size_t)nnz_g * sizeof(float)));
    EMG_HIP(hipMalloc((void **)&s->gt_vals32, (size_t)nnz_g * sizeof(float)));
    EMG_HIP(hipMalloc((void **)&s->dS, (size_t)np2 * 16));
    EMG_HIP(hipMalloc((void **)&s->dsinv64, (size_t)np2 * 16));
    EMG_HIP(hipMalloc((void **)&s->dsinv32, (size_t)np2 * 8));
    EMG_HIP(hipMalloc((void **)&s->tg, (size_t)np2 * 16));
    k_f64_to_f32<<<1024, RB, 0, st>>>(nnz_g, d_g_vals, s->g_vals32);
    k_f64_to_f32<<<1024, RB, 0, st>>>(nnz_g, d_gt_vals, s->gt_vals32);
    EMG_LAUNCH_CHECK();
    s->g_rowptr = d_g_rowptr; s->g_cols = d_g_cols; s->g_vals = d_g_vals;
    s->gt_rowptr = d_gt_rowptr; s->gt_cols = d_gt_cols; s->gt_vals = d_gt_vals;
    s->np2 = np2;
    return 0;
}

// ---- one restarted GMRES solve of K u = rhs (element type R), right preconditioned.
// On return u holds the solution; returns the iteration count through *iters and
// the final residual estimate relative to |rhs| through *rel.
// ---------------------------------------------------------------- device Arnoldi bookkeeping
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += sh[q];
    return t;
}

// g = (beta, 0, ...), residual estimates cleared; beta = *nrm.
__global__ void __launch_bounds__(RB) k_arn_init(int m, const double *__restrict__ nrm, double *__restrict__ arn,
                                                 arn_off o) {
    for (int i = threadIdx.x; i < 2 * (m + 1); i += RB) arn[o.g + i] = (i == 0) ? nrm[0] : 0.0;
    for (int i = threadIdx.x; i < m; i += RB) arn[o.res + i] = 0.0;
}

// Second Gram-Schmidt pass only if the first removed more than 99% of the
// vector's squared norm (||w after|| < 0.1 ||w before||, with
// ||w before||^2 = ||w after||^2 + ||h||^2).  The 0.707 threshold of the
// classical test reorthogonalised nearly every step of these solves without
// changing their iteration counts; 0.1 keeps the second pass for real
// cancellation only.
__global__ void __launch_bounds__(RB) k_arn_check(int nb, const double *__restrict__ pn, const double *__restrict__ h,
                                                  int nv, double *__restrict__ flag) {
    __shared__ double sh[RB / 64];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) a += pn[i];
    for (int i = threadIdx.x; i < 2 * nv; i += RB) b += h[i] * h[i];
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) flag[0] = (a < 0.01 * (a + b)) ? 1.0 : 0.0;
}

// Hessenberg column j = h1 + h2 with ||new vector||, previous rotations, the
// new rotation, g update, residual estimate |g[j+1]| and 1/||new vector||.
// The column and the rotations are staged in LDS by the whole block; the
// rotation chain itself is sequential and runs in registers of one lane.
__global__ void __launch_bounds__(RB) k_arn_givens(int j, int m, int nb, const double *__restrict__ pn,
                                                   const double *__restrict__ h1, const double *__restrict__ h2,
                                                   double *__restrict__ arn, arn_off o) {
    __shared__ double sh[RB / 64];
    __shared__ double2 col[512], rot[2 * 512];
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) a += pn[i];
    for (int i = threadIdx.x; i <= j; i += RB) col[i] = make_double2(h1[2 * i] + h2[2 * i], h1[2 * i + 1] + h2[2 * i + 1]);
    const double *cs = arn + o.cs, *sn = arn + o.sn;
    for (int i = threadIdx.x; i < j; i += RB) {
        rot[2 * i] = make_double2(cs[2 * i], cs[2 * i + 1]);
        rot[2 * i + 1] = make_double2(sn[2 * i], sn[2 * i + 1]);
    }
    a = block_sum(a, sh);          // contains __syncthreads: col / rot are visible after it
    if (threadIdx.x != 0) return;
    const double hn = sqrt(a);
    // carry = current entry i of the column after rotations 0..i-1
    double xr = col[0].x, xi = col[0].y;
    for (int i = 0; i < j; i++) {
        const double2 c = rot[2 * i], sv = rot[2 * i + 1], y = col[i + 1];
        // col[i] = conj(c) x + conj(s) y ; carry = -s x + c y
        col[i] = make_double2(c.x * xr + c.y * xi + sv.x * y.x + sv.y * y.y,
                              c.x * xi - c.y * xr + sv.x * y.y - sv.y * y.x);
        const double nr = -(sv.x * xr - sv.y * xi) + (c.x * y.x - c.y * y.y);
        const double ni = -(sv.x * xi + sv.y * xr) + (c.x * y.y + c.y * y.x);
        xr = nr;
        xi = ni;
    }
    // new rotation from (x, hn)
    const double den = sqrt(xr * xr + xi * xi + hn * hn);
    double cr = 1.0, ci = 0.0, sr = 0.0, si = 0.0;
    if (den > 0.0) { cr = xr / den; ci = xi / den; sr = hn / den; }
    double *Hc = arn + o.H + 2 * (int64_t)j * (m + 1);
    for (int i = 0; i < j; i++) { Hc[2 * i] = col[i].x; Hc[2 * i + 1] = col[i].y; }
    Hc[2 * j] = cr * xr + ci * xi + sr * hn;          // conj(c) x + conj(s) hn
    Hc[2 * j + 1] = cr * xi - ci * xr;
    Hc[2 * j + 2] = 0.0;
    Hc[2 * j + 3] = 0.0;
    double *csw = arn + o.cs, *snw = arn + o.sn, *g = arn + o.g;
    csw[2 * j] = cr; csw[2 * j + 1] = ci;
    snw[2 * j] = sr; snw[2 * j + 1] = si;
    const double gr = g[2 * j], gi = g[2 * j + 1];
    g[2 * j + 2] = -(sr * gr - si * gi);
    g[2 * j + 3] = -(sr * gi + si * gr);
    g[2 * j] = cr * gr + ci * gi;
    g[2 * j + 1] = cr * gi - ci * gr;
    arn[o.res + j] = sqrt(g[2 * j + 2] * g[2 * j + 2] + g[2 * j + 3] * g[2 * j + 3]);
    arn[o.inv] = hn > 0.0 ? 1.0 / hn : 0.0;
}

template <typename R>
__global__ void __launch_bounds__(RB) k_scale_inv(int64_t n, const double *__restrict__ inv, R *__restrict__ v) {
    const double s = inv[0];
    for (int64_t e = blockIdx.x * (int64_t)RB + threadIdx.x; e < n; e += (int64_t)gridDim.x * RB) {
        const cx<R> a = ld(v, e);
        st(v, e, cx<R>((R)(a.re * s), (R)(a.im * s)));
    }
}

// One Arnoldi step j: v_{j+1} = K M^{-1} v_j, CGS (second pass decided on the
// device), Hessenberg/Givens update, normalisation.  Only kernel launches, so
// it can be captured.
template <typename R>
static void enqueue_step(emg_solver *s, const R *vals, const R *inv, int j, hipStream_t st) {
    const int64_t n = s->n;
    const int m = s->restart, nb = s->nblk, nv = j + 1;
    R *V = reinterpret_cast<R *>(s->V);
    R *z = reinterpret_cast<R *>(s->z);
    R *vj = V + (size_t)j * n * 2, *vn = V + (size_t)(j + 1) * n * 2;
    double *h1 = s->h, *h2 = s->h + 2 * 512;
    const arn_off o(m);
    const int nbm = mdot_blocks(n, nv);
    const dim3 grid(nbm, (nv + DT - 1) / DT);
    apply_prec_poly<R>(s, vals, inv, vj, z, st);
    // the complex64 copy has its masked columns zeroed (k_to_c64_masked)
    launch_spmv_blk<R>(st, n, s->nnz, s->rowptr, s->cols, vals, s->mask, z, nullptr, vn, s->lanes[sizeof(R) == 4],
                       sizeof(R) == 8);
    k_mdot<R><<<grid, RB, 0, st>>>(n, nv, V, n, vn, s->partial, nullptr);
    k_mdot_finish<<<nv, 256, 0, st>>>(nv, nbm, s->partial, h1, 0, nullptr);
    k_maxpy_norm<R><<<nb, RB, 0, st>>>(n, nv, V, n, h1, vn, s->pn, nullptr);
    k_arn_check<<<1, RB, 0, st>>>(nb, s->pn, h1, nv, s->flag);
    k_mdot<R><<<grid, RB, 0, st>>>(n, nv, V, n, vn, s->partial, s->flag);
    k_mdot_finish<<<nv, 256, 0, st>>>(nv, nbm, s->partial, h2, 0, s->flag);
    k_maxpy_norm<R><<<nb, RB, 0, st>>>(n, nv, V, n, h2, vn, s->pn, s->flag);
    k_arn_givens<<<1, RB, 0, st>>>(j, m, nb, s->pn, h1, h2, s->arn, o);
    k_scale_inv<R><<<nb, RB, 0, st>>>(n, s->arn + o.inv, vn);
}

// Capture the m steps of a restart cycle as graphs of CHUNK_IT steps (once per
// solver, precision and matrix/preconditioner/mask buffers).
template <typename R>
static int ensure_graphs(emg_solver *s, const R *vals, const R *inv) {
    auto &g = s->gs[sizeof(R) == 8 ? 0 : 1];
    if (!g.exec.empty() && g.vals == vals && g.inv == inv && g.mask == s->mask && g.np2 == s->np2 &&
        g.poly == s->poly_deg)
        return 0;
    for (hipGraphExec_t e : g.exec) (void)hipGraphExecDestroy(e);
    g.exec.clear();
    const int m = s->restart;
    for (int c0 = 0; c0 < m; c0 += CHUNK_IT) {
        EMG_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
        for (int j = c0; j < std::min(m, c0 + CHUNK_IT); j++) enqueue_step<R>(s, vals, inv, j, s->cap);
        hipGraph_t graph = nullptr;
        EMG_HIP(hipStreamEndCapture(s->cap, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        EMG_HIP(e);
        g.exec.push_back(exec);
    }
    g.vals = vals;
    g.inv = inv;
    g.mask = s->mask;
    g.np2 = s->np2;
    g.poly = s->poly_deg;
    return 0;
}

template <typename R>
static int gmres_core(emg_solver *s, const R *vals, const R *inv, const R *rhs, R *u, double tol, int maxiter,
                      int *iters, double *rel, hipStream_t st) {
    const int64_t n = s->n;
    const int m = s->restart;
    const int nb = s->nblk;
    const arn_off o(m);
    R *V = reinterpret_cast<R *>(s->V);
    R *w = reinterpret_cast<R *>(s->w);
    R *z = reinterpret_cast<R *>(s->z);
    auto prec = [&](const R *x, R *y) { apply_prec_poly<R>(s, vals, inv, x, y, st); };
    double bnorm = 0;
    {
        k_norm2<R><<<nb, RB, 0, st>>>(n, rhs, s->partial);
        std::vector<double> p(nb);
        EMG_HIP(hipMemcpyAsync(p.data(), s->partial, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipStreamSynchronize(st));
        for (double v : p) bnorm += v;
        bnorm = std::sqrt(bnorm);
    }
    *iters = 0;
    *rel = 0;
    if (bnorm == 0) {
        EMG_HIP(hipMemsetAsync(u, 0, n * 2 * sizeof(R), st));
        return 0;
    }
    if (int rc = ensure_graphs<R>(s, vals, inv)) return rc;
    const auto &graphs = s->gs[sizeof(R) == 8 ? 0 : 1].exec;
    std::vector<double> Hh, gh(2 * (m + 1));
    std::vector<std::complex<double>> y(m);
    int total = 0;
    double resid = 1.0;
    while (true) {
        // r = rhs - K u  ->  v0 = r / |r|
        launch_spmv_blk<R>(st, n, s->nnz, s->rowptr, s->cols, vals, s->mask, u, rhs, w, s->lanes[sizeof(R) == 4],
                           sizeof(R) == 8);
        k_norm2<R><<<nb, RB, 0, st>>>(n, w, s->partial);
        k_scale_by_norm<R><<<nb, RB, 0, st>>>(n, s->partial, nb, w, V, s->nrm);
        double beta;
        EMG_HIP(hipMemcpyAsync(&beta, s->nrm, sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipStreamSynchronize(st));
        resid = beta / bnorm;
        if (resid <= tol || total >= maxiter || !(beta > 0) || !std::isfinite(beta)) break;
        k_arn_init<<<1, RB, 0, st>>>(m, s->nrm, s->arn, o);
        int k = 0;
        bool done = false;
        for (int c = 0; c < (int)graphs.size() && !done; c++) {
            const int j0 = c * CHUNK_IT, j1 = std::min(m, j0 + CHUNK_IT);
            EMG_HIP(hipGraphLaunch(graphs[c], st));
            EMG_HIP(hipMemcpyAsync(s->res_host + j0, s->arn + o.res + j0, (j1 - j0) * sizeof(double),
                                   hipMemcpyDeviceToHost, st));
            EMG_HIP(hipStreamSynchronize(st));
            for (int j = j0; j < j1; j++) {
                k = j + 1;
                resid = s->res_host[j] / bnorm;
                if (!std::isfinite(resid) || resid <= tol || total + k >= maxiter) { done = true; break; }
            }
        }
        total += k;
        // y = R^{-1} g (upper triangular after the rotations) ; u += M^{-1} V y
        Hh.resize(2 * (size_t)k * (m + 1));
        EMG_HIP(hipMemcpyAsync(Hh.data(), s->arn + o.H, Hh.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipMemcpyAsync(gh.data(), s->arn + o.g, 2 * (m + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipStreamSynchronize(st));
        auto Hc = [&](int i, int q) { return std::complex<double>(Hh[2 * ((size_t)q * (m + 1) + i)], Hh[2 * ((size_t)q * (m + 1) + i) + 1]); };
        for (int i = k - 1; i >= 0; i--) {
            std::complex<double> acc(gh[2 * i], gh[2 * i + 1]);
            for (int q = i + 1; q < k; q++) acc -= Hc(i, q) * y[q];
            y[i] = acc / Hc(i, i);
        }
        std::vector<double> yb(2 * k);
        for (int i = 0; i < k; i++) { yb[2 * i] = y[i].real(); yb[2 * i + 1] = y[i].imag(); }
        double *yd = s->h + 4 * 512;
        if (k > 0) {
            EMG_HIP(hipMemcpyAsync(yd, yb.data(), 2 * k * sizeof(double), hipMemcpyHostToDevice, st));
            k_mcomb<R><<<nb, RB, 0, st>>>(n, k, V, n, yd, w, 1);
            prec(w, z);
            k_add<R><<<nb, RB, 0, st>>>(n, z, u);
        }
        EMG_HIP(hipStreamSynchronize(st));
        if (!std::isfinite(resid)) break;
    }
    *iters = total;
    *rel = resid;
    return 0;
}

extern "C" int emg_solver_solve(emg_solver *s, const double *d_b, double *d_x, double tol, int maxiter, int max_refine,
                                emg_solve_info *info, void *stream) {
    EMG_CHECK(s && d_b && d_x && info, "emg_solver_solve: null argument");
    EMG_CHECK(s->vals != nullptr, "emg_solver_solve: emg_solver_set_matrix was not called");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = s->n;
    const int nb = s->nblk;
    memset(info, 0, sizeof(*info));
    auto true_resid = [&](double *out_rel, double *bn) -> int {
        // r = b - K x in complex128 (masked)
        launch_spmv_blk<double>(st, n, s->nnz, s->rowptr, s->cols, s->vals, s->mask, d_x, d_b, s->r, s->lanes[0]);
        k_norm2<double><<<nb, RB, 0, st>>>(n, s->r, s->partial);
        k_norm2<double><<<nb, RB, 0, st>>>(n, d_b, s->partial + 1024);
        std::vector<double> p(2048);
        EMG_HIP(hipMemcpyAsync(p.data(), s->partial, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipMemcpyAsync(p.data() + 1024, s->partial + 1024, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        EMG_HIP(hipStreamSynchronize(st));
        double rn = 0, b2 = 0;
        for (int i = 0; i < nb; i++) { rn += p[i]; b2 += p[1024 + i]; }
        *bn = std::sqrt(b2);
        *out_rel = *bn > 0 ? std::sqrt(rn) / *bn : 0.0;
        return 0;
    };
    double rel, bn;
    if (s->precision == 64) {
        int it;
        double est;
        int rc = gmres_core<double>(s, s->vals, s->inv64, d_b, d_x, tol, maxiter, &it, &est, st);
        if (rc) return rc;
        info->iterations = it;
        info->refinements = 0;
        rc = true_resid(&rel, &bn);
        if (rc) return rc;
        info->residual = rel;
    } else {
        // Mixed-precision restarted GMRES: every restart cycle runs in complex64
        // on K d = r / |r|, then x += d and r = b - K x are formed in FP64, so
        // complex64 never has to resolve more than one cycle's reduction.
        float *r32 = reinterpret_cast<float *>(s->d);
        float *d32 = reinterpret_cast<float *>(s->d) + 2 * n;
        int rc = true_resid(&rel, &bn);
        if (rc) return rc;
        int total = 0, ref = 0;
        const int limit = max_refine > 0 ? max_refine : (1 << 30);
        int stall = 0;
        while (rel > tol && ref < limit && total < maxiter) {
            const double rs = 1.0 / (rel * bn);
            k_c128_to_c64_scaled<<<1024, RB, 0, st>>>(n, s->r, rs, r32);
            EMG_HIP(hipMemsetAsync(d32, 0, n * 8, st));
            // A cycle stops once complex64 GMRES has reduced its residual 100-fold (or
            // reached what tol needs): each FP64 update then starts a short cycle
            // on a small Krylov basis (posts box, 201 points, polynomial degree 6:
            // sweep 24.5 s with a 1e-7 floor, 15.9 s with 1e-2).
            const double inner_tol = std::max(0.5 * tol / rel, 1e-2);
            int it;
            double est;
            rc = gmres_core<float>(s, s->vals32, s->inv32, r32, d32, inner_tol,
                                   std::min(s->restart, maxiter - total), &it, &est, st);
            if (rc) return rc;
            total += it;
            ref++;
            k_add_scaled<float><<<1024, RB, 0, st>>>(n, d32, 1.0 / rs, d_x);
            const double prev = rel;
            rc = true_resid(&rel, &bn);
            if (rc) return rc;
            if (it == 0) break;
            stall = (rel > 0.999 * prev) ? stall + 1 : 0;
            if (stall >= 3) break;   // no progress in complex64: stop (caller sees converged = 0)
        }
        info->iterations = total;
        info->refinements = ref;
        info->residual = rel;
    }
    info->converged = info->residual <= tol ? 1 : 0;
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_affine_combine(int64_t nnz, const double *d_E, const double *d_B, double s, double *d_K, void *stream) {
    EMG_CHECK(d_E && d_B && d_K, "emg_affine_combine: null argument");
    k_affine<<<2048, RB, 0, (hipStream_t)stream>>>(nnz, d_E, d_B, s, d_K);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_csr_find(int64_t m, const int32_t *d_rowptr, const int32_t *d_cols, const int32_t *d_qr,
                            const int32_t *d_qc, int32_t *d_pos, void *stream) {
    EMG_CHECK(d_rowptr && d_cols && d_qr && d_qc && d_pos, "emg_csr_find: null argument");
    if (m == 0) return 0;
    k_csr_find<<<emg::grid_for(m, RB), RB, 0, (hipStream_t)stream>>>(m, d_rowptr, d_cols, d_qr, d_qc, d_pos);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_scatter_add(int64_t m, const int32_t *d_pos, const double *d_v, double gre, double gim, double *d_K,
                               void *stream) {
    EMG_CHECK(d_pos && d_v && d_K, "emg_scatter_add: null argument");
    if (m == 0) return 0;
    k_scatter_add<<<emg::grid_for(m, RB), RB, 0, (hipStream_t)stream>>>(m, d_pos, d_v, gre, gim, d_K);
    EMG_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ vector primitives
// complex128 building blocks for the eigen driver (Arnoldi with CGS2).
extern "C" int emg_spmv(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_vals,
                        const uint8_t *d_mask, const double *d_x, double *d_y, void *stream) {
    EMG_CHECK(d_rowptr && d_cols && d_vals && d_x && d_y, "emg_spmv: null argument");
    launch_spmv<double>((hipStream_t)stream, n, d_rowptr, d_cols, d_vals, d_mask, d_x, nullptr, d_y);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_spmv_blocked(int64_t n, int64_t nnz, const int32_t *d_rowptr, const int32_t *d_cols,
                                const double *d_vals, const uint8_t *d_mask, const double *d_x, double *d_y,
                                void *stream) {
    EMG_CHECK(d_rowptr && d_cols && d_vals && d_x && d_y, "emg_spmv_blocked: null argument");
    EMG_CHECK(n > 0 && n % 2 == 0 && nnz % 2 == 0, "emg_spmv_blocked: n and nnz must be even");
    launch_spmv_blk<double>((hipStream_t)stream, n, nnz, d_rowptr, d_cols, d_vals, d_mask, d_x, nullptr, d_y,
                            spmv_blk_lanes());
    EMG_LAUNCH_CHECK();
    return 0;
}

// The product the solver's Arnoldi steps run: y = mask.(K (mask.x)) with the matrix of the last
// emg_solver_set_matrix, complex128 (precision 64) or the complex64 copy (precision 32, x and y
// complex64), PEC columns folded into that copy.
extern "C" int emg_solver_apply(emg_solver *s, int precision, const void *d_x, void *d_y, void *stream) {
    EMG_CHECK(s != nullptr && d_x && d_y, "emg_solver_apply: null argument");
    EMG_CHECK(s->vals != nullptr, "emg_solver_apply: no matrix set");
    hipStream_t st = (hipStream_t)stream;
    if (precision == 64) {
        launch_spmv_blk<double>(st, s->n, s->nnz, s->rowptr, s->cols, s->vals, s->mask,
                                static_cast<const double *>(d_x), nullptr, static_cast<double *>(d_y), s->lanes[0]);
    } else {
        EMG_CHECK(precision == 32 && s->precision == 32, "emg_solver_apply: precision 32 needs a precision-32 solver");
        launch_spmv_blk<float>(st, s->n, s->nnz, s->rowptr, s->cols, s->vals32, s->mask,
                               static_cast<const float *>(d_x), nullptr, static_cast<float *>(d_y), s->lanes[1], false);
    }
    EMG_LAUNCH_CHECK();
    return 0;
}

static int nblk_for(int64_t n) { return (int)std::min<int64_t>((n + RB - 1) / RB, 1024); }

// d_h[i] = sum_k conj(V_i[k]) w[k], i < nv; V holds nv vectors of n entries back to back.
// d_work: at least ceil(nv/8) * 1024 * 16 + 16 doubles.
extern "C" int emg_mdotc(int64_t n, int nv, const double *d_V, const double *d_w, double *d_h, double *d_work,
                         void *stream) {
    EMG_CHECK(d_V && d_w && d_h && d_work && nv > 0 && nv <= 512, "emg_mdotc: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const int nbm = mdot_blocks(n, nv);
    k_mdot<double><<<dim3(nbm, (nv + DT - 1) / DT), RB, 0, st>>>(n, nv, d_V, n, d_w, d_work, nullptr);
    k_mdot_finish<<<nv, 256, 0, st>>>(nv, nbm, d_work, d_h, 0, nullptr);
    EMG_LAUNCH_CHECK();
    return 0;
}

// w -= sum_i h_i V_i ; *d_norm (on device) = ||w|| after the update.
extern "C" int emg_maxpy(int64_t n, int nv, const double *d_V, const double *d_h, double *d_w, double *d_norm,
                         double *d_work, void *stream) {
    EMG_CHECK(d_V && d_h && d_w && d_norm && d_work && nv >= 0 && nv <= 512, "emg_maxpy: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const int nb = nblk_for(n);
    k_maxpy_norm<double><<<nb, RB, 0, st>>>(n, nv, d_V, n, d_h, d_w, d_work, nullptr);
    k_scale_by_norm<double><<<1, RB, 0, st>>>(0, d_work, nb, d_w, d_w, d_norm);   // n = 0: norm only
    EMG_LAUNCH_CHECK();
    return 0;
}

// y (+)= A x for a real CSR matrix A (float64 values) and complex128 x, y.
__global__ void __launch_bounds__(256) k_csr_real(int64_t nrows, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const double *__restrict__ av, const double *__restrict__ x,
                                                  double *__restrict__ y, int accumulate) {
    const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / SUB;
    const int lane = threadIdx.x % SUB;
    if (r >= nrows) return;
    double ar = 0.0, ai = 0.0;
    for (int32_t e = rp[r] + lane; e < rp[r + 1]; e += SUB) {
        ar += av[e] * x[2 * ci[e]];
        ai += av[e] * x[2 * ci[e] + 1];
    }
    for (int off = SUB / 2; off > 0; off >>= 1) {
        ar += __shfl_xor(ar, off, SUB);
        ai += __shfl_xor(ai, off, SUB);
    }
    if (lane == 0) {
        if (accumulate) { y[2 * r] += ar; y[2 * r + 1] += ai; }
        else { y[2 * r] = ar; y[2 * r + 1] = ai; }
    }
}

extern "C" int emg_csr_real_apply(int64_t nrows, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_vals,
                                  const double *d_x, double *d_y, int accumulate, void *stream) {
    EMG_CHECK(nrows >= 0 && d_rowptr && d_cols && d_vals && d_x && d_y, "emg_csr_real_apply: null argument");
    if (nrows == 0) return 0;
    k_csr_real<<<(unsigned)((nrows * SUB + 255) / 256), 256, 0, (hipStream_t)stream>>>(nrows, d_rowptr, d_cols, d_vals,
                                                                                     d_x, d_y, accumulate);
    EMG_LAUNCH_CHECK();
    return 0;
}

// x = sum_i y_i V_i  (y complex on device)
extern "C" int emg_mcomb(int64_t n, int nv, const double *d_V, const double *d_y, double *d_x, void *stream) {
    EMG_CHECK(d_V && d_y && d_x && nv > 0 && nv <= 512, "emg_mcomb: bad argument");
    k_mcomb<double><<<nblk_for(n), RB, 0, (hipStream_t)stream>>>(n, nv, d_V, n, d_y, d_x, 1);
    EMG_LAUNCH_CHECK();
    return 0;
}
