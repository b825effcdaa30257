// Nedelec-2 tetrahedral assembly on MI355X (gfx950).
//
// Replaces the reference's element loop and COO->CSR conversion
// (fem/physics/edm/optimized_assembly.py:43-116 with fem/mth/tet.py:854-1088).
//
// Design
//  * Entities = [edges | faces].  DOF = mode*n_ent + entity, which is the
//    reference's [e1 | f1 | e2 | f2] numbering (fem/elements/nedelec2.py:46-62),
//    so the global matrix is a 2x2 block matrix whose blocks share the
//    entity-adjacency pattern.  The symbolic step builds that pattern once per
//    mesh with a GPU radix sort of the 100 (entity, entity) keys of every tet.
//  * The sorted values of that sort form, for every adjacency entry (e, e'),
//    the list of tets containing both entities with their local slots.
//  * Numeric assembly is a row-owner gather: one thread per (e, e') computes
//    the 2x2 (mode, mode') sub-blocks of the curl-curl (E) and mass (B)
//    integrals for each shared tet and writes each CSR value exactly once.
//    No atomics, no zero fill, results are deterministic.
//  * Tets are canonical (vertices sorted by global id), which is exactly the
//    orientation the reference derives from its global-id local maps
//    (fem/mth/optimized.py:273-307), so all local tables are constants.
//
// Element integrals, per basis function N = s*(l_p l_q grad l_r - l_p l_r grad l_q):
//   mass   F_ij = s_i s_j /(6V*5040) * sum of 4 terms c4(..) g_x.Mm g_y
//   curl   curl N = s/(6V)^2 * sum_t cf_t l_{v_t} K_t  (K_t = g_x x g_y)
//          D_ij = s_i s_j /(120 (6V)^3) * sum_{t,t'} cf cf' (1+[v_t==v_t']) K_t.Ms K'_t'
// where g_k = 6V grad(lambda_k) and c4 = prod(multiplicity!) is the
// barycentric-monomial moment numerator (fem/mth/optimized.py:251-260).

#include <cstdio>
#include <cstring>
#include <thread>
#include <algorithm>
#include <type_traits>

#include <rocprim/rocprim.hpp>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

using emg::cplx;

static void hfree(void *p) { (void)hipFree(p); }

namespace emg {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace emg

extern "C" const char *emg_last_error(void) { return emg::g_err.c_str(); }

struct emg_pattern {
    int64_t nt = 0, n_ent = 0, nnz_ent = 0, n_inc = 0, nblk = 0;
    int bits = 0;
    int32_t *ent_ptr = nullptr;   // n_ent + 1
    int32_t *adj = nullptr;       // nnz_ent, column entity of each adjacency entry
    int32_t *pair_ptr = nullptr;  // nnz_ent + 1, into inc
    uint32_t *inc = nullptr;      // n_inc = 100*nt : (tet << 8) | (s << 4) | s'
    int32_t *blk_row0 = nullptr;  // nblk + 1, first row of each 256-entry block
    double *tet_blocks = nullptr; // nt * TET_BLOCK_DOUBLES, two-phase fallback workspace
    // fused path: per chunk of CHUNK adjacency entries, its sorted distinct tets
    int64_t nchunk = 0;
    int chunk = 0;                   // adjacency entries per chunk (CHUNK)
    int fused_ok = 0;
    int32_t *chunk_row0 = nullptr;   // nchunk + 1, row of each chunk's first entry (last: n_ent - 1)
    int32_t *chunk_tets = nullptr;   // nchunk * CHUNK_TMAX
    int32_t *chunk_ntets = nullptr;  // nchunk
    int chunk_tmax = 0;              // largest chunk_ntets of this mesh (sizes the fused kernel's LDS)
    uint16_t *inc_local = nullptr;   // n_inc: (local tet << 8) | (s << 4) | s'
    uint16_t *chunk_perm = nullptr;  // nchunk * CHUNK: chunk-local entries by descending incidence count
    double *tet_rec = nullptr;       // nt * REC: per-tet Gram records (fused-path workspace)
    double *tet_rec_batch = nullptr; // rec_batch_cap * nt * REC: the records of a batch of meshes
    int rec_batch_cap = 0;
    uint8_t *recipes = nullptr;      // 100 * RECIPE_BYTES
    int wg_per_cu = 0;               // fused launch: workgroups per CU, 0 until tuned (first fused assembly)
    float wg_ms[5] = {};             // measured ms per assembly for ASM_WG[i]
};

static constexpr int PAIR_BLOCK = 256;
// Per-tet element record of the isotropic path: 55 upper (s <= s') 2x2 blocks
// of D, 55 of F, and the 6 lower face-face blocks of F (the reference's face
// mass is not symmetric there), each block (m,m') row-major.
static constexpr int N_UPPER = 55;
static constexpr int TET_BLOCK_DOUBLES = 4 * (N_UPPER + N_UPPER + 6);
// Fused path: a workgroup owns `chunk` (at most CHUNK) consecutive adjacency
// entries and stages the Gram data of their (at most CHUNK_TMAX) distinct tets in LDS.
static constexpr int CHUNK = 512;
static constexpr int CHUNK_TMAX = 128;
static constexpr int CHUNK_NMAX = 2048;  // incidences per chunk handled by the tet-list builder
static constexpr int RECIPE_BYTES = 48;

static void build_recipes(uint8_t *out);
__global__ void __launch_bounds__(CHUNK) k_chunk_tets(const emg_pattern pat, int32_t *__restrict__ chunk_tets,
                                                      int32_t *__restrict__ chunk_ntets, uint16_t *__restrict__ inc_local,
                                                      uint16_t *__restrict__ chunk_perm, int *__restrict__ overflow);

// ------------------------------------------------------------------ symbolic
__global__ void k_make_keys(const int32_t *__restrict__ tet_ent, int64_t nt, int bits,
                            uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int32_t e[10];
#pragma unroll
    for (int s = 0; s < 10; s++) e[s] = tet_ent[t * 10 + s];
    int64_t base = t * 100;
#pragma unroll
    for (int s = 0; s < 10; s++)
#pragma unroll
        for (int s2 = 0; s2 < 10; s2++) {
            keys[base + s * 10 + s2] = ((uint64_t)(uint32_t)e[s] << bits) | (uint64_t)(uint32_t)e[s2];
            vals[base + s * 10 + s2] = ((uint32_t)t << 8) | (uint32_t)(s << 4) | (uint32_t)s2;
        }
}

__global__ void k_rows_from_keys(const uint64_t *__restrict__ uniq, int64_t nnz_ent, int bits, int64_t n_ent,
                                 int32_t *__restrict__ ent_ptr, int32_t *__restrict__ adj) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > nnz_ent) return;
    if (k == nnz_ent) { ent_ptr[n_ent] = (int32_t)nnz_ent; return; }
    uint64_t key = uniq[k];
    int64_t row = (int64_t)(key >> bits);
    adj[k] = (int32_t)(key & ((1ull << bits) - 1));
    if (k == 0 || (int64_t)(uniq[k - 1] >> bits) != row) ent_ptr[row] = (int32_t)k;
}

__global__ void k_block_rows(const uint64_t *__restrict__ uniq, int64_t nnz_ent, int bits, int64_t nblk,
                             int64_t n_ent, int block, int32_t *__restrict__ blk_row0) {
    int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b > nblk) return;
    blk_row0[b] = (b == nblk) ? (int32_t)(n_ent - 1) : (int32_t)(uniq[b * (int64_t)block] >> bits);
}

__global__ void k_set_last(int32_t *p, int64_t idx, int32_t v) { p[idx] = v; }

// Row entity of adjacency entry k, searched inside the range of its block.
__device__ __forceinline__ int find_row(const int32_t *__restrict__ ent_ptr, const int32_t *__restrict__ blk_row0,
                                        int64_t k) {
    int blk = (int)(k / PAIR_BLOCK);
    int lo = blk_row0[blk], hi = blk_row0[blk + 1];
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ent_ptr[mid] <= k) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_csr_cols(const emg_pattern pat, int32_t *__restrict__ cols) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= pat.nnz_ent) return;
    int e = find_row(pat.ent_ptr, pat.blk_row0, k);
    int32_t p0 = pat.ent_ptr[e];
    int32_t deg = pat.ent_ptr[e + 1] - p0;
    int32_t c = pat.adj[k];
    for (int m = 0; m < 2; m++) {
        int64_t base = (int64_t)m * 2 * pat.nnz_ent + p0 + k;  // = rowptr(m,e) + (k - p0)
        cols[base] = c;
        cols[base + deg] = c + (int32_t)pat.n_ent;
    }
}

__global__ void k_csr_rowptr(const emg_pattern pat, int32_t *__restrict__ rowptr) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t n = 2 * pat.n_ent;
    if (r > n) return;
    if (r == n) { rowptr[r] = (int32_t)(4 * pat.nnz_ent); return; }
    int64_t m = r / pat.n_ent, e = r - m * pat.n_ent;
    rowptr[r] = (int32_t)(m * 2 * pat.nnz_ent + 2 * (int64_t)pat.ent_ptr[e]);
}

static int bits_for(int64_t n) {
    int b = 1;
    while ((1ll << b) < n) b++;
    return b;
}

extern "C" int emg_pattern_build(const int32_t *d_tet_ent, int64_t nt, int64_t n_ent, void *stream, emg_pattern **out) {
    EMG_CHECK(out != nullptr && d_tet_ent != nullptr, "emg_pattern_build: null argument");
    EMG_CHECK(nt > 0 && n_ent > 0, "emg_pattern_build: empty mesh");
    EMG_CHECK(nt < (1ll << 24), "emg_pattern_build: more than 2^24 tets is not supported");
    hipStream_t st = (hipStream_t)stream;
    int bits = bits_for(n_ent);
    EMG_CHECK(2 * bits <= 64 && n_ent < (1ll << 30), "emg_pattern_build: too many entities");
    const int64_t n_inc = 100 * nt;
    uint64_t *keys = nullptr, *keys_s = nullptr, *uniq = nullptr;
    uint32_t *vals = nullptr, *counts = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0, need = 0;
    int64_t *d_nruns = nullptr;
    auto pat = new emg_pattern();
    pat->nt = nt;
    pat->n_ent = n_ent;
    pat->n_inc = n_inc;
    pat->bits = bits;
    int rc = 0;
    auto fail = [&](int code) {
        hfree(keys); hfree(keys_s); hfree(vals); hfree(uniq); hfree(counts); hfree(tmp); hfree(d_nruns);
        emg_pattern_destroy(pat);
        return code;
    };
#define TRY(call)                                                                 \
    do {                                                                          \
        hipError_t _e = (call);                                                   \
        if (_e != hipSuccess) {                                                   \
            emg::set_error(std::string(#call) + ": " + hipGetErrorString(_e));    \
            return fail(-2);                                                      \
        }                                                                         \
    } while (0)
    TRY(hipMalloc(&keys, n_inc * sizeof(uint64_t)));
    TRY(hipMalloc(&keys_s, n_inc * sizeof(uint64_t)));
    TRY(hipMalloc(&vals, n_inc * sizeof(uint32_t)));
    TRY(hipMalloc(&pat->inc, n_inc * sizeof(uint32_t)));
    k_make_keys<<<emg::grid_for(nt, 256), 256, 0, st>>>(d_tet_ent, nt, bits, keys, vals);
    TRY(hipGetLastError());
    TRY(rocprim::radix_sort_pairs(nullptr, need, keys, keys_s, vals, pat->inc, (size_t)n_inc, 0, 2 * bits, st));
    tmp_bytes = need;
    TRY(rocprim::run_length_encode(nullptr, need, keys_s, (size_t)n_inc, keys, vals, d_nruns, st));
    if (need > tmp_bytes) tmp_bytes = need;
    TRY(hipMalloc(&tmp, tmp_bytes));
    TRY(hipMalloc(&d_nruns, sizeof(int64_t)));
    need = tmp_bytes;
    TRY(rocprim::radix_sort_pairs(tmp, need, keys, keys_s, vals, pat->inc, (size_t)n_inc, 0, 2 * bits, st));
    // Unique keys go to `keys` (now free), run lengths to `vals`.
    need = tmp_bytes;
    TRY(rocprim::run_length_encode(tmp, need, keys_s, (size_t)n_inc, keys, vals, d_nruns, st));
    int64_t nnz_ent = 0;
    TRY(hipMemcpyAsync(&nnz_ent, d_nruns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    EMG_CHECK(4 * nnz_ent < (1ll << 31) || (fail(-1), false), "emg_pattern_build: nnz exceeds int32 CSR indices");
    pat->nnz_ent = nnz_ent;
    uniq = keys;
    keys = nullptr;
    counts = vals;
    vals = nullptr;
    TRY(hipMalloc(&pat->pair_ptr, (nnz_ent + 1) * sizeof(int32_t)));
    TRY(hipMalloc(&pat->ent_ptr, (n_ent + 1) * sizeof(int32_t)));
    TRY(hipMalloc(&pat->adj, nnz_ent * sizeof(int32_t)));
    pat->nblk = (nnz_ent + PAIR_BLOCK - 1) / PAIR_BLOCK;
    TRY(hipMalloc(&pat->blk_row0, (pat->nblk + 1) * sizeof(int32_t)));
    {
        size_t scan_need = 0;
        TRY(rocprim::exclusive_scan(nullptr, scan_need, counts, pat->pair_ptr, (int32_t)0, (size_t)nnz_ent,
                                    rocprim::plus<int32_t>(), st));
        if (scan_need > tmp_bytes) {
            hfree(tmp);
            tmp = nullptr;
            TRY(hipMalloc(&tmp, scan_need));
            tmp_bytes = scan_need;
        }
        scan_need = tmp_bytes;
        TRY(rocprim::exclusive_scan(tmp, scan_need, counts, pat->pair_ptr, (int32_t)0, (size_t)nnz_ent,
                                    rocprim::plus<int32_t>(), st));
    }
    k_set_last<<<1, 1, 0, st>>>(pat->pair_ptr, nnz_ent, (int32_t)n_inc);
    k_rows_from_keys<<<emg::grid_for(nnz_ent + 1, 256), 256, 0, st>>>(uniq, nnz_ent, bits, n_ent, pat->ent_ptr, pat->adj);
    k_block_rows<<<emg::grid_for(pat->nblk + 1, 256), 256, 0, st>>>(uniq, nnz_ent, bits, pat->nblk, n_ent, PAIR_BLOCK,
                                                                   pat->blk_row0);
    TRY(hipGetLastError());
    // fused-path chunk tables
    pat->chunk = CHUNK;
    pat->nchunk = (nnz_ent + pat->chunk - 1) / pat->chunk;
    TRY(hipMalloc(&pat->chunk_row0, (pat->nchunk + 1) * sizeof(int32_t)));
    k_block_rows<<<emg::grid_for(pat->nchunk + 1, 256), 256, 0, st>>>(uniq, nnz_ent, bits, pat->nchunk, n_ent, pat->chunk,
                                                                     pat->chunk_row0);
    {
        int *d_over = nullptr;
        TRY(hipMalloc(&pat->chunk_tets, (size_t)pat->nchunk * CHUNK_TMAX * sizeof(int32_t)));
        TRY(hipMalloc(&pat->chunk_ntets, (size_t)pat->nchunk * sizeof(int32_t)));
        // two spare codes: the fused kernel reads the codes as 32-bit words
        TRY(hipMalloc(&pat->inc_local, (size_t)(n_inc + 2) * sizeof(uint16_t)));
        TRY(hipMalloc(&pat->chunk_perm, (size_t)pat->nchunk * pat->chunk * sizeof(uint16_t)));
        TRY(hipMalloc(&pat->recipes, 100 * RECIPE_BYTES));
        TRY(hipMalloc(&d_over, sizeof(int)));
        TRY(hipMemsetAsync(d_over, 0, sizeof(int), st));
        uint8_t host_rec[100 * RECIPE_BYTES];
        build_recipes(host_rec);
        TRY(hipMemcpyAsync(pat->recipes, host_rec, sizeof(host_rec), hipMemcpyHostToDevice, st));
        k_chunk_tets<<<(unsigned)pat->nchunk, CHUNK, 0, st>>>(*pat, pat->chunk_tets, pat->chunk_ntets, pat->inc_local,
                                                             pat->chunk_perm, d_over);
        TRY(hipGetLastError());
        int over = 0;
        TRY(hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, st));
        std::vector<int32_t> cnt((size_t)pat->nchunk);
        TRY(hipMemcpyAsync(cnt.data(), pat->chunk_ntets, cnt.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        TRY(hipStreamSynchronize(st));
        hfree(d_over);
        pat->fused_ok = over ? 0 : 1;
        pat->chunk_tmax = 0;
        for (int32_t v : cnt) pat->chunk_tmax = std::max(pat->chunk_tmax, (int)v);
    }
    TRY(hipStreamSynchronize(st));
    hfree(uniq);
    hfree(counts);
    hfree(keys_s);
    hfree(tmp);
    hfree(d_nruns);
#undef TRY
    *out = pat;
    return rc;
}

extern "C" int emg_pattern_info(const emg_pattern *pat, int64_t *n_rows, int64_t *nnz, int64_t *nnz_ent, int64_t *n_inc) {
    EMG_CHECK(pat != nullptr, "emg_pattern_info: null pattern");
    if (n_rows) *n_rows = 2 * pat->n_ent;
    if (nnz) *nnz = 4 * pat->nnz_ent;
    if (nnz_ent) *nnz_ent = pat->nnz_ent;
    if (n_inc) *n_inc = pat->n_inc;
    return 0;
}

extern "C" int emg_pattern_csr(const emg_pattern *pat, int32_t *d_rowptr, int32_t *d_cols, void *stream) {
    EMG_CHECK(pat != nullptr && d_rowptr != nullptr && d_cols != nullptr, "emg_pattern_csr: null argument");
    hipStream_t st = (hipStream_t)stream;
    k_csr_rowptr<<<emg::grid_for(2 * pat->n_ent + 1, 256), 256, 0, st>>>(*pat, d_rowptr);
    k_csr_cols<<<(unsigned)pat->nblk, PAIR_BLOCK, 0, st>>>(*pat, d_cols);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" void emg_pattern_destroy(emg_pattern *pat) {
    if (!pat) return;
    hfree(pat->ent_ptr);
    hfree(pat->adj);
    hfree(pat->pair_ptr);
    hfree(pat->inc);
    hfree(pat->blk_row0);
    hfree(pat->tet_blocks);
    hfree(pat->chunk_tets);
    hfree(pat->chunk_ntets);
    hfree(pat->inc_local);
    hfree(pat->chunk_perm);
    hfree(pat->chunk_row0);
    hfree(pat->tet_rec);
    hfree(pat->tet_rec_batch);
    hfree(pat->recipes);
    delete pat;
}

// ------------------------------------------------------------------ numeric
// Local slot s (0..9) -> canonical vertices (a, b, c); edges repeat b as c.
// Edges (0,1)(0,2)(0,3)(1,2)(1,3)(2,3); faces (0,1,2)(0,2,3)(0,1,3)(1,2,3):
// the reference's local pairs/triples (fem/mesh3d.py:292,296) on sorted tets.
__device__ __forceinline__ void slot_vertices(int s, int &a, int &b, int &c) {
    constexpr uint64_t P =
        (0ull | 1ull << 2 | 1ull << 4) << 0 | (0ull | 2ull << 2 | 2ull << 4) << 6 | (0ull | 3ull << 2 | 3ull << 4) << 12 |
        (1ull | 2ull << 2 | 2ull << 4) << 18 | (1ull | 3ull << 2 | 3ull << 4) << 24 | (2ull | 3ull << 2 | 3ull << 4) << 30 |
        (0ull | 1ull << 2 | 2ull << 4) << 36 | (0ull | 2ull << 2 | 3ull << 4) << 42 | (0ull | 1ull << 2 | 3ull << 4) << 48 |
        (1ull | 2ull << 2 | 3ull << 4) << 54;
    uint32_t v = (uint32_t)(P >> (6 * s)) & 63u;
    a = v & 3;
    b = (v >> 2) & 3;
    c = (v >> 4) & 3;
}

// prod(multiplicity!) of the multiset {w, x, y, z} of local vertex ids.
__device__ __forceinline__ double c4(int w, int x, int y, int z) {
    uint32_t cnt = (1u << (4 * w)) + (1u << (4 * x)) + (1u << (4 * y)) + (1u << (4 * z));
    double r = 1.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t n = (cnt >> (4 * k)) & 15u;
        r *= (n == 4u) ? 24.0 : (n == 3u) ? 6.0 : (n == 2u) ? 2.0 : 1.0;
    }
    return r;
}

template <bool C>
struct Num;
template <>
struct Num<false> {
    using T = double;
};
template <>
struct Num<true> {
    using T = cplx;
};

__device__ __forceinline__ double dmul(double a, double b) { return a * b; }
__device__ __forceinline__ cplx dmul(cplx a, double b) { return a * b; }

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 crs(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double nrm(V3 a) { return sqrt(dot(a, a)); }

// a^T M b for a complex 3x3 tensor M (row-major, interleaved re/im).
__device__ __forceinline__ cplx qform(V3 a, const cplx *M, V3 b) {
    cplx r;
    double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r += (av[i] * bv[j]) * M[i * 3 + j];
    return r;
}

constexpr int ABLK = 128;

// Per-thread staging of the 4 vertices and 4 cofactor gradients in LDS, laid
// out component-major so dynamic vertex indices stay bank-conflict free.
struct Stage {
    double *s;
    int tid;
    __device__ __forceinline__ void put(int slot, V3 v) {
        s[(slot * 3 + 0) * ABLK + tid] = v.x;
        s[(slot * 3 + 1) * ABLK + tid] = v.y;
        s[(slot * 3 + 2) * ABLK + tid] = v.z;
    }
    __device__ __forceinline__ V3 get(int slot) const {
        return {s[(slot * 3 + 0) * ABLK + tid], s[(slot * 3 + 1) * ABLK + tid], s[(slot * 3 + 2) * ABLK + tid]};
    }
};

template <bool TENSOR>
__global__ void __launch_bounds__(ABLK) k_assemble(const emg_pattern pat, const double *__restrict__ nodes,
                                                   const int32_t *__restrict__ tets, const cplx *__restrict__ mat_m,
                                                   const cplx *__restrict__ mat_s, int exact_face_mass,
                                                   cplx *__restrict__ E, cplx *__restrict__ B) {
    using T = typename Num<TENSOR>::T;
    __shared__ double lds[24 * ABLK];
    Stage st{lds, (int)threadIdx.x};
    const int64_t k = blockIdx.x * (int64_t)ABLK + threadIdx.x;
    if (k >= pat.nnz_ent) return;
    const int e = find_row(pat.ent_ptr, pat.blk_row0, k);
    const int32_t p0 = pat.ent_ptr[e];
    const int32_t deg = pat.ent_ptr[e + 1] - p0;
    cplx accD[2][2], accF[2][2];
    const int32_t q0 = pat.pair_ptr[k], q1 = pat.pair_ptr[k + 1];
    for (int32_t q = q0; q < q1; q++) {
        const uint32_t v = pat.inc[q];
        const int t = (int)(v >> 8);
        int s = (int)((v >> 4) & 15u), s2 = (int)(v & 15u);
        // The reference mirrors its edge-face blocks by transposition
        // (fem/mth/tet.py:997-1004); with non-symmetric tensors the face-row /
        // edge-column entries are therefore the edge-row / face-column formula.
        const bool swp = TENSOR && (s >= 6) && (s2 < 6);
        if (swp) { int tmp = s; s = s2; s2 = tmp; }
        const int4 tv = reinterpret_cast<const int4 *>(tets)[t];
        V3 P[4];
        const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
        for (int i = 0; i < 4; i++) P[i] = {nodes[3 * (int64_t)vid[i]], nodes[3 * (int64_t)vid[i] + 1], nodes[3 * (int64_t)vid[i] + 2]};
        const V3 e1 = sub(P[1], P[0]), e2 = sub(P[2], P[0]), e3 = sub(P[3], P[0]);
        const V3 g1 = crs(e2, e3), g2 = crs(e3, e1), g3 = crs(e1, e2);
        const V3 g0 = {-(g1.x + g2.x + g3.x), -(g1.y + g2.y + g3.y), -(g1.z + g2.z + g3.z)};
        const double V6 = fabs(dot(e1, g1));
#pragma unroll
        for (int i = 0; i < 4; i++) st.put(i, P[i]);
        st.put(4, g0);
        st.put(5, g1);
        st.put(6, g2);
        st.put(7, g3);
        int ia, ib, ic, ja, jb, jc;
        slot_vertices(s, ia, ib, ic);
        slot_vertices(s2, ja, jb, jc);
        const bool fi = s >= 6, fj = s2 >= 6;
        const V3 Gi[3] = {st.get(4 + ia), st.get(4 + ib), st.get(4 + ic)};
        const V3 Gj[3] = {st.get(4 + ja), st.get(4 + jb), st.get(4 + jc)};
        const V3 Pia = st.get(ia), Pib = st.get(ib), Pic = st.get(ic);
        const V3 Pja = st.get(ja), Pjb = st.get(jb), Pjc = st.get(jc);
        // Scales: edges L_ab (both modes); faces -L_ac (mode 1), +L_ab (mode 2).
        const double Lia = nrm(sub(Pia, Pib)), Lja = nrm(sub(Pja, Pjb));
        const double si[2] = {fi ? -nrm(sub(Pia, Pic)) : Lia, Lia};
        const double sj[2] = {fj ? -nrm(sub(Pja, Pjc)) : Lja, Lja};
        // Curl terms: K = (g_a x g_b, g_a x g_c, g_b x g_c), vertices paired
        // with each K, and per-mode coefficients.
        const V3 Ki[3] = {crs(Gi[0], Gi[1]), crs(Gi[0], Gi[2]), crs(Gi[1], Gi[2])};
        const V3 Kj[3] = {crs(Gj[0], Gj[1]), crs(Gj[0], Gj[2]), crs(Gj[1], Gj[2])};
        const int tvi0 = fi ? ic : ia, tvi2 = fi ? ia : ib;
        const int tvj0 = fj ? jc : ja, tvj2 = fj ? ja : jb;
        const int tvi[3] = {tvi0, ib, tvi2};
        const int tvj[3] = {tvj0, jb, tvj2};
        const double cfi[2][3] = {{fi ? -1.0 : -3.0, fi ? -2.0 : 0.0, fi ? -1.0 : 0.0},
                                  {fi ? -2.0 : 0.0, fi ? -1.0 : -3.0, fi ? 1.0 : 0.0}};
        const double cfj[2][3] = {{fj ? -1.0 : -3.0, fj ? -2.0 : 0.0, fj ? -1.0 : 0.0},
                                  {fj ? -2.0 : 0.0, fj ? -1.0 : -3.0, fj ? 1.0 : 0.0}};
        T H[3][3], Gb[3][3];
        const cplx *Mm = TENSOR ? mat_m + 9 * (int64_t)t : nullptr;
        const cplx *Ms = TENSOR ? mat_s + 9 * (int64_t)t : nullptr;
        cplx mmt[9], mst[9];
        if constexpr (TENSOR) {
#pragma unroll
            for (int i = 0; i < 9; i++) { mmt[i] = Mm[i]; mst[i] = Ms[i]; }
        }
#pragma unroll
        for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 3; y++) {
                const double w = (tvi[x] == tvj[y]) ? 2.0 : 1.0;
                if constexpr (TENSOR) {
                    H[x][y] = qform(Ki[x], mst, Kj[y]) * w;
                    Gb[x][y] = qform(Gi[x], mmt, Gj[y]);
                } else {
                    H[x][y] = dot(Ki[x], Kj[y]) * w;
                    Gb[x][y] = dot(Gi[x], Gj[y]);
                }
            }
        // Basis (p, q, r): r = a always; edge mode 1 (a, b, a), edge mode 2
        // (b, b, a), face mode 1 (b, c, a), face mode 2 (c, b, a).  q sits at
        // position c only for face mode 1.
        const double dscale = 1.0 / (120.0 * V6 * V6 * V6);
        const double fscale = 1.0 / (5040.0 * V6);
        cplx cm = TENSOR ? cplx(1.0) : mat_m[t];
        cplx cs = TENSOR ? cplx(1.0) : mat_s[t];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int p = (m == 0) ? (fi ? ib : ia) : (fi ? ic : ib);
            const int qq = (m == 0 && fi) ? ic : ib;
            const int r = ia;
            const bool qc = (m == 0 && fi);
#pragma unroll
            for (int m2 = 0; m2 < 2; m2++) {
                const int p2 = (m2 == 0) ? (fj ? jb : ja) : (fj ? jc : jb);
                const int q2 = (m2 == 0 && fj) ? jc : jb;
                const int r2 = ja;
                const bool qc2 = (m2 == 0 && fj);
                const T g_rq = qc2 ? Gb[0][2] : Gb[0][1];
                const T g_qr = qc ? Gb[2][0] : Gb[1][0];
                const T g_qq = qc ? (qc2 ? Gb[2][2] : Gb[2][1]) : (qc2 ? Gb[1][2] : Gb[1][1]);
                // curl-curl
                T d = T(0.0);
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    T row = T(0.0);
#pragma unroll
                    for (int y = 0; y < 3; y++) row += dmul(H[x][y], cfj[m2][y]);
                    d += dmul(row, cfi[m][x]);
                }
                // mass (second coefficient: reference index choice unless exact)
                double c2 = c4(p, qq, p2, r2);
                if (!exact_face_mass && fi && fj && m2 == 0) c2 = c4(ib, r2, p2, q2);
                T f = dmul(Gb[0][0], c4(p, qq, p2, q2));
                f += dmul(g_rq, -c2);
                f += dmul(g_qr, -c4(p, r, p2, q2));
                f += dmul(g_qq, c4(p, r, p2, r2));
                const double sij = si[m] * sj[m2];
                cplx dv, fv;
                if constexpr (TENSOR) {
                    dv = d * (sij * dscale);
                    fv = f * (sij * fscale);
                } else {
                    dv = cs * (d * sij * dscale);
                    fv = cm * (f * sij * fscale);
                }
                const int a0 = swp ? m2 : m, a1 = swp ? m : m2;
                accD[a0][a1] += dv;
                accF[a0][a1] += fv;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int m2 = 0; m2 < 2; m2++) {
            const int64_t idx = (int64_t)m * 2 * pat.nnz_ent + p0 + (int64_t)m2 * deg + k;
            E[idx] = accD[m][m2];
            B[idx] = accF[m][m2];
        }
}

// ------------------------------------------------------ isotropic fast path
// Phase 1: one thread per canonical tet computes its real element blocks with
// every index resolved at compile time (slots, vertex coincidences, moment
// coefficients).  Phase 2: the row-owner gather scales them by the tet's
// complex eps_r and 1/mu_r and writes each CSR value once.

__host__ __device__ constexpr int slot_v(int s, int k) {
    // canonical vertices (a, b, c) of local slot s; edges repeat b as c
    return (s == 0) ? (k == 0 ? 0 : 1) : (s == 1) ? (k == 0 ? 0 : 2) : (s == 2) ? (k == 0 ? 0 : 3)
         : (s == 3) ? (k == 0 ? 1 : 2) : (s == 4) ? (k == 0 ? 1 : 3) : (s == 5) ? (k == 0 ? 2 : 3)
         : (s == 6) ? k : (s == 7) ? (k == 0 ? 0 : k + 1) : (s == 8) ? (k == 2 ? 3 : k) : k + 1;
}
__host__ __device__ constexpr int edge_index(int u, int v) {  // u < v
    return u == 0 ? v - 1 : (u == 1 ? v + 1 : 5);
}
__host__ __device__ constexpr int fact(int n) { return n <= 1 ? 1 : n * fact(n - 1); }
__host__ __device__ constexpr int moment(int w, int x, int y, int z) {
    int r = 1;
    for (int k = 0; k < 4; k++) r *= fact((w == k) + (x == k) + (y == k) + (z == k));
    return r;
}
__host__ __device__ constexpr int upper_index(int a, int b) { return a * 10 - a * (a - 1) / 2 + (b - a); }

template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn &&fn) {
    if constexpr (B < E) {
        fn(std::integral_constant<int, B>{});
        static_for<B + 1, E>(fn);
    }
}

struct TetGeo {
    double G[4][4];  // g_a . g_b / (5040 * 6V)
    double H[6][6];  // X_k . X_l / (120 * (6V)^3), X = g_a x g_b over canonical edges
    double L[6];     // canonical edge lengths
};

// Real 2x2 (mode, mode') blocks of D and F for local slots (S, S2).  fq holds
// F with the reference's face-face coefficient, fx with the exact one.
template <int S, int S2>
__device__ __forceinline__ void iso_block(const TetGeo &g, double d[2][2], double fq[2][2], double fx[2][2]) {
    constexpr bool fi = S >= 6, fj = S2 >= 6;
    constexpr int ia = slot_v(S, 0), ib = slot_v(S, 1), ic = slot_v(S, 2);
    constexpr int ja = slot_v(S2, 0), jb = slot_v(S2, 1), jc = slot_v(S2, 2);
    // curl terms: cross index, paired vertex, coefficient per mode
    constexpr int ki[3] = {edge_index(ia, ib), edge_index(ia, fi ? ic : ib), fi ? edge_index(ib, ic) : 0};
    constexpr int kj[3] = {edge_index(ja, jb), edge_index(ja, fj ? jc : jb), fj ? edge_index(jb, jc) : 0};
    constexpr int tvi[3] = {fi ? ic : ia, ib, fi ? ia : ib};
    constexpr int tvj[3] = {fj ? jc : ja, jb, fj ? ja : jb};
    constexpr int cfi[2][3] = {{fi ? -1 : -3, fi ? -2 : 0, fi ? -1 : 0}, {fi ? -2 : 0, fi ? -1 : -3, fi ? 1 : 0}};
    constexpr int cfj[2][3] = {{fj ? -1 : -3, fj ? -2 : 0, fj ? -1 : 0}, {fj ? -2 : 0, fj ? -1 : -3, fj ? 1 : 0}};
    const double si[2] = {fi ? -g.L[edge_index(ia, ic)] : g.L[edge_index(ia, ib)], g.L[edge_index(ia, ib)]};
    const double sj[2] = {fj ? -g.L[edge_index(ja, jc)] : g.L[edge_index(ja, jb)], g.L[edge_index(ja, jb)]};
#pragma unroll
    for (int m = 0; m < 2; m++) {
#pragma unroll
        for (int m2 = 0; m2 < 2; m2++) {
            double dv = 0.0;
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 3; y++) {
                    const int c = cfi[m][x] * cfj[m2][y] * ((tvi[x] == tvj[y]) ? 2 : 1);
                    if (c != 0) dv += (double)c * g.H[ki[x]][kj[y]];
                }
            // basis (p, q, r): edge m1 (a,b,a), edge m2 (b,b,a), face m1 (b,c,a), face m2 (c,b,a)
            const int p = (m == 0) ? (fi ? ib : ia) : (fi ? ic : ib), q = (m == 0 && fi) ? ic : ib, r = ia;
            const int p2 = (m2 == 0) ? (fj ? jb : ja) : (fj ? jc : jb), q2 = (m2 == 0 && fj) ? jc : jb, r2 = ja;
            const double base = moment(p, q, p2, q2) * g.G[r][r2] - moment(p, r, p2, q2) * g.G[q][r2] +
                                moment(p, r, p2, r2) * g.G[q][q2];
            const int c_exact = moment(p, q, p2, r2);
            const int c_ref = (fi && fj && m2 == 0) ? moment(ib, r2, p2, q2) : c_exact;
            const double sij = si[m] * sj[m2];
            d[m][m2] = sij * dv;
            fq[m][m2] = sij * (base - c_ref * g.G[r][q2]);
            fx[m][m2] = sij * (base - c_exact * g.G[r][q2]);
        }
    }
}

__global__ void __launch_bounds__(64) k_tet_blocks(int64_t nt, const double *__restrict__ nodes,
                                                   const int32_t *__restrict__ tets, int exact_face_mass,
                                                   double *__restrict__ rec) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int4 tv = reinterpret_cast<const int4 *>(tets)[t];
    const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
    V3 P[4];
#pragma unroll
    for (int i = 0; i < 4; i++) P[i] = {nodes[3 * (int64_t)vid[i]], nodes[3 * (int64_t)vid[i] + 1], nodes[3 * (int64_t)vid[i] + 2]};
    const V3 e1 = sub(P[1], P[0]), e2 = sub(P[2], P[0]), e3 = sub(P[3], P[0]);
    V3 gr[4];
    gr[1] = crs(e2, e3);
    gr[2] = crs(e3, e1);
    gr[3] = crs(e1, e2);
    gr[0] = {-(gr[1].x + gr[2].x + gr[3].x), -(gr[1].y + gr[2].y + gr[3].y), -(gr[1].z + gr[2].z + gr[3].z)};
    const double V6 = fabs(dot(e1, gr[1]));
    const double fscale = 1.0 / (5040.0 * V6);
    const double dscale = 1.0 / (120.0 * V6 * V6 * V6);
    TetGeo g;
    constexpr int EU[6] = {0, 0, 0, 1, 1, 2}, EV[6] = {1, 2, 3, 2, 3, 3};
    V3 X[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        X[k] = crs(gr[EU[k]], gr[EV[k]]);
        g.L[k] = nrm(sub(P[EV[k]], P[EU[k]]));
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = a; b < 4; b++) g.G[a][b] = g.G[b][a] = dot(gr[a], gr[b]) * fscale;
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int l = k; l < 6; l++) g.H[k][l] = g.H[l][k] = dot(X[k], X[l]) * dscale;
    double *out = rec + t * TET_BLOCK_DOUBLES;
    const bool exact = exact_face_mass != 0;
    static_for<0, 10>([&](auto si_) {
        constexpr int S = decltype(si_)::value;
        static_for<S, 10>([&](auto sj_) {
            constexpr int S2 = decltype(sj_)::value;
            double d[2][2], fq[2][2], fx[2][2];
            iso_block<S, S2>(g, d, fq, fx);
            constexpr int u = upper_index(S, S2);
            double2 *od = reinterpret_cast<double2 *>(out + 4 * u);
            double2 *of = reinterpret_cast<double2 *>(out + 4 * (N_UPPER + u));
            od[0] = make_double2(d[0][0], d[0][1]);
            od[1] = make_double2(d[1][0], d[1][1]);
            of[0] = make_double2(exact ? fx[0][0] : fq[0][0], exact ? fx[0][1] : fq[0][1]);
            of[1] = make_double2(exact ? fx[1][0] : fq[1][0], exact ? fx[1][1] : fq[1][1]);
        });
    });
    // lower face-face blocks of F (s > s' >= 6)
    static_for<7, 10>([&](auto si_) {
        constexpr int S = decltype(si_)::value;
        static_for<6, S>([&](auto sj_) {
            constexpr int S2 = decltype(sj_)::value;
            double d[2][2], fq[2][2], fx[2][2];
            iso_block<S, S2>(g, d, fq, fx);
            constexpr int x = (S - 6) * (S - 7) / 2 + (S2 - 6);
            double2 *of = reinterpret_cast<double2 *>(out + 4 * (2 * N_UPPER + x));
            of[0] = make_double2(exact ? fx[0][0] : fq[0][0], exact ? fx[0][1] : fq[0][1]);
            of[1] = make_double2(exact ? fx[1][0] : fq[1][0], exact ? fx[1][1] : fq[1][1]);
        });
    });
}

__global__ void __launch_bounds__(256) k_gather_iso(const emg_pattern pat, const double *__restrict__ rec,
                                                    const cplx *__restrict__ eps, const cplx *__restrict__ nu,
                                                    cplx *__restrict__ E, cplx *__restrict__ B) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= pat.nnz_ent) return;
    const int e = find_row(pat.ent_ptr, pat.blk_row0, k);
    const int32_t p0 = pat.ent_ptr[e];
    const int32_t deg = pat.ent_ptr[e + 1] - p0;
    cplx aD[2][2], aF[2][2];
    const int32_t q0 = pat.pair_ptr[k], q1 = pat.pair_ptr[k + 1];
    for (int32_t q = q0; q < q1; q++) {
        const uint32_t v = pat.inc[q];
        const int t = (int)(v >> 8);
        const int s = (int)((v >> 4) & 15u), s2 = (int)(v & 15u);
        const double *r = rec + (int64_t)t * TET_BLOCK_DOUBLES;
        const bool up = s <= s2;
        const int a = up ? s : s2, b = up ? s2 : s;
        const int u = upper_index(a, b);
        const double2 d0 = *reinterpret_cast<const double2 *>(r + 4 * u);
        const double2 d1 = *reinterpret_cast<const double2 *>(r + 4 * u + 2);
        const bool lowff = !up && s2 >= 6;   // stored lower face-face block of F
        const int fo = lowff ? 4 * (2 * N_UPPER + (s - 6) * (s - 7) / 2 + (s2 - 6)) : 4 * (N_UPPER + u);
        const double2 f0 = *reinterpret_cast<const double2 *>(r + fo);
        const double2 f1 = *reinterpret_cast<const double2 *>(r + fo + 2);
        const bool tf = !up && !lowff;       // F read from the transposed block
        const double D00 = d0.x, D11 = d1.y, D01 = up ? d0.y : d1.x, D10 = up ? d1.x : d0.y;
        const double F00 = f0.x, F11 = f1.y, F01 = tf ? f1.x : f0.y, F10 = tf ? f0.y : f1.x;
        const cplx ct = eps[t], nt_ = nu[t];
        aD[0][0] += nt_ * D00; aD[0][1] += nt_ * D01; aD[1][0] += nt_ * D10; aD[1][1] += nt_ * D11;
        aF[0][0] += ct * F00;  aF[0][1] += ct * F01;  aF[1][0] += ct * F10;  aF[1][1] += ct * F11;
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int m2 = 0; m2 < 2; m2++) {
            const int64_t idx = (int64_t)m * 2 * pat.nnz_ent + p0 + (int64_t)m2 * deg + k;
            E[idx] = aD[m][m2];
            B[idx] = aF[m][m2];
        }
}

// ------------------------------------------------------------ fused path
// Recipe of an ordered slot pair (s, s'), RECIPE_BYTES bytes:
//   [0..8]   H index (symmetric 6x6 over canonical edges) of block (x, y)
//   [9..17]  G index (symmetric 4x4 over vertices) of block (x, y)
//   [18..19] bit x*3+y: paired curl vertices coincide (integral weight 2)
//   [20..35] per (m, m') the moments of the 4 mass terms (exact)
//   [36..39] per (m, m') the second mass moment as the reference indexes it
//   [40..43] length index of i mode 0/1 and j mode 0/1
//   [44]     bit0 row is a face, bit1 column is a face
__host__ __device__ constexpr int sym4(int a, int b) { return a <= b ? a * 4 - a * (a - 1) / 2 + (b - a) : sym4(b, a); }
__host__ __device__ constexpr int sym6(int a, int b) { return a <= b ? a * 6 - a * (a - 1) / 2 + (b - a) : sym6(b, a); }

static void build_recipes(uint8_t *out) {
    for (int s = 0; s < 10; s++)
        for (int s2 = 0; s2 < 10; s2++) {
            uint8_t *R = out + (s * 10 + s2) * RECIPE_BYTES;
            for (int i = 0; i < RECIPE_BYTES; i++) R[i] = 0;
            const bool fi = s >= 6, fj = s2 >= 6;
            const int vi[3] = {slot_v(s, 0), slot_v(s, 1), slot_v(s, 2)};
            const int vj[3] = {slot_v(s2, 0), slot_v(s2, 1), slot_v(s2, 2)};
            const int ki[3] = {edge_index(vi[0], vi[1]), edge_index(vi[0], fi ? vi[2] : vi[1]), fi ? edge_index(vi[1], vi[2]) : 0};
            const int kj[3] = {edge_index(vj[0], vj[1]), edge_index(vj[0], fj ? vj[2] : vj[1]), fj ? edge_index(vj[1], vj[2]) : 0};
            const int tvi[3] = {fi ? vi[2] : vi[0], vi[1], fi ? vi[0] : vi[1]};
            const int tvj[3] = {fj ? vj[2] : vj[0], vj[1], fj ? vj[0] : vj[1]};
            int w = 0;
            for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) {
                    R[x * 3 + y] = (uint8_t)sym6(ki[x], kj[y]);
                    R[9 + x * 3 + y] = (uint8_t)sym4(vi[x], vj[y]);
                    if (tvi[x] == tvj[y]) w |= 1 << (x * 3 + y);
                }
            R[18] = (uint8_t)(w & 255);
            R[19] = (uint8_t)(w >> 8);
            for (int m = 0; m < 2; m++)
                for (int m2 = 0; m2 < 2; m2++) {
                    const int p = (m == 0) ? (fi ? vi[1] : vi[0]) : (fi ? vi[2] : vi[1]);
                    const int q = (m == 0 && fi) ? vi[2] : vi[1], r = vi[0];
                    const int p2 = (m2 == 0) ? (fj ? vj[1] : vj[0]) : (fj ? vj[2] : vj[1]);
                    const int q2 = (m2 == 0 && fj) ? vj[2] : vj[1], r2 = vj[0];
                    uint8_t *c = R + 20 + 4 * (m * 2 + m2);
                    c[0] = (uint8_t)moment(p, q, p2, q2);
                    c[1] = (uint8_t)moment(p, q, p2, r2);
                    c[2] = (uint8_t)moment(p, r, p2, q2);
                    c[3] = (uint8_t)moment(p, r, p2, r2);
                    R[36 + m * 2 + m2] = (fi && fj && m2 == 0) ? (uint8_t)moment(vi[1], r2, p2, q2) : c[1];
                }
            R[40] = (uint8_t)(fi ? edge_index(vi[0], vi[2]) : edge_index(vi[0], vi[1]));
            R[41] = (uint8_t)edge_index(vi[0], vi[1]);
            R[42] = (uint8_t)(fj ? edge_index(vj[0], vj[2]) : edge_index(vj[0], vj[1]));
            R[43] = (uint8_t)edge_index(vj[0], vj[1]);
            R[44] = (uint8_t)((fi ? 1 : 0) | (fj ? 2 : 0));
        }
}

// Distinct tets of every chunk, and chunk-local incidence codes.
__global__ void __launch_bounds__(CHUNK) k_chunk_tets(const emg_pattern pat, int32_t *__restrict__ chunk_tets,
                                                      int32_t *__restrict__ chunk_ntets, uint16_t *__restrict__ inc_local,
                                                      uint16_t *__restrict__ chunk_perm, int *__restrict__ overflow) {
    __shared__ uint32_t key[CHUNK_NMAX];
    __shared__ uint32_t uq[CHUNK_NMAX];
    __shared__ int hist[256], base[256];       // entries per key, and those with a smaller key
    __shared__ int wcnt[CHUNK / 64][256];      // entries per key in each wavefront
    __shared__ int nuniq;
    const int64_t c = blockIdx.x;
    const int tid = threadIdx.x;
    const int ch = pat.chunk;
    const int64_t k0 = c * ch, k1 = (k0 + ch < pat.nnz_ent) ? k0 + ch : pat.nnz_ent;
    const int32_t q0 = pat.pair_ptr[k0], q1 = pat.pair_ptr[k1];
    const int n = q1 - q0;
    // Lane order of the fused kernel: the chunk's entries sorted (stably) by
    // descending incidence count, so a wavefront's lanes loop a similar number
    // of times.  Position r of the chunk holds local entry chunk_perm[k0 + r].
    const int np = (int)(k1 - k0);
    const bool has = tid < np;
    int myk = 255;
    if (has) {
        const int cnt = pat.pair_ptr[k0 + tid + 1] - pat.pair_ptr[k0 + tid];
        myk = 255 - (cnt < 255 ? cnt : 255);
    }
    // Stable counting rank: the entries with a smaller key, the entries with my key in
    // earlier wavefronts, and those earlier in my wavefront.  A wavefront walks only the
    // few distinct keys it holds (one ballot each).
    for (int i = tid; i < 256; i += CHUNK) hist[i] = 0;
    for (int i = tid; i < (CHUNK / 64) * 256; i += CHUNK) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    int before = 0;
    unsigned long long rest = __ballot(has);
    while (rest) {
        const int leader = __ffsll(rest) - 1;
        const int k = __shfl(myk, leader);
        const unsigned long long m = __ballot(has && myk == k);
        if (has && myk == k) before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == leader) {
            wcnt[wave][k] = __popcll(m);
            atomicAdd(&hist[k], __popcll(m));
        }
        rest &= ~m;
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int i = 0; i < 256; i++) { base[i] = acc; acc += hist[i]; }
    }
    __syncthreads();
    if (has) {
        int rank = base[myk] + before;
        for (int w = 0; w < wave; w++) rank += wcnt[w][myk];
        if (rank < ch) chunk_perm[k0 + rank] = (uint16_t)tid;
    } else if (tid < ch) {
        chunk_perm[k0 + tid] = (uint16_t)tid;
    }
    if (n > CHUNK_NMAX * ch / CHUNK) {   // the fused kernel stages at most this many incidences
        if (tid == 0) { atomicOr(overflow, 1); chunk_ntets[c] = -1; }
        return;
    }
    for (int i = tid; i < CHUNK_NMAX; i += CHUNK) key[i] = (i < n) ? (pat.inc[q0 + i] >> 8) : 0xFFFFFFFFu;
    __syncthreads();
    for (int len = 2; len <= CHUNK_NMAX; len <<= 1)
        for (int stride = len >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < CHUNK_NMAX; i += CHUNK) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool asc = (i & len) == 0;
                    const uint32_t a = key[i], b = key[j];
                    if ((a > b) == asc) { key[i] = b; key[j] = a; }
                }
            }
            __syncthreads();
        }
    if (tid == 0) {
        int m = 0;
        for (int i = 0; i < n; i++)
            if (i == 0 || key[i] != key[i - 1]) uq[m++] = key[i];
        nuniq = m;
    }
    __syncthreads();
    const int m = nuniq;
    if (m > CHUNK_TMAX) {
        if (tid == 0) { atomicOr(overflow, 1); chunk_ntets[c] = -1; }
        return;
    }
    for (int i = tid; i < m; i += CHUNK) chunk_tets[c * CHUNK_TMAX + i] = (int32_t)uq[i];
    if (tid == 0) chunk_ntets[c] = m;
    for (int i = tid; i < n; i += CHUNK) {
        const uint32_t v = pat.inc[q0 + i];
        const uint32_t t = v >> 8;
        int lo = 0, hi = m - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (uq[mid] < t) lo = mid + 1; else hi = mid;
        }
        inc_local[q0 + i] = (uint16_t)((lo << 8) | (v & 255u));
    }
}

static constexpr int TD_FIELDS = 41;  // H(21) G(10) L(6) eps(2) nu(2)

__device__ __forceinline__ uint32_t rbyte(const uint32_t *w, int i) { return (w[i >> 2] >> ((i & 3) * 8)) & 255u; }

// Gram record of a tet (vertices P, materials ev, nv): field f at dst[f * S].
// Fields: H = (grad x grad) products of the 6 edges (0..20), G = grad.grad of
// the 4 vertices (21..30), the 6 edge lengths (31..36), eps (37, 38), nu (39, 40).
template <int S>
__device__ __forceinline__ void fused_tet_gram(const V3 (&P)[4], const V3 (&gr)[4], double V6, cplx ev, cplx nv,
                                               double *dst) {
    const double fscale = 1.0 / (5040.0 * V6);
    const double dscale = 1.0 / (120.0 * V6 * V6 * V6);
    constexpr int EU[6] = {0, 0, 0, 1, 1, 2}, EV[6] = {1, 2, 3, 2, 3, 3};
#pragma unroll
    for (int k = 0; k < 6; k++) dst[(31 + k) * S] = nrm(sub(P[EV[k]], P[EU[k]]));
    // H_kl = (g_a x g_b).(g_c x g_d) = G_ac G_bd - G_ad G_bc (Binet-Cauchy) with
    // G = the Gram matrix of the vertex gradients: no edge cross products held.
    double G[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = a; b < 4; b++) {
            G[a][b] = dot(gr[a], gr[b]);
            G[b][a] = G[a][b];
            dst[(21 + sym4(a, b)) * S] = G[a][b] * fscale;
        }
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int l = k; l < 6; l++) {
            const int a = EU[k], b = EV[k], c = EU[l], d = EV[l];
            dst[sym6(k, l) * S] = (G[a][c] * G[b][d] - G[a][d] * G[b][c]) * dscale;
        }
    dst[37 * S] = ev.re;
    dst[38 * S] = ev.im;
    dst[39 * S] = nv.re;
    dst[40 * S] = nv.im;
}

// One incidence v = (local tet << 8) | (s << 4) | s' of an adjacency entry:
// adds nu_t * D_t and eps_t * F_t of the slot pair (s, s') (2x2 each) into aD, aF.
template <int TS>
__device__ __forceinline__ void fused_incidence(const uint32_t *rcp, const double *td, uint32_t v,
                                                int exact_face_mass, cplx (&aD)[2][2], cplx (&aF)[2][2]) {
    const int tl = (int)(v >> 8);
    const int sp = (int)(((v >> 4) & 15u) * 10 + (v & 15u));
    uint32_t w[12];
    const uint4 *R4 = reinterpret_cast<const uint4 *>(rcp + sp * (RECIPE_BYTES / 4));
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const uint4 u = R4[i];
        w[4 * i] = u.x; w[4 * i + 1] = u.y; w[4 * i + 2] = u.z; w[4 * i + 3] = u.w;
    }
    const double *T = td + tl;
    const uint32_t wm = rbyte(w, 18) | (rbyte(w, 19) << 8);
    double h[3][3], g[3][3];
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const double hv = T[rbyte(w, x * 3 + y) * TS];
            h[x][y] = ((wm >> (x * 3 + y)) & 1u) ? 2.0 * hv : hv;
            g[x][y] = T[(21 + rbyte(w, 9 + x * 3 + y)) * TS];
        }
    const uint32_t fl = rbyte(w, 44);
    const bool fi = fl & 1u, fj = fl & 2u;
    const double cfi[2][3] = {{fi ? -1.0 : -3.0, fi ? -2.0 : 0.0, fi ? -1.0 : 0.0},
                              {fi ? -2.0 : 0.0, fi ? -1.0 : -3.0, fi ? 1.0 : 0.0}};
    const double cfj[2][3] = {{fj ? -1.0 : -3.0, fj ? -2.0 : 0.0, fj ? -1.0 : 0.0},
                              {fj ? -2.0 : 0.0, fj ? -1.0 : -3.0, fj ? 1.0 : 0.0}};
    const double si[2] = {(fi ? -1.0 : 1.0) * T[(31 + rbyte(w, 40)) * TS], T[(31 + rbyte(w, 41)) * TS]};
    const double sj[2] = {(fj ? -1.0 : 1.0) * T[(31 + rbyte(w, 42)) * TS], T[(31 + rbyte(w, 43)) * TS]};
    const cplx ct(T[37 * TS], T[38 * TS]), nt_(T[39 * TS], T[40 * TS]);
    double hc[3][2];
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int m2 = 0; m2 < 2; m2++)
            hc[x][m2] = h[x][0] * cfj[m2][0] + h[x][1] * cfj[m2][1] + h[x][2] * cfj[m2][2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const bool qc = (m == 0) && fi;
        const double g_q0 = qc ? g[2][0] : g[1][0];
#pragma unroll
        for (int m2 = 0; m2 < 2; m2++) {
            const bool qc2 = (m2 == 0) && fj;
            const double g_0q = qc2 ? g[0][2] : g[0][1];
            const double g_qq = qc ? (qc2 ? g[2][2] : g[2][1]) : (qc2 ? g[1][2] : g[1][1]);
            const int o = 20 + 4 * (m * 2 + m2);
            const double c1 = (double)(exact_face_mass ? rbyte(w, o + 1) : rbyte(w, 36 + m * 2 + m2));
            const double f = (double)rbyte(w, o) * g[0][0] - c1 * g_0q - (double)rbyte(w, o + 2) * g_q0 +
                             (double)rbyte(w, o + 3) * g_qq;
            const double d = cfi[m][0] * hc[0][m2] + cfi[m][1] * hc[1][m2] + cfi[m][2] * hc[2][m2];
            const double sij = si[m] * sj[m2];
            aD[m][m2] += nt_ * (sij * d);
            aF[m][m2] += ct * (sij * f);
        }
    }
}

// Gram record of every tet (REC doubles each, fields as fused_tet_gram, the
// rest padding): the first step of the numeric assembly, so each tet's geometry
// is computed once.  48 doubles = 384 bytes = three whole 128-byte lines.
static constexpr int REC = 48;
// The reference's vertex-gradient cofactors and volume from absolute coordinates
// (fem/mth/optimized.py:343-369), term for term and left to right with no fused
// multiply-adds, so the cancellation they suffer away from the origin is the
// reference's own.  Vertices in the reference's (the mesh's) tet order.
__device__ __attribute__((noinline)) void reference_bcd(const V3 (&Q)[4], double *bb, double *cc, double *dd,
                                                        double *V) {
#pragma clang fp contract(off)
    const double x1 = Q[0].x, x2 = Q[1].x, x3 = Q[2].x, x4 = Q[3].x;
    const double y1 = Q[0].y, y2 = Q[1].y, y3 = Q[2].y, y4 = Q[3].y;
    const double z1 = Q[0].z, z2 = Q[1].z, z3 = Q[2].z, z4 = Q[3].z;
    *V = fabs(-x1*y2*z3/6 + x1*y2*z4/6 + x1*y3*z2/6 - x1*y3*z4/6 - x1*y4*z2/6 + x1*y4*z3/6
              + x2*y1*z3/6 - x2*y1*z4/6 - x2*y3*z1/6 + x2*y3*z4/6 + x2*y4*z1/6 - x2*y4*z3/6
              - x3*y1*z2/6 + x3*y1*z4/6 + x3*y2*z1/6 - x3*y2*z4/6 - x3*y4*z1/6 + x3*y4*z2/6
              + x4*y1*z2/6 - x4*y1*z3/6 - x4*y2*z1/6 + x4*y2*z3/6 + x4*y3*z1/6 - x4*y3*z2/6);
    bb[0] = -y2*z3 + y2*z4 + y3*z2 - y3*z4 - y4*z2 + y4*z3;
    bb[1] = y1*z3 - y1*z4 - y3*z1 + y3*z4 + y4*z1 - y4*z3;
    bb[2] = -y1*z2 + y1*z4 + y2*z1 - y2*z4 - y4*z1 + y4*z2;
    bb[3] = y1*z2 - y1*z3 - y2*z1 + y2*z3 + y3*z1 - y3*z2;
    cc[0] = x2*z3 - x2*z4 - x3*z2 + x3*z4 + x4*z2 - x4*z3;
    cc[1] = -x1*z3 + x1*z4 + x3*z1 - x3*z4 - x4*z1 + x4*z3;
    cc[2] = x1*z2 - x1*z4 - x2*z1 + x2*z4 + x4*z1 - x4*z2;
    cc[3] = -x1*z2 + x1*z3 + x2*z1 - x2*z3 - x3*z1 + x3*z2;
    dd[0] = -x2*y3 + x2*y4 + x3*y2 - x3*y4 - x4*y2 + x4*y3;
    dd[1] = x1*y3 - x1*y4 - x3*y1 + x3*y4 + x4*y1 - x4*y3;
    dd[2] = -x1*y2 + x1*y4 + x2*y1 - x2*y4 - x4*y1 + x4*y2;
    dd[3] = x1*y2 - x1*y3 - x2*y1 + x2*y3 + x3*y1 - x3*y2;
}

// REF_GEOM = false: tets are canonical (vertices sorted by id); the vertex gradients
// (times 6V) are cross products of edge vectors relative to vertex 0.
// REF_GEOM = true: tets are in the reference's vertex order; the gradients and volume
// come from reference_bcd and are re-indexed to the canonical (sorted) order.
// blockIdx.y selects the mesh of a batch sharing the tets (strides in doubles / complex
// values between consecutive meshes' nodes, materials and records; 0 for one mesh).
template <bool REF_GEOM>
__global__ void __launch_bounds__(256) k_tet_gram(int64_t nt, const double *__restrict__ nodes,
                                                  const int32_t *__restrict__ tets, const cplx *__restrict__ eps,
                                                  const cplx *__restrict__ nu, double *__restrict__ rec,
                                                  int64_t node_stride, int64_t mat_stride, int64_t rec_stride) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    nodes += blockIdx.y * node_stride;
    eps += blockIdx.y * mat_stride;
    nu += blockIdx.y * mat_stride;
    rec += blockIdx.y * rec_stride;
    const int4 tv = reinterpret_cast<const int4 *>(tets)[t];
    int vid[4] = {tv.x, tv.y, tv.z, tv.w};
    V3 P[4], gr[4];
    double V6;
    if (REF_GEOM) {
        V3 Q[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            Q[j] = {nodes[3 * (int64_t)vid[j]], nodes[3 * (int64_t)vid[j] + 1], nodes[3 * (int64_t)vid[j] + 2]};
        double bb[4], cc[4], dd[4], V;
        reference_bcd(Q, bb, cc, dd, &V);
        V6 = 6.0 * V;
        int ord[4] = {0, 1, 2, 3};   // ord[k]: reference slot of the k-th smallest vertex id
#pragma unroll
        for (int i = 1; i < 4; i++)
#pragma unroll
            for (int j = i; j > 0; j--)
                if (vid[ord[j]] < vid[ord[j - 1]]) {
                    const int tmp = ord[j];
                    ord[j] = ord[j - 1];
                    ord[j - 1] = tmp;
                }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int o = ord[0];
#pragma unroll
            for (int j = 1; j < 4; j++) o = (k == j) ? ord[j] : o;
            double b = bb[0], c = cc[0], d = dd[0];
            V3 q = Q[0];
#pragma unroll
            for (int j = 1; j < 4; j++)
                if (o == j) { b = bb[j]; c = cc[j]; d = dd[j]; q = Q[j]; }
            gr[k] = {b, c, d};
            P[k] = q;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
            P[j] = {nodes[3 * (int64_t)vid[j]], nodes[3 * (int64_t)vid[j] + 1], nodes[3 * (int64_t)vid[j] + 2]};
        const V3 e1 = sub(P[1], P[0]), e2 = sub(P[2], P[0]), e3 = sub(P[3], P[0]);
        gr[1] = crs(e2, e3);
        gr[2] = crs(e3, e1);
        gr[3] = crs(e1, e2);
        gr[0] = {-(gr[1].x + gr[2].x + gr[3].x), -(gr[1].y + gr[2].y + gr[3].y), -(gr[1].z + gr[2].z + gr[3].z)};
        V6 = fabs(dot(e1, gr[1]));
    }
    fused_tet_gram<1>(P, gr, V6, eps[t], nu[t], rec + t * REC);
}

// Numeric assembly, one workgroup of CH lanes per chunk of CH consecutive
// adjacency entries (grid-stride over chunks).  Per chunk:
//  * its index data (entry offsets into the incidences, the offsets of the rows
//    it spans, the incidence codes, the lane order) and the Gram records of its
//    distinct tets (k_tet_gram) are loaded into registers while the previous
//    chunk writes its output (the tet ids one step earlier, while it runs its
//    incidence loop), and put into LDS when the chunk starts: no global load is
//    waited for on the chunk's critical path, and the incidence loop reads only LDS;
//  * lane r accumulates entry chunk_perm[r] (entries by descending incidence
//    count, so the lanes of a wavefront loop a similar number of times);
//  * the 2x2 blocks go back through LDS (reusing the Gram area) and are written
//    in entry order, coalesced.
// The wavefront taking the heaviest entries rotates with the chunk, so no SIMD
// always gets them.  Launch bound: 4 waves per SIMD (3 for the widest LDS layout).
template <int TS, int CH>
__global__ void __launch_bounds__(CH, TS <= 96 ? 4 : 3)
    k_assemble_staged(const emg_pattern pat, const double *__restrict__ rec, int exact_face_mass,
                      cplx *__restrict__ E, cplx *__restrict__ B, int64_t rec_stride, int64_t out_stride) {
    // blockIdx.y: the mesh of a batch (its records and output values at these strides)
    rec += blockIdx.y * rec_stride;
    E += blockIdx.y * out_stride;
    B += blockIdx.y * out_stride;
    constexpr int TSP = TS + 1;             // odd field stride of the Gram data: fewer LDS bank conflicts
    constexpr int TD = TD_FIELDS * TSP;     // doubles of Gram data
    constexpr int SD = 2 * 4 * CH;          // doubles of one staged matrix: 4 blocks x CH complex
    constexpr int NINC = CHUNK_NMAX * CH / CHUNK;
    constexpr int NID = NINC / 2 + 1;       // incidence codes as 32-bit words (first one may be half used)
    constexpr int RI = (NID + CH - 1) / CH; // words per lane
    constexpr int NW = CH / 64;
    constexpr int SPR = CH / 8;             // tet records per round: 8 lanes per record, 7 of them load 6 doubles
    constexpr int RR = (TS + SPR - 1) / SPR;
    static_assert(7 * 6 >= TD_FIELDS && 7 * 6 <= REC, "7 lanes of 6 doubles cover a record");
    static_assert(RR * SPR <= CHUNK_TMAX, "tet_ids reads slots below CHUNK_TMAX only");
    __shared__ uint32_t rcp[100 * RECIPE_BYTES / 4];
    __shared__ __align__(16) double td[TD > SD ? TD : SD];
    __shared__ int32_t s_pp[CH + 1];        // entry -> first incidence (absolute)
    __shared__ int32_t s_row[CH + 2];       // ent_ptr of the rows the chunk spans
    __shared__ uint32_t s_inc[NID];         // the chunk's incidence codes (16-bit), from an even index
    __shared__ uint16_t s_perm[CH];         // lane order
    __shared__ int2 s_out[CH];              // per entry: (row start, row length)
    const int tid = threadIdx.x;
    const int rs = tid >> 3, rsub = tid & 7;   // record slot within a round, 6-double part of the record
    for (int i = tid; i < 100 * RECIPE_BYTES / 4; i += CH) rcp[i] = reinterpret_cast<const uint32_t *>(pat.recipes)[i];
    cplx *stg = reinterpret_cast<cplx *>(td);
    // registers holding the next chunk's data; every one is assigned on every
    // path, so none stays live through the incidence loop
    int32_t f_q0 = 0, f_ninc = 0, f_nr = 0, f_nt = 0;              // uniform
    int32_t g_nt = 0;                                               // uniform
    int32_t g_u = 0;   // lanes 0-3 of each wave: the next chunk's first and end incidence, first and end row
    int32_t f_pp[2], f_row[2], g_t[RR];
    uint32_t f_inc[RI];
    int f_perm = 0;
    double2 f_rec[RR][3];
    // The chunk's tet ids and its uniform offsets, issued before the incidence loop.
    // Unconditional loads (the chunk index clamped, every slot below CHUNK_TMAX in
    // bounds): nothing here depends on a loaded value, so nothing waits for these
    // loads before prefetch uses them; slots at or past the tet count are never used.
    auto tet_ids = [&](int64_t cc) {
        const int64_t cl = cc < pat.nchunk ? cc : pat.nchunk - 1;
        const int64_t kl = cl * CH;
        g_nt = pat.chunk_ntets[cl];
        // one gather, lane by lane, so the values stay in vector registers until
        // prefetch reads them (a uniform load would be moved to a scalar register,
        // which waits for it at once)
        const int32_t *src = (tid & 3) == 0   ? pat.pair_ptr + kl
                             : (tid & 3) == 1 ? pat.pair_ptr + (pat.nnz_ent - kl < CH ? pat.nnz_ent : kl + CH)
                             : (tid & 3) == 2 ? pat.chunk_row0 + cl
                                              : pat.chunk_row0 + cl + 1;
        g_u = *src;
#pragma unroll
        for (int h = 0; h < RR; h++) g_t[h] = pat.chunk_tets[cl * CHUNK_TMAX + h * SPR + rs];
    };
    auto prefetch = [&](int64_t cc) {
        f_q0 = 0; f_ninc = 0; f_nr = 0; f_perm = 0; f_nt = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) { f_pp[h] = 0; f_row[h] = 0; }
#pragma unroll
        for (int h = 0; h < RI; h++) f_inc[h] = 0;
#pragma unroll
        for (int h = 0; h < RR; h++)
#pragma unroll
            for (int q = 0; q < 3; q++) f_rec[h][q] = make_double2(0.0, 0.0);
        if (cc >= pat.nchunk) return;
        f_nt = g_nt;
        const int64_t kk = cc * CH;
        const int npp = (int)(pat.nnz_ent - kk < CH ? pat.nnz_ent - kk : CH);
        f_q0 = __builtin_amdgcn_readlane(g_u, 0);
        f_ninc = __builtin_amdgcn_readlane(g_u, 1) - f_q0;
        const int32_t r0 = __builtin_amdgcn_readlane(g_u, 2);
        f_nr = __builtin_amdgcn_readlane(g_u, 3) - r0 + 2;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = tid + h * CH;
            if (i <= npp) f_pp[h] = pat.pair_ptr[kk + i];
            if (i < f_nr) f_row[h] = pat.ent_ptr[r0 + i];
        }
        const int64_t w0 = f_q0 >> 1;
        const int nwords = (int)(((int64_t)f_q0 + f_ninc + 1) / 2 - w0);
#pragma unroll
        for (int h = 0; h < RI; h++) {
            const int i = tid + h * CH;
            if (i < nwords) f_inc[h] = reinterpret_cast<const uint32_t *>(pat.inc_local)[w0 + i];
        }
        f_perm = pat.chunk_perm[kk + tid];
#pragma unroll
        for (int h = 0; h < RR; h++)
            if (h * SPR + rs < f_nt && rsub < 7) {
                const double2 *r = reinterpret_cast<const double2 *>(rec + (int64_t)g_t[h] * REC + rsub * 6);
#pragma unroll
                for (int q = 0; q < 3; q++) f_rec[h][q] = r[q];
            }
    };
    // the prefetched registers of chunk cc into LDS; the chunk's row count and first
    // incidence index stay in cur_nr / cur_qb
    int32_t cur_nr = 0, cur_qb = 0;
    auto stage = [&](int64_t cc) {
        const int64_t kk = cc * CH;
        const int npp = cc < pat.nchunk ? (int)(pat.nnz_ent - kk < CH ? pat.nnz_ent - kk : CH) : 0;
        const int32_t ioff = f_q0 & 1, nr = f_nr, nT = f_nt;
#pragma unroll
        for (int h = 0; h < RR; h++) {
            const int slot = h * SPR + rs;
            if (slot < nT && rsub < 7) {
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    const int f = rsub * 6 + q;
                    if (f < TD_FIELDS) td[f * TSP + slot] = (q & 1) ? f_rec[h][q >> 1].y : f_rec[h][q >> 1].x;
                }
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = tid + h * CH;
            if (i <= npp) s_pp[i] = f_pp[h];
            if (i < nr) s_row[i] = f_row[h];
        }
        {
            const int nwords = (ioff + f_ninc + 1) / 2;
#pragma unroll
            for (int h = 0; h < RI; h++) {
                const int i = tid + h * CH;
                if (i < nwords) s_inc[i] = f_inc[h];
            }
        }
        s_perm[tid] = (uint16_t)f_perm;
        cur_nr = nr;
        cur_qb = f_q0 - ioff;   // incidence index of s_inc's first 16-bit half
    };
    tet_ids(blockIdx.x);
    prefetch(blockIdx.x);
    stage(blockIdx.x);
    for (int64_t c = blockIdx.x; c < pat.nchunk; c += gridDim.x) {
        __syncthreads();   // the chunk's index data and Gram records are in LDS
        const int64_t k0 = c * CH;
        const int np = (int)(pat.nnz_ent - k0 < CH ? pat.nnz_ent - k0 : CH);
        const int32_t nr = cur_nr, qb = cur_qb;
        tet_ids(c + gridDim.x);
        const int hw = (int)((c + NW / 2) % NW);   // wave hw takes the first (heaviest) 64 lane-order positions
        const int lk = s_perm[(((tid >> 6) - hw + NW) % NW) * 64 + (tid & 63)];
        const bool act = lk < np;
        cplx aD[2][2], aF[2][2];
        int32_t p0 = 0, deg = 0;
        if (act) {
            const int64_t k = k0 + lk;
            int lo = 0, hi = nr - 2;   // the row of k: last i with s_row[i] <= k
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_row[mid] <= k) lo = mid; else hi = mid - 1;
            }
            p0 = s_row[lo];
            deg = s_row[lo + 1] - p0;
            const int q1 = s_pp[lk + 1] - qb;
            const uint16_t *inc16 = reinterpret_cast<const uint16_t *>(s_inc);
#pragma unroll 1
            for (int q = s_pp[lk] - qb; q < q1; q++) fused_incidence<TSP>(rcp, td, inc16[q], exact_face_mass, aD, aF);
        }
        __syncthreads();   // Gram data no longer read: td becomes the output stage
        prefetch(c + gridDim.x);
        if (act) {
#pragma unroll
            for (int m = 0; m < 4; m++) stg[m * CH + lk] = aD[m >> 1][m & 1];
            s_out[lk] = make_int2(p0, deg);
        }
        __syncthreads();
        // the blocks in entry order into registers, through the stage
        int2 o = make_int2(0, 0);
        cplx ev[4], bv[4];
        if (tid < np) {
            o = s_out[tid];
#pragma unroll
            for (int m = 0; m < 4; m++) ev[m] = stg[m * CH + tid];
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int m = 0; m < 4; m++) stg[m * CH + lk] = aF[m >> 1][m & 1];
        }
        __syncthreads();
        if (tid < np) {
#pragma unroll
            for (int m = 0; m < 4; m++) bv[m] = stg[m * CH + tid];
        }
        __syncthreads();   // LDS free: the next chunk's data goes in before this chunk's stores
        // are issued, so waiting for the prefetched loads does not wait for the stores
        // (loads and stores share one counter); the stores drain during the next chunk
        stage(c + gridDim.x);
        if (tid < np) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int64_t a = (int64_t)(m >> 1) * 2 * pat.nnz_ent + o.x + (int64_t)(m & 1) * o.y + k0 + tid;
                E[a] = ev[m];
                B[a] = bv[m];
            }
        }
    }
}

template <int CH>
static void launch_staged(const emg_pattern *pat, dim3 grid, hipStream_t st, const double *rec, int efm, cplx *oE,
                          cplx *oB, int64_t rec_stride = 0, int64_t out_stride = 0) {
    const int tm = pat->chunk_tmax;
    if (tm <= 32)
        k_assemble_staged<32, CH><<<grid, CH, 0, st>>>(*pat, rec, efm, oE, oB, rec_stride, out_stride);
    else if (tm <= 48)
        k_assemble_staged<48, CH><<<grid, CH, 0, st>>>(*pat, rec, efm, oE, oB, rec_stride, out_stride);
    else if (tm <= 64)
        k_assemble_staged<64, CH><<<grid, CH, 0, st>>>(*pat, rec, efm, oE, oB, rec_stride, out_stride);
    else if (tm <= 96)
        k_assemble_staged<96, CH><<<grid, CH, 0, st>>>(*pat, rec, efm, oE, oB, rec_stride, out_stride);
    else
        k_assemble_staged<CHUNK_TMAX, CH><<<grid, CH, 0, st>>>(*pat, rec, efm, oE, oB, rec_stride, out_stride);
}

static int g_num_cus = 0;

template <bool TENSOR>
static int assemble(const emg_pattern *pat, const double *d_nodes, const int32_t *d_tets, const double *d_m,
                    const double *d_s, int flags, double *d_E, double *d_B, void *stream) {
    EMG_CHECK(pat && d_nodes && d_tets && d_m && d_s && d_E && d_B, "emg_assemble: null argument");
    EMG_CHECK(!(flags & EMG_REFERENCE_GEOMETRY), "emg_assemble_tensor: EMG_REFERENCE_GEOMETRY is isotropic-only");
    hipStream_t st = (hipStream_t)stream;
    unsigned grid = (unsigned)((pat->nnz_ent + ABLK - 1) / ABLK);
    k_assemble<TENSOR><<<grid, ABLK, 0, st>>>(*pat, d_nodes, d_tets, reinterpret_cast<const cplx *>(d_m),
                                              reinterpret_cast<const cplx *>(d_s), (flags & EMG_EXACT_FACE_MASS) ? 1 : 0,
                                              reinterpret_cast<cplx *>(d_E), reinterpret_cast<cplx *>(d_B));
    EMG_LAUNCH_CHECK();
    return 0;
}

// Autotuner of the fused launch's grid: workgroups per CU.  A grid-stride workgroup owns
// whole chunks, so every grid gives the same values bit for bit; the grid only decides how
// evenly the chunks' uneven costs spread over the resident workgroups.  The first fused
// assembly on a pattern runs each candidate once untimed and twice between HIP events (so it
// synchronises the stream once) and keeps the fastest for the pattern's life;
// EMG_ASM_WG_PER_CU (one of the candidates) overrides the search.
static const int ASM_WG[5] = {4, 8, 16, 32, 64};

template <typename RUN>
static int tune_assembly(emg_pattern *mp, RUN &run, hipStream_t st) {
    if (const char *env = getenv("EMG_ASM_WG_PER_CU")) {
        const int v = atoi(env);
        mp->wg_per_cu = 32;
        for (int w : ASM_WG)
            if (v == w) mp->wg_per_cu = v;
        return 0;
    }
    hipEvent_t ev[2];
    EMG_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        emg::set_error("tune_assembly: hipEventCreate failed");
        return -2;
    }
    hipError_t e = hipSuccess;
    float best = 3.4e38f;
    int bw = 32;
    for (int i = 0; i < 5 && e == hipSuccess; i++) {
        run(ASM_WG[i]);
        e = hipEventRecord(ev[0], st);
        for (int r = 0; r < 2 && e == hipSuccess; r++) run(ASM_WG[i]);
        if (e == hipSuccess) e = hipEventRecord(ev[1], st);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (e == hipSuccess) e = hipGetLastError();
        ms /= 2.f;
        mp->wg_ms[i] = ms;
        if (e == hipSuccess && ms < best) { best = ms; bw = ASM_WG[i]; }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    EMG_HIP(e);
    mp->wg_per_cu = bw;
    return 0;
}

extern "C" int emg_pattern_assembly_tuning(const emg_pattern *pat, int *h_wg_per_cu, float *h_ms) {
    EMG_CHECK(pat && h_wg_per_cu && h_ms, "emg_pattern_assembly_tuning: null argument");
    *h_wg_per_cu = pat->wg_per_cu;
    for (int i = 0; i < 5; i++) h_ms[i] = pat->wg_ms[i];
    return 0;
}

extern "C" int emg_assemble_iso(const emg_pattern *pat, const double *d_nodes, const int32_t *d_tets, const double *d_eps,
                                const double *d_nu, int flags, double *d_E, double *d_B, void *stream) {
    EMG_CHECK(pat && d_nodes && d_tets && d_eps && d_nu && d_E && d_B, "emg_assemble_iso: null argument");
    hipStream_t st = (hipStream_t)stream;
    emg_pattern *mp = const_cast<emg_pattern *>(pat);
    if (pat->fused_ok && !(flags & EMG_FORCE_TWO_PHASE)) {
        if (g_num_cus == 0) {
            int dev = 0;
            EMG_HIP(hipGetDevice(&dev));
            EMG_HIP(hipDeviceGetAttribute(&g_num_cus, hipDeviceAttributeMultiprocessorCount, dev));
        }
        const int efm = (flags & EMG_EXACT_FACE_MASS) ? 1 : 0;
        const cplx *ce = reinterpret_cast<const cplx *>(d_eps), *cn = reinterpret_cast<const cplx *>(d_nu);
        cplx *oE = reinterpret_cast<cplx *>(d_E), *oB = reinterpret_cast<cplx *>(d_B);
        if (!mp->tet_rec) EMG_HIP(hipMalloc(&mp->tet_rec, (size_t)pat->nt * REC * sizeof(double)));
        auto run = [&](int wg) {
            if (flags & EMG_REFERENCE_GEOMETRY)
                k_tet_gram<true><<<emg::grid_for(pat->nt, 256), 256, 0, st>>>(pat->nt, d_nodes, d_tets, ce, cn,
                                                                             mp->tet_rec, 0, 0, 0);
            else
                k_tet_gram<false><<<emg::grid_for(pat->nt, 256), 256, 0, st>>>(pat->nt, d_nodes, d_tets, ce, cn,
                                                                              mp->tet_rec, 0, 0, 0);
            const int64_t grid = std::min<int64_t>(pat->nchunk, (int64_t)g_num_cus * wg);
            launch_staged<CHUNK>(pat, dim3((unsigned)grid), st, mp->tet_rec, efm, oE, oB);
        };
        if (mp->wg_per_cu == 0)
            if (int rc = tune_assembly(mp, run, st)) return rc;
        run(mp->wg_per_cu);
        EMG_LAUNCH_CHECK();
        return 0;
    }
    EMG_CHECK(!(flags & EMG_REFERENCE_GEOMETRY),
              "emg_assemble_iso: EMG_REFERENCE_GEOMETRY needs the fused path (not EMG_FORCE_TWO_PHASE)");
    if (!mp->tet_blocks) EMG_HIP(hipMalloc(&mp->tet_blocks, (size_t)pat->nt * TET_BLOCK_DOUBLES * sizeof(double)));
    k_tet_blocks<<<emg::grid_for(pat->nt, 64), 64, 0, st>>>(pat->nt, d_nodes, d_tets,
                                                            (flags & EMG_EXACT_FACE_MASS) ? 1 : 0, mp->tet_blocks);
    k_gather_iso<<<emg::grid_for(pat->nnz_ent, 256), 256, 0, st>>>(*pat, mp->tet_blocks,
                                                                   reinterpret_cast<const cplx *>(d_eps),
                                                                   reinterpret_cast<const cplx *>(d_nu),
                                                                   reinterpret_cast<cplx *>(d_E),
                                                                   reinterpret_cast<cplx *>(d_B));
    EMG_LAUNCH_CHECK();
    return 0;
}

// nbatch meshes with the pattern's tets and their own node coordinates and materials, in
// one pair of launches: mesh b's nodes at d_nodes + b * 3 nv (nv x 3 each), eps and nu at
// + b * nt, E and B values at + b * nnz.  The batch's records live in a buffer of the
// pattern that grows to the largest batch seen.
extern "C" int emg_assemble_iso_batched(const emg_pattern *pat, int nbatch, int64_t nv, const double *d_nodes,
                                        const int32_t *d_tets, const double *d_eps, const double *d_nu, int flags,
                                        double *d_E, double *d_B, void *stream) {
    EMG_CHECK(pat && d_nodes && d_tets && d_eps && d_nu && d_E && d_B, "emg_assemble_iso_batched: null argument");
    EMG_CHECK(nbatch >= 1 && nbatch <= 65535, "emg_assemble_iso_batched: nbatch must be in [1, 65535]");
    EMG_CHECK(nv > 0, "emg_assemble_iso_batched: nv must be positive");
    EMG_CHECK(pat->fused_ok, "emg_assemble_iso_batched: the pattern has no fused-path chunk tables");
    EMG_CHECK(!(flags & ~(EMG_EXACT_FACE_MASS | EMG_REFERENCE_GEOMETRY)), "emg_assemble_iso_batched: unsupported flag");
    hipStream_t st = (hipStream_t)stream;
    emg_pattern *mp = const_cast<emg_pattern *>(pat);
    if (g_num_cus == 0) {
        int dev = 0;
        EMG_HIP(hipGetDevice(&dev));
        EMG_HIP(hipDeviceGetAttribute(&g_num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    if (mp->rec_batch_cap < nbatch) {
        if (mp->tet_rec_batch) EMG_HIP(hipFree(mp->tet_rec_batch));
        mp->tet_rec_batch = nullptr;
        mp->rec_batch_cap = 0;
        EMG_HIP(hipMalloc(&mp->tet_rec_batch, (size_t)nbatch * pat->nt * REC * sizeof(double)));
        mp->rec_batch_cap = nbatch;
    }
    // every mesh gets the single-mesh grid (the tuned workgroups per CU, else 32): the batch is the
    // single launches back to back without their gaps, and small meshes, whose own launch
    // cannot fill the GPU, fill it together
    const int64_t gx = std::max<int64_t>(1, std::min<int64_t>(pat->nchunk,
                                                              (int64_t)g_num_cus * (pat->wg_per_cu ? pat->wg_per_cu : 32)));
    const int efm = (flags & EMG_EXACT_FACE_MASS) ? 1 : 0;
    const cplx *ce = reinterpret_cast<const cplx *>(d_eps), *cn = reinterpret_cast<const cplx *>(d_nu);
    cplx *oE = reinterpret_cast<cplx *>(d_E), *oB = reinterpret_cast<cplx *>(d_B);
    const int64_t nvs = 3 * nv, rs = pat->nt * (int64_t)REC, os = 4 * pat->nnz_ent;
    const dim3 gg(emg::grid_for(pat->nt, 256), (unsigned)nbatch);
    if (flags & EMG_REFERENCE_GEOMETRY)
        k_tet_gram<true><<<gg, 256, 0, st>>>(pat->nt, d_nodes, d_tets, ce, cn, mp->tet_rec_batch, nvs, pat->nt, rs);
    else
        k_tet_gram<false><<<gg, 256, 0, st>>>(pat->nt, d_nodes, d_tets, ce, cn, mp->tet_rec_batch, nvs, pat->nt, rs);
    launch_staged<CHUNK>(pat, dim3((unsigned)gx, (unsigned)nbatch), st, mp->tet_rec_batch, efm, oE, oB, rs, os);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_assemble_tensor(const emg_pattern *pat, const double *d_nodes, const int32_t *d_tets,
                                   const double *d_mm, const double *d_ms, int flags, double *d_E, double *d_B,
                                   void *stream) {
    return assemble<true>(pat, d_nodes, d_tets, d_mm, d_ms, flags, d_E, d_B, stream);
}
