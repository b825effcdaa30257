// Mesh topology on the GPU: the arrays of Mesh3D.update (fem/mesh3d.py:224-355)
// and the entity table of the Nedelec-2 DOF map (fem/elements/nedelec2.py:32-70),
// from the tets' vertex ids.
//
//  * every tet's vertices are sorted (canonical tets); its 6 edges and 4 faces
//    get 64-bit keys a*nv + b and (a*nv + b)*nv + c of their sorted vertex ids,
//    laid out tet-major (slot t*6 + k, t*4 + k);
//  * a stable rocPRIM radix sort of (key, slot) pairs, a flag at every new key and
//    an inclusive scan number the unique edges and faces in ascending key order
//    (lexicographic in the sorted vertex ids); the ids scatter back to the slots
//    (tet_ent), and because the sort is stable the tets sharing a face come out in
//    ascending tet order, which gives tri_to_tet as the reference fills it;
//  * tet_to_edge / tet_to_tri in the reference's local orders on the tet's own
//    vertex order, tri_to_edge and the edge lengths come from binary searches in
//    the sorted unique keys and from the node coordinates.
// All integer work is exact; the edge lengths are sqrt(dx^2 + dy^2 + dz^2) summed
// in that order with no fused multiply-add, as numpy evaluates them.

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

namespace {

constexpr int TB = 256;

__device__ __forceinline__ void sort4(int64_t (&v)[4]) {
#pragma unroll
    for (int i = 1; i < 4; i++)
#pragma unroll
        for (int j = i; j > 0; j--)
            if (v[j] < v[j - 1]) {
                const int64_t t = v[j];
                v[j] = v[j - 1];
                v[j - 1] = t;
            }
}

// canonical local edges (0,1)(0,2)(0,3)(1,2)(1,3)(2,3), faces (0,1,2)(0,2,3)(0,1,3)(1,2,3)
__constant__ int CE[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int CF[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 1, 3}, {1, 2, 3}};

__global__ void __launch_bounds__(TB) k_topo_keys(int64_t nt, int64_t nv, const int64_t *__restrict__ tets,
                                                  int32_t *__restrict__ canon, uint64_t *__restrict__ ekey,
                                                  int32_t *__restrict__ eslot, uint64_t *__restrict__ fkey,
                                                  int32_t *__restrict__ fslot, int *__restrict__ bad) {
    const int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= nt) return;
    int64_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = tets[k * nt + t];
    sort4(v);
    if (v[0] < 0 || v[3] >= nv || v[0] == v[1] || v[1] == v[2] || v[2] == v[3]) atomicOr(bad, 1);
#pragma unroll
    for (int k = 0; k < 4; k++) canon[4 * t + k] = (int32_t)v[k];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        ekey[6 * t + k] = (uint64_t)v[CE[k][0]] * (uint64_t)nv + (uint64_t)v[CE[k][1]];
        eslot[6 * t + k] = (int32_t)(6 * t + k);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        fkey[4 * t + k] = ((uint64_t)v[CF[k][0]] * (uint64_t)nv + (uint64_t)v[CF[k][1]]) * (uint64_t)nv + (uint64_t)v[CF[k][2]];
        fslot[4 * t + k] = (int32_t)(4 * t + k);
    }
}

__global__ void __launch_bounds__(TB) k_new_key(int64_t n, const uint64_t *__restrict__ k, int32_t *__restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) f[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

// at each first occurrence: the unique key and its start position in the sorted order
__global__ void __launch_bounds__(TB) k_unique(int64_t n, const uint64_t *__restrict__ k, const int32_t *__restrict__ id1,
                                               uint64_t *__restrict__ uniq, int32_t *__restrict__ start) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || k[i] != k[i - 1]) {
        uniq[id1[i] - 1] = k[i];
        start[id1[i] - 1] = (int32_t)i;
    }
}

// slot -> entity id (written into the tet_ent column of the slot), and for faces
// tri_to_tet[rank, id] = tet with rank = position within the run of equal keys
__global__ void __launch_bounds__(TB) k_scatter_ids(int64_t n, int per_tet, int col0, int64_t id_off,
                                                    const int32_t *__restrict__ slot, const int32_t *__restrict__ id1,
                                                    const int32_t *__restrict__ start, int32_t *__restrict__ tet_ent,
                                                    int64_t *__restrict__ tri_to_tet, int64_t nf,
                                                    int *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int32_t s = slot[i];
    const int64_t id = id1[i] - 1;
    const int64_t t = s / per_tet;
    const int k = s - (int)(t * per_tet);
    tet_ent[10 * t + col0 + k] = (int32_t)(id + id_off);
    if (tri_to_tet != nullptr) {
        const int64_t r = i - start[id];
        if (r > 1) atomicOr(bad, 2);
        else tri_to_tet[r * nf + id] = t;
    }
}

__device__ __forceinline__ int64_t find_key(const uint64_t *__restrict__ u, int64_t n, uint64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (u[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < n && u[lo] == key) ? lo : -1;
}

__device__ __forceinline__ int64_t sgn(int64_t x) { return (x > 0) - (x < 0); }

// the reference's local orders on the tet's own vertex order (fem/mesh3d.py:292-305):
// edges (1,2),(1,3),(1,4),(2,3),(4,2),(3,4), faces (1,2,3),(1,3,4),(1,4,2),(2,3,4) (1-based)
__constant__ int RE[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
__constant__ int RF[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};

__global__ void __launch_bounds__(TB) k_tet_maps(int64_t nt, int64_t nv, const int64_t *__restrict__ tets,
                                                 const uint64_t *__restrict__ ue, int64_t ne,
                                                 const uint64_t *__restrict__ uf, int64_t nf,
                                                 int64_t *__restrict__ t2e, int64_t *__restrict__ t2e_sign,
                                                 int64_t *__restrict__ t2f, int64_t *__restrict__ t2f_sign) {
    const int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= nt) return;
    int64_t T[4];
#pragma unroll
    for (int k = 0; k < 4; k++) T[k] = tets[k * nt + t];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int64_t a = T[RE[k][0]], b = T[RE[k][1]];
        const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
        t2e[k * nt + t] = find_key(ue, ne, (uint64_t)lo * (uint64_t)nv + (uint64_t)hi);
        t2e_sign[k * nt + t] = a > b ? -1 : 1;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int64_t v[3] = {T[RF[k][0]], T[RF[k][1]], T[RF[k][2]]};
        // _tri_ordering (fem/mesh3d.py:37-51) as written: sign(sign(i2 - 1) + sign(i3 - i2) + sign(i1 - i3))
        t2f_sign[k * nt + t] = sgn(sgn(v[1] - 1) + sgn(v[2] - v[1]) + sgn(v[0] - v[2]));
        if (v[1] < v[0]) { const int64_t x = v[0]; v[0] = v[1]; v[1] = x; }
        if (v[2] < v[1]) { const int64_t x = v[1]; v[1] = v[2]; v[2] = x; }
        if (v[1] < v[0]) { const int64_t x = v[0]; v[0] = v[1]; v[1] = x; }
        t2f[k * nt + t] = find_key(uf, nf, ((uint64_t)v[0] * (uint64_t)nv + (uint64_t)v[1]) * (uint64_t)nv + (uint64_t)v[2]);
    }
}

__global__ void __launch_bounds__(TB) k_tri_maps(int64_t nf, int64_t nv, const uint64_t *__restrict__ uf,
                                                 const uint64_t *__restrict__ ue, int64_t ne,
                                                 int64_t *__restrict__ tris, int64_t *__restrict__ tri_to_edge) {
    const int64_t f = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (f >= nf) return;
    const uint64_t k = uf[f];
    const uint64_t nvu = (uint64_t)nv;
    const int64_t a = (int64_t)(k / (nvu * nvu)), b = (int64_t)((k / nvu) % nvu), c = (int64_t)(k % nvu);
    tris[f] = a;
    tris[nf + f] = b;
    tris[2 * nf + f] = c;
    tri_to_edge[f] = find_key(ue, ne, (uint64_t)a * nvu + (uint64_t)b);
    tri_to_edge[nf + f] = find_key(ue, ne, (uint64_t)b * nvu + (uint64_t)c);
    tri_to_edge[2 * nf + f] = find_key(ue, ne, (uint64_t)a * nvu + (uint64_t)c);
}

__global__ void __launch_bounds__(TB) k_edge_maps(int64_t ne, int64_t nv, const uint64_t *__restrict__ ue,
                                                  const double *__restrict__ nodes, int64_t *__restrict__ edges,
                                                  double *__restrict__ len) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (e >= ne) return;
    const int64_t a = (int64_t)(ue[e] / (uint64_t)nv), b = (int64_t)(ue[e] % (uint64_t)nv);
    edges[e] = a;
    edges[ne + e] = b;
    if (nodes != nullptr) {
        const double dx = nodes[a] - nodes[b], dy = nodes[nv + a] - nodes[nv + b], dz = nodes[2 * nv + a] - nodes[2 * nv + b];
        const double sx = dx * dx, sy = dy * dy, sz = dz * dz;
        len[e] = sqrt((sx + sy) + sz);
    }
}

}  // namespace

struct emg_topology {
    int64_t nv = 0, nt = 0, ne = 0, nf = 0;
    int32_t *canon = nullptr, *tet_ent = nullptr;
    int64_t *edges = nullptr, *tris = nullptr, *t2e = nullptr, *t2e_sign = nullptr, *t2f = nullptr,
            *t2f_sign = nullptr, *tri_to_tet = nullptr, *tri_to_edge = nullptr;
    double *len = nullptr;
};

static void topo_free(emg_topology *h) {
    if (!h) return;
    void *ps[] = {h->canon, h->tet_ent, h->edges, h->tris, h->t2e, h->t2e_sign, h->t2f, h->t2f_sign,
                  h->tri_to_tet, h->tri_to_edge, h->len};
    for (void *p : ps) (void)hipFree(p);
    delete h;
}

extern "C" void emg_topology_destroy(emg_topology *h) { topo_free(h); }

// unique keys of n (key, slot) pairs: sorted, numbered; returns the count in *nu
static hipError_t unique_keys(int64_t n, int bits, uint64_t *keys, int32_t *slots, uint64_t *keys_s, int32_t *slots_s,
                              int32_t *id1, int32_t *flag, uint64_t *uniq, int32_t *start, int64_t *nu,
                              hipStream_t st) {
    size_t need = 0, need2 = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, need, keys, keys_s, slots, slots_s, (size_t)n, 0, bits, st);
    if (e != hipSuccess) return e;
    e = rocprim::inclusive_scan(nullptr, need2, flag, id1, (size_t)n, rocprim::plus<int32_t>(), st);
    if (e != hipSuccess) return e;
    if (need2 > need) need = need2;
    void *tmp = nullptr;
    e = hipMallocAsync(&tmp, need, st);
    if (e != hipSuccess) return e;
    size_t got = need;
    e = rocprim::radix_sort_pairs(tmp, got, keys, keys_s, slots, slots_s, (size_t)n, 0, bits, st);
    if (e == hipSuccess) {
        k_new_key<<<emg::grid_for(n, TB), TB, 0, st>>>(n, keys_s, flag);
        got = need;
        e = rocprim::inclusive_scan(tmp, got, flag, id1, (size_t)n, rocprim::plus<int32_t>(), st);
    }
    (void)hipFreeAsync(tmp, st);
    if (e != hipSuccess) return e;
    k_unique<<<emg::grid_for(n, TB), TB, 0, st>>>(n, keys_s, id1, uniq, start);
    int32_t last = 0;
    e = hipMemcpyAsync(&last, id1 + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    *nu = last;
    return e;
}

extern "C" int emg_topology_build(int64_t nv, int64_t nt, const int64_t *d_tets, const double *d_nodes,
                                  emg_topology **out, void *stream) {
    EMG_CHECK(out && d_tets, "emg_topology_build: null argument");
    *out = nullptr;
    EMG_CHECK(nt > 0 && nv > 0, "emg_topology_build: empty mesh");
    // face keys need nv^3 < 2^64 and slots int32
    EMG_CHECK(nv < (1ll << 21), "emg_topology_build: more than 2^21 vertices (face keys are 64-bit)");
    EMG_CHECK(6 * nt < (1ll << 31), "emg_topology_build: more than 357M tets");
    hipStream_t st = (hipStream_t)stream;
    auto h = new emg_topology();
    h->nv = nv;
    h->nt = nt;
    int bits_v = 1;
    while ((1ll << bits_v) < nv) bits_v++;
    uint64_t *ekey = nullptr, *ekey_s = nullptr, *fkey = nullptr, *fkey_s = nullptr, *ue = nullptr, *uf = nullptr;
    int32_t *eslot = nullptr, *eslot_s = nullptr, *fslot = nullptr, *fslot_s = nullptr, *id1 = nullptr,
            *flag = nullptr, *start = nullptr;
    int *d_bad = nullptr;
    auto cleanup = [&]() {
        void *ps[] = {ekey, ekey_s, fkey, fkey_s, ue, uf, eslot, eslot_s, fslot, fslot_s, id1, flag, start, d_bad};
        for (void *p : ps) (void)hipFree(p);
    };
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, bytes); };
    const int64_t n6 = 6 * nt, n4 = 4 * nt;
    A((void **)&ekey, n6 * 8); A((void **)&ekey_s, n6 * 8); A((void **)&eslot, n6 * 4); A((void **)&eslot_s, n6 * 4);
    A((void **)&fkey, n4 * 8); A((void **)&fkey_s, n4 * 8); A((void **)&fslot, n4 * 4); A((void **)&fslot_s, n4 * 4);
    A((void **)&id1, n6 * 4); A((void **)&flag, n6 * 4); A((void **)&start, n6 * 4);
    A((void **)&ue, n6 * 8); A((void **)&uf, n4 * 8); A((void **)&d_bad, sizeof(int));
    A((void **)&h->canon, n4 * 4); A((void **)&h->tet_ent, 10 * nt * 4);
    A((void **)&h->t2e, n6 * 8); A((void **)&h->t2e_sign, n6 * 8); A((void **)&h->t2f, n4 * 8);
    A((void **)&h->t2f_sign, n4 * 8);
    if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, sizeof(int), st);
    if (e == hipSuccess) {
        k_topo_keys<<<emg::grid_for(nt, TB), TB, 0, st>>>(nt, nv, d_tets, h->canon, ekey, eslot, fkey, fslot, d_bad);
        e = hipGetLastError();
    }
    // invalid vertex ids stop here, before any key is decoded into a node index
    int bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && bad) {
        cleanup();
        topo_free(h);
        emg::set_error("emg_topology_build: a tet has a vertex id out of range or repeated");
        return -1;
    }
    if (e == hipSuccess) e = unique_keys(n6, 2 * bits_v, ekey, eslot, ekey_s, eslot_s, id1, flag, ue, start, &h->ne, st);
    if (e == hipSuccess) {
        k_scatter_ids<<<emg::grid_for(n6, TB), TB, 0, st>>>(n6, 6, 0, 0, eslot_s, id1, start, h->tet_ent, nullptr, 0,
                                                           d_bad);
        e = hipGetLastError();
    }
    // faces: the edge sort's scratch is free again
    if (e == hipSuccess) e = unique_keys(n4, 3 * bits_v, fkey, fslot, fkey_s, fslot_s, id1, flag, uf, start, &h->nf, st);
    if (e == hipSuccess) {
        A((void **)&h->tri_to_tet, 2 * h->nf * 8); A((void **)&h->tris, 3 * h->nf * 8);
        A((void **)&h->tri_to_edge, 3 * h->nf * 8); A((void **)&h->edges, 2 * h->ne * 8);
        A((void **)&h->len, h->ne * 8);
    }
    if (e == hipSuccess) e = hipMemsetAsync(h->tri_to_tet, 0xff, 2 * h->nf * 8, st);   // -1 padded
    if (e == hipSuccess) {
        k_scatter_ids<<<emg::grid_for(n4, TB), TB, 0, st>>>(n4, 4, 6, h->ne, fslot_s, id1, start, h->tet_ent,
                                                           h->tri_to_tet, h->nf, d_bad);
        k_tet_maps<<<emg::grid_for(nt, TB), TB, 0, st>>>(nt, nv, d_tets, ue, h->ne, uf, h->nf, h->t2e, h->t2e_sign,
                                                        h->t2f, h->t2f_sign);
        k_tri_maps<<<emg::grid_for(h->nf, TB), TB, 0, st>>>(h->nf, nv, uf, ue, h->ne, h->tris, h->tri_to_edge);
        k_edge_maps<<<emg::grid_for(h->ne, TB), TB, 0, st>>>(h->ne, nv, ue, d_nodes, h->edges, h->len);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) {
        topo_free(h);
        emg::set_error(std::string("emg_topology_build: ") + hipGetErrorString(e));
        return -2;
    }
    if (bad) {
        topo_free(h);
        emg::set_error("emg_topology_build: non-manifold mesh: a triangle is shared by more than two tets");
        return -1;
    }
    if (!d_nodes) {
        (void)hipFree(h->len);
        h->len = nullptr;
    }
    *out = h;
    return 0;
}

extern "C" int emg_topology_sizes(const emg_topology *h, int64_t *ne, int64_t *nf) {
    EMG_CHECK(h && ne && nf, "emg_topology_sizes: null argument");
    *ne = h->ne;
    *nf = h->nf;
    return 0;
}

extern "C" int emg_topology_get(const emg_topology *h, int64_t *d_edges, int64_t *d_tris, int32_t *d_canon,
                                int32_t *d_tet_ent, int64_t *d_tet_to_edge, int64_t *d_tet_to_edge_sign,
                                int64_t *d_tet_to_tri, int64_t *d_tet_to_tri_sign, int64_t *d_tri_to_tet,
                                int64_t *d_tri_to_edge, double *d_edge_lengths, void *stream) {
    EMG_CHECK(h != nullptr, "emg_topology_get: null handle");
    EMG_CHECK(!d_edge_lengths || h->len, "emg_topology_get: no node coordinates were given to the build");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nt = h->nt, ne = h->ne, nf = h->nf;
    struct C { void *dst; const void *src; size_t bytes; } cs[] = {
        {d_edges, h->edges, (size_t)(2 * ne * 8)},         {d_tris, h->tris, (size_t)(3 * nf * 8)},
        {d_canon, h->canon, (size_t)(4 * nt * 4)},         {d_tet_ent, h->tet_ent, (size_t)(10 * nt * 4)},
        {d_tet_to_edge, h->t2e, (size_t)(6 * nt * 8)},     {d_tet_to_edge_sign, h->t2e_sign, (size_t)(6 * nt * 8)},
        {d_tet_to_tri, h->t2f, (size_t)(4 * nt * 8)},      {d_tet_to_tri_sign, h->t2f_sign, (size_t)(4 * nt * 8)},
        {d_tri_to_tet, h->tri_to_tet, (size_t)(2 * nf * 8)}, {d_tri_to_edge, h->tri_to_edge, (size_t)(3 * nf * 8)},
        {d_edge_lengths, h->len, (size_t)(ne * 8)}};
    for (const C &c : cs)
        if (c.dst) EMG_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToDevice, st));
    return 0;
}
