// Boundary terms and field evaluation for the frequency-domain path on MI355X.
//
//  * k_robin_tri: per port / absorbing-boundary triangle, the 8x8 Robin mass
//    block (with gamma = 1; the caller scales by gamma(f)) and the 8-entry
//    source vector.  Replaces ned2_tri_stiff_force / ned2_tri_stiff
//    (fem/mth/tri.py:575-796) as driven by assemble_robin_bc(_excited)
//    (fem/physics/edm/assembler.py:100-144).
//  * k_interp: the Nedelec-2 field of one or more solution vectors at points
//    given by (tet, barycentric coordinates).  Replaces ned2_tet_interp
//    (fem/mth/tet.py:371-497), whose brute-force point-in-tet search over
//    all port tets is unnecessary because every quadrature point knows its tet.
//
// Triangle basis, local vertices 0,1,2 (sorted global ids), slots
// [e(0,1) e(1,2) e(0,2) f | same, mode 2], N = s (l_p l_q grad l_r - l_p l_r grad l_q):
//   edge (a,b) mode 1 (a,b,a) s=L_ab, mode 2 (b,b,a) s=L_ab,
//   face mode 1 (1,2,0) s=-L_02, mode 2 (2,1,0) s=L_01.

#include <cmath>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

using emg::cplx;

__device__ __forceinline__ int tri_moment(int w, int x, int y, int z) {
    int r = 1;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int c = (w == k) + (x == k) + (y == k) + (z == k);
        r *= (c == 4) ? 24 : (c == 3) ? 6 : (c == 2) ? 2 : 1;
    }
    return r;
}

__global__ void __launch_bounds__(64) k_robin_tri(int64_t ntri, const double *__restrict__ lxy, const double *__restrict__ elen,
                                                  const cplx *__restrict__ uinc, int nq, const double *__restrict__ dpts,
                                                  double *__restrict__ Bout, cplx *__restrict__ bout) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const double *p = lxy + 6 * t;
    const double xs[3] = {p[0], p[1], p[2]}, ys[3] = {p[3], p[4], p[5]};
    const double bb[3] = {ys[1] - ys[2], ys[2] - ys[0], ys[0] - ys[1]};
    const double cc[3] = {xs[2] - xs[1], xs[0] - xs[2], xs[1] - xs[0]};
    const double det = (xs[1] - xs[0]) * (ys[2] - ys[0]) - (xs[2] - xs[0]) * (ys[1] - ys[0]);
    const double A = 0.5 * fabs(det);
    auto dist = [&](int i, int j) { return sqrt((xs[i] - xs[j]) * (xs[i] - xs[j]) + (ys[i] - ys[j]) * (ys[i] - ys[j])); };
    double Le[3];
    if (elen) { Le[0] = elen[3 * t]; Le[1] = elen[3 * t + 1]; Le[2] = elen[3 * t + 2]; }
    else { Le[0] = dist(0, 1); Le[1] = dist(1, 2); Le[2] = dist(0, 2); }
    // slot descriptors
    int P[8], Q[8], Rr[8];
    double S[8];
    constexpr int EA[3] = {0, 1, 0}, EB[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        P[k] = EA[k]; Q[k] = EB[k]; Rr[k] = EA[k]; S[k] = Le[k];
        P[4 + k] = EB[k]; Q[4 + k] = EB[k]; Rr[4 + k] = EA[k]; S[4 + k] = Le[k];
    }
    P[3] = 1; Q[3] = 2; Rr[3] = 0; S[3] = -dist(2, 0);
    P[7] = 2; Q[7] = 1; Rr[7] = 0; S[7] = dist(1, 0);
    double G[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) G[i][j] = bb[i] * bb[j] + cc[i] * cc[j];
    // integral l^k.. over the triangle = 2A prod(k!)/(sum k + 2)!, grad l = (b, c)/(2A)
    const double scale = 2.0 * A / 720.0 / (4.0 * A * A);
    double *Bt = Bout + 64 * t;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int p0 = P[i], q0 = Q[i], r0 = Rr[i], p1 = P[j], q1 = Q[j], r1 = Rr[j];
            const double v = tri_moment(p0, q0, p1, q1) * G[r0][r1] - tri_moment(p0, q0, p1, r1) * G[r0][q1] -
                             tri_moment(p0, r0, p1, q1) * G[q0][r1] + tri_moment(p0, r0, p1, r1) * G[q0][q1];
            Bt[i * 8 + j] = S[i] * S[j] * v * scale;
        }
    if (bout == nullptr) return;
    // source: -A * sum_q w_q N_i(x_q) . U(x_q)  (the reference's sign convention)
    cplx acc[8];
    for (int q = 0; q < nq; q++) {
        const double w = dpts[q];
        const double lam[3] = {dpts[nq + q], dpts[2 * nq + q], dpts[3 * nq + q]};
        const cplx ux = uinc[(t * nq + q) * 2], uy = uinc[(t * nq + q) * 2 + 1];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const double gx_r = bb[Rr[i]] / det, gy_r = cc[Rr[i]] / det;
            const double gx_q = bb[Q[i]] / det, gy_q = cc[Q[i]] / det;
            const double f1 = lam[P[i]] * lam[Q[i]], f2 = lam[P[i]] * lam[Rr[i]];
            const double nx = S[i] * (f1 * gx_r - f2 * gx_q), ny = S[i] * (f1 * gy_r - f2 * gy_q);
            acc[i] += (w * nx) * ux + (w * ny) * uy;
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) bout[8 * t + i] = acc[i] * (-A);
}

extern "C" int emg_robin_blocks(int64_t ntri, const double *d_lxy, const double *d_elen, const double *d_uinc, int nq,
                                const double *d_dpts, double *d_B, double *d_b, void *stream) {
    EMG_CHECK(d_lxy && d_B && (d_b == nullptr || (d_uinc && d_dpts)), "emg_robin_blocks: null argument");
    if (ntri == 0) return 0;
    k_robin_tri<<<emg::grid_for(ntri, 64), 64, 0, (hipStream_t)stream>>>(
        ntri, d_lxy, d_elen, reinterpret_cast<const cplx *>(d_uinc), nq, d_dpts, d_B, reinterpret_cast<cplx *>(d_b));
    EMG_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ interpolation
__host__ __device__ constexpr int slot_vertex(int s, int k) {
    return (s == 0) ? (k == 0 ? 0 : 1) : (s == 1) ? (k == 0 ? 0 : 2) : (s == 2) ? (k == 0 ? 0 : 3)
         : (s == 3) ? (k == 0 ? 1 : 2) : (s == 4) ? (k == 0 ? 1 : 3) : (s == 5) ? (k == 0 ? 2 : 3)
         : (s == 6) ? k : (s == 7) ? (k == 0 ? 0 : k + 1) : (s == 8) ? (k == 2 ? 3 : k) : k + 1;
}

// E at point `pt` of solution `sol`; thread per (point, solution).
__global__ void __launch_bounds__(128) k_interp(int64_t npts, const int32_t *__restrict__ ptet, const double *__restrict__ bary,
                                                const double *__restrict__ nodes, const int32_t *__restrict__ tets,
                                                const int32_t *__restrict__ tet_ent, int64_t n_ent, int nsol,
                                                const cplx *__restrict__ x, cplx *__restrict__ E) {
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (gid >= npts * nsol) return;
    const int64_t pt = gid % npts;
    const int sol = (int)(gid / npts);
    const int t = ptet[pt];
    if (t < 0) {                                    // outside the mesh: zero field, as the reference
        cplx *z = E + ((int64_t)sol * npts + pt) * 3;
        z[0] = cplx(); z[1] = cplx(); z[2] = cplx();
        return;
    }
    const int4 tv = reinterpret_cast<const int4 *>(tets)[t];
    const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
    double P[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) P[i][c] = nodes[3 * (int64_t)vid[i] + c];
    double e1[3], e2[3], e3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { e1[c] = P[1][c] - P[0][c]; e2[c] = P[2][c] - P[0][c]; e3[c] = P[3][c] - P[0][c]; }
    double g[4][3];
    g[1][0] = e2[1] * e3[2] - e2[2] * e3[1]; g[1][1] = e2[2] * e3[0] - e2[0] * e3[2]; g[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
    g[2][0] = e3[1] * e1[2] - e3[2] * e1[1]; g[2][1] = e3[2] * e1[0] - e3[0] * e1[2]; g[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
    g[3][0] = e1[1] * e2[2] - e1[2] * e2[1]; g[3][1] = e1[2] * e2[0] - e1[0] * e2[2]; g[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double det = e1[0] * g[1][0] + e1[1] * g[1][1] + e1[2] * g[1][2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        g[0][c] = -(g[1][c] + g[2][c] + g[3][c]) / det;
        g[1][c] /= det; g[2][c] /= det; g[3][c] /= det;
    }
    const double lam[4] = {bary[4 * pt], bary[4 * pt + 1], bary[4 * pt + 2], bary[4 * pt + 3]};
    const cplx *xs = x + (int64_t)sol * 2 * n_ent;
    cplx ex, ey, ez;
#pragma unroll
    for (int s = 0; s < 10; s++) {
        const int a = slot_vertex(s, 0), b = slot_vertex(s, 1), c = slot_vertex(s, 2);
        const bool face = s >= 6;
        const int ent = tet_ent[(int64_t)t * 10 + s];
        auto len = [&](int u, int v) {
            return sqrt((P[u][0] - P[v][0]) * (P[u][0] - P[v][0]) + (P[u][1] - P[v][1]) * (P[u][1] - P[v][1]) +
                        (P[u][2] - P[v][2]) * (P[u][2] - P[v][2]));
        };
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int p = (m == 0) ? (face ? b : a) : (face ? c : b);
            const int q = (m == 0 && face) ? c : b;
            const int r = a;
            const double sc = (m == 0 && face) ? -len(a, c) : len(a, b);
            const cplx coef = xs[(int64_t)m * n_ent + ent];
            const double f1 = sc * lam[p] * lam[q], f2 = sc * lam[p] * lam[r];
            ex += coef * (f1 * g[r][0] - f2 * g[q][0]);
            ey += coef * (f1 * g[r][1] - f2 * g[q][1]);
            ez += coef * (f1 * g[r][2] - f2 * g[q][2]);
        }
    }
    cplx *o = E + ((int64_t)sol * npts + pt) * 3;
    o[0] = ex; o[1] = ey; o[2] = ez;
}

// curl E at point `pt` of solution `sol` times c[tet] (c complex per tet, NULL = 1): for
// N = s l_p (l_q grad l_r - l_r grad l_q),
// curl N = s (l_q grad l_p x grad l_r - l_r grad l_p x grad l_q + 2 l_p grad l_q x grad l_r).
__global__ void __launch_bounds__(128) k_interp_curl(int64_t npts, const int32_t *__restrict__ ptet,
                                                     const double *__restrict__ bary, const double *__restrict__ nodes,
                                                     const int32_t *__restrict__ tets, const int32_t *__restrict__ tet_ent,
                                                     int64_t n_ent, int nsol, const cplx *__restrict__ x,
                                                     const cplx *__restrict__ cscale, cplx *__restrict__ H) {
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (gid >= npts * nsol) return;
    const int64_t pt = gid % npts;
    const int sol = (int)(gid / npts);
    const int t = ptet[pt];
    cplx *o = H + ((int64_t)sol * npts + pt) * 3;
    if (t < 0) {
        o[0] = cplx(); o[1] = cplx(); o[2] = cplx();
        return;
    }
    const int4 tv = reinterpret_cast<const int4 *>(tets)[t];
    const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
    double P[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) P[i][c] = nodes[3 * (int64_t)vid[i] + c];
    double e1[3], e2[3], e3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { e1[c] = P[1][c] - P[0][c]; e2[c] = P[2][c] - P[0][c]; e3[c] = P[3][c] - P[0][c]; }
    double g[4][3];
    g[1][0] = e2[1] * e3[2] - e2[2] * e3[1]; g[1][1] = e2[2] * e3[0] - e2[0] * e3[2]; g[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
    g[2][0] = e3[1] * e1[2] - e3[2] * e1[1]; g[2][1] = e3[2] * e1[0] - e3[0] * e1[2]; g[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
    g[3][0] = e1[1] * e2[2] - e1[2] * e2[1]; g[3][1] = e1[2] * e2[0] - e1[0] * e2[2]; g[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double det = e1[0] * g[1][0] + e1[1] * g[1][1] + e1[2] * g[1][2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        g[0][c] = -(g[1][c] + g[2][c] + g[3][c]) / det;
        g[1][c] /= det; g[2][c] /= det; g[3][c] /= det;
    }
    auto crs = [&](int u, int v, int c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        return g[u][c1] * g[v][c2] - g[u][c2] * g[v][c1];
    };
    const double lam[4] = {bary[4 * pt], bary[4 * pt + 1], bary[4 * pt + 2], bary[4 * pt + 3]};
    const cplx *xs = x + (int64_t)sol * 2 * n_ent;
    cplx hx, hy, hz;
#pragma unroll
    for (int s = 0; s < 10; s++) {
        const int a = slot_vertex(s, 0), b = slot_vertex(s, 1), c = slot_vertex(s, 2);
        const bool face = s >= 6;
        const int ent = tet_ent[(int64_t)t * 10 + s];
        auto len = [&](int u, int v) {
            return sqrt((P[u][0] - P[v][0]) * (P[u][0] - P[v][0]) + (P[u][1] - P[v][1]) * (P[u][1] - P[v][1]) +
                        (P[u][2] - P[v][2]) * (P[u][2] - P[v][2]));
        };
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int p = (m == 0) ? (face ? b : a) : (face ? c : b);
            const int q = (m == 0 && face) ? c : b;
            const int r = a;
            const double sc = (m == 0 && face) ? -len(a, c) : len(a, b);
            const cplx coef = xs[(int64_t)m * n_ent + ent];
            double cv[3];
#pragma unroll
            for (int k = 0; k < 3; k++)
                cv[k] = sc * (lam[q] * crs(p, r, k) - lam[r] * crs(p, q, k) + 2.0 * lam[p] * crs(q, r, k));
            hx += coef * cv[0];
            hy += coef * cv[1];
            hz += coef * cv[2];
        }
    }
    const cplx cs = cscale ? cscale[t] : cplx(1.0, 0.0);
    o[0] = hx * cs; o[1] = hy * cs; o[2] = hz * cs;
}

// Point location on a uniform grid of cells: each point tests the tets listed for its cell
// (ascending tet id) and keeps the last one that holds it, with the reference's inside test
// (fem/mth/tet.py:423-426: local coordinates >= -1e-6, their sum <= 1.00000001).  tet -1
// and zero barycentric coordinates for points outside every candidate.
__global__ void __launch_bounds__(256) k_locate(int64_t npts, const double *__restrict__ pts, const double *__restrict__ grid,
                                                const int32_t *__restrict__ cell_ptr, const int32_t *__restrict__ cell_tets,
                                                const double *__restrict__ nodes, const int32_t *__restrict__ tets,
                                                int32_t *__restrict__ out_tet, double *__restrict__ out_bary) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= npts) return;
    const double px = pts[3 * k], py = pts[3 * k + 1], pz = pts[3 * k + 2];
    // grid: origin (3), inverse cell size (3), dims (3, as doubles)
    const int nx = (int)grid[6], ny = (int)grid[7], nz = (int)grid[8];
    const double fx = (px - grid[0]) * grid[3], fy = (py - grid[1]) * grid[4], fz = (pz - grid[2]) * grid[5];
    int best = -1;
    double bl[4] = {0.0, 0.0, 0.0, 0.0};
    if (fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < nx && fy < ny && fz < nz) {
        const int64_t cell = ((int64_t)fz * ny + (int)fy) * nx + (int)fx;
        for (int32_t i = cell_ptr[cell]; i < cell_ptr[cell + 1]; i++) {
            const int t = cell_tets[i];
            const int4 tv = reinterpret_cast<const int4 *>(tets)[t];
            const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
            double P[4][3];
#pragma unroll
            for (int v = 0; v < 4; v++)
#pragma unroll
                for (int c = 0; c < 3; c++) P[v][c] = nodes[3 * (int64_t)vid[v] + c];
            double e[3][3], r[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                e[0][c] = P[1][c] - P[0][c]; e[1][c] = P[2][c] - P[0][c]; e[2][c] = P[3][c] - P[0][c];
            }
            r[0] = px - P[0][0]; r[1] = py - P[0][1]; r[2] = pz - P[0][2];
            // local coordinates u = T^{-1} r with T = [e0 e1 e2] as columns (Cramer's rule)
            auto det3 = [](const double *a, const double *b, const double *c) {
                return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                       a[2] * (b[0] * c[1] - b[1] * c[0]);
            };
            const double D = det3(e[0], e[1], e[2]);
            if (D == 0.0) continue;
            const double u0 = det3(r, e[1], e[2]) / D, u1 = det3(e[0], r, e[2]) / D, u2 = det3(e[0], e[1], r) / D;
            if (u0 >= -1e-6 && u1 >= -1e-6 && u2 >= -1e-6 && u0 + u1 + u2 <= 1.00000001) {
                best = t;
                bl[0] = 1.0 - u0 - u1 - u2; bl[1] = u0; bl[2] = u1; bl[3] = u2;
            }
        }
    }
    out_tet[k] = best;
#pragma unroll
    for (int j = 0; j < 4; j++) out_bary[4 * k + j] = bl[j];
}

extern "C" int emg_interp_curl_points(int64_t npts, const int32_t *d_ptet, const double *d_bary, const double *d_nodes,
                                      const int32_t *d_tets, const int32_t *d_tet_ent, int64_t n_ent, int nsol,
                                      const double *d_x, const double *d_cscale, double *d_H, void *stream) {
    EMG_CHECK(d_ptet && d_bary && d_nodes && d_tets && d_tet_ent && d_x && d_H, "emg_interp_curl_points: null argument");
    if (npts == 0 || nsol == 0) return 0;
    k_interp_curl<<<emg::grid_for(npts * nsol, 128), 128, 0, (hipStream_t)stream>>>(
        npts, d_ptet, d_bary, d_nodes, d_tets, d_tet_ent, n_ent, nsol, reinterpret_cast<const cplx *>(d_x),
        reinterpret_cast<const cplx *>(d_cscale), reinterpret_cast<cplx *>(d_H));
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_locate_points(int64_t npts, const double *d_pts, const double *d_grid, const int32_t *d_cell_ptr,
                                 const int32_t *d_cell_tets, const double *d_nodes, const int32_t *d_tets,
                                 int32_t *d_tet, double *d_bary, void *stream) {
    EMG_CHECK(d_pts && d_grid && d_cell_ptr && d_cell_tets && d_nodes && d_tets && d_tet && d_bary,
              "emg_locate_points: null argument");
    if (npts == 0) return 0;
    k_locate<<<emg::grid_for(npts, 256), 256, 0, (hipStream_t)stream>>>(npts, d_pts, d_grid, d_cell_ptr, d_cell_tets,
                                                                       d_nodes, d_tets, d_tet, d_bary);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_interp_points(int64_t npts, const int32_t *d_ptet, const double *d_bary, const double *d_nodes,
                                 const int32_t *d_tets, const int32_t *d_tet_ent, int64_t n_ent, int nsol,
                                 const double *d_x, double *d_E, void *stream) {
    EMG_CHECK(d_ptet && d_bary && d_nodes && d_tets && d_tet_ent && d_x && d_E, "emg_interp_points: null argument");
    if (npts == 0 || nsol == 0) return 0;
    k_interp<<<emg::grid_for(npts * nsol, 128), 128, 0, (hipStream_t)stream>>>(
        npts, d_ptet, d_bary, d_nodes, d_tets, d_tet_ent, n_ent, nsol, reinterpret_cast<const cplx *>(d_x),
        reinterpret_cast<cplx *>(d_E));
    EMG_LAUNCH_CHECK();
    return 0;
}
