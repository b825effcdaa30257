// Port-face modal analysis on the GPU (fem/physics/edm/emfreq3d.py:201-364).
//
// The face of a port is a triangle mesh carrying the 14-DOF NedelecLegrange2
// field of each triangle: 8 second-order Nedelec tangential functions (two per
// edge, two on the face) and the 6 quadratic Lagrange functions of the normal
// component (fem/elements/nedleg2.py:30-126).  The modal problem is the pencil
//     A e = lambda B e,   lambda = -beta^2
// of generalized_matrix_GQ (fem/physics/edm/nedeleclegrange2.py:224-374), which
// the reference solves densely with LAPACK (fem/solver.py:311-341).  Here:
//
//  * k_port_element: one thread per (triangle, 14 x 14 entry).  Each thread
//    evaluates its two basis functions at the six quadrature points, forms the
//    entry of A and B and adds it into the dense column-major matrices (FP64
//    atomics) at the caller's reduced DOF numbers; DOFs numbered -1 (PEC) are
//    dropped.  The reference's quirks are kept: its 8-digit quadrature table,
//    face rows of Att/Btt/Dtt copied from the face columns, and the face-face
//    entries of Dtt left zero unless EMG_PORT_FACE_BLOCK asks for them.
//  * emg_port_factor / emg_port_operator / emg_port_krylov: shift-invert Arnoldi on
//    T = (A - sigma B)^{-1} B, formed once from a rocSOLVER LU (zgetrf, then one
//    zgetrs with n right-hand sides), each step one rocBLAS zgemv; classical
//    Gram-Schmidt with one re-orthogonalisation.
//  * k_port_field: the face field (and the reference's curl expression) at
//    points, each point taking the last triangle that contains it, as
//    ned2_tri_interp_full / _curl do (fem/mth/tri.py:228-401).

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

using emg::cplx;

namespace {

// nedeleclegrange2.py:201-204, as written there: weight, then the three barycentric weights
__constant__ double GQ_W[6] = {0.22338159, 0.22338159, 0.22338159, 0.10995174, 0.10995174, 0.10995174};
__constant__ double GQ_L1[6] = {0.10810302, 0.44594849, 0.44594849, 0.81684757, 0.09157621, 0.09157621};
__constant__ double GQ_L2[6] = {0.44594849, 0.44594849, 0.10810302, 0.09157621, 0.09157621, 0.81684757};
__constant__ double GQ_L3[6] = {0.44594849, 0.10810302, 0.44594849, 0.09157621, 0.81684757, 0.09157621};

// local edges of a triangle with ascending vertex numbers: (0,1), (1,2), (0,2)
__device__ __forceinline__ int edge_a(int k) { return k == 1 ? 1 : 0; }
__device__ __forceinline__ int edge_b(int k) { return k == 0 ? 1 : 2; }

struct Tri {
    double a[3], b[3], c[3];   // l_i = (a_i + b_i x + c_i y) / (2 area), signs for a positive area
    double area;               // |signed area|
    double L[3];               // edge lengths (0,1), (1,2), (0,2)
};

__device__ inline Tri tri_setup(const double *xs, const double *ys) {
    Tri T;
    const double x1 = xs[0], x2 = xs[1], x3 = xs[2], y1 = ys[0], y2 = ys[1], y3 = ys[2];
    const double sA = 0.5 * ((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1));
    const double s = sA < 0 ? -1.0 : (sA > 0 ? 1.0 : 0.0);
    T.a[0] = s * (x2 * y3 - y2 * x3); T.a[1] = s * (x3 * y1 - y3 * x1); T.a[2] = s * (x1 * y2 - y1 * x2);
    T.b[0] = s * (y2 - y3); T.b[1] = s * (y3 - y1); T.b[2] = s * (y1 - y2);
    T.c[0] = s * (x3 - x2); T.c[1] = s * (x1 - x3); T.c[2] = s * (x2 - x1);
    T.area = fabs(sA);
    T.L[0] = hypot(x1 - x2, y1 - y2);
    T.L[1] = hypot(x2 - x3, y2 - y3);
    T.L[2] = hypot(x1 - x3, y1 - y3);
    return T;
}

// DOF scale: edge DOFs by their edge length, face DOF 3 by |(0,2)|, face DOF 7 by |(0,1)|
__device__ __forceinline__ double dof_scale(const Tri &T, int k) {
    if (k < 3) return T.L[k];
    if (k == 3) return T.L[2];
    if (k < 7) return T.L[k - 4];
    if (k == 7) return T.L[0];
    return 1.0;
}

struct Pt {
    double l[3];
    double gx[3], gy[3];
};

__device__ inline Pt point_at(const Tri &T, double x, double y) {
    Pt P;
    const double k = 1.0 / (2.0 * T.area);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        P.l[i] = (T.a[i] + T.b[i] * x + T.c[i] * y) * k;
        P.gx[i] = T.b[i] * k;
        P.gy[i] = T.c[i] * k;
    }
    return P;
}

__device__ __forceinline__ double cr(const Pt &P, int u, int v) { return P.gx[u] * P.gy[v] - P.gy[u] * P.gx[v]; }

// l_p (l_q grad l_r - l_r grad l_q) and its z-curl
__device__ inline void wfun(const Pt &P, int p, int q, int r, double s, double &wx, double &wy, double &wc) {
    wx = s * P.l[p] * (P.l[q] * P.gx[r] - P.l[r] * P.gx[q]);
    wy = s * P.l[p] * (P.l[q] * P.gy[r] - P.l[r] * P.gy[q]);
    wc = s * (P.l[q] * cr(P, p, r) - P.l[r] * cr(P, p, q) + 2.0 * P.l[p] * cr(P, q, r));
}

// tangential function k of [e1(01), e1(12), e1(02), f1, e2(01), e2(12), e2(02), f2]
__device__ inline void tangential(const Pt &P, int k, double &wx, double &wy, double &wc) {
    if (k < 3) {
        const int i = edge_a(k), j = edge_b(k);
        wfun(P, i, j, i, 1.0, wx, wy, wc);
    } else if (k == 3) {
        wfun(P, 1, 2, 0, -1.0, wx, wy, wc);
    } else if (k < 7) {
        const int i = edge_a(k - 4), j = edge_b(k - 4);
        wfun(P, j, j, i, 1.0, wx, wy, wc);
    } else {
        wfun(P, 2, 1, 0, 1.0, wx, wy, wc);
    }
}

// quadratic Lagrange function k (vertices 0..2, then edges (01), (12), (02)) and its gradient
__device__ inline void nodal(const Pt &P, int k, double &v, double &gx, double &gy) {
    if (k < 3) {
        v = P.l[k] * (2.0 * P.l[k] - 1.0);
        gx = P.gx[k] * (4.0 * P.l[k] - 1.0);
        gy = P.gy[k] * (4.0 * P.l[k] - 1.0);
    } else {
        const int i = edge_a(k - 3), j = edge_b(k - 3);
        v = 4.0 * P.l[i] * P.l[j];
        gx = 4.0 * (P.gx[i] * P.l[j] + P.l[i] * P.gx[j]);
        gy = 4.0 * (P.gy[i] * P.l[j] + P.l[i] * P.gy[j]);
    }
}

// u . (M v) for a real 2-vector pair and the upper-left 2 x 2 block of a complex 3 x 3 M
__device__ __forceinline__ cplx dot_m(const cplx *M, double ux, double uy, double vx, double vy) {
    return ux * (M[0] * vx + M[1] * vy) + uy * (M[3] * vx + M[4] * vy);
}

constexpr int EL_BLOCK = 256;

__global__ void __launch_bounds__(EL_BLOCK) k_port_element(int64_t ntri, const double *__restrict__ verts,
                                                           const cplx *__restrict__ Ms, const cplx *__restrict__ Mm,
                                                           double k0, int face_block, const int64_t *__restrict__ t2r,
                                                           int64_t n, double *__restrict__ Ad, double *__restrict__ Bd,
                                                           cplx *__restrict__ Ael, cplx *__restrict__ Bel) {
    const int64_t gid = (int64_t)blockIdx.x * EL_BLOCK + threadIdx.x;
    if (gid >= ntri * 196) return;
    const int64_t t = gid / 196;
    const int ij = (int)(gid - t * 196);
    const int i = ij / 14, j = ij - (ij / 14) * 14;
    const Tri T = tri_setup(verts + 6 * t, verts + 6 * t + 3);
    const cplx *ms = Ms + 9 * t;
    const cplx *mm = Mm + 9 * t;
    const double ox[3] = {verts[6 * t], verts[6 * t + 1], verts[6 * t + 2]};
    const double oy[3] = {verts[6 * t + 3], verts[6 * t + 4], verts[6 * t + 5]};
    cplx A, B;
    const bool fi = (i == 3 || i == 7), fj = (j == 3 || j == 7);
    for (int q = 0; q < 6; q++) {
        const double px = ox[0] * GQ_L1[q] + ox[1] * GQ_L2[q] + ox[2] * GQ_L3[q];
        const double py = oy[0] * GQ_L1[q] + oy[1] * GQ_L2[q] + oy[2] * GQ_L3[q];
        const Pt P = point_at(T, px, py);
        const double W = GQ_W[q];
        if (i < 8 && j < 8) {
            // the reference computes [edge, face] and copies it to [face, edge]
            const int ii = (fi && !fj) ? j : i, jj = (fi && !fj) ? i : j;
            double ux, uy, uc, vx, vy, vc;
            tangential(P, ii, ux, uy, uc);
            tangential(P, jj, vx, vy, vc);
            const cplx att = (uc * vc) * ms[8];
            const cplx btt = dot_m(mm, ux, uy, vx, vy);
            A += W * (att - (k0 * k0) * btt);
            if (!(fi && fj) || face_block) B += W * dot_m(ms, ux, uy, vx, vy);
        } else if (i >= 8 && j < 8) {
            double nv, ngx, ngy, vx, vy, vc;
            nodal(P, i - 8, nv, ngx, ngy);
            tangential(P, j, vx, vy, vc);
            B += W * dot_m(ms, ngx, ngy, vx, vy);
        } else if (i < 8) {
            double nv, ngx, ngy, vx, vy, vc;
            nodal(P, j - 8, nv, ngx, ngy);
            tangential(P, i, vx, vy, vc);
            B += W * dot_m(ms, ngx, ngy, vx, vy);
        } else {
            double nu, ugx, ugy, nv, vgx, vgy;
            nodal(P, i - 8, nu, ugx, ugy);
            nodal(P, j - 8, nv, vgx, vgy);
            B += W * (dot_m(ms, ugx, ugy, vgx, vgy) - (k0 * k0 * nu * nv) * mm[8]);
        }
    }
    const double s = dof_scale(T, i) * dof_scale(T, j) * T.area;
    A = s * A;
    B = s * B;
    if (Ael) {
        Ael[t * 196 + ij] = A;
        Bel[t * 196 + ij] = B;
    }
    if (Ad) {
        const int64_t r = t2r[t * 14 + i], c = t2r[t * 14 + j];
        if (r >= 0 && c >= 0 && r < n && c < n) {
            const int64_t o = 2 * (r + c * n);
            atomicAdd(Ad + o, A.re);
            atomicAdd(Ad + o + 1, A.im);
            atomicAdd(Bd + o, B.re);
            atomicAdd(Bd + o + 1, B.im);
        }
    }
}

constexpr int FLD_BLOCK = 128;

__global__ void __launch_bounds__(FLD_BLOCK) k_port_field(int64_t np, const double *__restrict__ pts, int64_t ntri,
                                                          const int64_t *__restrict__ tris, int64_t nnodes,
                                                          const double *__restrict__ nodes,
                                                          const int64_t *__restrict__ t2f, int64_t nfield,
                                                          const cplx *__restrict__ field, double beta,
                                                          const cplx *__restrict__ diadic,
                                                          const double *__restrict__ basis, cplx hconst,
                                                          cplx *__restrict__ Eout, cplx *__restrict__ Hout) {
    const int64_t p = (int64_t)blockIdx.x * FLD_BLOCK + threadIdx.x;
    if (p >= np) return;
    const double x = pts[2 * p], y = pts[2 * p + 1];
    int64_t hit = -1;
    bool bad = false;
    for (int64_t t = 0; t < ntri; t++) {
        const int64_t v0 = tris[3 * t], v1 = tris[3 * t + 1], v2 = tris[3 * t + 2];
        if (v0 < 0 || v1 < 0 || v2 < 0 || v0 >= nnodes || v1 >= nnodes || v2 >= nnodes) { bad = true; break; }
        const double x0 = nodes[2 * v0], y0 = nodes[2 * v0 + 1];
        const double t00 = nodes[2 * v1] - x0, t01 = nodes[2 * v2] - x0;
        const double t10 = nodes[2 * v1 + 1] - y0, t11 = nodes[2 * v2 + 1] - y0;
        const double id = 1.0 / (t00 * t11 - t01 * t10);
        const double dx = x - x0, dy = y - y0;
        const double s = (t11 * dx - t01 * dy) * id, u = (-t10 * dx + t00 * dy) * id;
        if (s + u <= 1.00000001 && s >= -1e-6 && u >= -1e-6) hit = t;    // the last containing triangle wins
    }
    cplx E[3], H[3];
    if (!bad && hit >= 0) {
        for (int k = 0; k < 14; k++) {
            const int64_t f = t2f[14 * hit + k];
            if (f < 0 || f >= nfield) bad = true;
        }
    }
    if (bad) {
        // an index outside the given sizes: NaN outputs instead of an out-of-bounds read
        const double nan = __builtin_nan("");
        for (int r = 0; r < 3; r++) E[r] = H[r] = cplx(nan, nan);
    } else if (hit >= 0) {
        const int64_t *tv = tris + 3 * hit;
        const double xs[3] = {nodes[2 * tv[0]], nodes[2 * tv[1]], nodes[2 * tv[2]]};
        const double ys[3] = {nodes[2 * tv[0] + 1], nodes[2 * tv[1] + 1], nodes[2 * tv[2] + 1]};
        const Tri T = tri_setup(xs, ys);
        const Pt P = point_at(T, x, y);
        cplx ex, ey, cz;
        for (int k = 0; k < 8; k++) {
            double wx, wy, wc;
            tangential(P, k, wx, wy, wc);
            const cplx e = dof_scale(T, k) * field[t2f[14 * hit + k]];
            ex += wx * e;
            ey += wy * e;
            cz += wc * e;
        }
        const int64_t *f = t2f + 14 * hit;
        const cplx S9 = field[f[8]] + field[f[9]] + field[f[10]];
        const cplx S12 = field[f[11]] + field[f[12]] + field[f[13]];
        const double l0 = P.l[0], l1 = P.l[1];
        // Ez and its gradient as ned2_tri_interp_full writes them (tri.py:303)
        const cplx ez = (l0 * (2.0 * l0 - 1.0)) * S9 + (2.0 * l0 * l1) * S12;
        const cplx gx = P.gx[0] * ((2.0 * l0 - 1.0) * S9 + (2.0 * l1) * S12) + l0 * (2.0 * P.gx[0] * S9 + 2.0 * P.gx[1] * S12);
        const cplx gy = P.gy[0] * ((2.0 * l0 - 1.0) * S9 + (2.0 * l1) * S12) + l0 * (2.0 * P.gy[0] * S9 + 2.0 * P.gy[1] * S12);
        const cplx jb(0.0, beta);
        const cplx hl[3] = {gy + jb * ey, gx - jb * ex, cz};
        const cplx el[3] = {ex, ey, ez};
#pragma unroll
        for (int r = 0; r < 3; r++) {
            E[r] = basis[3 * r] * el[0] + basis[3 * r + 1] * el[1] + basis[3 * r + 2] * el[2];
        }
        if (Hout) {
            const cplx *dg = diadic + 9 * hit;
            const cplx hd[3] = {hl[0] * dg[0], hl[1] * dg[4], hl[2] * dg[8]};
#pragma unroll
            for (int r = 0; r < 3; r++)
                H[r] = hconst * (basis[3 * r] * hd[0] + basis[3 * r + 1] * hd[1] + basis[3 * r + 2] * hd[2]);
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) Eout[3 * p + r] = E[r];
    if (Hout) {
#pragma unroll
        for (int r = 0; r < 3; r++) Hout[3 * p + r] = H[r];
    }
}

__global__ void k_shift(int64_t nn, const cplx *__restrict__ A, const cplx *__restrict__ B, cplx sigma,
                        cplx *__restrict__ M) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nn) M[i] = A[i] - sigma * B[i];
}

// w /= nrm, with nrm = H[j + 1, j]; a zero norm (Krylov breakdown) leaves w zero
__global__ void k_scale_by(int64_t n, const double *__restrict__ nrm, cplx *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const double s = *nrm;
    if (i < n) w[i] = s > 1e-300 ? (1.0 / s) * w[i] : cplx();
}

// H[0..j, j] += h, and the subdiagonal entry from the real norm
__global__ void k_acc(int j, const cplx *__restrict__ h, cplx *__restrict__ Hcol) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i <= j) Hcol[i] += h[i];
}

__global__ void k_put_real(const double *__restrict__ v, cplx *__restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = cplx(*v, 0.0);
}

struct Blas {
    std::mutex mu;
    std::vector<rocblas_handle> h;
};
Blas g_blas;

// one rocBLAS handle per device, bound to the caller's stream on every use; the caller
// holds g_blas.mu for as long as it uses the handle (entry points on one device from
// several host threads run one at a time)
rocblas_handle blas_for(hipStream_t stream, int *err) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { *err = 1; return nullptr; }
    if ((int)g_blas.h.size() <= dev) g_blas.h.resize(dev + 1, nullptr);
    if (!g_blas.h[dev] && rocblas_create_handle(&g_blas.h[dev]) != rocblas_status_success) { *err = 2; return nullptr; }
    if (rocblas_set_stream(g_blas.h[dev], stream) != rocblas_status_success) { *err = 3; return nullptr; }
    rocblas_set_pointer_mode(g_blas.h[dev], rocblas_pointer_mode_host);
    *err = 0;
    return g_blas.h[dev];
}

#define EMG_BLAS(call)                                                                          \
    do {                                                                                        \
        rocblas_status _s = (call);                                                             \
        if (_s != rocblas_status_success) {                                                     \
            emg::set_error(std::string(#call) + ": " + rocblas_status_to_string(_s));          \
            return -2;                                                                          \
        }                                                                                       \
    } while (0)

inline rocblas_operation op_of(int t) {
    return t == 0 ? rocblas_operation_none : (t == 1 ? rocblas_operation_transpose : rocblas_operation_conjugate_transpose);
}

}  // namespace

extern "C" int emg_port_assemble(int64_t ntri, const double *d_verts, const double *d_Ms, const double *d_Mm, double k0,
                                 int flags, const int64_t *d_t2r, int64_t n, double *d_A, double *d_B, double *d_Ael,
                                 double *d_Bel, void *stream) {
    EMG_CHECK(d_verts && d_Ms && d_Mm, "emg_port_assemble: null geometry or material");
    EMG_CHECK(ntri >= 0 && ntri < (1ll << 40) / 196, "emg_port_assemble: bad triangle count");
    EMG_CHECK((d_A == nullptr) == (d_B == nullptr) && (d_Ael == nullptr) == (d_Bel == nullptr),
              "emg_port_assemble: A and B (and Ael and Bel) go together");
    EMG_CHECK(!d_A || (d_t2r && n > 0), "emg_port_assemble: dense output needs t2r and n > 0");
    EMG_CHECK((flags & ~EMG_PORT_FACE_BLOCK) == 0, "emg_port_assemble: unknown flag");
    if (ntri == 0) return 0;
    const int64_t nt = ntri * 196;
    k_port_element<<<emg::grid_for(nt, EL_BLOCK), EL_BLOCK, 0, (hipStream_t)stream>>>(
        ntri, d_verts, reinterpret_cast<const cplx *>(d_Ms), reinterpret_cast<const cplx *>(d_Mm), k0,
        (flags & EMG_PORT_FACE_BLOCK) ? 1 : 0, d_t2r, n, d_A, d_B, reinterpret_cast<cplx *>(d_Ael),
        reinterpret_cast<cplx *>(d_Bel));
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_port_field(int64_t np, const double *d_pts, int64_t ntri, const int64_t *d_tris, int64_t nnodes,
                              const double *d_nodes, const int64_t *d_t2f, int64_t nfield, const double *d_field,
                              double beta,
                              const double *d_diadic, const double *d_basis, double hconst_re, double hconst_im,
                              double *d_E, double *d_H, void *stream) {
    EMG_CHECK(d_pts && d_tris && d_nodes && d_t2f && d_field && d_basis && d_E, "emg_port_field: null argument");
    EMG_CHECK(!d_H || d_diadic, "emg_port_field: H needs the diadic");
    EMG_CHECK(np >= 0 && ntri >= 0 && nnodes >= 0 && nfield >= 0, "emg_port_field: bad sizes");
    if (np == 0) return 0;
    k_port_field<<<emg::grid_for(np, FLD_BLOCK), FLD_BLOCK, 0, (hipStream_t)stream>>>(
        np, d_pts, ntri, d_tris, nnodes, d_nodes, d_t2f, nfield, reinterpret_cast<const cplx *>(d_field), beta,
        reinterpret_cast<const cplx *>(d_diadic), d_basis, cplx(hconst_re, hconst_im), reinterpret_cast<cplx *>(d_E),
        reinterpret_cast<cplx *>(d_H));
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_port_factor(int64_t n, const double *d_A, const double *d_B, double sigma_re, double sigma_im,
                               double *d_M, int *d_ipiv, void *stream) {
    EMG_CHECK(d_A && d_B && d_M && d_ipiv, "emg_port_factor: null argument");
    EMG_CHECK(n > 0 && n < (1 << 30) && n * n < (1ll << 40), "emg_port_factor: bad size");
    hipStream_t s = (hipStream_t)stream;
    int err = 0;
    std::lock_guard<std::mutex> blas_lock(g_blas.mu);
    rocblas_handle h = blas_for(s, &err);
    EMG_CHECK(h, "emg_port_factor: no rocBLAS handle");
    k_shift<<<emg::grid_for(n * n, 256), 256, 0, s>>>(n * n, reinterpret_cast<const cplx *>(d_A),
                                                     reinterpret_cast<const cplx *>(d_B), cplx(sigma_re, sigma_im),
                                                     reinterpret_cast<cplx *>(d_M));
    EMG_LAUNCH_CHECK();
    int *d_info = nullptr;
    EMG_HIP(hipMallocAsync((void **)&d_info, sizeof(int), s));
    rocblas_status st = rocsolver_zgetrf(h, (rocblas_int)n, (rocblas_int)n,
                                         reinterpret_cast<rocblas_double_complex *>(d_M), (rocblas_int)n, d_ipiv, d_info);
    int info = 0;
    hipError_t e1 = hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFreeAsync(d_info, s);
    EMG_BLAS(st);
    EMG_HIP(e1);
    EMG_HIP(e2);
    EMG_CHECK(info == 0, "emg_port_factor: A - sigma B is singular at the shift");
    return 0;
}

extern "C" int emg_port_operator(int64_t n, const double *d_M, const int *d_ipiv, const double *d_B, double *d_T,
                                 void *stream) {
    EMG_CHECK(d_M && d_ipiv && d_B && d_T && d_T != d_B && d_T != d_M, "emg_port_operator: null or aliased argument");
    EMG_CHECK(n > 0 && n < (1 << 30) && n * n < (1ll << 40), "emg_port_operator: bad size");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> blas_lock(g_blas.mu);
    int err = 0;
    rocblas_handle h = blas_for(s, &err);
    EMG_CHECK(h, "emg_port_operator: no rocBLAS handle");
    using Z = rocblas_double_complex;
    EMG_HIP(hipMemcpyAsync(d_T, d_B, sizeof(Z) * n * n, hipMemcpyDeviceToDevice, s));
    EMG_BLAS(rocsolver_zgetrs(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                              const_cast<Z *>(reinterpret_cast<const Z *>(d_M)), (rocblas_int)n, d_ipiv,
                              reinterpret_cast<Z *>(d_T), (rocblas_int)n));
    return 0;
}

extern "C" int emg_port_krylov(int64_t n, const double *d_T, int m, double *d_V, double *d_H, void *stream) {
    EMG_CHECK(d_T && d_V && d_H, "emg_port_krylov: null argument");
    EMG_CHECK(n > 0 && n < (1 << 30) && m >= 1 && m < n + 1 && m < 4096, "emg_port_krylov: bad sizes");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> blas_lock(g_blas.mu);
    int err = 0;
    rocblas_handle h = blas_for(s, &err);
    EMG_CHECK(h, "emg_port_krylov: no rocBLAS handle");
    using Z = rocblas_double_complex;
    const Z *T = reinterpret_cast<const Z *>(d_T);
    Z *V = reinterpret_cast<Z *>(d_V);
    Z *H = reinterpret_cast<Z *>(d_H);
    const int rows = m + 1;
    double *d_nrm = nullptr;
    Z *d_h = nullptr;
    EMG_HIP(hipMallocAsync((void **)&d_nrm, sizeof(double), s));
    hipError_t ea = hipMallocAsync((void **)&d_h, sizeof(Z) * rows, s);
    if (ea != hipSuccess) { (void)hipFreeAsync(d_nrm, s); EMG_HIP(ea); }
    int rc = 0;
    const Z one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
    auto fail = [&](rocblas_status st, const char *what) {
        if (st != rocblas_status_success && rc == 0) {
            emg::set_error(std::string("emg_port_krylov: ") + what + ": " + rocblas_status_to_string(st));
            rc = -2;
        }
    };
    auto nrm_into = [&](const Z *x, double *dst) {
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_device);
        fail(rocblas_dznrm2(h, (rocblas_int)n, x, 1, dst), "dznrm2");
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
    };
    if (hipMemsetAsync(H, 0, sizeof(Z) * rows * m, s) != hipSuccess) fail(rocblas_status_internal_error, "clearing H");
    nrm_into(V, d_nrm);
    k_scale_by<<<emg::grid_for(n, 256), 256, 0, s>>>(n, d_nrm, reinterpret_cast<cplx *>(V));
    for (int j = 0; j < m && rc == 0; j++) {
        Z *w = V + (int64_t)(j + 1) * n;
        // w = T v_j, T = (A - sigma B)^{-1} B formed once by emg_port_operator
        fail(rocblas_zgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, &one, T, (rocblas_int)n,
                           V + (int64_t)j * n, 1, &zero, w, 1), "zgemv T v");
        for (int pass = 0; pass < 2; pass++) {
            fail(rocblas_zgemv(h, rocblas_operation_conjugate_transpose, (rocblas_int)n, j + 1, &one, V, (rocblas_int)n,
                               w, 1, &zero, d_h, 1), "zgemv V^H w");
            fail(rocblas_zgemv(h, rocblas_operation_none, (rocblas_int)n, j + 1, &mone, V, (rocblas_int)n, d_h, 1, &one,
                               w, 1), "zgemv w - V h");
            k_acc<<<emg::grid_for(j + 1, 256), 256, 0, s>>>(j, reinterpret_cast<const cplx *>(d_h),
                                                            reinterpret_cast<cplx *>(H + (int64_t)j * rows));
        }
        nrm_into(w, d_nrm);
        k_put_real<<<1, 64, 0, s>>>(d_nrm, reinterpret_cast<cplx *>(H + (int64_t)j * rows + j + 1));
        k_scale_by<<<emg::grid_for(n, 256), 256, 0, s>>>(n, d_nrm, reinterpret_cast<cplx *>(w));
    }
    hipError_t le = hipGetLastError();
    (void)hipFreeAsync(d_nrm, s);
    (void)hipFreeAsync(d_h, s);
    if (rc) return rc;
    if (le != hipSuccess) {
        emg::set_error(std::string("emg_port_krylov: ") + hipGetErrorString(le));
        return -2;
    }
    return 0;
}

extern "C" int emg_port_inverse_step(int64_t n, const double *d_M, const int *d_ipiv, const double *d_B, int64_t k,
                              const double *d_X, double *d_Y, void *stream) {
    EMG_CHECK(d_M && d_ipiv && d_B && d_X && d_Y && d_X != d_Y, "emg_port_inverse_step: null or aliased argument");
    EMG_CHECK(n > 0 && n < (1 << 30) && k >= 0 && k < (1 << 30), "emg_port_inverse_step: bad sizes");
    if (k == 0) return 0;
    int err = 0;
    std::lock_guard<std::mutex> blas_lock(g_blas.mu);
    rocblas_handle h = blas_for((hipStream_t)stream, &err);
    EMG_CHECK(h, "emg_port_inverse_step: no rocBLAS handle");
    using Z = rocblas_double_complex;
    const Z one(1.0, 0.0), zero(0.0, 0.0);
    EMG_BLAS(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)k,
                           (rocblas_int)n, &one, reinterpret_cast<const Z *>(d_B), (rocblas_int)n,
                           reinterpret_cast<const Z *>(d_X), (rocblas_int)n, &zero, reinterpret_cast<Z *>(d_Y),
                           (rocblas_int)n));
    EMG_BLAS(rocsolver_zgetrs(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)k,
                              const_cast<Z *>(reinterpret_cast<const Z *>(d_M)), (rocblas_int)n, d_ipiv,
                              reinterpret_cast<Z *>(d_Y), (rocblas_int)n));
    return 0;
}

extern "C" int emg_zgemm(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha_re, double alpha_im,
                         const double *d_A, int64_t lda, const double *d_B, int64_t ldb, double beta_re, double beta_im,
                         double *d_C, int64_t ldc, void *stream) {
    EMG_CHECK(d_A && d_B && d_C, "emg_zgemm: null argument");
    EMG_CHECK(transa >= 0 && transa <= 2 && transb >= 0 && transb <= 2, "emg_zgemm: trans is 0 (N), 1 (T) or 2 (C)");
    EMG_CHECK(m >= 0 && n >= 0 && k >= 0 && m < (1 << 30) && n < (1 << 30) && k < (1 << 30), "emg_zgemm: bad sizes");
    EMG_CHECK(lda >= 1 && ldb >= 1 && ldc >= (m > 1 ? m : 1), "emg_zgemm: bad leading dimension");
    if (m == 0 || n == 0) return 0;
    int err = 0;
    std::lock_guard<std::mutex> blas_lock(g_blas.mu);
    rocblas_handle h = blas_for((hipStream_t)stream, &err);
    EMG_CHECK(h, "emg_zgemm: no rocBLAS handle");
    using Z = rocblas_double_complex;
    const Z alpha(alpha_re, alpha_im), beta(beta_re, beta_im);
    EMG_BLAS(rocblas_zgemm(h, op_of(transa), op_of(transb), (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &alpha,
                           reinterpret_cast<const Z *>(d_A), (rocblas_int)lda, reinterpret_cast<const Z *>(d_B),
                           (rocblas_int)ldb, &beta, reinterpret_cast<Z *>(d_C), (rocblas_int)ldc));
    return 0;
}
