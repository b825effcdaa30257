// Far field by the Stratton-Chu radiation integral (fem/physics/edm/sc.py:27-180).
//
// For every direction r = (cos t cos p, cos t sin p, sin t) and the sampled surface
// currents at points v_j with area-weighted normals w_j:
//   e_j = (w_j x E_j - Z0 r x (w_j x H_j)) exp(j k0 r.v_j)
//   E(r) = -j k0 / (4 pi) sum_j r x e_j,        H(r) = r x E(r) / Z0.
// One workgroup per direction; its lanes stride over the source points and the
// partial sums are reduced in LDS.  FP64 throughout (the reference evaluates the
// same sum in complex64).

#include <cmath>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

using emg::cplx;

namespace {

constexpr int FF_BLOCK = 256;
constexpr double Z0_FF = 376.73031366857;   // the reference's constant (sc.py:97)

__global__ void __launch_bounds__(FF_BLOCK) k_farfield(int64_t ns, const double *__restrict__ pts,
                                                       const double *__restrict__ wn, const cplx *__restrict__ E,
                                                       const cplx *__restrict__ H, const double *__restrict__ theta,
                                                       const double *__restrict__ phi, double k0,
                                                       cplx *__restrict__ Eout, cplx *__restrict__ Hout) {
    const int64_t d = blockIdx.x;
    const double ct = cos(theta[d]), st = sin(theta[d]), cp = cos(phi[d]), sp = sin(phi[d]);
    const double rx = ct * cp, ry = ct * sp, rz = st;
    double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;   // sum of e_j: re (a), im (b)
    for (int64_t j = threadIdx.x; j < ns; j += FF_BLOCK) {
        const double vx = pts[3 * j], vy = pts[3 * j + 1], vz = pts[3 * j + 2];
        const double nx = wn[3 * j], ny = wn[3 * j + 1], nz = wn[3 * j + 2];
        const cplx ex = E[3 * j], ey = E[3 * j + 1], ez = E[3 * j + 2];
        const cplx hx = H[3 * j], hy = H[3 * j + 1], hz = H[3 * j + 2];
        // n x E and n x H
        const cplx nEx = ny * ez - nz * ey, nEy = nz * ex - nx * ez, nEz = nx * ey - ny * ex;
        const cplx nHx = ny * hz - nz * hy, nHy = nz * hx - nx * hz, nHz = nx * hy - ny * hx;
        // r x (n x H)
        const cplx rHx = ry * nHz - rz * nHy, rHy = rz * nHx - rx * nHz, rHz = rx * nHy - ry * nHx;
        const cplx jx = nEx - Z0_FF * rHx, jy = nEy - Z0_FF * rHy, jz = nEz - Z0_FF * rHz;
        double s, c;
        sincos(k0 * (rx * vx + ry * vy + rz * vz), &s, &c);
        const cplx g(c, s);
        const cplx gx = jx * g, gy = jy * g, gz = jz * g;
        ax += gx.re; bx += gx.im;
        ay += gy.re; by += gy.im;
        az += gz.re; bz += gz.im;
    }
    __shared__ double red[6][FF_BLOCK];
    red[0][threadIdx.x] = ax; red[1][threadIdx.x] = bx;
    red[2][threadIdx.x] = ay; red[3][threadIdx.x] = by;
    red[4][threadIdx.x] = az; red[5][threadIdx.x] = bz;
    __syncthreads();
    for (int o = FF_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int q = 0; q < 6; q++) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const cplx Sx(red[0][0], red[1][0]), Sy(red[2][0], red[3][0]), Sz(red[4][0], red[5][0]);
        const cplx Q(0.0, -k0 / (4.0 * M_PI));
        const cplx Ex = Q * (ry * Sz - rz * Sy), Ey = Q * (rz * Sx - rx * Sz), Ez = Q * (rx * Sy - ry * Sx);
        Eout[3 * d] = Ex; Eout[3 * d + 1] = Ey; Eout[3 * d + 2] = Ez;
        const double iz = 1.0 / Z0_FF;
        Hout[3 * d] = (ry * Ez - rz * Ey) * iz;
        Hout[3 * d + 1] = (rz * Ex - rx * Ez) * iz;
        Hout[3 * d + 2] = (rx * Ey - ry * Ex) * iz;
    }
}

}  // namespace

extern "C" int emg_farfield(int64_t ns, const double *d_pts, const double *d_wn, const double *d_E, const double *d_H,
                            int64_t ndir, const double *d_theta, const double *d_phi, double k0, double *d_Eout,
                            double *d_Hout, void *stream) {
    EMG_CHECK(d_pts && d_wn && d_E && d_H && d_theta && d_phi && d_Eout && d_Hout, "emg_farfield: null argument");
    EMG_CHECK(ns >= 0 && ndir >= 0 && ndir < (1ll << 31), "emg_farfield: bad sizes");
    if (ndir == 0) return 0;
    k_farfield<<<(unsigned)ndir, FF_BLOCK, 0, (hipStream_t)stream>>>(
        ns, d_pts, d_wn, reinterpret_cast<const cplx *>(d_E), reinterpret_cast<const cplx *>(d_H), d_theta, d_phi, k0,
        reinterpret_cast<cplx *>(d_Eout), reinterpret_cast<cplx *>(d_Hout));
    EMG_LAUNCH_CHECK();
    return 0;
}
