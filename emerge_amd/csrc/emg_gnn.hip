// Learned smoother weights: a graph network over the mesh's vertices and edges.
//
// The network maps per-vertex features (boundary flags, k0 * local edge
// length, material) through an encoder and message-passing layers over the
// mesh edges to per-entity weights of the hybrid smoother:
//   z = w_e . BJ(r) + G (w_p . diag(G^T K G)^{-1} . G^T r),
// w_e one complex weight per edge / face entity, w_p one per P2 node (vertex or
// edge).  Weights multiply the smoother's tables once per frequency
// (emg_solver_set_weights), so the learned preconditioner costs nothing per
// GMRES iteration.  The reference has no counterpart: it factorises directly
// (fem/solver.py:294-309).
//
// Layout: activations float32 row-major (rows x channels), layer weights
// float16 (k x m, row-major) with float32 accumulation, gradients float32.
// Every kernel here is used both for inference and for training (forward and
// backward); the host drives the layer sequence.

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "../../include/emerge_amd.h"
#include "emg_common.h"

namespace {

constexpr int GB = 256;
constexpr int GK_MAX = 128;   // max input / output width of a dense layer

// C[n x m] = A[n x k] W[k x m] + bias + beta * C   (W float16, staged in LDS as float)
__global__ void __launch_bounds__(GB) k_gemm_nn(int64_t n, int k, int m, const float *__restrict__ A,
                                                const __half *__restrict__ W, const float *__restrict__ bias,
                                                float beta, float *__restrict__ C) {
    __shared__ float sw[GK_MAX * 64];
    const int m0 = blockIdx.y * 64;
    const int mw = min(64, m - m0);
    for (int t = threadIdx.x; t < k * 64; t += GB) {
        const int i = t / 64, j = t % 64;
        sw[t] = j < mw ? __half2float(W[(int64_t)i * m + m0 + j]) : 0.0f;
    }
    __syncthreads();
    const int64_t row = blockIdx.x * (int64_t)(GB / 64) + threadIdx.x / 64;
    const int j = threadIdx.x % 64;
    if (row >= n || j >= mw) return;
    const float *a = A + row * k;
    float acc = bias ? bias[m0 + j] : 0.0f;
    for (int i = 0; i < k; i++) acc += a[i] * sw[i * 64 + j];
    float *c = C + row * m + m0 + j;
    *c = beta != 0.0f ? acc + beta * (*c) : acc;
}

// out[n x k] = D[n x m] W^T + beta * out   (W k x m float16)
__global__ void __launch_bounds__(GB) k_gemm_nt(int64_t n, int m, int k, const float *__restrict__ D,
                                                const __half *__restrict__ W, float beta, float *__restrict__ out) {
    __shared__ float sw[GK_MAX * 64];
    const int k0 = blockIdx.y * 64;
    const int kw = min(64, k - k0);
    for (int t = threadIdx.x; t < 64 * m; t += GB) {
        const int i = t / m, j = t % m;          // i: output column (row of W), j: reduction
        sw[j * 64 + i] = i < kw ? __half2float(W[(int64_t)(k0 + i) * m + j]) : 0.0f;
    }
    __syncthreads();
    const int64_t row = blockIdx.x * (int64_t)(GB / 64) + threadIdx.x / 64;
    const int i = threadIdx.x % 64;
    if (row >= n || i >= kw) return;
    const float *d = D + row * m;
    float acc = 0.0f;
    for (int j = 0; j < m; j++) acc += d[j] * sw[j * 64 + i];
    float *o = out + row * k + k0 + i;
    *o = beta != 0.0f ? acc + beta * (*o) : acc;
}

// dW[k x m] += A^T D over rows (A n x k, D n x m); dbias[m] += column sums of D.
// Each block reduces a 256-row slab in registers and adds its partial sums once.
constexpr int TN_ROWS = 256;
__global__ void __launch_bounds__(GB) k_gemm_tn(int64_t n, int k, int m, const float *__restrict__ A,
                                                const float *__restrict__ D, float *__restrict__ dW,
                                                float *__restrict__ dbias) {
    const int64_t r0 = blockIdx.x * (int64_t)TN_ROWS;
    const int64_t r1 = min<int64_t>(n, r0 + TN_ROWS);
    for (int o = threadIdx.x; o < k * m + (dbias ? m : 0); o += GB) {
        float acc = 0.0f;
        if (o < k * m) {
            const int i = o / m, j = o % m;
            for (int64_t r = r0; r < r1; r++) acc += A[r * k + i] * D[r * m + j];
            atomicAdd(&dW[o], acc);
        } else {
            const int j = o - k * m;
            for (int64_t r = r0; r < r1; r++) acc += D[r * m + j];
            atomicAdd(&dbias[j], acc);
        }
    }
}

// Y[v] = rs[v] * sum_{u in N(v)} cs[u] X[u]   (C channels; rs / cs may be null = 1)
__global__ void __launch_bounds__(GB) k_agg(int64_t nv, int C, const int32_t *__restrict__ ptr,
                                            const int32_t *__restrict__ idx, const float *__restrict__ rs,
                                            const float *__restrict__ cs, const float *__restrict__ X,
                                            float *__restrict__ Y) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t >= nv * C) return;
    const int64_t v = t / C;
    const int c = (int)(t % C);
    float acc = 0.0f;
    for (int32_t e = ptr[v]; e < ptr[v + 1]; e++) {
        const int32_t u = idx[e];
        acc += (cs ? cs[u] : 1.0f) * X[(int64_t)u * C + c];
    }
    Y[t] = (rs ? rs[v] : 1.0f) * acc;
}

// H = relu(P) + Hprev
__global__ void __launch_bounds__(GB) k_relu_res(int64_t n, const float *__restrict__ P, const float *__restrict__ Hp,
                                                 float *__restrict__ H) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t < n) H[t] = fmaxf(P[t], 0.0f) + Hp[t];
}

// dP = dH . (P > 0)
__global__ void __launch_bounds__(GB) k_relu_bwd(int64_t n, const float *__restrict__ P, const float *__restrict__ dH,
                                                 float *__restrict__ dP) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t < n) dP[t] = P[t] > 0.0f ? dH[t] : 0.0f;
}

// O[e, c] = bias[c] + sum_q Hp[v_q(e), q * C + c]   (vertex projections, ordered by slot q)
__global__ void __launch_bounds__(GB) k_readout(int64_t ne, int K, int C, const int32_t *__restrict__ ev,
                                                const float *__restrict__ Hp, int ldh, int off,
                                                const float *__restrict__ bias, float *__restrict__ O) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t >= ne * C) return;
    const int64_t e = t / C;
    const int c = (int)(t % C);
    float acc = bias ? bias[c] : 0.0f;
    for (int q = 0; q < K; q++) acc += Hp[(int64_t)ev[e * K + q] * ldh + off + q * C + c];
    O[t] = acc;
}

// dHp[v_q(e), q * C + c] += dO[e, c];  dbias[c] += sum_e dO[e, c]
__global__ void __launch_bounds__(GB) k_readout_bwd(int64_t ne, int K, int C, const int32_t *__restrict__ ev,
                                                    const float *__restrict__ dO, int ldh, int off,
                                                    float *__restrict__ dHp, float *__restrict__ dbias) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t >= ne * C) return;
    const int64_t e = t / C;
    const int c = (int)(t % C);
    const float g = dO[t];
    for (int q = 0; q < K; q++) atomicAdd(&dHp[(int64_t)ev[e * K + q] * ldh + off + q * C + c], g);
    if (dbias) atomicAdd(&dbias[c], g);
}

// Weights from the readout, complex128: mode 0 w = 1 + O0 + j O1, mode 1 w = exp(O0) (real, positive),
// mode 2 w = exp(0.5 tanh(O0)) (real, in [0.61, 1.65]).
__global__ void __launch_bounds__(GB) k_weights(int64_t n, int mode, const float *__restrict__ O, double *__restrict__ w) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t >= n) return;
    if (mode == 1) {
        w[2 * t] = exp((double)O[2 * t]);
        w[2 * t + 1] = 0.0;
    } else if (mode == 2) {
        w[2 * t] = exp(0.5 * tanh((double)O[2 * t]));
        w[2 * t + 1] = 0.0;
    } else {
        w[2 * t] = 1.0 + (double)O[2 * t];
        w[2 * t + 1] = (double)O[2 * t + 1];
    }
}

// dO from d(Re w, Im w) (2 floats per row): mode 0 identity, mode 1 dO0 = dRe(w) exp(O0), dO1 = 0.
__global__ void __launch_bounds__(GB) k_weights_bwd(int64_t n, int mode, const float *__restrict__ O,
                                                    const float *__restrict__ dw, float *__restrict__ dO) {
    const int64_t t = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (t >= n) return;
    if (mode == 1) {
        dO[2 * t] = dw[2 * t] * expf(O[2 * t]);
        dO[2 * t + 1] = 0.0f;
    } else if (mode == 2) {
        const float th = tanhf(O[2 * t]);
        dO[2 * t] = dw[2 * t] * expf(0.5f * th) * 0.5f * (1.0f - th * th);
        dO[2 * t + 1] = 0.0f;
    } else {
        dO[2 * t] = dw[2 * t];
        dO[2 * t + 1] = dw[2 * t + 1];
    }
}

// dO[i] (2 floats) += scale * 2 (Re, Im)(sum_s conj(a_s[i]) b_s[i]) for ns samples laid out
// back to back; entity i of a DOF vector owns DOFs i and i + stride (stride = 0: one DOF).
__global__ void __launch_bounds__(GB) k_weight_grad(int64_t n, int ns, int64_t ld, int64_t stride,
                                                    const double *__restrict__ a, const double *__restrict__ b,
                                                    const double *__restrict__ scale, float *__restrict__ dO) {
    const int64_t i = blockIdx.x * (int64_t)GB + threadIdx.x;
    if (i >= n) return;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < ns; s++) {
        const double *as = a + 2 * ld * s, *bs = b + 2 * ld * s;
        for (int q = 0; q < (stride ? 2 : 1); q++) {
            const int64_t d = i + q * stride;
            const double ar = as[2 * d], ai = as[2 * d + 1], br = bs[2 * d], bi = bs[2 * d + 1];
            re += scale[s] * (ar * br + ai * bi);
            im += scale[s] * (ar * bi - ai * br);
        }
    }
    dO[2 * i] += (float)(2.0 * re);
    dO[2 * i + 1] += (float)(2.0 * im);
}

inline unsigned g1(int64_t n) { return emg::grid_for(n, GB); }

}  // namespace

extern "C" int emg_gnn_gemm(int64_t n, int k, int m, const float *d_A, const void *d_W, const float *d_bias,
                            float beta, float *d_C, void *stream) {
    EMG_CHECK(n >= 0 && k > 0 && k <= GK_MAX && m > 0 && d_A && d_W && d_C, "emg_gnn_gemm: bad argument");
    if (n == 0) return 0;
    dim3 grid((unsigned)((n + GB / 64 - 1) / (GB / 64)), (unsigned)((m + 63) / 64));
    k_gemm_nn<<<grid, GB, 0, (hipStream_t)stream>>>(n, k, m, d_A, (const __half *)d_W, d_bias, beta, d_C);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_gemm_nt(int64_t n, int m, int k, const float *d_D, const void *d_W, float beta, float *d_out,
                               void *stream) {
    EMG_CHECK(n >= 0 && m > 0 && m <= GK_MAX && k > 0 && d_D && d_W && d_out, "emg_gnn_gemm_nt: bad argument");
    if (n == 0) return 0;
    dim3 grid((unsigned)((n + GB / 64 - 1) / (GB / 64)), (unsigned)((k + 63) / 64));
    k_gemm_nt<<<grid, GB, 0, (hipStream_t)stream>>>(n, m, k, d_D, (const __half *)d_W, beta, d_out);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_gemm_tn(int64_t n, int k, int m, const float *d_A, const float *d_D, float *d_dW,
                               float *d_dbias, void *stream) {
    EMG_CHECK(n >= 0 && k > 0 && m > 0 && d_A && d_D && d_dW, "emg_gnn_gemm_tn: bad argument");
    if (n == 0) return 0;
    k_gemm_tn<<<(unsigned)((n + TN_ROWS - 1) / TN_ROWS), GB, 0, (hipStream_t)stream>>>(n, k, m, d_A, d_D, d_dW, d_dbias);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_agg(int64_t nv, int C, const int32_t *d_ptr, const int32_t *d_idx, const float *d_rs,
                           const float *d_cs, const float *d_X, float *d_Y, void *stream) {
    EMG_CHECK(nv >= 0 && C > 0 && d_ptr && d_idx && d_X && d_Y, "emg_gnn_agg: bad argument");
    if (nv == 0) return 0;
    k_agg<<<g1(nv * C), GB, 0, (hipStream_t)stream>>>(nv, C, d_ptr, d_idx, d_rs, d_cs, d_X, d_Y);
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_relu(int64_t n, const float *d_P, const float *d_Hprev, float *d_H, const float *d_dH,
                            float *d_dP, void *stream) {
    EMG_CHECK(n >= 0 && d_P, "emg_gnn_relu: bad argument");
    if (n == 0) return 0;
    if (d_dH) {
        EMG_CHECK(d_dP != nullptr, "emg_gnn_relu: backward needs d_dP");
        k_relu_bwd<<<g1(n), GB, 0, (hipStream_t)stream>>>(n, d_P, d_dH, d_dP);
    } else {
        EMG_CHECK(d_Hprev && d_H, "emg_gnn_relu: forward needs d_Hprev and d_H");
        k_relu_res<<<g1(n), GB, 0, (hipStream_t)stream>>>(n, d_P, d_Hprev, d_H);
    }
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_readout(int64_t ne, int K, int C, const int32_t *d_ev, const float *d_Hp, int ldh, int off,
                               const float *d_bias, float *d_O, const float *d_dO, float *d_dHp, float *d_dbias,
                               void *stream) {
    EMG_CHECK(ne >= 0 && K >= 1 && K <= 3 && C > 0 && d_ev && off + K * C <= ldh, "emg_gnn_readout: bad argument");
    if (ne == 0) return 0;
    if (d_dO) {
        EMG_CHECK(d_dHp != nullptr, "emg_gnn_readout: backward needs d_dHp");
        k_readout_bwd<<<g1(ne * C), GB, 0, (hipStream_t)stream>>>(ne, K, C, d_ev, d_dO, ldh, off, d_dHp, d_dbias);
    } else {
        EMG_CHECK(d_Hp && d_O, "emg_gnn_readout: forward needs d_Hp and d_O");
        k_readout<<<g1(ne * C), GB, 0, (hipStream_t)stream>>>(ne, K, C, d_ev, d_Hp, ldh, off, d_bias, d_O);
    }
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_weights(int64_t n, int mode, const float *d_O, double *d_w, const float *d_dw, float *d_dO,
                               void *stream) {
    EMG_CHECK(n >= 0 && mode >= 0 && mode <= 2 && d_O, "emg_gnn_weights: bad argument");
    if (n == 0) return 0;
    if (d_dw) {
        EMG_CHECK(d_dO != nullptr, "emg_gnn_weights: backward needs d_dO");
        k_weights_bwd<<<g1(n), GB, 0, (hipStream_t)stream>>>(n, mode, d_O, d_dw, d_dO);
    } else {
        EMG_CHECK(d_w != nullptr, "emg_gnn_weights: forward needs d_w");
        k_weights<<<g1(n), GB, 0, (hipStream_t)stream>>>(n, mode, d_O, d_w);
    }
    EMG_LAUNCH_CHECK();
    return 0;
}

extern "C" int emg_gnn_weight_grad(int64_t n, int ns, int64_t ld, int64_t stride, const double *d_a, const double *d_b,
                                   const double *d_scale, float *d_dO, void *stream) {
    EMG_CHECK(n >= 0 && ns > 0 && d_a && d_b && d_scale && d_dO, "emg_gnn_weight_grad: bad argument");
    if (n == 0) return 0;
    k_weight_grad<<<g1(n), GB, 0, (hipStream_t)stream>>>(n, ns, ld, stride, d_a, d_b, d_scale, d_dO);
    EMG_LAUNCH_CHECK();
    return 0;
}
