// Shared helpers for the emerge_amd HIP library: error reporting, launch
// checks and a small complex type for device code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace emg {

void set_error(const std::string &msg);

#define EMG_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            emg::set_error(std::string(#call) + ": " + hipGetErrorString(_e));         \
            return -2;                                                                 \
        }                                                                              \
    } while (0)

#define EMG_CHECK(cond, msg)                                                           \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            emg::set_error(msg);                                                       \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

#define EMG_LAUNCH_CHECK()                                                             \
    do {                                                                               \
        hipError_t _e = hipGetLastError();                                             \
        if (_e != hipSuccess) {                                                        \
            emg::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));    \
            return -2;                                                                 \
        }                                                                              \
    } while (0)

struct cplx {
    double re, im;
    __host__ __device__ cplx() : re(0), im(0) {}
    __host__ __device__ cplx(double r, double i = 0) : re(r), im(i) {}
};
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return cplx(a.re + b.re, a.im + b.im); }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return cplx(a.re - b.re, a.im - b.im); }
__host__ __device__ inline cplx operator*(cplx a, cplx b) { return cplx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__host__ __device__ inline cplx operator*(double s, cplx a) { return cplx(s * a.re, s * a.im); }
__host__ __device__ inline cplx operator*(cplx a, double s) { return cplx(s * a.re, s * a.im); }
__host__ __device__ inline cplx &operator+=(cplx &a, cplx b) { a.re += b.re; a.im += b.im; return a; }

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace emg
