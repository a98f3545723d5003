// norm_device.hpp - the Simpson norm of psi from the z sweep's column
// partials (calcnorm, imag3d.cu:1194-1233), as a device function of one
// workgroup: used by the norm kernel and, folded into the next step, by the
// first workgroup of the forward z pass.
#pragma once

#include "internal.hpp"

namespace dbec {

// norm^2 = sum_x wx[x] sum_t rows[t][x] with NT threads (any count), sh[NT]
// LDS scratch; thread 0 writes scal[NORM] (the reference's pre-rescale
// norm), scal[SCALE] = 1 / norm, scal[LASTNORM] and, with the exponent shift
// c of the mixed form, scal[SHIFT] = c - log(norm): psi carried an extra
// factor exp(c), so the reference's norm is norm * exp(-c), and the new shift
// brings the next step's norm to about 1 (c follows dt * mu).
template <int NT>
__device__ __forceinline__ void norm_from_cols(const double *__restrict__ rows, int tiles,
                                               const double *__restrict__ wx, int Nx, double *__restrict__ scal,
                                               int shift, double *sh) {
    const int tid = threadIdx.x;
    double p = 0.0;
    for (int x = tid; x < Nx; x += NT) {
        double cs = 0.0;
        for (int t = 0; t < tiles; t++) cs += rows[(long)t * Nx + x];
        p += wx[x] * cs;
    }
    sh[tid] = p;
    __syncthreads();
    constexpr int P2 = NT <= 1 ? 1 : (NT <= 64 ? 64 : (NT <= 128 ? 128 : (NT <= 256 ? 256 : (NT <= 512 ? 512 : 1024))));
    for (int o = P2 / 2; o > 0; o >>= 1) {
        if (tid < o && tid + o < NT) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid == 0) {
        const double norm = sqrt(sh[0]);
        if (shift) {
            const double c = scal[SC_SHIFT];
            scal[SC_NORM] = norm * exp(-c);
            scal[SC_SHIFT] = c - log(norm);
        } else {
            scal[SC_NORM] = norm;
        }
        scal[SC_SCALE] = 1.0 / norm;
        scal[SC_LASTNORM] = scal[SC_NORM];
    }
}

}  // namespace dbec
