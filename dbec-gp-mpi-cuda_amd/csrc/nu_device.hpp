// nu_device.hpp - the nonlinear + trap exponential (calcnu) as device
// functions, shared by the standalone calcnu kernel and the fused
// inverse-FFT kernels.
//
// Reference: calcnu_kernel, src/imag3d-mpicuda/imag3d-mpicuda.cu:1759-1783
// (imaginary time) and src/real3d-mpicuda/real3d-mpicuda.cu:1856-1883 (real
// time); trap potential initpot imag3d-mpicuda.cu:872-889.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_device.hpp"

namespace dbec {

struct TrapConst {
    double half_par, g2, n2, l2;  // 0.5*par, gamma^2, nu^2, lambda^2
};
struct NuConst {
    double dt, g, gd, nd;  // nd = Nx*Ny*Nz
    double nd_odd, inv_odd, pow2_inv;  // nd = nd_odd 2^e with nd_odd odd; RN(1 / nd_odd); 2^-e (nd_split)
    double shift;          // imaginary time: exponent shift (0 = the reference's factor)
    double *capped;        // set to 1 where the exponent was capped
    int cap;               // cap the exponent (mixed form)
    float dtf, gf, gdnf, shiftf;  // mixed form: dt, g, gd / nd, shift in FP32 (nu_prepare)
};

// Fills the FP32 constants of the mixed form from dt, g, gd, nd and shift.
__host__ __device__ inline void nu_prepare(NuConst &nc) {
    nc.dtf = (float)nc.dt;
    nc.gf = (float)nc.g;
    nc.gdnf = (float)(nc.gd / nc.nd);
    nc.shiftf = (float)nc.shift;
}

// Splits nd = nd_odd 2^e for div_nd.
__host__ __device__ inline void nd_split(NuConst &nc, long nd) {
    double p = 1.0;
    while (nd % 2 == 0) {
        nd /= 2;
        p *= 0.5;
    }
    nc.nd_odd = (double)nd;
    nc.inv_odd = 1.0 / (double)nd;
    nc.pow2_inv = p;
}

// phi / nd rounded to nearest, equal to the IEEE quotient, without a divide
// instruction sequence.  With b = nd_odd < 2^50 and y = RN(1/b):
// q = RN(phi y) is within 1.5 ulp u of phi / b; the remainder r = phi - b q is
// a multiple of u below 2^51 u, so the FMA forms it exactly; q + r y differs
// from phi / b by |r| |y - 1/b| <= 1.5 2^-53 u, while phi / b lies at least
// u / (2 b) from every midpoint between doubles (2 phi is an even multiple
// of u and b k u an odd one for odd k), so RN(q + r y) = RN(phi / b).  The
// factor 2^-e is exact in the normal range.  For b = 1 this is phi 2^-e.
__device__ __forceinline__ double div_nd(double phi, const NuConst &nc) {
    const double q = phi * nc.inv_odd;
    const double r = fma(-q, nc.nd_odd, phi);
    return fma(r, nc.inv_odd, q) * nc.pow2_inv;
}

// 0.5*par*(g^2 x^2 + n^2 y^2 + l^2 z^2) in the reference's evaluation order
// with no contraction, i.e. the number the reference's pot table holds.  It
// splits into a per-row part g^2 x^2 + n^2 y^2, a per-z part l^2 z^2 and
// their combination, each rounded as in the whole expression, so that
// kernels can stage the first two and form the same number per element.
__device__ __forceinline__ double trap_xy(const TrapConst &tc, double x2, double y2) {
#pragma clang fp contract(off)
    return tc.g2 * x2 + tc.n2 * y2;
}
__device__ __forceinline__ double trap_z(const TrapConst &tc, double z2) { return tc.l2 * z2; }
__device__ __forceinline__ double trap_sum(const TrapConst &tc, double pxy, double pz) {
#pragma clang fp contract(off)
    return tc.half_par * (pxy + pz);
}
__device__ __forceinline__ double trap(const TrapConst &tc, double x2, double y2, double z2) {
    return trap_sum(tc, trap_xy(tc, x2, y2), trap_z(tc, z2));
}

// imaginary time: psi * exp(-dt (pot + g psi^2 + gd phi / N)), times
// exp(shift): a constant factor that the renormalisation after the sweeps
// removes again (the sweeps are linear); the mixed form uses it to keep FP32
// psi away from overflow, the FP64 form passes 0 (exact: x + 0 = x).
__device__ __forceinline__ double nu_imag(double psitmp, double phi, double pot, const NuConst &nc) {
    const double psi2lin = psitmp * psitmp * nc.g;
    const double psidd2lin = div_nd(phi, nc) * nc.gd;
    const double pottmp = nc.dt * (pot + psi2lin + psidd2lin);
    double e = -pottmp + nc.shift;
    // FP32 storage: cap the growth factor at e^80 so psi stays finite, and
    // flag it (FP64: no cap).  Only a collapsing state gets there: the
    // shift follows dt * mu, so e stays O(1) on a converging state.
    if (nc.cap && e > 80.0) {
        e = 80.0;
        *nc.capped = 1.0;
    }
    return psitmp * exp(e);
}

// sin t and cos t of the real-time phase t = -dt V_eff, without a branch:
// Cody and Waite's reduction t = k pi/2 + r with k = rint(2t/pi) and pi/2
// in two parts (two FMAs, each one rounding), the Taylor series of sin r and
// cos r on |r| <= pi/4 in Horner form in r^2, and the quadrant k mod 4
// selecting and negating them.  Through r^17/17! and r^16/16! the omitted
// terms are below 2^-60 of the result; the values stay within about 1 ulp of
// the correctly rounded ones for |t| < 2^30.  A library sincos in the kernel,
// even as a rarely taken branch, cost about a third of the fused inverse
// pass at 256x256x384 complex128.
__device__ __forceinline__ void sincos_phase(double t, double *sn, double *cs) {
    const double k = rint(t * 0.63661977236758134);  // 2/pi
    double r = fma(-k, 1.5707963267948966, t);        // pi/2, high part
    r = fma(-k, 6.123233995736766e-17, r);            // pi/2, low part
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;          // 1/17!
    ps = fma(ps, -r2, 1.0 / 1307674368000.0);     // 1/15!
    ps = fma(ps, -r2, 1.0 / 6227020800.0);        // 1/13!
    ps = fma(ps, -r2, 1.0 / 39916800.0);          // 1/11!
    ps = fma(ps, -r2, 1.0 / 362880.0);            // 1/9!
    ps = fma(ps, -r2, 1.0 / 5040.0);              // 1/7!
    ps = fma(ps, -r2, 1.0 / 120.0);               // 1/5!
    ps = fma(ps, -r2, 1.0 / 6.0);                 // 1/3!
    const double s = fma(-r * r2, ps, r);
    double pc = 1.0 / 20922789888000.0;           // 1/16!
    pc = fma(pc, -r2, 1.0 / 87178291200.0);       // 1/14!
    pc = fma(pc, -r2, 1.0 / 479001600.0);         // 1/12!
    pc = fma(pc, -r2, 1.0 / 3628800.0);           // 1/10!
    pc = fma(pc, -r2, 1.0 / 40320.0);             // 1/8!
    pc = fma(pc, -r2, 1.0 / 720.0);               // 1/6!
    pc = fma(pc, -r2, 1.0 / 24.0);                // 1/4!
    pc = fma(pc, -r2, 0.5);                       // 1/2!
    const double c = fma(-r2, pc, 1.0);
    // t = r + q pi/2: sin t = s, c, -s, -c and cos t = c, -s, -c, s for q = 0..3
    // (k clamped so that the conversion is defined for any t, NaN included;
    // s and c are NaN then anyway)
    const int q = (int)(long long)fmin(fmax(k, -4.0e15), 4.0e15) & 3;
    const double a = (q & 1) ? c : s, bq = (q & 1) ? s : c;
    *sn = (q & 2) ? -a : a;
    *cs = ((q + 1) & 2) ? -bq : bq;
}
// The same in FP32 (the mixed form): pi/2 in two float parts, the series
// through r^9/9! and r^10/10! (omitted terms below 2^-26 of the result),
// within about 1 ulp for |t| < 2^12.
__device__ __forceinline__ void sincos_phase(float t, float *sn, float *cs) {
    const float k = rintf(t * 0.636619772f);
    float r = fmaf(-k, 1.57079637f, t);
    r = fmaf(-k, -4.37113883e-8f, r);
    const float r2 = r * r;
    float ps = 1.0f / 362880.0f;                  // 1/9!
    ps = fmaf(ps, -r2, 1.0f / 5040.0f);           // 1/7!
    ps = fmaf(ps, -r2, 1.0f / 120.0f);            // 1/5!
    ps = fmaf(ps, -r2, 1.0f / 6.0f);              // 1/3!
    const float s = fmaf(-r * r2, ps, r);
    float pc = 1.0f / 3628800.0f;                 // 1/10!
    pc = fmaf(pc, -r2, 1.0f / 40320.0f);          // 1/8!
    pc = fmaf(pc, -r2, 1.0f / 720.0f);            // 1/6!
    pc = fmaf(pc, -r2, 1.0f / 24.0f);             // 1/4!
    pc = fmaf(pc, -r2, 0.5f);                     // 1/2!
    const float c = fmaf(-r2, pc, 1.0f);
    const int q = (int)fminf(fmaxf(k, -1.0e9f), 1.0e9f) & 3;
    const float a = (q & 1) ? c : s, bq = (q & 1) ? s : c;
    *sn = (q & 2) ? -a : a;
    *cs = ((q + 1) & 2) ? -bq : bq;
}

// real time: psi * e^{-i dt (pot + g |psi|^2 + gd phi / N)}
__device__ __forceinline__ cpx<double> nu_real(cpx<double> v, double phi, double pot, const NuConst &nc) {
    const double psi2lin = fma(v.x, v.x, v.y * v.y) * nc.g;  // |psi|^2 (the reference squares cabs(psi))
    const double psidd2lin = div_nd(phi, nc) * nc.gd;
    const double pottmp = nc.dt * (pot + psi2lin + psidd2lin);
    double sn, cs;
    sincos_phase(-pottmp, &sn, &cs);
    return {v.x * cs - v.y * sn, v.x * sn + v.y * cs};
}

// Generic over the storage of psi: double / float (imaginary time, with the
// pending renormalisation s) and cpx<double> / cpx<float> (real time).
__device__ __forceinline__ double nu_apply(double v, double s, double phi, double pot, const NuConst &nc) {
    return nu_imag(v * s, phi, pot, nc);
}
__device__ __forceinline__ cpx<double> nu_apply(cpx<double> v, double, double phi, double pot, const NuConst &nc) {
    return nu_real(v, phi, pot, nc);
}
// The mixed form evaluates the nonlinear factor in FP32 (single-precision
// exp / sincos) from FP32 phi and the FP32 constants of nu_prepare; the
// renormalisation and the trap potential are formed in FP64 first.
__device__ __forceinline__ float nu_apply(float v, double s, float phi, double pot, const NuConst &nc) {
    const float p = (float)((double)v * s);
    const float lin = (float)pot + p * p * nc.gf + phi * nc.gdnf;
    float e = -nc.dtf * lin + nc.shiftf;
    if (nc.cap && e > 80.0f) {
        e = 80.0f;
        *nc.capped = 1.0;
    }
    return p * expf(e);
}
__device__ __forceinline__ cpx<float> nu_apply(cpx<float> v, double, float phi, double pot, const NuConst &nc) {
    const float a2 = v.x * v.x + v.y * v.y;
    const float lin = (float)pot + a2 * nc.gf + phi * nc.gdnf;
    float sn, cs;
    sincos_phase(-nc.dtf * lin, &sn, &cs);
    return {v.x * cs - v.y * sn, v.x * sn + v.y * cs};
}

}  // namespace dbec
