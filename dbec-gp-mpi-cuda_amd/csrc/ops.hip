// ops.hip - pointwise, sweep and reduction kernels of the split step.
//
// Reference aliases: imag3d.cu = src/imag3d-mpicuda/imag3d-mpicuda.cu,
// real3d.cu = src/real3d-mpicuda/real3d-mpicuda.cu, diffint.cu =
// src/utils/diffint.cu.
#include <cstdlib>

#include "ops.hpp"

namespace dbec {

// ---------------------------------------------------------------------------
// Periodic face copies of the dipolar potential after the inverse transform
// (calcpsidd2_kernel2a/2b, imag3d.cu:1491-1505, 1591-1625, at one rank):
//   A: phi[Nx-1][j][k] = phi[0][j][k]  (j < Ny-1, k < Nz-1)
//   B: phi[i][Ny-1][k] = phi[i][0][k]  (i < Nx-1, k < Nz-1)
//   C: phi[i][j][Nz-1] = phi[i][j][0]  (i < Nx-1, j < Ny-1)
// No part reads an element another part writes, so one launch is race free.
template <typename R>
__global__ void k_facefix(R *__restrict__ phi, int Nx, int Ny, int Nz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nA = (long)(Ny - 1) * (Nz - 1), nB = (long)(Nx - 1) * (Nz - 1), nC = (long)(Nx - 1) * (Ny - 1);
    const long NyNz = (long)Ny * Nz;
    if (t < nA) {
        const long j = t / (Nz - 1), k = t % (Nz - 1);
        phi[(long)(Nx - 1) * NyNz + j * Nz + k] = phi[j * Nz + k];
    }
    if (t < nB) {
        const long i = t / (Nz - 1), k = t % (Nz - 1);
        phi[i * NyNz + (long)(Ny - 1) * Nz + k] = phi[i * NyNz + k];
    }
    if (t < nC) {
        const long i = t / (Ny - 1), j = t % (Ny - 1);
        phi[i * NyNz + j * Nz + Nz - 1] = phi[i * NyNz + j * Nz];
    }
}

hipError_t launch_facefix(hipStream_t s, void *phi, bool fp32, long Nx, long Ny, long Nz) {
    long n = std::max((Ny - 1) * (Nz - 1), std::max((Nx - 1) * (Nz - 1), (Nx - 1) * (Ny - 1)));
    dim3 block(256), grid((unsigned)((n + 255) / 256));
    if (fp32) k_facefix<float><<<grid, block, 0, s>>>(static_cast<float *>(phi), (int)Nx, (int)Ny, (int)Nz);
    else k_facefix<double><<<grid, block, 0, s>>>(static_cast<double *>(phi), (int)Nx, (int)Ny, (int)Nz);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// calcnu (calcnu_kernel imag3d.cu:1759-1783; real3d.cu:1856-1883), with psi
// read as psi * scale when a renormalisation is pending (imaginary time).  The
// trap is recomputed from the grid (nu_device.hpp).
template <typename P, typename R>
__global__ void __launch_bounds__(256)
k_calcnu(P *__restrict__ psi, const R *__restrict__ phi, const double *__restrict__ scale,
         const double *__restrict__ sq, TrapConst tc, NuConst nc, int Nx, int Ny, int Nz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Nz) return;
    const long row = blockIdx.y;  // x * Ny + y
    const int i = (int)(row / Ny), j = (int)(row % Ny);
    const long n = row * Nz + k;
    const double s = scale ? *scale : 1.0;
    const double pot = trap(tc, sq[i], sq[Nx + j], sq[Nx + Ny + k]);
    psi[n] = nu_apply(psi[n], s, (double)phi[n], pot, nc);
}

TrapConst trap_const(const dbec_ctx *c) {
    TrapConst tc;
    tc.half_par = 0.5 * c->p.par;
    tc.g2 = c->p.vgamma * c->p.vgamma;
    tc.n2 = c->p.vnu * c->p.vnu;
    tc.l2 = c->p.vlambda * c->p.vlambda;
    return tc;
}

NuConst nu_const(const dbec_ctx *c) {
    NuConst nc;
    nc.dt = c->p.dt;
    nc.g = c->g;
    nc.gd = c->gd;
    const long nd = c->Nx * c->Ny * c->Nz;
    nc.nd = (double)nd;
    nc.inv_nd = 1.0 / (double)nd;
    nc.nd_pow2 = (nd & (nd - 1)) == 0;
    return nc;
}

hipError_t launch_calcnu(hipStream_t s, const dbec_ctx *c, const double *scale) {
    const auto &b = c->b;
    const TrapConst tc = trap_const(c);
    const NuConst nc = nu_const(c);
    dim3 block(256), grid((unsigned)((c->Nz + 255) / 256), (unsigned)(c->Nx * c->Ny));
    const int Nx = (int)c->Nx, Ny = (int)c->Ny, Nz = (int)c->Nz;
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) {
        if (!f32)
            k_calcnu<double, double><<<grid, block, 0, s>>>(static_cast<double *>(b.psi), static_cast<const double *>(b.phi),
                                                            scale, b.sq, tc, nc, Nx, Ny, Nz);
        else
            k_calcnu<float, float><<<grid, block, 0, s>>>(static_cast<float *>(b.psi), static_cast<const float *>(b.phi),
                                                          scale, b.sq, tc, nc, Nx, Ny, Nz);
    } else {
        if (!f32)
            k_calcnu<cpx<double>, double><<<grid, block, 0, s>>>(static_cast<cpx<double> *>(b.psi),
                                                                 static_cast<const double *>(b.phi), nullptr, b.sq, tc,
                                                                 nc, Nx, Ny, Nz);
        else
            k_calcnu<cpx<float>, float><<<grid, block, 0, s>>>(static_cast<cpx<float> *>(b.psi),
                                                               static_cast<const float *>(b.phi), nullptr, b.sq, tc, nc,
                                                               Nx, Ny, Nz);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// psi *= scale (the renormalisation of calcnorm_kernel, imag3d.cu:1220-1233).
template <typename P>
__global__ void k_scale(P *__restrict__ psi, const double *__restrict__ scale, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double s = *scale;
    psi[t] = (P)((double)psi[t] * s);
}

hipError_t launch_scale(hipStream_t s, const dbec_ctx *c, const double *scale) {
    dim3 block(256);
    if (c->mode == DBEC_IMAG) {
        dim3 grid((unsigned)((c->N + 255) / 256));
        if (c->prec == DBEC_MIXED) k_scale<float><<<grid, block, 0, s>>>(static_cast<float *>(c->b.psi), scale, c->N);
        else k_scale<double><<<grid, block, 0, s>>>(static_cast<double *>(c->b.psi), scale, c->N);
    } else {
        dim3 grid((unsigned)((2 * c->N + 255) / 256));
        if (c->prec == DBEC_MIXED) k_scale<float><<<grid, block, 0, s>>>(static_cast<float *>(c->b.psi), scale, 2 * c->N);
        else k_scale<double><<<grid, block, 0, s>>>(static_cast<double *>(c->b.psi), scale, 2 * c->N);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Thomas sweeps (calclux/calcluy/calcluz kernels, imag3d.cu:1802-1942;
// real3d.cu:1902-2040).  A workgroup stages a tile of L lines in LDS with
// coalesced loads (x and y lines: L consecutive z per row; z lines: whole
// rows), one lane per line runs the reference recurrence on the LDS copy, and
// the tile is written back.  cbeta lives in the LDS slots of psi: the
// backward recurrence overwrites slot i-1 with cbeta[i-1] right after reading
// psi[i-1], and the forward substitution overwrites slot i with the new
// psi[i] after reading cbeta[i], so one LDS array per line suffices and no
// intermediate touches HBM.  The LDS reads of each recurrence do not depend
// on it, so they are issued in batches of U ahead of the dependent
// arithmetic.  With COLS (z sweep, imaginary time) the lane also accumulates
// the Simpson z integral of the new psi^2 and the wavefront sums the tile's
// rows with the y weights: colpart[tile][x] (calcnorm, imag3d.cu:1194-1209).
__device__ __forceinline__ double m_mul(double a, double b) { return a * b; }
__device__ __forceinline__ double m_add(double a, double b) { return a + b; }
__device__ __forceinline__ double m_sub(double a, double b) { return a - b; }
__device__ __forceinline__ double m_fma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ double m_zero(double) { return 0.0; }
__device__ __forceinline__ double m_mod2(double a) { return a * a; }
__device__ __forceinline__ cpx<double> m_mul(cpx<double> a, cpx<double> b) { return cmul(a, b); }
__device__ __forceinline__ cpx<double> m_add(cpx<double> a, cpx<double> b) { return a + b; }
__device__ __forceinline__ cpx<double> m_sub(cpx<double> a, cpx<double> b) { return a - b; }
__device__ __forceinline__ cpx<double> m_fma(cpx<double> a, cpx<double> b, cpx<double> c) {
    double re = fma(a.x, b.x, c.x);
    re = fma(-a.y, b.y, re);
    double im = fma(a.x, b.y, c.y);
    im = fma(a.y, b.x, im);
    return {re, im};
}
__device__ __forceinline__ cpx<double> m_zero(cpx<double>) { return {0.0, 0.0}; }
__device__ __forceinline__ double m_mod2(cpx<double> a) { return a.x * a.x + a.y * a.y; }

// Storage <-> arithmetic conversions (FP32 storage computes in FP64 inside the line).
template <typename S> struct Arith;
template <> struct Arith<double> { using T = double; };
template <> struct Arith<float> { using T = double; };
template <> struct Arith<cpx<double>> { using T = cpx<double>; };
template <> struct Arith<cpx<float>> { using T = cpx<double>; };
__device__ __forceinline__ double to_a(double v) { return v; }
__device__ __forceinline__ double to_a(float v) { return (double)v; }
__device__ __forceinline__ cpx<double> to_a(cpx<double> v) { return v; }
__device__ __forceinline__ cpx<double> to_a(cpx<float> v) { return {(double)v.x, (double)v.y}; }
template <typename S> __device__ __forceinline__ S from_a(double v);
template <> __device__ __forceinline__ double from_a<double>(double v) { return v; }
template <> __device__ __forceinline__ float from_a<float>(double v) { return (float)v; }
template <typename S> __device__ __forceinline__ S from_a(cpx<double> v);
template <> __device__ __forceinline__ cpx<double> from_a<cpx<double>>(cpx<double> v) { return v; }
template <> __device__ __forceinline__ cpx<float> from_a<cpx<float>>(cpx<double> v) { return {(float)v.x, (float)v.y}; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double m_one(double) { return 1.0; }
__device__ __forceinline__ cpx<double> m_one(cpx<double>) { return {1.0, 0.0}; }

// Partitioned Thomas solve of L lines held in LDS (element e of line l at
// buf[e * LS + l]) by 256 threads: thread (l, s) owns the steps
// i in [ilo, ihi] of segment s of line l (P = 256 / L segments per line).
//
// Backward (reference: cbeta[i-1] = cgamma[i] (A cbeta[i] - c_i), cbeta[N-2] =
// psi[N-1]): each segment runs the same recurrence from a zero carry,
// Q[i-1], and the product of its coefficients cgamma[m] A, P[i-1]; then
// cbeta[i-1] = Q[i-1] + P[i-1] x with x = cbeta[ihi] handed over from the
// segment on the right (psi[N-1] for the last one).  Forward (psi'[0] = 0,
// psi'[i+1] = calpha[i] psi'[i] + cbeta[i]): the same with the carry psi'[ilo-1]
// from the left.  The carries are folded through a small LDS exchange.  The
// algebra is exact; the results differ from the sequential sweep only by
// round-off.  cbeta[i-1] and the new psi'[i] are stored in slot i.
template <typename A, int L, int LS, bool COLS>
__device__ __forceinline__ void thomas_tile(A *__restrict__ buf, A *__restrict__ exa, A *__restrict__ exb,
                                            double *__restrict__ exn, int n, int nl,
                                            const A *__restrict__ calpha, const A *__restrict__ cgamma,
                                            const SweepConst<A> &k, const double *__restrict__ wz,
                                            const double *__restrict__ wy, int y0, double *colval) {
    constexpr int P = 256 / L;
    const int tid = threadIdx.x;
    const int l = tid % L, s = tid / L;
    const int m = n - 2;
    const int ilo = 1 + (s * m) / P, ihi = ((s + 1) * m) / P;
    A *f = buf + l;
    const A zero = m_zero(k.A), one = m_one(k.A);
    // halos: psi[ilo-1], psi[ihi], psi[ihi+1] before any slot is overwritten
    A h_lo = zero, p_hi = zero, h_hi = zero;
    if (ilo <= ihi) {
        h_lo = f[(ilo - 1) * LS];
        p_hi = f[ihi * LS];
        h_hi = f[(ihi + 1) * LS];
    }
    __syncthreads();
    // backward, zero carry
    A bq = zero, bp = one;
    {
        A pn = h_hi, pc = p_hi;
        for (int i = ihi; i >= ilo; --i) {
            const A pp = (i - 1 >= ilo) ? f[(i - 1) * LS] : h_lo;
            const A c = m_sub(m_add(m_mul(k.mA, pn), m_mul(k.A0r, pc)), m_mul(k.A, pp));
            const A g = cgamma[i];
            bq = m_mul(g, m_sub(m_mul(k.A, bq), c));
            bp = m_mul(m_mul(g, k.A), bp);
            f[i * LS] = bq;
            pn = pc;
            pc = pp;
        }
    }
    exa[s * L + l] = bq;
    exb[s * L + l] = bp;
    __syncthreads();
    A x = f[(n - 1) * LS];  // cbeta[N-2] = psi[N-1]
    for (int t = P - 1; t > s; --t) x = m_add(exa[t * L + l], m_mul(exb[t * L + l], x));
    {
        A pr = one;
        for (int i = ihi; i >= ilo; --i) {
            pr = m_mul(m_mul(cgamma[i], k.A), pr);
            f[i * LS] = m_add(f[i * LS], m_mul(pr, x));
        }
    }
    // forward, zero carry: slot i+1 holds cbeta[i] for i in [ilo-1, ihi-1]
    A fy = zero, fr = one;
    for (int i = ilo - 1; i <= ihi - 1; ++i) {
        const A a = calpha[i];
        fy = m_fma(a, fy, f[(i + 1) * LS]);
        fr = m_mul(a, fr);
        f[(i + 1) * LS] = fy;
    }
    __syncthreads();  // everyone has read its backward carries from exa/exb
    exa[s * L + l] = fy;
    exb[s * L + l] = fr;
    __syncthreads();
    A y = zero;  // psi'[0] = 0
    for (int t = 0; t < s; ++t) y = m_add(exa[t * L + l], m_mul(exb[t * L + l], y));
    double acc = 0.0;
    {
        A pr = one;
        for (int i = ilo - 1; i <= ihi - 1; ++i) {
            pr = m_mul(calpha[i], pr);
            const A v = m_add(f[(i + 1) * LS], m_mul(pr, y));
            f[(i + 1) * LS] = v;
            if (COLS) acc += wz[i + 1] * m_mod2(v);
        }
    }
    if (s == 0) f[0] = zero;
    if (s == P - 1) f[(n - 1) * LS] = zero;
    if constexpr (COLS) {
        exn[s * L + l] = acc;
        __syncthreads();
        if (tid < 64) {
            double w = 0.0;
            if (tid < L && tid < nl) {
                double row = 0.0;
                for (int t = 0; t < P; ++t) row += exn[t * L + tid];
                w = wy[y0 + tid] * row;
            }
            w = wave_sum(w);
            if (tid == 0) *colval = w;
        }
    }
}

template <typename S, int AXIS, int L, bool COLS>
__global__ void __launch_bounds__(256)
k_sweep(S *__restrict__ psi, int Nx, int Ny, int Nz, const typename Arith<S>::T *__restrict__ calpha,
        const typename Arith<S>::T *__restrict__ cgamma, SweepConst<typename Arith<S>::T> k,
        double *__restrict__ colpart, const double *__restrict__ wz, const double *__restrict__ wy) {
    using A = typename Arith<S>::T;
    extern __shared__ char smem_raw[];
    constexpr int LS = (AXIS == 2) ? L + 1 : L;
    const int n = AXIS == 0 ? Nx : (AXIS == 1 ? Ny : Nz);
    A *buf = reinterpret_cast<A *>(smem_raw);
    A *exa = buf + n * LS;
    A *exb = exa + 256;
    double *exn = reinterpret_cast<double *>(exb + 256);
    const int tid = threadIdx.x;
    long base;
    int nl, xrow = 0, y0 = 0, tile = 0;
    if constexpr (AXIS == 0) {
        const int zt = (Nz + L - 1) / L;
        const int y = blockIdx.x / zt, z0 = (blockIdx.x % zt) * L;
        nl = min(L, Nz - z0);
        base = (long)y * Nz + z0;
    } else if constexpr (AXIS == 1) {
        const int zt = (Nz + L - 1) / L;
        const int x = blockIdx.x / zt, z0 = (blockIdx.x % zt) * L;
        nl = min(L, Nz - z0);
        base = (long)x * Ny * Nz + z0;
    } else {
        const int yt = (Ny + L - 1) / L;
        xrow = blockIdx.x / yt;
        tile = blockIdx.x % yt;
        y0 = tile * L;
        nl = min(L, Ny - y0);
        base = ((long)xrow * Ny + y0) * Nz;
    }
    const long estride = AXIS == 0 ? (long)Ny * Nz : (AXIS == 1 ? (long)Nz : 1L);
    if constexpr (AXIS == 2) {
        for (int i = tid; i < n * L; i += 256) {
            const int e = i % n, l = i / n;
            buf[e * LS + l] = (l < nl) ? to_a(psi[base + (long)l * Nz + e]) : m_zero(A{});
        }
    } else {
        for (int i = tid; i < n * L; i += 256) {
            const int l = i % L, e = i / L;
            buf[e * LS + l] = (l < nl) ? to_a(psi[base + e * estride + l]) : m_zero(A{});
        }
    }
    __syncthreads();
    double colval = 0.0;
    thomas_tile<A, L, LS, COLS>(buf, exa, exb, exn, n, nl, calpha, cgamma, k, wz, wy, y0, &colval);
    if constexpr (COLS) {
        if (tid == 0) colpart[(long)tile * Nx + xrow] = colval;
    }
    __syncthreads();
    if constexpr (AXIS == 2) {
        for (int i = tid; i < n * L; i += 256) {
            const int e = i % n, l = i / n;
            if (l < nl) psi[base + (long)l * Nz + e] = from_a<S>(buf[e * LS + l]);
        }
    } else {
        for (int i = tid; i < n * L; i += 256) {
            const int l = i % L, e = i / L;
            if (l < nl) psi[base + e * estride + l] = from_a<S>(buf[e * LS + l]);
        }
    }
}

// x and y sweeps with register-resident segments.  A workgroup covers L
// consecutive z (lanes, coalesced) of the lines of one (y) or (x) plane;
// thread (l, s) holds the M elements [s*M, (s+1)*M) of line l in registers,
// read straight from HBM, and runs the partitioned recurrence of thomas_tile
// on them; only the segment carries go through LDS.  P = n / M segments.
template <typename S, int AXIS, int L, int M>
__global__ void __launch_bounds__(512)
k_sweep_reg(S *__restrict__ psi, int Nx, int Ny, int Nz, const typename Arith<S>::T *__restrict__ calpha,
            const typename Arith<S>::T *__restrict__ cgamma, SweepConst<typename Arith<S>::T> k) {
    using A = typename Arith<S>::T;
    __shared__ A exa[512], exb[512], exf[512], exg[512], exl[64];
    const int n = AXIS == 0 ? Nx : Ny;
    const int P = n / M;
    const int tid = threadIdx.x;
    const int l = tid % L, s = tid / L;
    const int zt = (Nz + L - 1) / L;
    const int pl = blockIdx.x / zt, z0 = (blockIdx.x % zt) * L;
    const int nl = min(L, Nz - z0);
    const long base = (AXIS == 0 ? (long)pl * Nz : (long)pl * Ny * Nz) + z0 + l;
    const long es = AXIS == 0 ? (long)Ny * Nz : (long)Nz;
    const bool act = l < nl;
    const int e0 = s * M;
    const A zero = m_zero(k.A), one = m_one(k.A);
    A v[M];
#pragma unroll
    for (int j = 0; j < M; ++j) v[j] = act ? to_a(psi[base + (long)(e0 + j) * es]) : zero;
    const A hl = (act && e0 > 0) ? to_a(psi[base + (long)(e0 - 1) * es]) : zero;
    const A hr = (act && e0 + M < n) ? to_a(psi[base + (long)(e0 + M) * es]) : zero;
    // backward, zero carry; cbeta[i-1] replaces psi[i] in v[i - e0]
    A bq = zero, bp = one;
    {
        A pn = hr, pc = v[M - 1];
#pragma unroll
        for (int j = M - 1; j >= 0; --j) {
            const int i = e0 + j;
            const A pp = (j > 0) ? v[j - 1] : hl;
            if (i >= 1 && i <= n - 2) {
                const A c = m_sub(m_add(m_mul(k.mA, pn), m_mul(k.A0r, pc)), m_mul(k.A, pp));
                const A g = cgamma[i];
                bq = m_mul(g, m_sub(m_mul(k.A, bq), c));
                bp = m_mul(m_mul(g, k.A), bp);
                v[j] = bq;
            }
            pn = pc;
            pc = pp;
        }
    }
    exa[s * L + l] = bq;
    exb[s * L + l] = bp;
    if (s == P - 1) exl[l] = v[M - 1];  // cbeta[N-2] = psi[N-1]
    __syncthreads();
    A x = exl[l];
    for (int t = P - 1; t > s; --t) x = m_add(exa[t * L + l], m_mul(exb[t * L + l], x));
    {
        A pr = one;
#pragma unroll
        for (int j = M - 1; j >= 0; --j) {
            const int i = e0 + j;
            if (i >= 1 && i <= n - 2) {
                pr = m_mul(m_mul(cgamma[i], k.A), pr);
                v[j] = m_add(v[j], m_mul(pr, x));
            }
        }
    }
    // forward, zero carry: v[j] holds cbeta[e0 + j - 1]
    A fy = zero, fr = one;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int i = e0 + j - 1;
        if (i + 1 >= 1 && i + 1 <= n - 2) {
            const A a = calpha[i];
            fy = m_fma(a, fy, v[j]);
            fr = m_mul(a, fr);
            v[j] = fy;
        }
    }
    exf[s * L + l] = fy;
    exg[s * L + l] = fr;
    __syncthreads();
    A y = zero;
    for (int t = 0; t < s; ++t) y = m_add(exf[t * L + l], m_mul(exg[t * L + l], y));
    {
        A pr = one;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int i = e0 + j - 1;
            if (i + 1 >= 1 && i + 1 <= n - 2) {
                pr = m_mul(calpha[i], pr);
                v[j] = m_add(v[j], m_mul(pr, y));
            }
        }
    }
    if (s == 0) v[0] = zero;
    if (s == P - 1) v[M - 1] = zero;
    if (act) {
#pragma unroll
        for (int j = 0; j < M; ++j) psi[base + (long)(e0 + j) * es] = from_a<S>(v[j]);
    }
}

template <typename S, int AXIS>
static hipError_t sweep_reg_dispatch(hipStream_t s, S *psi, int Nx, int Ny, int Nz,
                                     const typename Arith<S>::T *ca, const typename Arith<S>::T *cg,
                                     const SweepConst<typename Arith<S>::T> &k) {
    using A = typename Arith<S>::T;
    constexpr int L = sizeof(A) == 8 ? 32 : 16;  // 256-byte rows of z per load
    const int n = AXIS == 0 ? Nx : Ny;
    const int M = (sizeof(A) == 8 && n % 32 == 0 && n / 32 >= 2) ? 32 : 16;
    const int P = n / M;
    if (n % M != 0 || L * P > 512) return hipErrorInvalidValue;
    const long nblocks = (long)((Nz + L - 1) / L) * (AXIS == 0 ? Ny : Nx);
    dim3 grid((unsigned)nblocks), block(L * P);
    if constexpr (sizeof(A) == 8) {
        if (M == 32) {
            k_sweep_reg<S, AXIS, L, 32><<<grid, block, 0, s>>>(psi, Nx, Ny, Nz, ca, cg, k);
            return hipGetLastError();
        }
    }
    k_sweep_reg<S, AXIS, L, 16><<<grid, block, 0, s>>>(psi, Nx, Ny, Nz, ca, cg, k);
    return hipGetLastError();
}

// LDS bytes of a sweep workgroup: the tile plus the carry exchange.
static size_t sweep_lds(long n, int L, size_t elem) {
    return (size_t)n * (L + 1) * elem + 2 * 256 * elem + 256 * sizeof(double);
}

// Lines per tile: the largest power of two <= 64 whose LDS fits the budget
// (so several workgroups share a CU while others load or store), but with
// segments of at least two steps.  DBEC_SWEEP_LDS_KB overrides the budget.
static int sweep_lines(long n, size_t elem) {
    static long budget = [] {
        const char *e = getenv("DBEC_SWEEP_LDS_KB");
        long kb = e ? atol(e) : 48;
        return (kb > 8 && kb <= 160 ? kb : 48) * 1024;
    }();
    int L = 64;
    while (L > 4 && (long)sweep_lds(n, L, elem) > budget) L /= 2;
    while (L < 64 && 256 / L > (n - 2) / 2) L *= 2;
    return L;
}

static size_t arith_elem(const dbec_ctx *c) { return c->mode == DBEC_REAL ? 16 : 8; }

long sweep_z_tiles(const dbec_ctx *c) {
    const int L = sweep_lines(c->Nz, arith_elem(c));
    return (c->Ny + L - 1) / L;
}

template <typename S, int AXIS, bool COLS>
static hipError_t sweep_dispatch(hipStream_t s, S *psi, int Nx, int Ny, int Nz, const void *ca, const void *cg,
                                 const SweepConst<typename Arith<S>::T> &k, double *colpart, const double *wz,
                                 const double *wy) {
    using A = typename Arith<S>::T;
    const int n = AXIS == 0 ? Nx : (AXIS == 1 ? Ny : Nz);
    const int other = AXIS == 2 ? Ny : Nz;
    const int L = sweep_lines(n, sizeof(A));
    const int tiles = (other + L - 1) / L;
    const long nblocks = (long)tiles * (AXIS == 0 ? Ny : Nx);
    const size_t lds = sweep_lds(n, L, sizeof(A));
    dim3 grid((unsigned)nblocks), block(256);
    const A *cal = static_cast<const A *>(ca), *cgm = static_cast<const A *>(cg);
    switch (L) {
    case 64: k_sweep<S, AXIS, 64, COLS><<<grid, block, lds, s>>>(psi, Nx, Ny, Nz, cal, cgm, k, colpart, wz, wy); break;
    case 32: k_sweep<S, AXIS, 32, COLS><<<grid, block, lds, s>>>(psi, Nx, Ny, Nz, cal, cgm, k, colpart, wz, wy); break;
    case 16: k_sweep<S, AXIS, 16, COLS><<<grid, block, lds, s>>>(psi, Nx, Ny, Nz, cal, cgm, k, colpart, wz, wy); break;
    case 8: k_sweep<S, AXIS, 8, COLS><<<grid, block, lds, s>>>(psi, Nx, Ny, Nz, cal, cgm, k, colpart, wz, wy); break;
    default: k_sweep<S, AXIS, 4, COLS><<<grid, block, lds, s>>>(psi, Nx, Ny, Nz, cal, cgm, k, colpart, wz, wy); break;
    }
    return hipGetLastError();
}

template <typename S, typename A>
static hipError_t sweep_axis(hipStream_t s, S *psi, const dbec_ctx *c, int axis, const A *ca, const A *cg,
                             const SweepConst<A> &k, double *colpart) {
    const int Nx = (int)c->Nx, Ny = (int)c->Ny, Nz = (int)c->Nz;
    const double *wy = c->b.wgt + c->Nx, *wz = c->b.wgt + c->Nx + c->Ny;
    if (axis == 0) return sweep_reg_dispatch<S, 0>(s, psi, Nx, Ny, Nz, ca, cg, k);
    if (axis == 1) return sweep_reg_dispatch<S, 1>(s, psi, Nx, Ny, Nz, ca, cg, k);
    if (colpart) return sweep_dispatch<S, 2, true>(s, psi, Nx, Ny, Nz, ca, cg, k, colpart, wz, wy);
    return sweep_dispatch<S, 2, false>(s, psi, Nx, Ny, Nz, ca, cg, k, nullptr, wz, wy);
}

hipError_t launch_sweep(hipStream_t s, const dbec_ctx *c, int axis, double *colpart) {
    const auto &b = c->b;
    const long off[3] = {0, c->Nx - 1, c->Nx - 1 + c->Ny - 1};
    const long tot = (c->Nx - 1) + (c->Ny - 1) + (c->Nz - 1);
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) {
        const double *ca = b.coef + off[axis];
        const double *cg = b.coef + tot + off[axis];
        SweepConst<double> k;
        k.A0r = c->A0r[axis][0];
        k.A = c->A[axis][0];
        k.mA = -c->A[axis][0];
        return f32 ? sweep_axis(s, static_cast<float *>(b.psi), c, axis, ca, cg, k, colpart)
                   : sweep_axis(s, static_cast<double *>(b.psi), c, axis, ca, cg, k, colpart);
    }
    const cpx<double> *cc = reinterpret_cast<const cpx<double> *>(b.coef);
    const cpx<double> *ca = cc + off[axis];
    const cpx<double> *cg = cc + tot + off[axis];
    SweepConst<cpx<double>> k;
    k.A0r = {c->A0r[axis][0], c->A0r[axis][1]};
    k.A = {c->A[axis][0], c->A[axis][1]};
    k.mA = {c->mA[axis][0], c->mA[axis][1]};
    return f32 ? sweep_axis(s, static_cast<cpx<float> *>(b.psi), c, axis, ca, cg, k, colpart)
               : sweep_axis(s, static_cast<cpx<double> *>(b.psi), c, axis, ca, cg, k, colpart);
}

// norm^2 = sum_x wx[x] sum_t colpart[t][x]; scal[NORM] = norm, scal[SCALE] = 1/norm.
__global__ void __launch_bounds__(256)
k_norm_cols(const double *__restrict__ colpart, int tiles, const double *__restrict__ wx, int Nx,
            double *__restrict__ scal) {
    __shared__ double sh[256];
    const int tid = threadIdx.x;
    double p = 0.0;
    for (int x = tid; x < Nx; x += 256) {
        double cs = 0.0;
        for (int t = 0; t < tiles; t++) cs += colpart[(long)t * Nx + x];
        p += wx[x] * cs;
    }
    sh[tid] = p;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid == 0) {
        const double norm = sqrt(sh[0]);
        scal[SC_NORM] = norm;
        scal[SC_SCALE] = 1.0 / norm;
    }
}

hipError_t launch_norm_from_cols(hipStream_t s, const dbec_ctx *c, const double *colpart) {
    k_norm_cols<<<1, 256, 0, s>>>(colpart, (int)sweep_z_tiles(c), c->b.wgt, (int)c->Nx, c->b.scal);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reductions.  Every integral is the Simpson sum of the reference
// (diffint.cu:37-51), written as weights w[i] (h/3 times 1,4,2,...,4,1 plus the
// even-N end correction) so that integral = sum_x wx sum_y wy sum_z wz f.
// Rows are summed by one wavefront each (lane-strided, then a fixed
// shuffle tree), columns and the total by fixed-order loops and a fixed LDS
// tree: the result is reproducible bit for bit from run to run.

__device__ __forceinline__ double mod2_of(double v) { return v * v; }
__device__ __forceinline__ double mod2_of(float v) { return (double)v * (double)v; }
__device__ __forceinline__ double mod2_of(cpx<double> v) {
    double a = hypot(v.x, v.y);
    return a * a;
}
__device__ __forceinline__ double mod2_of(cpx<float> v) {
    double a = hypot((double)v.x, (double)v.y);
    return a * a;
}
__device__ __forceinline__ double amp_of(double v) { return v; }
__device__ __forceinline__ double amp_of(float v) { return (double)v; }
__device__ __forceinline__ double amp_of(cpx<double> v) { return hypot(v.x, v.y); }
__device__ __forceinline__ double amp_of(cpx<float> v) { return hypot((double)v.x, (double)v.y); }

// rows0[y][x] = sum_z wz |psi|^2 ; rowsz[y][x] = sum_z wz z^2 |psi|^2
template <typename P>
__global__ void __launch_bounds__(256)
k_rows_mod2(const P *__restrict__ psi, const double *__restrict__ wz, const double *__restrict__ z2,
            double *__restrict__ rows0, double *__restrict__ rowsz, int Nx, int Ny, int Nz) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * 4 + wave;
    if (row >= (long)Nx * Ny) return;
    const P *r = psi + row * Nz;
    double a0 = 0.0, az = 0.0;
    for (int k = lane; k < Nz; k += 64) {
        const double m = mod2_of(r[k]) * wz[k];
        a0 += m;
        az += m * z2[k];
    }
    a0 = wave_sum(a0);
    az = wave_sum(az);
    if (lane == 0) {
        const int x = (int)(row / Ny), y = (int)(row % Ny);
        rows0[(long)y * Nx + x] = a0;
        if (rowsz) rowsz[(long)y * Nx + x] = az;
    }
}

hipError_t launch_rows_mod2(hipStream_t s, const dbec_ctx *c, double *rows0, double *rowsz) {
    const auto &b = c->b;
    const double *wz = b.wgt + c->Nx + c->Ny, *z2 = b.sq + c->Nx + c->Ny;
    dim3 block(256), grid((unsigned)((c->Nx * c->Ny + 3) / 4));
    const int Nx = (int)c->Nx, Ny = (int)c->Ny, Nz = (int)c->Nz;
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) {
        if (f32) k_rows_mod2<float><<<grid, block, 0, s>>>(static_cast<const float *>(b.psi), wz, z2, rows0, rowsz, Nx, Ny, Nz);
        else k_rows_mod2<double><<<grid, block, 0, s>>>(static_cast<const double *>(b.psi), wz, z2, rows0, rowsz, Nx, Ny, Nz);
    } else {
        if (f32)
            k_rows_mod2<cpx<float>><<<grid, block, 0, s>>>(static_cast<const cpx<float> *>(b.psi), wz, z2, rows0, rowsz, Nx, Ny, Nz);
        else
            k_rows_mod2<cpx<double>><<<grid, block, 0, s>>>(static_cast<const cpx<double> *>(b.psi), wz, z2, rows0, rowsz, Nx, Ny, Nz);
    }
    return hipGetLastError();
}

// Five-point derivative with the boundary rule of calcmuen_kernel1x
// (imag3d.cu:1303-1332): 0 at both ends, centred 3-point at 1 and N-2.
template <typename P>
__device__ __forceinline__ double deriv5(const P *__restrict__ f, int i, int n, long st, double h) {
    if (i == 0 || i == n - 1) return 0.0;
    if (i == 1) return (amp_of(f[2 * st]) - amp_of(f[0])) / (2.0 * h);
    if (i == n - 2) return (amp_of(f[(long)(n - 1) * st]) - amp_of(f[(long)(n - 3) * st])) / (2.0 * h);
    return (amp_of(f[(long)(i - 2) * st]) - 8.0 * amp_of(f[(long)(i - 1) * st]) + 8.0 * amp_of(f[(long)(i + 1) * st]) -
            amp_of(f[(long)(i + 2) * st])) / (12.0 * h);
}

// Per-row z integrals of the chemical potential and energy densities
// (calcmuen_kernel1x/1y/1z and kernel2, imag3d.cu:1293-1450; real3d.cu
// 1382-1543, where real time differentiates |psi|).
template <typename P, typename R>
__global__ void __launch_bounds__(256)
k_rows_muen(const P *__restrict__ psi, const R *__restrict__ phi, const double *__restrict__ sq,
            const double *__restrict__ wz, TrapConst tc, MuenConst mc, double *__restrict__ rmu,
            double *__restrict__ ren, int Nx, int Ny, int Nz) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * 4 + wave;
    if (row >= (long)Nx * Ny) return;
    const int i = (int)(row / Ny), j = (int)(row % Ny);
    const long NyNz = (long)Ny * Nz;
    double amu = 0.0, aen = 0.0;
    for (int k = lane; k < Nz; k += 64) {
        const long n = row * Nz + k;
        const double dxv = deriv5(psi + (long)j * Nz + k, i, Nx, NyNz, mc.dx);
        double d = (dxv * dxv) / mc.kin;
        const double dyv = deriv5(psi + (long)i * NyNz + k, j, Ny, (long)Nz, mc.dy);
        d += (dyv * dyv) / mc.kin;
        const double dzv = deriv5(psi + row * Nz, k, Nz, 1L, mc.dz);
        d += (dzv * dzv) / mc.kin;
        double psi2 = amp_of(psi[n]);
        psi2 *= psi2;
        const double psi2lin = psi2 * mc.g;
        const double psidd2lin = ((double)phi[n] / mc.nd) * mc.gd;
        const double pot = trap(tc, sq[i], sq[Nx + j], sq[Nx + Ny + k]);
        const double tmu = (pot + psi2lin + psidd2lin) * psi2 + d;
        const double ten = (pot + 0.5 * psi2lin + 0.5 * psidd2lin) * psi2 + d;
        amu += wz[k] * tmu;
        aen += wz[k] * ten;
    }
    amu = wave_sum(amu);
    aen = wave_sum(aen);
    if (lane == 0) {
        rmu[(long)j * Nx + i] = amu;
        ren[(long)j * Nx + i] = aen;
    }
}

hipError_t launch_rows_muen(hipStream_t s, const dbec_ctx *c, double *rmu, double *ren) {
    const auto &b = c->b;
    TrapConst tc;
    tc.half_par = 0.5 * c->p.par;
    tc.g2 = c->p.vgamma * c->p.vgamma;
    tc.n2 = c->p.vnu * c->p.vnu;
    tc.l2 = c->p.vlambda * c->p.vlambda;
    MuenConst mc;
    mc.dx = c->p.dx;
    mc.dy = c->p.dy;
    mc.dz = c->p.dz;
    mc.kin = 3.0 - c->p.par;
    mc.g = c->g;
    mc.gd = c->gd;
    mc.nd = (double)(c->Nx * c->Ny * c->Nz);
    const double *wz = b.wgt + c->Nx + c->Ny;
    dim3 block(256), grid((unsigned)((c->Nx * c->Ny + 3) / 4));
    const int Nx = (int)c->Nx, Ny = (int)c->Ny, Nz = (int)c->Nz;
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) {
        if (f32)
            k_rows_muen<float, float><<<grid, block, 0, s>>>(static_cast<const float *>(b.psi), static_cast<const float *>(b.phi),
                                                             b.sq, wz, tc, mc, rmu, ren, Nx, Ny, Nz);
        else
            k_rows_muen<double, double><<<grid, block, 0, s>>>(static_cast<const double *>(b.psi),
                                                               static_cast<const double *>(b.phi), b.sq, wz, tc, mc, rmu,
                                                               ren, Nx, Ny, Nz);
    } else {
        if (f32)
            k_rows_muen<cpx<float>, float><<<grid, block, 0, s>>>(static_cast<const cpx<float> *>(b.psi),
                                                                  static_cast<const float *>(b.phi), b.sq, wz, tc, mc,
                                                                  rmu, ren, Nx, Ny, Nz);
        else
            k_rows_muen<cpx<double>, double><<<grid, block, 0, s>>>(static_cast<const cpx<double> *>(b.psi),
                                                                    static_cast<const double *>(b.phi), b.sq, wz, tc,
                                                                    mc, rmu, ren, Nx, Ny, Nz);
    }
    return hipGetLastError();
}

// Column and total sums over the row integrals, one workgroup.
//   NORM:    scal[NORM] = sqrt(I0), scal[SCALE] = 1 / scal[NORM]        (I0 from A)
//   MOMENTS: scal[NORM], scal[RX/RY/RZ] = sqrt(<x^2>, <y^2>, <z^2>), scal[R0]
//   MUEN:    scal[MU] = integral of A, scal[EN] = integral of B
template <int MODE>
__global__ void __launch_bounds__(256)
k_finalize(const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ wx,
           const double *__restrict__ wy, const double *__restrict__ x2, const double *__restrict__ y2,
           double *__restrict__ scal, int Nx, int Ny) {
    __shared__ double s0[256], s1[256], s2[256], s3[256];
    const int tid = threadIdx.x;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    for (int x = tid; x < Nx; x += 256) {
        double c0 = 0.0, c1 = 0.0, c2 = 0.0;
        for (int y = 0; y < Ny; y++) {
            const double a = A[(long)y * Nx + x];
            c0 += wy[y] * a;
            if (MODE == 1) {
                c1 += wy[y] * y2[y] * a;
                c2 += wy[y] * B[(long)y * Nx + x];
            }
            if (MODE == 2) c1 += wy[y] * B[(long)y * Nx + x];
        }
        p0 += wx[x] * c0;
        if (MODE == 1) {
            p1 += wx[x] * x2[x] * c0;
            p2 += wx[x] * c1;
            p3 += wx[x] * c2;
        }
        if (MODE == 2) p1 += wx[x] * c1;
    }
    s0[tid] = p0;
    s1[tid] = p1;
    s2[tid] = p2;
    s3[tid] = p3;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            s0[tid] += s0[tid + o];
            s1[tid] += s1[tid + o];
            s2[tid] += s2[tid + o];
            s3[tid] += s3[tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (MODE == 0) {
            const double norm = sqrt(s0[0]);
            scal[SC_NORM] = norm;
            scal[SC_SCALE] = 1.0 / norm;
        } else if (MODE == 1) {
            scal[SC_NORM] = sqrt(s0[0]);
            const double rx = sqrt(s1[0]), ry = sqrt(s2[0]), rz = sqrt(s3[0]);
            scal[SC_RX] = rx;
            scal[SC_RY] = ry;
            scal[SC_RZ] = rz;
            scal[SC_R0] = sqrt(rx * rx + ry * ry + rz * rz);
        } else {
            scal[SC_MU] = s0[0];
            scal[SC_EN] = s1[0];
        }
    }
}

hipError_t launch_finalize(hipStream_t s, const dbec_ctx *c, int mode, const double *A, const double *B) {
    const auto &b = c->b;
    const double *wx = b.wgt, *wy = b.wgt + c->Nx, *x2 = b.sq, *y2 = b.sq + c->Nx;
    const int Nx = (int)c->Nx, Ny = (int)c->Ny;
    if (mode == 0) k_finalize<0><<<1, 256, 0, s>>>(A, B, wx, wy, x2, y2, b.scal, Nx, Ny);
    else if (mode == 1) k_finalize<1><<<1, 256, 0, s>>>(A, B, wx, wy, x2, y2, b.scal, Nx, Ny);
    else k_finalize<2><<<1, 256, 0, s>>>(A, B, wx, wy, x2, y2, b.scal, Nx, Ny);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Density extracts of the output writers (outden*, outpsi2*, imag3d.cu:1944-2399).
// out[a][b] = sum over `axis` of w * |psi|^2, remaining indices in [x][y][z] order.
template <typename P>
__global__ void k_reduce3(const P *__restrict__ psi, int axis, const double *__restrict__ w, double *__restrict__ out,
                          int Nx, int Ny, int Nz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long NyNz = (long)Ny * Nz;
    double acc = 0.0;
    if (axis == 2) {
        if (t >= (long)Nx * Ny) return;
        const P *r = psi + t * Nz;
        for (int k = 0; k < Nz; k++) acc += w[k] * mod2_of(r[k]);
    } else if (axis == 1) {
        if (t >= (long)Nx * Nz) return;
        const long i = t / Nz, k = t % Nz;
        for (int j = 0; j < Ny; j++) acc += w[j] * mod2_of(psi[i * NyNz + (long)j * Nz + k]);
    } else {
        if (t >= (long)Ny * Nz) return;
        for (int i = 0; i < Nx; i++) acc += w[i] * mod2_of(psi[(long)i * NyNz + t]);
    }
    out[t] = acc;
}

// out[a] = sum_b w[b] in[a][b] (axis 1) or out[b] = sum_a w[a] in[a][b] (axis 0)
__global__ void k_reduce2(const double *__restrict__ in, int n0, int n1, int axis, const double *__restrict__ w,
                          double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (axis == 1) {
        if (t >= n0) return;
        for (int b = 0; b < n1; b++) acc += w[b] * in[(long)t * n1 + b];
    } else {
        if (t >= n1) return;
        for (int a = 0; a < n0; a++) acc += w[a] * in[(long)a * n1 + t];
    }
    out[t] = acc;
}

template <typename P>
__global__ void k_mod2(const P *__restrict__ psi, double *__restrict__ out, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = mod2_of(psi[t]);
}

template <typename P>
static hipError_t reduce3_t(hipStream_t s, const P *psi, int axis, const double *w, double *out, long Nx, long Ny, long Nz) {
    const long n = axis == 2 ? Nx * Ny : (axis == 1 ? Nx * Nz : Ny * Nz);
    k_reduce3<P><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(psi, axis, w, out, (int)Nx, (int)Ny, (int)Nz);
    return hipGetLastError();
}

hipError_t launch_reduce3(hipStream_t s, const dbec_ctx *c, int axis, double *out) {
    const double *w = c->b.wgt + (axis == 0 ? 0 : (axis == 1 ? c->Nx : c->Nx + c->Ny));
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG)
        return f32 ? reduce3_t(s, static_cast<const float *>(c->b.psi), axis, w, out, c->Nx, c->Ny, c->Nz)
                   : reduce3_t(s, static_cast<const double *>(c->b.psi), axis, w, out, c->Nx, c->Ny, c->Nz);
    return f32 ? reduce3_t(s, static_cast<const cpx<float> *>(c->b.psi), axis, w, out, c->Nx, c->Ny, c->Nz)
               : reduce3_t(s, static_cast<const cpx<double> *>(c->b.psi), axis, w, out, c->Nx, c->Ny, c->Nz);
}

hipError_t launch_reduce2(hipStream_t s, const double *in, long n0, long n1, int axis, const double *w, double *out) {
    const long n = axis == 1 ? n0 : n1;
    k_reduce2<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, (int)n0, (int)n1, axis, w, out);
    return hipGetLastError();
}

hipError_t launch_mod2(hipStream_t s, const dbec_ctx *c, double *out) {
    const unsigned g = (unsigned)((c->N + 255) / 256);
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) {
        if (f32) k_mod2<float><<<g, 256, 0, s>>>(static_cast<const float *>(c->b.psi), out, c->N);
        else k_mod2<double><<<g, 256, 0, s>>>(static_cast<const double *>(c->b.psi), out, c->N);
    } else {
        if (f32) k_mod2<cpx<float>><<<g, 256, 0, s>>>(static_cast<const cpx<float> *>(c->b.psi), out, c->N);
        else k_mod2<cpx<double>><<<g, 256, 0, s>>>(static_cast<const cpx<double> *>(c->b.psi), out, c->N);
    }
    return hipGetLastError();
}

}  // namespace dbec
