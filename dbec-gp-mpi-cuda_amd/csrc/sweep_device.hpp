// sweep_device.hpp - the partitioned Thomas solve of the Crank-Nicolson
// sweeps as device code: value-type helpers, the coefficient tables in LDS and
// the per-segment solve.  Used by the sweep kernels (ops.hip) and by the z
// sweep fused with the next step's forward z transform (fft_kernels.hpp).
#pragma once

#include <type_traits>

#include "fft_device.hpp"
#include "ops.hpp"

namespace dbec {

// Thomas sweeps (calclux/calcluy/calcluz kernels, imag3d.cu:1802-1942;
// real3d.cu:1902-2040), as a partitioned solve.  Every line is split into
// P = n / M segments of M elements held in the registers of one thread.
//
// Backward (reference: cbeta[N-2] = psi[N-1]; cbeta[i-1] = cgamma[i] (A
// cbeta[i] - c_i), c_i = minusA psi[i+1] + A0r psi[i] - A psi[i-1]): each
// segment runs the same recurrence from a zero carry, Q, and the product of
// its coefficients cgamma[m] A, Pb; then cbeta[i-1] = Q[i-1] + Pb[i-1] x with
// x = cbeta at the segment's top, handed over from the segment on the right
// (psi[N-1] for the last one).  Forward (psi'[0] = 0; psi'[i+1] = calpha[i]
// psi'[i] + cbeta[i]; psi'[N-1] = 0): the same with the carry psi' coming
// from the left.  Carries are folded through a small LDS exchange.  The
// algebra is exact; results differ from the sequential sweep by round-off
// only.  The coefficient tables sit in LDS so no step of the dependent
// chains waits on a cache.  cbeta[i-1] replaces psi[i] in its register.
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
// FP32 (mixed form): the recurrences in single precision, the Simpson
// partials of |psi|^2 in double.
__device__ __forceinline__ float m_mul(float a, float b) { return a * b; }
__device__ __forceinline__ float m_add(float a, float b) { return a + b; }
__device__ __forceinline__ float m_sub(float a, float b) { return a - b; }
__device__ __forceinline__ float m_fma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ float m_zero(float) { return 0.0f; }
__device__ __forceinline__ double m_mod2(float a) { return (double)a * (double)a; }
__device__ __forceinline__ cpx<float> m_mul(cpx<float> a, cpx<float> b) { return cmul(a, b); }
__device__ __forceinline__ cpx<float> m_add(cpx<float> a, cpx<float> b) { return a + b; }
__device__ __forceinline__ cpx<float> m_sub(cpx<float> a, cpx<float> b) { return a - b; }
__device__ __forceinline__ cpx<float> m_fma(cpx<float> a, cpx<float> b, cpx<float> c) {
    float re = fmaf(a.x, b.x, c.x);
    re = fmaf(-a.y, b.y, re);
    float im = fmaf(a.x, b.y, c.y);
    im = fmaf(a.y, b.x, im);
    return {re, im};
}
__device__ __forceinline__ cpx<float> m_zero(cpx<float>) { return {0.0f, 0.0f}; }
__device__ __forceinline__ double m_mod2(cpx<float> a) { return (double)a.x * a.x + (double)a.y * a.y; }

// Storage and arithmetic types: FP64 lines in FP64, FP32 lines (mixed form)
// in FP32.  The coefficient tables are FP64 in device memory for every line
// type (TabOf); the mixed form's sweeps round each entry to FP32 once, as they
// stage it into LDS.
template <typename S> struct Arith { using T = S; };
template <typename A> struct TabOf { using T = A; };
template <> struct TabOf<float> { using T = double; };
template <> struct TabOf<cpx<float>> { using T = cpx<double>; };
__device__ __forceinline__ double coef_round(double, double v) { return v; }
__device__ __forceinline__ float coef_round(float, double v) { return (float)v; }
__device__ __forceinline__ cpx<double> coef_round(cpx<double>, cpx<double> v) { return v; }
__device__ __forceinline__ cpx<float> coef_round(cpx<float>, cpx<double> v) { return {(float)v.x, (float)v.y}; }
template <typename S> __device__ __forceinline__ S to_a(S v) { return v; }
template <typename S> __device__ __forceinline__ S from_a(S v) { return v; }

// Cross-lane moves of a double as two 32-bit halves.  DPP (data-parallel
// primitives) reads another lane of the same 16-lane row inside the VALU
// instruction; ds_swizzle permutes within 32 lanes through the LDS crossbar
// without touching LDS memory.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int PATTERN>
__device__ __forceinline__ double swizzle_d(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)b, PATTERN);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)(b >> 32), PATTERN);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum over the 64 lanes of a wavefront (all lanes active), returned in every
// lane: DPP quad permutations (lane ^ 1, lane ^ 2), DPP row half-mirror and
// row mirror (the 8- and 16-lane sums), a ds_swizzle bit-mask step (lane ^
// 16: the 32-lane sums) and the two 32-lane halves read as scalars.
__device__ __forceinline__ double wave_sum(double v) {
    constexpr int kQuadXor1 = 0xB1;       // quad_perm [1, 0, 3, 2]
    constexpr int kQuadXor2 = 0x4E;       // quad_perm [2, 3, 0, 1]
    constexpr int kRowHalfMirror = 0x141;  // lane i <-> 7 - i within each 8 lanes
    constexpr int kRowMirror = 0x140;      // lane i <-> 15 - i within each row of 16
    constexpr int kSwizzleXor16 = 0x1F | (0x10 << 10);  // bit-mask mode: and 0x1f, or 0, xor 0x10
    v += dpp_d<kQuadXor1>(v);
    v += dpp_d<kQuadXor2>(v);
    v += dpp_d<kRowHalfMirror>(v);
    v += dpp_d<kRowMirror>(v);
    v += swizzle_d<kSwizzleXor16>(v);
    return readlane_d(v, 0) + readlane_d(v, 32);
}

// Device coefficient tables of one axis: calpha and cgamma (n - 1 each), and
// for the segment length M of the launch the carry multipliers of the
// partitioned solve (n each):
//   gb[i] = prod_{t = i .. top} cgamma[t] A,   fa[e] = prod_{t = bot .. e} calpha[t - 1]
// over the in-range indices (1..n-2) of the segment holding i / e, top and
// bot its last and first in-range index.  They are products of coefficients
// only, formed on the host in FP64 and kept in FP64 (TabOf<A>).
template <typename A>
struct SweepTab {
    const typename TabOf<A>::T *ca, *cg, *gb, *fa;
};

// LDS layout shared by the sweep kernels with segments of M: the four
// tables, the carry exchange (one per thread), the last elements, the row
// partials (one per thread, only with the z integrals; placed elsewhere when
// the kernel passes exn_at).  A table keeps the M entries of segment s in
// slots s (M + PAD) .. s (M + PAD) + M - 1 (slot()).  The lanes of a
// wavefront read the same entry of several segments at once: unpadded,
// segments of 256 bytes (16 complex doubles, 32 complex floats) all start in
// one bank and segments of 128 bytes in two, and those reads serialise.
// Complex segments of 128 bytes or more are therefore followed by 16 bytes
// of padding, which puts the starts of up to 16 consecutive segments in
// different banks (measured: the fused z kernel at 192^3 complex64 32.9 ->
// 31.7 us, the z sweep at 256x256x384 complex128 205 -> 199 us).  Real
// lines take none: FP32 segments of 16 are 64 bytes, and padded FP64 ones
// made the fused z kernel at 256^3 slower (75 -> 89 us).  With padding, ca
// holds calpha shifted by one slot (CSH), so a segment reads only its own
// slots.
template <typename A, int M>
struct SweepShared {
    static constexpr int PAD = std::is_floating_point<A>::value || M * sizeof(A) < 128 ? 0
                               : sizeof(A) >= 16                                     ? 1
                                                                                     : 16 / (int)sizeof(A);
    static constexpr int CSH = PAD > 0 ? 1 : 0;
    A *ca, *cg, *gb, *fa, *exa, *exb, *exl;
    double *exn;
    __host__ __device__ static constexpr int tab_len(int n) { return n + (n / M) * PAD; }
    __host__ __device__ static constexpr int slot(int i) { return i + (i / M) * PAD; }
    __device__ SweepShared(char *raw, int n, int nthreads, int L, double *exn_at = nullptr) {
        const int tl = tab_len(n);
        ca = reinterpret_cast<A *>(raw);
        cg = ca + tl;
        gb = cg + tl;
        fa = gb + tl;
        exa = fa + tl;
        exb = exa + nthreads;
        exl = exb + nthreads;
        exn = exn_at ? exn_at : reinterpret_cast<double *>(exl + L);
    }
    __host__ __device__ static constexpr size_t bytes(int n, int nthreads, int L, bool with_exn) {
        return (4 * (size_t)tab_len(n) + 2 * (size_t)nthreads + L) * sizeof(A) +
               (with_exn ? (size_t)nthreads * sizeof(double) : 0);
    }
};

// The tables are staged into LDS in two parts: the first slot of each thread
// is loaded (coef_pre) before the thread's psi loads, so writing it to LDS
// (coef_put) waits only for that older load while the psi loads are still
// in flight; slots beyond the first thread count follow.
template <typename A>
struct CoefPre {
    A ca, cg, gb, fa;
};
template <typename A>
__device__ __forceinline__ CoefPre<A> coef_pre(const SweepTab<A> &t, int n, int tid) {
    const int i = tid < n ? tid : 0, ic = i < n - 1 ? i : 0;
    const A z{};
    return {coef_round(z, t.ca[ic]), coef_round(z, t.cg[ic]), coef_round(z, t.gb[i]), coef_round(z, t.fa[i])};
}
template <typename A, int M>
__device__ __forceinline__ void coef_put(const SweepShared<A, M> &sh, const CoefPre<A> &c, const SweepTab<A> &t, int n,
                                         int tid, int nthreads) {
    using SH = SweepShared<A, M>;
    if (tid < n) {
        const int p = SH::slot(tid);
        if (tid < n - 1) {
            sh.ca[SH::slot(tid + SH::CSH)] = c.ca;
            sh.cg[p] = c.cg;
        }
        sh.gb[p] = c.gb;
        sh.fa[p] = c.fa;
    }
    const A z{};
    for (int i = tid + nthreads; i < n; i += nthreads) {
        const int p = SH::slot(i);
        if (i < n - 1) {
            sh.ca[SH::slot(i + SH::CSH)] = coef_round(z, t.ca[i]);
            sh.cg[p] = coef_round(z, t.cg[i]);
        }
        sh.gb[p] = coef_round(z, t.gb[i]);
        sh.fa[p] = coef_round(z, t.fa[i]);
    }
}

// The backward recurrence's right-hand side c_i = mA psi[i+1] + A0r psi[i] -
// A psi[i-1] and update cbeta[i-1] = cgamma[i] (A cbeta[i] - c_i).  Complex
// lines (real time) have A = i a, mA = -A and A0r = 1 + i b (gencoef,
// real3d.cu:991-1035; launch_sweep checks it), so there
//   c = -i a (psi[i+1] + psi[i-1]) + (1 + i b) psi[i],  A cbeta - c = i a cbeta - c.
// Real lines (imaginary time): mA = -A exactly, so c = mA (psi[i+1] +
// psi[i-1]) + A0r psi[i]; both forms are written as explicit FMAs.
template <typename T>
__device__ __forceinline__ T sweep_rhs(const SweepConst<T> &k, T pn, T pc, T pp) {
    return fma(k.mA, pn + pp, k.A0r * pc);
}
template <typename T>
__device__ __forceinline__ T sweep_back(const SweepConst<T> &k, T cg, T bq, T c) {
    return cg * fma(k.A, bq, -c);
}
template <typename R>
__device__ __forceinline__ cpx<R> sweep_rhs(const SweepConst<cpx<R>> &k, cpx<R> pn, cpx<R> pc, cpx<R> pp) {
    const R a = k.A.y, b = k.A0r.y;
    const R sx = pn.x + pp.x, sy = pn.y + pp.y;
    return {fma(a, sy, fma(-b, pc.y, pc.x)), fma(-a, sx, fma(b, pc.x, pc.y))};
}
template <typename R>
__device__ __forceinline__ cpx<R> sweep_back(const SweepConst<cpx<R>> &k, cpx<R> cg, cpx<R> bq, cpx<R> c) {
    const R a = k.A.y;
    return cmul(cg, cpx<R>{fma(-a, bq.y, -c.x), fma(a, bq.x, -c.y)});
}

// The partitioned solve of one segment (registers v, halos hl / hr).  Every
// thread of the workgroup calls it.  Returns, with COLS, the Simpson z
// partial sum_j wz[e] |psi'|^2 of the segment.  Of the line's elements only
// 0 (first segment, j = 0) and n-1 (last segment, j = M-1) fall outside the
// recurrences' range 1..n-2, so with PEEL each pass tests only its first and
// last iteration and runs the M-2 between them unconditionally; without it
// every iteration keeps the test.  The x / y sweeps of complex lines use the
// latter: peeled, the compiler schedules them with more registers, which
// measured slower there (and faster in every other sweep).
// PC > 0: the line has exactly PC segments, known at compile time (P == PC);
// the carry folds then load every segment's carry at once and run the same
// FMA chain predicated, instead of one LDS round trip per step.
template <typename A, int M, bool COLS, bool PEEL, int PC = 0>
__device__ __forceinline__ double seg_solve(A (&v)[M], A hl, A hr, int e0, int n, int s, int l, int L, int P,
                                            const SweepShared<A, M> &sh, const SweepConst<A> &k,
                                            const double *__restrict__ wz) {
    static_assert(M >= 2, "segment length");
    // carries per group in the run-time folds: 4, or 2 for complex FP32
    // lines, whose x / y sweeps lose a wave of occupancy with 4
    constexpr int FB = std::is_same<A, cpx<float>>::value ? 2 : 4;
    const bool in0 = s > 0, inM = s < P - 1;  // j = 0 / j = M-1 inside 1..n-2
    auto inr = [&](int j) { return PEEL || (e0 + j >= 1 && e0 + j <= n - 2); };  // middle iterations
    const int bot = e0 > 1 ? e0 : 1, top = e0 + M - 1 < n - 2 ? e0 + M - 1 : n - 2;
    using SH = SweepShared<A, M>;
    const int tb = SH::PAD ? s * (M + SH::PAD) : e0;  // the segment's first table slot
    const A zero = m_zero(k.A);
    A bq = zero;
    {
        A pn = hr, pc = v[M - 1];
        auto back = [&](int j, A pp) {
            bq = sweep_back(k, sh.cg[tb + j], bq, sweep_rhs(k, pn, pc, pp));
            v[j] = bq;
        };
        if (inM) back(M - 1, v[M - 2]);
        pn = pc;
        pc = v[M - 2];
#pragma unroll
        for (int j = M - 2; j >= 1; --j) {
            const A pp = v[j - 1];
            if (inr(j)) back(j, pp);
            pn = pc;
            pc = pp;
        }
        if (in0) back(0, hl);
    }
    sh.exa[s * L + l] = bq;
    sh.exb[s * L + l] = sh.gb[tb + (bot - e0)];  // the segment's product of cgamma A
    if (s == P - 1) sh.exl[l] = v[M - 1];  // cbeta[N-2] = psi[N-1]
    __syncthreads();
    A x = sh.exl[l];
    if constexpr (PC > 0) {
        A eb[PC], ea[PC];
#pragma unroll
        for (int t = 0; t < PC; ++t) {
            eb[t] = sh.exb[t * L + l];
            ea[t] = sh.exa[t * L + l];
        }
#pragma unroll
        for (int t = PC - 1; t > 0; --t)
            if (t > s) x = m_fma(eb[t], x, ea[t]);
    } else {
        // in groups of FB carries loaded at once (the same FMA order as one
        // at a time, without an LDS round trip per step)
        int t = P - 1;
        for (; t - (FB - 1) > s; t -= FB) {
            A eb[FB], ea[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                eb[u] = sh.exb[(t - u) * L + l];
                ea[u] = sh.exa[(t - u) * L + l];
            }
#pragma unroll
            for (int u = 0; u < FB; ++u) x = m_fma(eb[u], x, ea[u]);
        }
        for (; t > s; --t) x = m_fma(sh.exb[t * L + l], x, sh.exa[t * L + l]);
    }
    {
        auto fixb = [&](int j) { v[j] = m_fma(sh.gb[tb + j], x, v[j]); };
        if (inM) fixb(M - 1);
#pragma unroll
        for (int j = M - 2; j >= 1; --j)
            if (inr(j)) fixb(j);
        if (in0) fixb(0);
    }
    __syncthreads();  // carries read; the exchange is reused below
    A fy = zero;
    {
        auto fwd = [&](int j) {
            fy = m_fma(sh.ca[tb + j - 1 + SH::CSH], fy, v[j]);
            v[j] = fy;
        };
        if (in0) fwd(0);
#pragma unroll
        for (int j = 1; j <= M - 2; ++j)
            if (inr(j)) fwd(j);
        if (inM) fwd(M - 1);
    }
    sh.exa[s * L + l] = fy;
    sh.exb[s * L + l] = sh.fa[tb + (top - e0)];  // the segment's product of calpha
    __syncthreads();
    A y = zero;  // psi'[0] = 0
    if constexpr (PC > 0) {
        A eb[PC], ea[PC];
#pragma unroll
        for (int t = 0; t < PC - 1; ++t) {
            eb[t] = sh.exb[t * L + l];
            ea[t] = sh.exa[t * L + l];
        }
#pragma unroll
        for (int t = 0; t < PC - 1; ++t)
            if (t < s) y = m_fma(eb[t], y, ea[t]);
    } else {
        int t = 0;
        for (; t + (FB - 1) < s; t += FB) {
            A eb[FB], ea[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                eb[u] = sh.exb[(t + u) * L + l];
                ea[u] = sh.exa[(t + u) * L + l];
            }
#pragma unroll
            for (int u = 0; u < FB; ++u) y = m_fma(eb[u], y, ea[u]);
        }
        for (; t < s; ++t) y = m_fma(sh.exb[t * L + l], y, sh.exa[t * L + l]);
    }
    double acc = 0.0;
    {
        auto fixf = [&](int j) {
            v[j] = m_fma(sh.fa[tb + j], y, v[j]);
            if (COLS) acc = fma(wz[e0 + j], m_mod2(v[j]), acc);
        };
        if (in0) fixf(0);
#pragma unroll
        for (int j = 1; j <= M - 2; ++j)
            if (inr(j)) fixf(j);
        if (inM) fixf(M - 1);
    }
    if (s == 0) v[0] = zero;
    if (s == P - 1) v[M - 1] = zero;
    return acc;
}

}  // namespace dbec
