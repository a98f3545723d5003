// fft_kernels.hpp - the passes of the dipolar convolution
//   phi = IDFT3( potdd * DFT3( |psi|^2 ) )      (unnormalised)
// over the half spectrum spec[kz][x][y], kz = 0..Nz/2.
//
//   zfwd : reads psi rows [x][y][:], forms |s psi|^2 (s = pending
//          renormalisation), transforms two real rows at once as the real and
//          imaginary parts of one complex line, splits the result into the two
//          half spectra and writes spec[kz][x][y-tile] (coalesced over y).
//   ypass: contiguous y lines of spec, forward or inverse.
//   xconv: x lines of spec (stride Ny, coalesced over y tiles): forward FFT,
//          multiplication by potdd(|kx|, |ky|, kz), inverse FFT, one pass.
//   zinv : rebuilds two Hermitian full spectra per complex line, inverse
//          transform, real and imaginary parts are the two real rows of phi.
//   zinv_nu: zinv fused with calcnu (the composite step).
//   sweepz_zfwd: the z sweep of a step fused with zfwd of the next step.
//
// Included by fft_inst.hip with DBEC_FFT_N set; every kernel is a template on
// the length N, the lines per workgroup L and the real type R of spec.
#pragma once

#include <cstdlib>

#include "fft_launch.hpp"
#include "internal.hpp"
#include "norm_device.hpp"
#include "sweep_device.hpp"

namespace dbec {

template <int N, int LM = 1>
struct FftTile {
    static constexpr int T = FftPlan<N>::T;
    // lines per workgroup: at least 8, and at least one wavefront of threads,
    // times the tile multiplier LM (wider tiles: longer contiguous HBM runs
    // on the strided side, fewer workgroups per CU)
    static constexpr int L0 = (64 + T - 1) / T;
    static constexpr int L = LM * (L0 <= 8 ? 8 : (L0 <= 16 ? 16 : (L0 <= 32 ? 32 : 64)));
    static constexpr int WG = L * T;
};

template <typename R>
__device__ __forceinline__ void stage_twiddles(cpx<R> *tw_s, const cpx<R> *tw_g, int n, int tid, int nthreads) {
    for (int i = tid; i < n; i += nthreads) tw_s[i] = tw_g[i];
}

template <typename P>
struct PsiTraits;
template <>
struct PsiTraits<double> {
    static __device__ __forceinline__ double mod2(double v, double s) {
        double t = v * s;
        return t * t;
    }
};
template <>
struct PsiTraits<float> {
    static __device__ __forceinline__ double mod2(float v, double s) {
        double t = (double)v * s;
        return t * t;
    }
};
// complex psi: x^2 + y^2 as one FMA, as calcnu forms it (the reference's
// cabs(psi)^2 differs by rounding only; the library hypot costs tens of
// instructions per point)
template <>
struct PsiTraits<cpx<double>> {
    static __device__ __forceinline__ double mod2(cpx<double> v, double) { return fma(v.x, v.x, v.y * v.y); }
};
// mixed form: |psi|^2 of complex64 psi in FP32 (the spectrum is FP32 too)
template <>
struct PsiTraits<cpx<float>> {
    static __device__ __forceinline__ double mod2(cpx<float> v, double) { return v.x * v.x + v.y * v.y; }
};

// ---------------------------------------------------------------------------
template <int N, typename R, typename P, int LM>
__global__ void __launch_bounds__((FftTile<N, LM>::WG))
k_zfwd(Solves v, int Nx, int Ny, const cpx<R> *__restrict__ twg, const double *__restrict__ wx, int tiles) {
    constexpr int L = FftTile<N, LM>::L, WG = FftTile<N, LM>::WG;
    using Pl = FftPlan<N>;
    __shared__ cpx<R> buf[N * (L + 1)];
    __shared__ cpx<R> tw[N];
    const SolveAt sr = solve_at(v);
    const P *__restrict__ psi = reinterpret_cast<const P *>(sr.psi);
    cpx<R> *__restrict__ spec = reinterpret_cast<cpx<R> *>(sr.spec);
    const int tid = threadIdx.x;
    const int ytiles = (Ny + 2 * L - 1) / (2 * L);
    const int x = blockIdx.x / ytiles;
    const int y0 = (blockIdx.x % ytiles) * 2 * L;
    const int nrows = min(2 * L, Ny - y0);
    const double s = (v.use_scale && !v.fold) ? sr.scal[SC_SCALE] : 1.0;
    stage_twiddles(tw, twg, N, tid, WG);
    // Rows 2p and 2p+1 of the tile become the real and imaginary parts of
    // line p.  A thread takes VE consecutive z of both rows, so its LDS
    // writes are whole complex slots; all of its loads are issued before the
    // first LDS write.
    constexpr int VE = sizeof(P) == 4 ? 2 : 1, NVR = N / VE, KV = L * NVR / WG;
    static_assert((L * NVR) % WG == 0, "zfwd tile");
    struct alignas(VE * sizeof(P)) Pv {
        P e[VE];
    };
    const P *src = psi + ((long)x * Ny + y0) * N;
    Pv ra[KV], rb[KV];
#pragma unroll
    for (int k = 0; k < KV; ++k) {
        const int iv = tid + k * WG;
        const int p = iv / NVR, zv = iv % NVR;
        ra[k] = *reinterpret_cast<const Pv *>(src + (unsigned)(min(2 * p, nrows - 1) * N + zv * VE));
        rb[k] = *reinterpret_cast<const Pv *>(src + (unsigned)(min(2 * p + 1, nrows - 1) * N + zv * VE));
    }
#pragma unroll
    for (int k = 0; k < KV; ++k) {
        const int iv = tid + k * WG;
        const int p = iv / NVR, zv = iv % NVR;
#pragma unroll
        for (int j = 0; j < VE; ++j) {
            const R va = 2 * p < nrows ? (R)PsiTraits<P>::mod2(ra[k].e[j], s) : R(0);
            const R vb = 2 * p + 1 < nrows ? (R)PsiTraits<P>::mod2(rb[k].e[j], s) : R(0);
            buf[Pl::template at<L, R>(zv * VE + j, p)] = {va, vb};
        }
    }
    __syncthreads();
    Pl::template run<-1, L>(buf, tid % L, tid / L, tw);
    // split Z = A + iB into the spectra of the two real rows:
    //   A[k] = (Z[k] + conj Z[N-k]) / 2,  B[k] = (Z[k] - conj Z[N-k]) / (2i)
    constexpr int nkz = N / 2 + 1;
#pragma unroll
    for (int q = 0; q < (nkz * 2 * L + WG - 1) / WG; ++q) {
        const int i = tid + q * WG;
        const int row = i % (2 * L), k = i / (2 * L);
        if (k < nkz && row < nrows) {
            const int p = row >> 1;
            const cpx<R> z = buf[Pl::template at<L, R>(k, p)];
            const cpx<R> w = buf[Pl::template at<L, R>((N - k) % N, p)];
            cpx<R> o;
            if ((row & 1) == 0) o = {R(0.5) * (z.x + w.x), R(0.5) * (z.y - w.y)};
            else o = {R(0.5) * (z.y + w.y), R(0.5) * (w.x - z.x)};
            spec[(unsigned)((k * Nx + x) * Ny + y0 + row)] = o;
        }
    }
    if (v.fold == 1 && blockIdx.x == 0) {  // the last step's norm, for the x convolution and calcnu
        __syncthreads();
        norm_from_cols<WG>(sr.rows, tiles, wx, Nx, sr.scal, v.shift, reinterpret_cast<double *>(buf));
    }
}

// ---------------------------------------------------------------------------
// The z sweep of one step fused with the forward z pass of the next
// (calcluz, imag3d.cu:1895-1942, then the first transform of calcpsidd2):
// both walk the same contiguous z rows.  A workgroup takes L rows of one x
// plane through the LDS tile of k_sweep_z (partitioned Thomas solve, Simpson
// column partials with COLS), stores psi, then reuses the tile's LDS for the
// |psi|^2 of its rows, two rows per complex line, and runs the forward
// transform and split of k_zfwd on them, so psi is read once for both.  The
// rows transform unscaled: the next step's norm comes from the column
// partials (v.fold == 2, first workgroup of the forward y pass).  Built for
// plans with T = N / 8, where L / 2 lines of T threads are the sweep's L * P
// threads (P = N / 16 segments per row).
template <int N>
struct FusedZ {
    static constexpr bool ok = FftPlan<N>::T * 8 == N;
    // Two LDS layouts.  FP64 (R = double): the sweep tables and exchange,
    // the tile / transform buffer (padded), the twiddles; the row partials
    // of the z integrals sit in the buffer's part past the tile when that is
    // free (they are summed before the buffer is filled), else with the
    // exchange.  FP32 spectra (compact: R = float): the tables, which the
    // twiddles overwrite once the solve is done; the rows' last elements;
    // the row partials; the tile region, which holds the carry exchange
    // while the tile sits in registers and the transform buffer after the
    // tile is stored (swizzled and unpadded where the padded one would not
    // fit the tile's region: FP32 real lines; padded for complex FP32 rows,
    // which are wide enough).  Measured at 256^3 and 192^3: the
    // compact layout is 4-5 % faster with FP32 spectra (more workgroups per
    // CU) and 8 % slower in FP64 (its two extra barriers).
    template <typename R>
    static constexpr bool compact() { return std::is_same<R, float>::value; }
    // row partials of the z integrals: real lines (imaginary time) only
    template <typename S>
    static constexpr bool cols() { return std::is_floating_point<S>::value; }
    template <typename S, typename R, int L>
    static constexpr size_t tile_b() { return ((size_t)(N + Pk<S>::n) * L * sizeof(S) + 15) & ~size_t(15); }
    // transform buffer swizzled (unpadded) only in the compact layout and
    // only where the padded one would not fit in the tile's region (FP32
    // real lines); complex FP32 rows are wide enough for the padded buffer
    template <typename S, typename R, int L>
    static constexpr bool swz() {
        return compact<R>() && (size_t)N * (L / 2 + 1) * sizeof(cpx<R>) > tile_b<S, R, L>();
    }
    template <typename S, typename R, int L>
    static constexpr size_t x_b() {
        constexpr size_t tile = tile_b<S, R, L>();
        constexpr size_t buf = (size_t)N * (swz<S, R, L>() ? L / 2 : L / 2 + 1) * sizeof(cpx<R>);
        constexpr size_t ex = compact<R>() ? 2 * (size_t)L * (N / 16) * sizeof(typename Arith<S>::T) : 0;
        constexpr size_t m = tile > buf ? tile : buf;
        return ((m > ex ? m : ex) + 15) & ~size_t(15);
    }
    // FP64 layout
    template <typename S, typename R, int L>
    static constexpr bool exn_in_x() {
        return x_b<S, R, L>() - tile_b<S, R, L>() >= (size_t)L * (N / 16) * sizeof(double);
    }
    template <typename S, typename R, int L>
    static constexpr size_t sh_b() {
        using A = typename Arith<S>::T;
        return (SweepShared<A, 16>::bytes(N, L * (N / 16), L, cols<S>() && !exn_in_x<S, R, L>()) + 15) & ~size_t(15);
    }
    // compact layout: tables (or twiddles), last elements, row partials
    template <typename S, typename R, int L>
    static constexpr size_t tab_b() {
        using A = typename Arith<S>::T;
        constexpr size_t t = 4 * (size_t)SweepShared<A, 16>::tab_len(N) * sizeof(A), w = (size_t)N * sizeof(cpx<R>);
        return ((t > w ? t : w) + 15) & ~size_t(15);
    }
    template <typename S, typename R, int L>
    static constexpr size_t exl_b() { return ((size_t)L * sizeof(typename Arith<S>::T) + 15) & ~size_t(15); }
    template <typename S, typename R, int L>
    static constexpr size_t exn_b() {
        return cols<S>() ? ((size_t)L * (N / 16) * sizeof(double) + 15) & ~size_t(15) : 0;
    }
    // offset of the tile region
    template <typename S, typename R, int L>
    static constexpr size_t x_off() {
        return compact<R>() ? tab_b<S, R, L>() + exl_b<S, R, L>() + exn_b<S, R, L>() : sh_b<S, R, L>();
    }
    template <typename S, typename R, int L>
    static constexpr size_t lds() {
        return x_off<S, R, L>() + x_b<S, R, L>() + (compact<R>() ? 0 : N * sizeof(cpx<R>));
    }
    // the kernel exists for rows L: whole transform lines, at most 512
    // threads, LDS within 160 KiB
    template <typename S, typename R, int L>
    static constexpr bool fits() {
        return ok && (L / 2) * FftPlan<N>::T == L * (N / 16) && L * (N / 16) <= 512 && lds<S, R, L>() <= 160 * 1024;
    }
};

template <int N, typename S, typename R, int L, bool COLS>
__global__ void __launch_bounds__(L *(N / 16))
k_sweepz_zfwd(Solves v, int Nx, int Ny, SweepTab<typename Arith<S>::T> tab, SweepConst<typename Arith<S>::T> k,
              const double *__restrict__ wz, const double *__restrict__ wy, const cpx<R> *__restrict__ twg) {
    using A = typename Arith<S>::T;
    // psi tile rows of N + VE elements: 16-byte aligned, and consecutive rows
    // 4 banks apart, so the packet accesses below are conflict-free
    constexpr int M = 16, VE = Pk<S>::n, KV = M / VE;
    constexpr int P = N / M, WG = L * P, LF = L / 2, RS = N + VE;
    using Pl = FftPlan<N>;
    static_assert(FftPlan<N>::T * LF == WG, "one transform thread per sweep thread");
    using F = FusedZ<N>;
    static_assert(COLS == F::template cols<S>(), "z integrals for real lines");
    constexpr bool CMP = F::template compact<R>();  // compact layout
    constexpr size_t X_OFF = F::template x_off<S, R, L>(), X_B = F::template x_b<S, R, L>();
    constexpr bool SWZ = F::template swz<S, R, L>();  // swizzled transform buffer
    static_assert(X_B >= (size_t)RS * L * sizeof(S) && X_B >= (size_t)N * (SWZ ? LF : LF + 1) * sizeof(cpx<R>),
                  "fused z LDS");
    __shared__ __align__(16) char smem[F::template lds<S, R, L>()];
    const SolveAt sr = solve_at(v);
    S *__restrict__ psi = reinterpret_cast<S *>(sr.psi);
    cpx<R> *__restrict__ spec = reinterpret_cast<cpx<R> *>(sr.spec);
    SweepShared<A, M> sh(smem, N, WG, L,
                      (!CMP && F::template exn_in_x<S, R, L>()) ? reinterpret_cast<double *>(smem + X_OFF + F::template tile_b<S, R, L>())
                                                                 : nullptr);
    cpx<R> *tw = reinterpret_cast<cpx<R> *>(smem + X_OFF + X_B);
    if constexpr (CMP) {
        constexpr size_t TAB_B = F::template tab_b<S, R, L>(), EXL_B = F::template exl_b<S, R, L>();
        sh.exa = reinterpret_cast<A *>(smem + X_OFF);  // in the tile region while the tile is in registers
        sh.exb = sh.exa + WG;
        sh.exl = reinterpret_cast<A *>(smem + TAB_B);
        sh.exn = reinterpret_cast<double *>(smem + TAB_B + EXL_B);
        tw = reinterpret_cast<cpx<R> *>(smem);  // over the tables, after the solve
    }
    S *tile = reinterpret_cast<S *>(smem + X_OFF);
    cpx<R> *buf = reinterpret_cast<cpx<R> *>(smem + X_OFF);  // after the tile is stored
    const int tid = threadIdx.x;
    const int l = tid % L, s = tid / L;
    const int yt = (Ny + L - 1) / L;
    const int xrow = blockIdx.x / yt, tno = blockIdx.x % yt, y0 = tno * L;
    const int nl = min(L, Ny - y0);
    const long base = ((long)xrow * Ny + y0) * N;
    const A zero = m_zero(k.A);
    const int npk = nl * N / VE;
    const CoefPre<A> cp = coef_pre(tab, N, tid);
    // compact layout: the twiddles wait in registers until the tables are free
    constexpr int KT = CMP ? (N + WG - 1) / WG : 1;
    cpx<R> twr[KT];
    {
        typename Pk<S>::type r[KV];
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            const int iv = tid + q * WG;
            r[q] = ld_pk(psi + base + (unsigned)((iv < npk ? iv : 0) * VE));
        }
        if constexpr (CMP) {
#pragma unroll
            for (int q = 0; q < KT; ++q) {
                const int i = tid + q * WG;
                twr[q] = twg[i < N ? i : 0];
            }
        }
        coef_put(sh, cp, tab, N, tid, WG);
        if constexpr (!CMP) stage_twiddles(tw, twg, N, tid, WG);
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            const int iv = tid + q * WG;
            if (iv < npk) {
                const int e = (iv * VE) % N, row = (iv * VE) / N;
                st_pk(tile + row * RS + e, r[q]);
            }
        }
    }
    __syncthreads();
    const int e0 = s * M;
    const bool act = l < nl;
    A vv[M];
    {
        const S *row = tile + l * RS;
#pragma unroll
        for (int m = 0; m < KV; ++m) {
            const typename Pk<S>::type pk = ld_pk(row + e0 + m * VE);
#pragma unroll
            for (int j = 0; j < VE; ++j) vv[m * VE + j] = act ? to_a(pk.e[j]) : zero;
        }
        const A hl = (act && e0 > 0) ? to_a(row[e0 - 1]) : zero;
        const A hr = (act && e0 + M < N) ? to_a(row[e0 + M]) : zero;
        if constexpr (CMP) __syncthreads();  // the tile is in registers; its region takes the carry exchange
        // batched carry folds for real lines (measured 1.3 us faster at 256 in
        // both precisions); complex lines keep the loop (slower batched: the
        // extra registers cost occupancy)
        constexpr int PC = std::is_floating_point<A>::value ? P : 0;
        const double acc = seg_solve<A, M, COLS, true, PC>(vv, hl, hr, e0, N, s, l, L, P, sh, k, wz);
        if constexpr (CMP) __syncthreads();  // every carry is read; the region holds the tile again
        if (act) {
            S *orow = tile + l * RS;
#pragma unroll
            for (int m = 0; m < KV; ++m) {
                typename Pk<S>::type o;
#pragma unroll
                for (int j = 0; j < VE; ++j) o.e[j] = from_a<S>(vv[m * VE + j]);
                st_pk(orow + e0 + m * VE, o);
            }
        }
        if constexpr (COLS) {
            sh.exn[s * L + l] = acc;
            __syncthreads();
            if (tid < 64) {
                double w = 0.0;
                if (tid < L && tid < nl) {
                    double rsum = 0.0;
                    for (int t = 0; t < P; ++t) rsum += sh.exn[t * L + tid];
                    w = wy[y0 + tid] * rsum;
                }
                w = wave_sum(w);
                if (tid == 0) sr.rows[(long)tno * Nx + xrow] = w;
            }
        } else {
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < KV; ++q) {
        const int iv = tid + q * WG;
        if (iv < npk) {
            const int e = (iv * VE) % N, row = (iv * VE) / N;
            st_pk(psi + base + (unsigned)(iv * VE), ld_pk(tile + row * RS + e));
        }
    }
    __syncthreads();  // the tile is stored; its LDS becomes the transform buffer (and the tables the twiddles)
    if constexpr (CMP) {
#pragma unroll
        for (int q = 0; q < KT; ++q) {
            const int i = tid + q * WG;
            if (i < N) tw[i] = twr[q];
        }
    }
    // |psi|^2 of the new rows (zero past the plane's last row): row 2p is the
    // real part, row 2p + 1 the imaginary part of line p
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const R m2 = act ? (R)PsiTraits<S>::mod2(from_a<S>(vv[j]), 1.0) : R(0);
        R *slot = reinterpret_cast<R *>(&buf[Pl::template at<LF, R, SWZ>(e0 + j, l >> 1)]);
        slot[l & 1] = m2;
    }
    __syncthreads();
    Pl::template run<-1, LF, SWZ>(buf, tid % LF, tid / LF, tw);
    // split Z = A + iB into the spectra of the two real rows (as k_zfwd)
    constexpr int nkz = N / 2 + 1;
#pragma unroll
    for (int q = 0; q < (nkz * L + WG - 1) / WG; ++q) {
        const int i = tid + q * WG;
        const int row = i % L, kz = i / L;
        if (kz < nkz && row < nl) {
            const int p = row >> 1;
            const cpx<R> z = buf[Pl::template at<LF, R, SWZ>(kz, p)];
            const cpx<R> w = buf[Pl::template at<LF, R, SWZ>((N - kz) % N, p)];
            cpx<R> o;
            if ((row & 1) == 0) o = {R(0.5) * (z.x + w.x), R(0.5) * (z.y - w.y)};
            else o = {R(0.5) * (z.y + w.y), R(0.5) * (w.x - z.x)};
            spec[(unsigned)((kz * Nx + xrow) * Ny + y0 + row)] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// With v.fold == 2 (the forward z pass ran fused with the last z sweep) the
// first workgroup of each solve's forward pass also forms the last step's
// norm from the z sweep's column partials, for the x convolution and calcnu.
template <int N, int DIR, typename R, int LM>
__global__ void __launch_bounds__((FftTile<N, LM>::WG))
k_ypass(Solves v, long nlines, const cpx<R> *__restrict__ twg, const double *__restrict__ wx, int tiles, int Nx) {
    constexpr int L = FftTile<N, LM>::L, WG = FftTile<N, LM>::WG;
    using Pl = FftPlan<N>;
    __shared__ cpx<R> buf[N * (L + 1)];
    __shared__ cpx<R> tw[N];
    cpx<R> *__restrict__ spec = reinterpret_cast<cpx<R> *>(solve_at(v).spec);
    const int tid = threadIdx.x;
    stage_twiddles(tw, twg, N, tid, WG);
    const long line0 = (long)blockIdx.x * L;
    const int nl = (int)min((long)L, nlines - line0);
    cpx<R> *g = spec + line0 * N;
    constexpr int K = L * N / WG;
    static_assert((L * N) % WG == 0, "ypass tile");
    cpx<R> r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid + k * WG;
        r[k] = g[i < nl * N ? i : 0];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid + k * WG;
        buf[Pl::template at<L, R>(i % N, i / N)] = r[k];
    }
    __syncthreads();
    Pl::template run<DIR, L>(buf, tid % L, tid / L, tw);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid + k * WG;
        if (i < nl * N) g[i] = buf[Pl::template at<L, R>(i % N, i / N)];
    }
    if (DIR < 0 && v.fold == 2 && blockIdx.x == 0) {
        const SolveAt sr = solve_at(v);
        __syncthreads();  // buf becomes the reduction scratch
        norm_from_cols<WG>(sr.rows, tiles, wx, Nx, sr.scal, v.shift, reinterpret_cast<double *>(buf));
    }
}

// ---------------------------------------------------------------------------
template <int N, typename R, int LM>
__global__ void __launch_bounds__((FftTile<N, LM>::WG))
k_xconv(Solves v, const R *__restrict__ potddq, int Ny, const cpx<R> *__restrict__ twg) {
    constexpr int L = FftTile<N, LM>::L, WG = FftTile<N, LM>::WG;
    constexpr int T = FftPlan<N>::T, K = N / T;  // x values per thread
    static_assert(L * T == WG, "xconv tile");
    constexpr int NXH = N / 2 + 1;
    __shared__ cpx<R> buf[N * (L + 1)];
    __shared__ cpx<R> tw[N];
    const SolveAt sr = solve_at(v);
    cpx<R> *__restrict__ spec = reinterpret_cast<cpx<R> *>(sr.spec);
    // folded norm: |psi|^2 was transformed unscaled, so potdd takes 1/norm^2
    const R s2 = v.fold ? (R)(sr.scal[SC_SCALE] * sr.scal[SC_SCALE]) : R(1);
    const int tid = threadIdx.x;
    const int ytiles = (Ny + L - 1) / L;
    const int kz = blockIdx.x / ytiles;
    // y tiles of one kz: the first half in order, the second half reversed,
    // so tile t and its mirror tile (ytiles - 1 - t, the same |ky| table
    // rows) are h workgroups apart.  Workgroups go to the 8 XCDs round
    // robin; with h a multiple of 8 (256 lines in tiles of 16) the pair runs
    // on one XCD at about the same time and the second reads the table rows
    // from that XCD's L2.
    const int p = blockIdx.x % ytiles, h = (ytiles + 1) / 2;
    const int y0 = (p < h ? p : ytiles - 1 - (p - h)) * L;
    const int nl = min(L, Ny - y0);
    stage_twiddles(tw, twg, N, tid, WG);
    // Thread (l, t) holds x = t + k T of y line l, lanes along y (coalesced
    // rows of L values): the layout in which the transforms take and leave
    // their line in registers, so the potdd product is applied in registers.
    const int l = tid % L, t = tid / L;
    const int y = y0 + (l < nl ? l : 0);
    cpx<R> *g = spec + (long)kz * N * Ny + y;
    cpx<R> r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = g[(unsigned)((t + k * T) * Ny)];
    // potdd(|kx|, |ky|, kz): the table rows [kz][|ky|] are contiguous in |kx|
    const int Ny2 = Ny / 2, Nyh = Ny2 + 1;
    const R *pk = potddq + ((long)kz * Nyh + (y <= Ny2 ? y : Ny - y)) * NXH;
    R pw[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int e = t + k * T;
        pw[k] = pk[e <= N / 2 ? e : N - e];
    }
    FftPlan<N>::template run_reg<-1, L>(buf, l, t, tw, r);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const R w = pw[k] * s2;
        r[k] = {r[k].x * w, r[k].y * w};
    }
    __syncthreads();  // the forward transform's last LDS reads are done
    FftPlan<N>::template run_reg<+1, L>(buf, l, t, tw, r);
    if (l < nl) {
#pragma unroll
        for (int k = 0; k < K; ++k) g[(unsigned)((t + k * T) * Ny)] = r[k];
    }
}

// ---------------------------------------------------------------------------
// Inverse z transform of two real rows per complex line: line p holds
//   Z[k] = A[k] + i B[k] (k <= N/2) and Z[N-k] = conj A[k] + i conj B[k],
// A, B the half spectra of the rows; the imaginary parts of the
// self-conjugate bins k = 0, N/2 are dropped.  SpecIdx(p, k, &ia, &ib) sets
// the spec indices of A and B and returns which of them exist (bit 0: A,
// bit 1: B); a missing one reads as zero.  All loads are issued before the
// first LDS write.
template <int N, int L, int WG, bool SWZ, typename R, typename SpecIdx>
__device__ __forceinline__ void load_hermitian_pairs(cpx<R> *buf, const cpx<R> *__restrict__ spec, int tid,
                                                     const SpecIdx &idx) {
    constexpr int nkz = N / 2 + 1;
    constexpr int K = (nkz * L + WG - 1) / WG;
    cpx<R> ra[K], rb[K];
    int m[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int i = tid + q * WG;
        const int p = i % L, k = i / L;
        unsigned ia = 0, ib = 0;
        m[q] = i < nkz * L ? idx(p, k, ia, ib) : 0;
        ra[q] = spec[(m[q] & 1) ? ia : 0u];
        rb[q] = spec[(m[q] & 2) ? ib : 0u];
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int i = tid + q * WG;
        const int p = i % L, k = i / L;
        if (i < nkz * L) {
            cpx<R> a = ra[q], b = rb[q];
            if (!(m[q] & 1)) a = {R(0), R(0)};
            if (!(m[q] & 2)) b = {R(0), R(0)};
            if (k == 0 || k == N / 2) {
                a.y = R(0);
                b.y = R(0);
            }
            buf[FftPlan<N>::template at<L, R, SWZ>(k, p)] = {a.x - b.y, a.y + b.x};
            if (k > 0 && k < N / 2) buf[FftPlan<N>::template at<L, R, SWZ>(N - k, p)] = {a.x + b.y, b.x - a.y};
        }
    }
}

// The same pairs straight into the registers of the first transform pass:
// v[r] = Z[t + r T] of line p = l (T threads per line), each bin above N/2
// formed from the spectrum rows of its mirror N - k.
template <int N, typename R, typename SpecIdx>
__device__ __forceinline__ void load_hermitian_reg(cpx<R> *v, const cpx<R> *__restrict__ spec, int l, int t,
                                                   const SpecIdx &idx) {
    constexpr int T = FftPlan<N>::T, K = N / T;
    cpx<R> ra[K], rb[K];
    int m[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const int k = t + r * T, kk = k <= N / 2 ? k : N - k;
        unsigned ia = 0, ib = 0;
        m[r] = idx(l, kk, ia, ib);
        ra[r] = spec[(m[r] & 1) ? ia : 0u];
        rb[r] = spec[(m[r] & 2) ? ib : 0u];
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const int k = t + r * T, kk = k <= N / 2 ? k : N - k;
        cpx<R> a = ra[r], b = rb[r];
        if (!(m[r] & 1)) a = {R(0), R(0)};
        if (!(m[r] & 2)) b = {R(0), R(0)};
        if (kk == 0 || kk == N / 2) {
            a.y = R(0);
            b.y = R(0);
        }
        v[r] = k <= N / 2 ? cpx<R>{a.x - b.y, a.y + b.x} : cpx<R>{a.x + b.y, b.x - a.y};
    }
}

template <int N, typename R, int LM>
__global__ void __launch_bounds__((FftTile<N, LM>::WG))
k_zinv(const cpx<R> *__restrict__ spec, R *__restrict__ phi, int Nx, int Ny, const cpx<R> *__restrict__ twg) {
    constexpr int L = FftTile<N, LM>::L, WG = FftTile<N, LM>::WG;
    using Pl = FftPlan<N>;
    __shared__ cpx<R> buf[N * (L + 1)];
    __shared__ cpx<R> tw[N];
    const int tid = threadIdx.x;
    const int ytiles = (Ny + 2 * L - 1) / (2 * L);
    const int x = blockIdx.x / ytiles;
    const int y0 = (blockIdx.x % ytiles) * 2 * L;
    const int nrows = min(2 * L, Ny - y0);
    stage_twiddles(tw, twg, N, tid, WG);
    load_hermitian_pairs<N, L, WG, false>(buf, spec, tid, [&](int p, int k, unsigned &ia, unsigned &ib) {
        ia = (unsigned)((k * Nx + x) * Ny + y0 + 2 * p);
        ib = ia + 1;
        return (2 * p < nrows ? 1 : 0) | (2 * p + 1 < nrows ? 2 : 0);
    });
    __syncthreads();
    Pl::template run<+1, L>(buf, tid % L, tid / L, tw);
    R *dst = phi + ((long)x * Ny + y0) * N;
    constexpr int Q = 2 * L * N / WG;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int i = tid + q * WG;
        const int z = i % N, row = i / N;
        if (row < nrows) {
            const cpx<R> c = buf[Pl::template at<L, R>(z, row >> 1)];
            dst[(unsigned)(row * N + z)] = (row & 1) ? c.y : c.x;
        }
    }
}


// ---------------------------------------------------------------------------
// zinv fused with calcnu (calcnu_kernel, imag3d.cu:1759-1783; real3d.cu
// 1856-1883).  Blocks [0, nfaceb) take the face rows x = Nx-1 or y = Ny-1,
// whose phi comes from other rows after the periodic face copies
// (calcpsidd2_kernel2a/2b, imag3d.cu:1591-1625):
//   row (Nx-1, y):          z < Nz-1 and y < Ny-1 from row (0, y), else own
//   row (x, Ny-1), x<Nx-1:  z < Nz-1 from row (x, 0), z = Nz-1 own.
// A face row is transformed together with its source row as the real and
// imaginary parts of one complex line; face f < Ny is (Nx-1, f), face f >= Ny
// is (f - Ny, Ny-1).  They go first so that they overlap the interior
// blocks instead of trailing them.  The remaining blocks take the interior
// rows x < Nx-1, y < Ny-1: within such a row the face copy of
// calcpsidd2_kernel2b (imag3d.cu:1618-1624) takes phi[.][.][Nz-1] from
// phi[.][.][0] of the same row, so the row's potential never leaves the chip.
// The tile's psi is fetched into registers before the transform when it is at
// most 256 bytes per thread (every plan with T = N / 8 or N / 4), so its
// latency overlaps the spectrum load and the passes; otherwise after it in
// batches of 8.  The per-row and per-z parts of the trap
// potential (trap_xy, trap_z) are staged in LDS.  With ZREG the spectrum
// pairs go straight into the registers of the first transform pass
// (load_hermitian_reg) instead of through LDS; with SWZ the transform buffer
// is the swizzled one (lds_at).  Neither changes the results.
template <int N, typename R, typename P, int LM, bool ZREG, bool SWZ>
__global__ void __launch_bounds__((FftTile<N, LM>::WG))
k_zinv_nu(Solves v, const double *__restrict__ sq, NuConst nc, int Nx, int Ny, int nfaceb,
          const cpx<R> *__restrict__ twg) {
    constexpr int L = FftTile<N, LM>::L, WG = FftTile<N, LM>::WG;
    using Pl = FftPlan<N>;
    constexpr int Q = 2 * L * N / WG;  // psi elements per thread (interior rows)
    // the tile's psi (at most 256 bytes per thread) is fetched before the
    // transform, so its latency overlaps the spectrum load and the passes
    constexpr bool PREFETCH = Q * sizeof(P) <= 256;
    constexpr int C = PREFETCH ? Q : (Q < 8 ? Q : 8);  // psi batch
    static_assert((2 * L * N) % WG == 0 && Q % C == 0, "zinv_nu tile");
    __shared__ cpx<R> buf[N * lds_slots_per_row<L, SWZ>()];
    // complex FP64 psi with the swizzled buffer: the passes read the
    // twiddles from the global table (L1 / L2), which leaves LDS for a third
    // workgroup per CU
    constexpr bool TWG = SWZ && std::is_same<P, cpx<double>>::value;
    __shared__ cpx<R> tw_s[TWG ? 1 : N];
    const cpx<R> *tw = TWG ? twg : tw_s;
    __shared__ double sqz[N];      // trap_z of each z
    __shared__ double sqr[2 * L];  // trap_xy of the tile's rows (faces: of each face row)
    const SolveAt sr = solve_at(v);
    const cpx<R> *__restrict__ spec = reinterpret_cast<const cpx<R> *>(sr.spec);
    P *__restrict__ psi = reinterpret_cast<P *>(sr.psi);
    const TrapConst tc = sr.p.tc;
    nc.g = sr.p.g;
    nc.gd = sr.p.gd;
    nc.shift = v.shift ? sr.scal[SC_SHIFT] : 0.0;
    nc.cap = v.shift;
    nc.capped = sr.scal + SC_CAPPED;
    nu_prepare(nc);
    const int tid = threadIdx.x;
    const double s = v.use_scale ? sr.scal[SC_SCALE] : 1.0;
    if constexpr (!TWG) stage_twiddles(tw_s, twg, N, tid, WG);
    for (int i = tid; i < N; i += WG) sqz[i] = trap_z(tc, sq[Nx + Ny + i]);
    if ((int)blockIdx.x >= nfaceb) {
        const int b = (int)blockIdx.x - nfaceb;
        const int ytiles = (Ny + 2 * L - 1) / (2 * L);
        const int x = b / ytiles;  // x < Nx - 1
        const int y0 = (b % ytiles) * 2 * L;
        const int nrows = min(2 * L, Ny - y0);
        P *dst = psi + ((long)x * Ny + y0) * N;
        // 2L rows; a workgroup of L * T threads covers them in one pass when
        // T >= 2, in two when T = 1 (N = 16)
        for (int r = tid; r < 2 * L; r += WG) sqr[r] = trap_xy(tc, sq[x], sq[Nx + min(y0 + r, Ny - 1)]);
        P pv[C];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int i = tid + q * WG;
                pv[q] = dst[(unsigned)(min(i / N, nrows - 1) * N + i % N)];
            }
        }
        const auto idx = [&](int p, int k, unsigned &ia, unsigned &ib) {
            ia = (unsigned)((k * Nx + x) * Ny + y0 + 2 * p);
            ib = ia + 1;
            return (2 * p < nrows ? 1 : 0) | (2 * p + 1 < nrows ? 2 : 0);
        };
        if constexpr (ZREG) {
            cpx<R> zv[N / FftPlan<N>::T];
            load_hermitian_reg<N>(zv, spec, tid % L, tid / L, idx);
            Pl::template run_from_reg<+1, L, SWZ>(buf, tid % L, tid / L, tw, zv);
        } else {
            load_hermitian_pairs<N, L, WG, SWZ>(buf, spec, tid, idx);
            __syncthreads();
            Pl::template run<+1, L, SWZ>(buf, tid % L, tid / L, tw);
        }
#pragma unroll
        for (int q0 = 0; q0 < Q; q0 += C) {
            if constexpr (!PREFETCH) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int i = tid + (q0 + c) * WG;
                    pv[c] = dst[(unsigned)(min(i / N, nrows - 1) * N + i % N)];
                }
            }
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int i = tid + (q0 + c) * WG;
                const int z = i % N, row = i / N;
                if (row < nrows && y0 + row < Ny - 1) {
                    const cpx<R> cv = buf[Pl::template at<L, R, SWZ>(z == N - 1 ? 0 : z, row >> 1)];
                    const R phi = (row & 1) ? cv.y : cv.x;
                    const double pot = trap_sum(tc, sqr[row], sqz[z]);
                    dst[(unsigned)(row * N + z)] = nu_apply(pv[PREFETCH ? q0 + c : c], s, phi, pot, nc);
                }
            }
        }
        return;
    }
    // face rows
    const int nface = Ny + Nx - 1;
    const int f0 = (int)blockIdx.x * L;
    const int nl = min(L, nface - f0);
    if (tid < L) {
        const int f = f0 + min(tid, nl - 1);
        sqr[tid] = trap_xy(tc, sq[f < Ny ? Nx - 1 : f - Ny], sq[Nx + (f < Ny ? f : Ny - 1)]);
    }
    const auto fidx = [&](int p, int k, unsigned &ia, unsigned &ib) {
        if (p >= nl) return 0;
        const int f = f0 + p;
        int xo, yo, xs, ys;
        if (f < Ny) { xo = Nx - 1; yo = f; xs = 0; ys = f; }
        else { xo = f - Ny; yo = Ny - 1; xs = f - Ny; ys = 0; }
        ia = (unsigned)((k * Nx + xo) * Ny + yo);
        ib = (unsigned)((k * Nx + xs) * Ny + ys);
        return 3;
    };
    if constexpr (ZREG) {
        cpx<R> zv[N / FftPlan<N>::T];
        load_hermitian_reg<N>(zv, spec, tid % L, tid / L, fidx);
        Pl::template run_from_reg<+1, L, SWZ>(buf, tid % L, tid / L, tw, zv);
    } else {
        load_hermitian_pairs<N, L, WG, SWZ>(buf, spec, tid, fidx);
        __syncthreads();
        Pl::template run<+1, L, SWZ>(buf, tid % L, tid / L, tw);
    }
    constexpr int QF = L * N / WG;
    constexpr int CF = PREFETCH ? QF : (QF < 8 ? QF : 8);
    static_assert((L * N) % WG == 0 && QF % CF == 0, "zinv_nu face tile");
#pragma unroll
    for (int q0 = 0; q0 < QF; q0 += CF) {
        P fv[CF];
        unsigned fn[CF];
#pragma unroll
        for (int c = 0; c < CF; ++c) {
            const int i = tid + (q0 + c) * WG;
            const int z = i % N, f = f0 + min(i / N, nl - 1);
            const int xo = f < Ny ? Nx - 1 : f - Ny, yo = f < Ny ? f : Ny - 1;
            fn[c] = (unsigned)((xo * Ny + yo) * N + z);
            fv[c] = psi[fn[c]];
        }
#pragma unroll
        for (int c = 0; c < CF; ++c) {
            const int i = tid + (q0 + c) * WG;
            const int z = i % N, p = i / N;
            if (p < nl) {
                const int f = f0 + p;
                const bool from_src = (f < Ny) ? (f < Ny - 1) && (z < N - 1) : (z < N - 1);
                const cpx<R> cv = buf[Pl::template at<L, R, SWZ>(z, p)];
                const R phi = from_src ? cv.y : cv.x;
                const double pot = trap_sum(tc, sqr[p], sqz[z]);
                psi[fn[c]] = nu_apply(fv[c], s, phi, pot, nc);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers for one length.  Each pass has a tile multiplier LM (1 or 2)
// chosen per precision and plan from measurements; the environment variables
// DBEC_FFT_LM_{ZFWD,YPASS,XCONV,FUSED}{64,32} override them (tuning runs).
inline int fft_lm(const char *name, int dflt) {
    const char *e = getenv(name);
    if (!e) return dflt;
    int v = atoi(e);
    return (v == 1 || v == 2) ? v : dflt;
}
// DBEC_FFT_ZINV_SWZ: transform buffer layout of the FP64 fused inverse pass
// (0 padded, 1 swizzled); -1 when unset
inline int fft_swz_env() {
    const char *e = getenv("DBEC_FFT_ZINV_SWZ");
    if (!e) return -1;
    const int v = atoi(e);
    return (v == 0 || v == 1) ? v : -1;
}

template <int N>
struct FftLaunchImpl {
    // Plans whose workgroup at LM = 1 has fewer than 192 threads take LM = 2
    // where wide tiles measured faster (FP32 passes, FP64 fused with real
    // psi); plans of 192 threads or more keep LM = 1 there (measured at
    // 256^3 and 192^3).  FP32 xconv always takes LM = 2, the other FP64
    // passes LM = 1.
    static constexpr int WIDE = FftTile<N, 1>::WG < 192 ? 2 : 1;

    template <int LM>
    static hipError_t zfwd_t(const FftArgs &a) {
        using Tl = FftTile<N, LM>;
        const int ytiles = (int)((a.Ny + 2 * Tl::L - 1) / (2 * Tl::L));
        dim3 grid((unsigned)(a.Nx * ytiles), (unsigned)a.v.nb), block(Tl::WG);
        const int Nx = (int)a.Nx, Ny = (int)a.Ny;
        if (!a.fp32) {
            auto *tw = static_cast<const cpx<double> *>(a.tw);
            if (a.cpsi) k_zfwd<N, double, cpx<double>, LM><<<grid, block, 0, a.stream>>>(a.v, Nx, Ny, tw, a.wx, a.tiles);
            else k_zfwd<N, double, double, LM><<<grid, block, 0, a.stream>>>(a.v, Nx, Ny, tw, a.wx, a.tiles);
        } else {
            auto *tw = static_cast<const cpx<float> *>(a.tw);
            if (a.cpsi) k_zfwd<N, float, cpx<float>, LM><<<grid, block, 0, a.stream>>>(a.v, Nx, Ny, tw, a.wx, a.tiles);
            else k_zfwd<N, float, float, LM><<<grid, block, 0, a.stream>>>(a.v, Nx, Ny, tw, a.wx, a.tiles);
        }
        return hipGetLastError();
    }
    // The tile multiplier of a pass: the environment variable, else the
    // context's tuned choice, else the rule.
    static int pick_lm(const FftArgs &a, int env, int pass, int rule) {
        if (env) return env;
        return a.tune.lm[pass] ? a.tune.lm[pass] : rule;
    }
    static hipError_t zfwd(const FftArgs &a) {
        static const int l64 = fft_lm("DBEC_FFT_LM_ZFWD64", 0), l32 = fft_lm("DBEC_FFT_LM_ZFWD32", 0);
        return pick_lm(a, a.fp32 ? l32 : l64, 0, a.fp32 ? WIDE : 1) == 2 ? zfwd_t<2>(a) : zfwd_t<1>(a);
    }
    // FP32 spectra use tiles twice as wide, so every strided access still
    // moves 128-256 contiguous bytes.

    template <int LM>
    static hipError_t ypass_t(int dir, const FftArgs &a) {
        using Tl = FftTile<N, LM>;
        const long nlines = a.Nkz * a.Nx;
        dim3 grid((unsigned)((nlines + Tl::L - 1) / Tl::L), (unsigned)a.v.nb), block(Tl::WG);
        if (!a.fp32) {
            auto *tw = static_cast<const cpx<double> *>(a.tw);
            if (dir < 0) k_ypass<N, -1, double, LM><<<grid, block, 0, a.stream>>>(a.v, nlines, tw, a.wx, a.tiles, (int)a.Nx);
            else k_ypass<N, +1, double, LM><<<grid, block, 0, a.stream>>>(a.v, nlines, tw, a.wx, a.tiles, (int)a.Nx);
        } else {
            auto *tw = static_cast<const cpx<float> *>(a.tw);
            if (dir < 0) k_ypass<N, -1, float, LM><<<grid, block, 0, a.stream>>>(a.v, nlines, tw, a.wx, a.tiles, (int)a.Nx);
            else k_ypass<N, +1, float, LM><<<grid, block, 0, a.stream>>>(a.v, nlines, tw, a.wx, a.tiles, (int)a.Nx);
        }
        return hipGetLastError();
    }
    static hipError_t ypass(int dir, const FftArgs &a) {
        static const int l64 = fft_lm("DBEC_FFT_LM_YPASS64", 0), l32 = fft_lm("DBEC_FFT_LM_YPASS32", 0);
        return pick_lm(a, a.fp32 ? l32 : l64, 1, a.fp32 ? WIDE : 1) == 2 ? ypass_t<2>(dir, a) : ypass_t<1>(dir, a);
    }

    template <int LM>
    static hipError_t xconv_t(const FftArgs &a) {
        using Tl = FftTile<N, LM>;
        const int ytiles = (int)((a.Ny + Tl::L - 1) / Tl::L);
        dim3 grid((unsigned)(a.Nkz * ytiles), (unsigned)a.v.nb), block(Tl::WG);
        if (!a.fp32)
            k_xconv<N, double, LM><<<grid, block, 0, a.stream>>>(a.v, static_cast<const double *>(a.potddq),
                                                                 (int)a.Ny, static_cast<const cpx<double> *>(a.tw));
        else
            k_xconv<N, float, LM><<<grid, block, 0, a.stream>>>(a.v, static_cast<const float *>(a.potddq),
                                                                (int)a.Ny, static_cast<const cpx<float> *>(a.tw));
        return hipGetLastError();
    }
    static hipError_t xconv(const FftArgs &a) {
        static const int l64 = fft_lm("DBEC_FFT_LM_XCONV64", 0), l32 = fft_lm("DBEC_FFT_LM_XCONV32", 0);
        return pick_lm(a, a.fp32 ? l32 : l64, 2, a.fp32 ? 2 : 1) == 2 ? xconv_t<2>(a) : xconv_t<1>(a);
    }

    template <typename S, typename R, bool COLS, int L>
    static hipError_t sweepz_zfwd_t(const FftArgs &a) {
        using A = typename Arith<S>::T;
        if constexpr (!FusedZ<N>::template fits<S, R, L>()) {
            return hipErrorInvalidValue;
        } else {
            const int yt = (int)((a.Ny + L - 1) / L);
            dim3 grid((unsigned)(a.Nx * yt), (unsigned)a.v.nb), block(L * (N / 16));
            using TT = typename TabOf<A>::T;
            const SweepTab<A> tab{static_cast<const TT *>(a.ztab[0]), static_cast<const TT *>(a.ztab[1]),
                                  static_cast<const TT *>(a.ztab[2]), static_cast<const TT *>(a.ztab[3])};
            SweepConst<A> k;
            zconst(a, k);
            k_sweepz_zfwd<N, S, R, L, COLS><<<grid, block, 0, a.stream>>>(a.v, (int)a.Nx, (int)a.Ny, tab, k, a.wz, a.wy,
                                                                         static_cast<const cpx<R> *>(a.tw));
            return hipGetLastError();
        }
    }
    static void zconst(const FftArgs &a, SweepConst<double> &k) {
        k.A0r = a.zk[0][0];
        k.A = a.zk[1][0];
        k.mA = a.zk[2][0];
    }
    static void zconst(const FftArgs &a, SweepConst<float> &k) {
        k.A0r = (float)a.zk[0][0];
        k.A = (float)a.zk[1][0];
        k.mA = (float)a.zk[2][0];
    }
    static void zconst(const FftArgs &a, SweepConst<cpx<double>> &k) {
        k.A0r = {a.zk[0][0], a.zk[0][1]};
        k.A = {a.zk[1][0], a.zk[1][1]};
        k.mA = {a.zk[2][0], a.zk[2][1]};
    }
    static void zconst(const FftArgs &a, SweepConst<cpx<float>> &k) {
        k.A0r = {(float)a.zk[0][0], (float)a.zk[0][1]};
        k.A = {(float)a.zk[1][0], (float)a.zk[1][1]};
        k.mA = {(float)a.zk[2][0], (float)a.zk[2][1]};
    }
    template <typename S, typename R, bool COLS>
    static hipError_t sweepz_zfwd_rows(const FftArgs &a) {
        switch (a.zrows) {
        case 64: return sweepz_zfwd_t<S, R, COLS, 64>(a);
        case 32: return sweepz_zfwd_t<S, R, COLS, 32>(a);
        case 16: return sweepz_zfwd_t<S, R, COLS, 16>(a);
        case 8: return sweepz_zfwd_t<S, R, COLS, 8>(a);
        case 4: return sweepz_zfwd_t<S, R, COLS, 4>(a);
        default: return hipErrorInvalidValue;
        }
    }
    static hipError_t sweepz_zfwd(const FftArgs &a) {
        if (!a.fp32) {
            if (a.cpsi) return sweepz_zfwd_rows<cpx<double>, double, false>(a);
            return sweepz_zfwd_rows<double, double, true>(a);
        }
        if (a.cpsi) return sweepz_zfwd_rows<cpx<float>, float, false>(a);
        return sweepz_zfwd_rows<float, float, true>(a);
    }
    template <typename S, typename R>
    static bool fits_rows(int zrows) {
        switch (zrows) {
        case 64: return FusedZ<N>::template fits<S, R, 64>();
        case 32: return FusedZ<N>::template fits<S, R, 32>();
        case 16: return FusedZ<N>::template fits<S, R, 16>();
        case 8: return FusedZ<N>::template fits<S, R, 8>();
        case 4: return FusedZ<N>::template fits<S, R, 4>();
        default: return false;
        }
    }
    static bool fusable(bool fp32, bool cpsi, int zrows) {
        if (!fp32) return cpsi ? fits_rows<cpx<double>, double>(zrows) : fits_rows<double, double>(zrows);
        return cpsi ? fits_rows<cpx<float>, float>(zrows) : fits_rows<float, float>(zrows);
    }

    // single solve: spectrum of `one` -> phi
    static hipError_t zinv(const FftArgs &a) {
        using Tl = FftTile<N, 1>;
        const int ytiles = (int)((a.Ny + 2 * Tl::L - 1) / (2 * Tl::L));
        dim3 grid((unsigned)(a.Nx * ytiles)), block(Tl::WG);
        if (!a.fp32)
            k_zinv<N, double, 1><<<grid, block, 0, a.stream>>>(static_cast<const cpx<double> *>(a.v.spec),
                                                               static_cast<double *>(a.phi), (int)a.Nx, (int)a.Ny,
                                                               static_cast<const cpx<double> *>(a.tw));
        else
            k_zinv<N, float, 1><<<grid, block, 0, a.stream>>>(static_cast<const cpx<float> *>(a.v.spec),
                                                              static_cast<float *>(a.phi), (int)a.Nx, (int)a.Ny,
                                                              static_cast<const cpx<float> *>(a.tw));
        return hipGetLastError();
    }

    template <typename P, typename R, int LM>
    static hipError_t fused_t(const FftArgs &a) {
        using Tl = FftTile<N, LM>;
        const auto *tw = static_cast<const cpx<R> *>(a.tw);
        const int ytiles = (int)((a.Ny + 2 * Tl::L - 1) / (2 * Tl::L));
        const int nmain = (int)((a.Nx - 1) * ytiles);
        const int nfaceb = (int)((a.Ny + a.Nx - 1 + Tl::L - 1) / Tl::L);
        dim3 grid((unsigned)(nmain + nfaceb), (unsigned)a.v.nb), block(Tl::WG);
        // FP32 spectra load straight into the first pass's registers (faster at
        // every measured size); FP64 ones through LDS (the extra registers
        // cost more occupancy than the LDS round trip saves).  FP64 has both
        // transform buffer layouts (swizzled: real psi faster at 256, slower
        // at 160 and 192; complex psi, twiddles from the global table, 234 ->
        // 215 us at 256x256x384, slower at 192): the environment variable,
        // else the tuned choice, else padded.
        constexpr bool ZREG = std::is_same<R, float>::value;
        if constexpr (std::is_same<R, double>::value) {
            static const int env = fft_swz_env();
            if (env >= 0 ? env == 1 : a.tune.zinv_swz == 1) {
                k_zinv_nu<N, R, P, LM, ZREG, true><<<grid, block, 0, a.stream>>>(a.v, a.sq, a.nc, (int)a.Nx, (int)a.Ny,
                                                                                  nfaceb, tw);
                return hipGetLastError();
            }
        }
        k_zinv_nu<N, R, P, LM, ZREG, false><<<grid, block, 0, a.stream>>>(a.v, a.sq, a.nc, (int)a.Nx, (int)a.Ny,
                                                                           nfaceb, tw);
        return hipGetLastError();
    }
    template <int LM>
    static hipError_t fused_lm(const FftArgs &a) {
        if (!a.fp32) return a.cpsi ? fused_t<cpx<double>, double, LM>(a) : fused_t<double, double, LM>(a);
        return a.cpsi ? fused_t<cpx<float>, float, LM>(a) : fused_t<float, float, LM>(a);
    }
    static hipError_t fused(const FftArgs &a) {
        // FP64: narrow tiles for complex psi
        static const int l64 = fft_lm("DBEC_FFT_LM_FUSED64", 0), l32 = fft_lm("DBEC_FFT_LM_FUSED32", 0);
        const int lm = pick_lm(a, a.fp32 ? l32 : l64, 3, (a.fp32 || !a.cpsi) ? WIDE : 1);
        return lm == 2 ? fused_lm<2>(a) : fused_lm<1>(a);
    }
};

template <int N>
hipError_t FftLaunch<N>::zfwd(const FftArgs &a) { return FftLaunchImpl<N>::zfwd(a); }
template <int N>
hipError_t FftLaunch<N>::ypass(int dir, const FftArgs &a) { return FftLaunchImpl<N>::ypass(dir, a); }
template <int N>
hipError_t FftLaunch<N>::xconv(const FftArgs &a) { return FftLaunchImpl<N>::xconv(a); }
template <int N>
hipError_t FftLaunch<N>::zinv(const FftArgs &a) { return FftLaunchImpl<N>::zinv(a); }
template <int N>
hipError_t FftLaunch<N>::fused(const FftArgs &a) { return FftLaunchImpl<N>::fused(a); }
template <int N>
hipError_t FftLaunch<N>::sweepz_zfwd(const FftArgs &a) { return FftLaunchImpl<N>::sweepz_zfwd(a); }
template <int N>
bool FftLaunch<N>::fusable(bool fp32, bool cpsi, int zrows) { return FftLaunchImpl<N>::fusable(fp32, cpsi, zrows); }

}  // namespace dbec
