// fft_device.hpp - LDS-resident Stockham FFT building blocks for gfx950.
//
// A workgroup holds L lines of length N in LDS, element e of line l in slot
// lds_at(e, l): e-major, line index fastest, padded or XOR-swizzled against
// bank conflicts.  Every line is transformed by T threads; thread (l, t) has
// l = tid % L, t = tid / L.
// Pass p of radix R (span NS = product of the earlier radices) computes
// butterfly j from inputs j + r*N/R and writes outputs to
// (j / NS) * NS * R + j % NS + r * NS (Stockham autosort: natural order in and
// out, no bit reversal).  Twiddles come from an N-entry table of
// exp(-2 pi i t / N) staged in LDS; DIR = -1 is the forward transform and +1
// the (unnormalised) inverse.
#pragma once

#include <hip/hip_runtime.h>

namespace dbec {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R x, y;
};

// 16-byte packets of storage elements: one global_load/store_dwordx4 per lane.
// Staging loops load every packet of a tile into registers first (addresses
// clamped into the array instead of branching), so a wavefront has all its
// loads in flight at once, and only then write the tile to LDS.
template <typename T>
struct Pk {
    static constexpr int n = sizeof(T) >= 16 ? 1 : 16 / (int)sizeof(T);
    struct alignas(16) type {
        T e[n];
    };
};
template <typename T>
__device__ __forceinline__ typename Pk<T>::type ld_pk(const T *p) {
    return *reinterpret_cast<const typename Pk<T>::type *>(p);
}
template <typename T>
__device__ __forceinline__ void st_pk(T *p, const typename Pk<T>::type &v) {
    *reinterpret_cast<typename Pk<T>::type *>(p) = v;
}

template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }

// a * (DIR * i)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> mul_i(cpx<R> a) {
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}
// a * exp(DIR * i * theta) for the constant (c, s) = (cos theta, sin theta)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> rot(cpx<R> a, R c, R s) {
    if constexpr (DIR < 0) return {a.x * c + a.y * s, a.y * c - a.x * s};
    else return {a.x * c - a.y * s, a.y * c + a.x * s};
}

// ---- small DFTs: X[k] = sum_n v[n] exp(DIR 2 pi i n k / R) -------------------

template <int DIR, typename R>
__device__ __forceinline__ void dft2(cpx<R> *v) {
    cpx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4(cpx<R> &v0, cpx<R> &v1, cpx<R> &v2, cpx<R> &v3) {
    cpx<R> t0 = v0 + v2, t1 = v0 - v2, t2 = v1 + v3, t3 = mul_i<DIR>(v1 - v3);
    v0 = t0 + t2;
    v2 = t0 - t2;
    v1 = t1 + t3;
    v3 = t1 - t3;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft3(cpx<R> *v) {
    const R c1 = R(-0.5), s1 = R(0.86602540378443864676372317075294);
    cpx<R> t1 = v[1] + v[2];
    cpx<R> t2 = mul_i<DIR>(v[1] - v[2]);
    cpx<R> m = {v[0].x + c1 * t1.x, v[0].y + c1 * t1.y};
    v[0] = v[0] + t1;
    v[1] = {m.x + s1 * t2.x, m.y + s1 * t2.y};
    v[2] = {m.x - s1 * t2.x, m.y - s1 * t2.y};
}

template <int DIR, typename R>
__device__ __forceinline__ void dft5(cpx<R> *v) {
    const R c1 = R(0.30901699437494742410229341718282);   // cos(2 pi / 5)
    const R c2 = R(-0.80901699437494742410229341718282);  // cos(4 pi / 5)
    const R s1 = R(0.95105651629515357211643933337938);   // sin(2 pi / 5)
    const R s2 = R(0.58778525229247312916870595463907);   // sin(4 pi / 5)
    cpx<R> a1 = v[1] + v[4], b1 = v[1] - v[4], a2 = v[2] + v[3], b2 = v[2] - v[3];
    cpx<R> p1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
    cpx<R> p2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
    cpx<R> q1 = mul_i<DIR>(cpx<R>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
    cpx<R> q2 = mul_i<DIR>(cpx<R>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
    v[0] = v[0] + a1 + a2;
    v[1] = p1 + q1;
    v[4] = p1 - q1;
    v[2] = p2 + q2;
    v[3] = p2 - q2;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft8(cpx<R> *v) {
    const R h = R(0.70710678118654752440084436210485);
    // even / odd DFT4s, then the radix-2 combination with W8^k
    cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = rot<DIR>(o1, h, h);
    o2 = mul_i<DIR>(o2);
    o3 = rot<DIR>(o3, -h, h);
    v[0] = e0 + o0; v[4] = e0 - o0;
    v[1] = e1 + o1; v[5] = e1 - o1;
    v[2] = e2 + o2; v[6] = e2 - o2;
    v[3] = e3 + o3; v[7] = e3 - o3;
}

// 16 = 4 x 4: n = n1 + 4 n2, k = 4 k1 + k2.
template <int DIR, typename R>
__device__ __forceinline__ void dft16(cpx<R> *v) {
    const R c1 = R(0.92387953251128675612818318939679);  // cos(pi/8)
    const R s1 = R(0.38268343236508977172845998403040);  // sin(pi/8)
    const R h = R(0.70710678118654752440084436210485);
    // step 1: DFT4 over n2 for each n1
    dft4<DIR>(v[0], v[4], v[8], v[12]);
    dft4<DIR>(v[1], v[5], v[9], v[13]);
    dft4<DIR>(v[2], v[6], v[10], v[14]);
    dft4<DIR>(v[3], v[7], v[11], v[15]);
    // step 2: twiddle Y[n1][k2] (stored at v[n1 + 4 k2]) by W16^(n1 k2)
    v[5] = rot<DIR>(v[5], c1, s1);    // n1=1,k2=1: W^1
    v[9] = rot<DIR>(v[9], h, h);      // n1=1,k2=2: W^2
    v[13] = rot<DIR>(v[13], s1, c1);  // n1=1,k2=3: W^3
    v[6] = rot<DIR>(v[6], h, h);      // n1=2,k2=1: W^2
    v[10] = mul_i<DIR>(v[10]);        // n1=2,k2=2: W^4
    v[14] = rot<DIR>(v[14], -h, h);   // n1=2,k2=3: W^6
    v[7] = rot<DIR>(v[7], s1, c1);    // n1=3,k2=1: W^3
    v[11] = rot<DIR>(v[11], -h, h);   // n1=3,k2=2: W^6
    v[15] = rot<DIR>(v[15], -c1, -s1);  // n1=3,k2=3: W^9
    // step 3: DFT4 over n1 for each k2; result X[4 k1 + k2] sits at v[k1 + 4 k2]
    dft4<DIR>(v[0], v[1], v[2], v[3]);
    dft4<DIR>(v[4], v[5], v[6], v[7]);
    dft4<DIR>(v[8], v[9], v[10], v[11]);
    dft4<DIR>(v[12], v[13], v[14], v[15]);
    // transpose the 4x4 index map back to natural order
    cpx<R> t;
    t = v[1]; v[1] = v[4]; v[4] = t;
    t = v[2]; v[2] = v[8]; v[8] = t;
    t = v[3]; v[3] = v[12]; v[12] = t;
    t = v[6]; v[6] = v[9]; v[9] = t;
    t = v[7]; v[7] = v[13]; v[13] = t;
    t = v[11]; v[11] = v[14]; v[14] = t;
}

template <int RADIX, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R> *v) {
    if constexpr (RADIX == 2) dft2<DIR>(v);
    else if constexpr (RADIX == 3) dft3<DIR>(v);
    else if constexpr (RADIX == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (RADIX == 5) dft5<DIR>(v);
    else if constexpr (RADIX == 8) dft8<DIR>(v);
    else if constexpr (RADIX == 16) dft16<DIR>(v);
    else static_assert(RADIX == 2, "unsupported radix");
}

constexpr int ilog2(int v) { return v > 1 ? 1 + ilog2(v / 2) : 0; }

// LDS slot of element e of line l in a buffer of L transform lines for the
// plan of length N with T threads per line.  Padded (SWZ false):
// e (L + 1) + l, N (L + 1) slots.  Swizzled (SWZ true, L a power of two):
//   (e ^ ((e >> log2 R1) & 3)) * L + (l ^ ((e >> S1) & (L - 1))),
// R1 = N / T the first radix, S1 = log2 of the lines' rows per 256 bytes, a
// permutation of [0, N L).  Its row swizzle spreads the first pass's outputs
// (rows R1 t + r of four consecutive t) over four bank groups, its column
// swizzle spreads a run of consecutive e of one line (the coalesced row
// copies) over all banks, and the other passes touch four consecutive rows of
// every line at once; on 64 banks of 4 bytes, serving 256 bytes per cycle,
// that removes the bank conflicts the padded layout has in the passes.  It
// costs address registers: measured, it pays only in the FP64 fused inverse
// pass (real psi at 256; complex psi at 384, where dropping the padding and
// the LDS twiddles lets a third workgroup onto each CU).
template <int L, bool SWZ>
constexpr int lds_slots_per_row() { return SWZ ? L : L + 1; }

template <int N, int T, int L, typename R, bool SWZ>
__device__ __forceinline__ int lds_at(int e, int l) {
    if constexpr (!SWZ) {
        return e * (L + 1) + l;
    } else {
        static_assert((L & (L - 1)) == 0, "lines per buffer: a power of two");
        constexpr int SH = ilog2(N / T);
        constexpr int RB = 256 / (L * (int)sizeof(cpx<R>));
        constexpr int S1 = RB > 1 ? ilog2(RB) : 0;
        return (e ^ ((e >> SH) & 3)) * L + (l ^ ((e >> S1) & (L - 1)));
    }
}

// One Stockham pass over the L lines held in LDS.
template <int N, int RADIX, int NS, int T, int DIR, int L, bool SWZ, typename R>
__device__ __forceinline__ void stockham_pass(cpx<R> *buf, int l, int t, const cpx<R> *tw) {
    constexpr int M = N / RADIX;
    constexpr int Q = (M + T - 1) / T;
    cpx<R> v[Q][RADIX];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + q * T;
        if (M % T == 0 || j < M) {
#pragma unroll
            for (int r = 0; r < RADIX; ++r) v[q][r] = buf[lds_at<N, T, L, R, SWZ>(j + r * M, l)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + q * T;
        if (M % T == 0 || j < M) {
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RADIX; ++r) {
                    const cpx<R> w = tw[r * k * (N / (NS * RADIX))];
                    v[q][r] = (DIR < 0) ? cmul(v[q][r], w) : cmulc(v[q][r], w);
                }
            }
            dft<RADIX, DIR>(v[q]);
            const int base = (j / NS) * NS * RADIX + k;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) buf[lds_at<N, T, L, R, SWZ>(base + r * NS, l)] = v[q][r];
        }
    }
    __syncthreads();
}

template <int N, int T, int DIR, int L, bool SWZ, int NS, int RADIX, int... REST, typename R>
__device__ __forceinline__ void stockham(cpx<R> *buf, int l, int t, const cpx<R> *tw) {
    stockham_pass<N, RADIX, NS, T, DIR, L, SWZ>(buf, l, t, tw);
    if constexpr (sizeof...(REST) > 0) stockham<N, T, DIR, L, SWZ, NS * RADIX, REST...>(buf, l, t, tw);
}

// The same transform with the line in registers at both ends: on entry and on
// return v[k] holds element t + k T of line l (N / T values per thread).  The
// first pass (T = N / RADIX: one butterfly per thread, no twiddles) takes its
// inputs from v, and when its butterflies divide evenly among the T threads
// the last pass (NS = N / RADIX) leaves its outputs in v: butterfly j = t + q T
// yields elements j + r NS = t + (q + r Q) T.  The arithmetic is that of
// stockham(); only LDS round trips and barriers go.  buf must be free on entry
// (the caller's barrier); on return other threads may still be reading it.
template <int N, int RADIX, int T, int DIR, int L, bool SWZ, typename R>
__device__ __forceinline__ void stockham_first_reg(cpx<R> *buf, int l, int t, cpx<R> *v) {
    static_assert(N / RADIX == T, "first pass: one butterfly per thread");
    dft<RADIX, DIR>(v);
#pragma unroll
    for (int r = 0; r < RADIX; ++r) buf[lds_at<N, T, L, R, SWZ>(t * RADIX + r, l)] = v[r];
    __syncthreads();
}

template <int N, int RADIX, int NS, int T, int DIR, int L, bool SWZ, typename R>
__device__ __forceinline__ void stockham_last_reg(const cpx<R> *buf, int l, int t, const cpx<R> *tw, cpx<R> *out) {
    constexpr int M = N / RADIX, Q = M / T;
    static_assert(NS == M && M % T == 0, "last pass: whole butterflies per thread");
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + q * T;
        cpx<R> v[RADIX];
#pragma unroll
        for (int r = 0; r < RADIX; ++r) v[r] = buf[lds_at<N, T, L, R, SWZ>(j + r * M, l)];
#pragma unroll
        for (int r = 1; r < RADIX; ++r) {
            const cpx<R> w = tw[r * j];
            v[r] = (DIR < 0) ? cmul(v[r], w) : cmulc(v[r], w);
        }
        dft<RADIX, DIR>(v);
#pragma unroll
        for (int r = 0; r < RADIX; ++r) out[q + r * Q] = v[r];
    }
}

template <int N, int T, int DIR, int L, bool SWZ, int NS, int RADIX, int... REST, typename R>
__device__ __forceinline__ void stockham_reg_tail(cpx<R> *buf, int l, int t, const cpx<R> *tw, cpx<R> *v) {
    if constexpr (sizeof...(REST) > 0) {
        stockham_pass<N, RADIX, NS, T, DIR, L, SWZ>(buf, l, t, tw);
        stockham_reg_tail<N, T, DIR, L, SWZ, NS * RADIX, REST...>(buf, l, t, tw, v);
    } else if constexpr ((N / RADIX) % T == 0) {
        stockham_last_reg<N, RADIX, NS, T, DIR, L, SWZ>(buf, l, t, tw, v);
    } else {
        stockham_pass<N, RADIX, NS, T, DIR, L, SWZ>(buf, l, t, tw);
#pragma unroll
        for (int k = 0; k < N / T; ++k) v[k] = buf[lds_at<N, T, L, R, SWZ>(t + k * T, l)];
    }
}

template <int N, int T, int DIR, int L, bool SWZ, int RADIX, int... REST, typename R>
__device__ __forceinline__ void stockham_reg(cpx<R> *buf, int l, int t, const cpx<R> *tw, cpx<R> *v) {
    if constexpr (sizeof...(REST) == 0) {
        dft<RADIX, DIR>(v);  // one pass, T = 1: the whole line in one thread
    } else {
        stockham_first_reg<N, RADIX, T, DIR, L, SWZ>(buf, l, t, v);
        stockham_reg_tail<N, T, DIR, L, SWZ, RADIX, REST...>(buf, l, t, tw, v);
    }
}

// Registers in (v as for stockham_reg), result in LDS as stockham() leaves it.
template <int N, int T, int DIR, int L, bool SWZ, int RADIX, int... REST, typename R>
__device__ __forceinline__ void stockham_from_reg(cpx<R> *buf, int l, int t, const cpx<R> *tw, cpx<R> *v) {
    stockham_first_reg<N, RADIX, T, DIR, L, SWZ>(buf, l, t, v);
    if constexpr (sizeof...(REST) > 0) stockham<N, T, DIR, L, SWZ, RADIX, REST...>(buf, l, t, tw);
}

// Radix plan per supported length.  T = threads per line = N / (first radix).
// First radices of 4 or 8 give T >= 12 threads per line: fewer values per
// thread (fewer registers in the fused passes, whose psi tile is 2N / T
// values per thread) and workgroups of whole wavefronts; at every length but
// 384 these measured faster than plans that start with radix 16.  at<L, R>()
// is the LDS slot of element e of line l (lds_at).
template <int N>
struct FftPlan;

#define DBEC_FFT_PLAN(NN, TT, ...)                                                           \
    template <>                                                                              \
    struct FftPlan<NN> {                                                                     \
        static constexpr int T = TT;                                                         \
        template <int L, typename R, bool SWZ = false>                                       \
        static __device__ __forceinline__ int at(int e, int l) {                             \
            return lds_at<NN, TT, L, R, SWZ>(e, l);                                          \
        }                                                                                    \
        template <int DIR, int L, bool SWZ = false, typename R>                              \
        static __device__ __forceinline__ void run(cpx<R> *buf, int l, int t, const cpx<R> *tw) { \
            stockham<NN, TT, DIR, L, SWZ, 1, __VA_ARGS__>(buf, l, t, tw);                    \
        }                                                                                    \
        template <int DIR, int L, bool SWZ = false, typename R>                              \
        static __device__ __forceinline__ void run_reg(cpx<R> *buf, int l, int t,            \
                                                       const cpx<R> *tw, cpx<R> *v) {        \
            stockham_reg<NN, TT, DIR, L, SWZ, __VA_ARGS__>(buf, l, t, tw, v);                \
        }                                                                                    \
        template <int DIR, int L, bool SWZ = false, typename R>                              \
        static __device__ __forceinline__ void run_from_reg(cpx<R> *buf, int l, int t,      \
                                                            const cpx<R> *tw, cpx<R> *v) {   \
            stockham_from_reg<NN, TT, DIR, L, SWZ, __VA_ARGS__>(buf, l, t, tw, v);           \
        }                                                                                    \
    };

DBEC_FFT_PLAN(16, 1, 16)
DBEC_FFT_PLAN(32, 4, 8, 4)
DBEC_FFT_PLAN(48, 12, 4, 4, 3)
DBEC_FFT_PLAN(64, 16, 4, 4, 4)
DBEC_FFT_PLAN(80, 20, 4, 4, 5)
DBEC_FFT_PLAN(96, 12, 8, 4, 3)
DBEC_FFT_PLAN(128, 16, 8, 16)
DBEC_FFT_PLAN(160, 20, 8, 4, 5)
DBEC_FFT_PLAN(192, 24, 8, 8, 3)
DBEC_FFT_PLAN(256, 32, 8, 8, 4)
DBEC_FFT_PLAN(320, 40, 8, 8, 5)
DBEC_FFT_PLAN(384, 24, 16, 8, 3)
DBEC_FFT_PLAN(512, 64, 8, 8, 8)

#undef DBEC_FFT_PLAN

// Apply F(N) to every supported length.
#define DBEC_FOR_EACH_FFT_LEN(F) F(16) F(32) F(48) F(64) F(80) F(96) F(128) F(160) F(192) F(256) F(320) F(384) F(512)

inline bool fft_len_supported(long n) {
    switch (n) {
    case 16: case 32: case 48: case 64: case 80: case 96: case 128: case 160: case 192:
    case 256: case 320: case 384: case 512:
        return true;
    default:
        return false;
    }
}

}  // namespace dbec
