// ops.hip - pointwise, sweep and reduction kernels of the split step.
//
// Reference aliases: imag3d.cu = src/imag3d-mpicuda/imag3d-mpicuda.cu,
// real3d.cu = src/real3d-mpicuda/real3d-mpicuda.cu, diffint.cu =
// src/utils/diffint.cu.
#include <cstdlib>
#include <type_traits>

#include "norm_device.hpp"
#include "ops.hpp"
#include "sweep_device.hpp"

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
    psi[n] = nu_apply(psi[n], s, phi[n], pot, nc);
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
    nd_split(nc, nd);
    nc.shift = 0.0;
    nc.cap = 0;
    nc.capped = nullptr;
    nu_prepare(nc);
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

__global__ void k_set_scalar(double *p, double v) { *p = v; }

hipError_t launch_set_scalar(hipStream_t s, double *p, double v) {
    k_set_scalar<<<1, 1, 0, s>>>(p, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// x and y sweeps: a workgroup covers L consecutive z (lanes; one 256-byte row
// per load) of the lines of one y (x sweep) or x (y sweep) plane; thread
// (l, s) loads its segment straight from HBM into registers.
template <typename S, int AXIS, int L, int M, bool PEEL>
__global__ void __launch_bounds__(512)
k_sweep_xy(Solves sv, int Nx, int Ny, int Nz, SweepTab<typename Arith<S>::T> tab,
           SweepConst<typename Arith<S>::T> k) {
    using A = typename Arith<S>::T;
    extern __shared__ char smem_raw[];
    S *__restrict__ psi = reinterpret_cast<S *>(solve_at(sv).psi);
    const int n = AXIS == 0 ? Nx : Ny;
    const int P = n / M;
    const int nt = L * P;
    SweepShared<A, M> sh(smem_raw, n, nt, L);
    const int tid = threadIdx.x;
    const int l = tid % L, s = tid / L;
    const int zt = (Nz + L - 1) / L;
    const int pl = blockIdx.x / zt, z0 = (blockIdx.x % zt) * L;
    const int nl = min(L, Nz - z0);
    const bool act = l < nl;
    const long base = (AXIS == 0 ? (long)pl * Nz : (long)pl * Ny * Nz) + z0 + l;
    const long es = AXIS == 0 ? (long)Ny * Nz : (long)Nz;
    const int e0 = s * M;
    const A zero = m_zero(k.A);
    const CoefPre<A> cp = coef_pre(tab, n, tid);
    // The segment's elements at 32-bit byte offsets j * es from its first one
    // (at most 16 * 512^2 * 16 bytes): one 64-bit address per thread, then a
    // 32-bit step per element instead of a 64-bit multiply-add.
    char *const seg = reinterpret_cast<char *>(psi + base + (long)e0 * es);
    const unsigned bes = (unsigned)(es * (long)sizeof(S));
    auto at = [&](int j) -> S & { return *reinterpret_cast<S *>(seg + (unsigned)j * bes); };
    // Predicated loads converted in place: fewer registers (higher occupancy)
    // than issuing them into a separate array, which measured faster here.
    A v[M];
#pragma unroll
    for (int j = 0; j < M; ++j) v[j] = act ? to_a(at(j)) : zero;
    const A hl = (act && e0 > 0) ? to_a(*reinterpret_cast<const S *>(seg - bes)) : zero;
    const A hr = (act && e0 + M < n) ? to_a(at(M)) : zero;
    coef_put(sh, cp, tab, n, tid, nt);
    __syncthreads();
    seg_solve<A, M, false, PEEL>(v, hl, hr, e0, n, s, l, L, P, sh, k, nullptr);
    if (act) {
#pragma unroll
        for (int j = 0; j < M; ++j) at(j) = from_a<S>(v[j]);
    }
}

// Row and column of packet tid + q * nt (VE elements each) of a run of rows
// of n elements, for q = 0, 1, ...: one division per thread, then a step and
// a wrap per packet (n is a run-time value: a division per packet costs
// tens of instructions).
struct RunPos {
    int row, e, dr, de;
    __device__ RunPos(int tid, int nt, int VE, int n)
        : row(tid * VE / n), e(tid * VE % n), dr(nt * VE / n), de(nt * VE % n) {}
    __device__ void next(int n) {
        row += dr;
        e += de;
        if (e >= n) {
            e -= n;
            ++row;
        }
    }
};

// z sweep: lines are contiguous rows.  A workgroup covers L rows of one x
// plane.  Staged (DIRECT false): the rows go through an LDS tile in the
// storage type (row-major, rows padded by one 16-byte packet: coalesced row
// loads and stores, and 16-byte LDS accesses without bank conflicts both for
// the row copies and for the segments), and each thread moves its segment
// between the tile and its registers.  Direct (DIRECT true): each thread
// loads its segment, M contiguous and 16-byte aligned values of its row,
// straight into registers as packets and stores it back the same way; the
// lanes of a wavefront read whole 64-128-byte runs of different rows, which
// the caches merge.  With COLS (imaginary time) the Simpson z integrals of
// the new psi^2 are summed per row and the rows of the tile with the y
// weights: colpart[tile][x] (calcnorm, imag3d.cu:1194-1209).
template <typename S, int L, int M, bool COLS, bool DIRECT>
__global__ void __launch_bounds__(512)
k_sweep_z(Solves sv, int Nx, int Ny, int Nz, SweepTab<typename Arith<S>::T> tab,
          SweepConst<typename Arith<S>::T> k, const double *__restrict__ wz, const double *__restrict__ wy) {
    using A = typename Arith<S>::T;
    extern __shared__ char smem_raw[];
    const SolveAt sr = solve_at(sv);
    S *__restrict__ psi = reinterpret_cast<S *>(sr.psi);
    double *__restrict__ colpart = sr.rows;
    const int n = Nz;
    constexpr int VE = Pk<S>::n, KV = M / VE;
    static_assert(M % VE == 0, "z sweep packets");
    // row stride of the tile: n + VE elements keeps every row 16-byte aligned
    // and consecutive rows 4 banks apart, so the packet accesses are
    // conflict-free
    const int RS = n + VE;
    const int P = n / M;
    const int nt = L * P;
    SweepShared<A, M> sh(smem_raw, n, nt, L);
    S *tile = reinterpret_cast<S *>(smem_raw + ((SweepShared<A, M>::bytes(n, nt, L, COLS) + 15) & ~size_t(15)));
    const int tid = threadIdx.x;
    const int l = tid % L, s = tid / L;
    const int yt = (Ny + L - 1) / L;
    const int xrow = blockIdx.x / yt, tno = blockIdx.x % yt, y0 = tno * L;
    const int nl = min(L, Ny - y0);
    const long base = ((long)xrow * Ny + y0) * Nz;
    const A zero = m_zero(k.A);
    const int npk = nl * n / VE;
    const int e0 = s * M;
    const bool act = l < nl;
    const CoefPre<A> cp = coef_pre(tab, n, tid);
    A v[M], hl, hr;
    if constexpr (DIRECT) {
        // rows past the tile's last read that row (clamped) and are not stored
        const S *src = psi + base + (long)min(l, nl - 1) * n + e0;
        typename Pk<S>::type r[KV];
#pragma unroll
        for (int q = 0; q < KV; ++q) r[q] = ld_pk(src + q * VE);
        const S sl = src[e0 > 0 ? -1 : 0], sr_ = src[e0 + M < n ? M : M - 1];
        coef_put(sh, cp, tab, n, tid, nt);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KV; ++q)
#pragma unroll
            for (int j = 0; j < VE; ++j) v[q * VE + j] = act ? to_a(r[q].e[j]) : zero;
        hl = (act && e0 > 0) ? to_a(sl) : zero;
        hr = (act && e0 + M < n) ? to_a(sr_) : zero;
    } else {
        // The tile's nl rows are one contiguous run of nl * n elements; each
        // thread moves M / VE packets of it (all loads first).
        typename Pk<S>::type r[KV];
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            const int iv = tid + q * nt;
            r[q] = ld_pk(psi + base + (unsigned)((iv < npk ? iv : 0) * VE));
        }
        coef_put(sh, cp, tab, n, tid, nt);
        RunPos rp(tid, nt, VE, n);
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            if (tid + q * nt < npk) st_pk(tile + rp.row * RS + rp.e, r[q]);
            rp.next(n);
        }
        __syncthreads();
        const S *row = tile + l * RS;
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            const typename Pk<S>::type pk = ld_pk(row + e0 + q * VE);
#pragma unroll
            for (int j = 0; j < VE; ++j) v[q * VE + j] = act ? to_a(pk.e[j]) : zero;
        }
        hl = (act && e0 > 0) ? to_a(row[e0 - 1]) : zero;
        hr = (act && e0 + M < n) ? to_a(row[e0 + M]) : zero;
    }
    const double acc = seg_solve<A, M, COLS, true>(v, hl, hr, e0, n, s, l, L, P, sh, k, wz);
    // every thread read its segment before the first barrier inside seg_solve
    if constexpr (DIRECT) {
        if (act) {
            S *dst = psi + base + (long)l * n + e0;
#pragma unroll
            for (int q = 0; q < KV; ++q) {
                typename Pk<S>::type o;
#pragma unroll
                for (int j = 0; j < VE; ++j) o.e[j] = from_a<S>(v[q * VE + j]);
                st_pk(dst + q * VE, o);
            }
        }
    } else if (act) {
        S *orow = tile + l * RS;
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            typename Pk<S>::type o;
#pragma unroll
            for (int j = 0; j < VE; ++j) o.e[j] = from_a<S>(v[q * VE + j]);
            st_pk(orow + e0 + q * VE, o);
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
            if (tid == 0) colpart[(long)tno * Nx + xrow] = w;
        }
    } else if constexpr (!DIRECT) {
        __syncthreads();
    }
    if constexpr (!DIRECT) {
        RunPos rp(tid, nt, VE, n);
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            const int iv = tid + q * nt;
            if (iv < npk) st_pk(psi + base + (unsigned)(iv * VE), ld_pk(tile + rp.row * RS + rp.e));
            rp.next(n);
        }
    }
}

// Segment length of the x / y sweeps for a line of n values of S: 32 for
// complex FP32 lines that split into at least 7 segments of 32 (measured at
// 96^3 to 256^3; at 192, 16-element segments are faster since the carry
// multipliers are tabulated: x 21.7 -> 20.9 us, y 21.8 -> 20.6 us at 192^3),
// for FP64 real lines of 512 or more (measured at 512^3: the
// x sweep 532 -> 407 us, the y sweep unchanged) and for lines whose
// 16-element segments would need more than 512 threads even with 128-byte
// rows; else 16.  It fixes the rounding of the partitioned solve, so it
// depends on the line only (never on a tuned choice).
template <typename S>
static int seg_len(long n) {
    using A = typename Arith<S>::T;
    if (!std::is_floating_point<A>::value && sizeof(A) == 8 && n % 32 == 0 && n / 32 >= 7) return 32;
    if (std::is_same<S, double>::value && n >= 512 && n % 32 == 0) return 32;
    if ((n / 16) * (long)(128 / sizeof(S)) > 512) return 32;
    return 16;
}

// z sweep: 16-element segments (more threads per tile row); rows per
// workgroup: a power of two >= 4 whose tile fits 64 KiB of LDS, with at most
// 256 threads (measured at 256^3 and 256x256x384).
static int z_seg_len(long) { return 16; }

static int env_int(const char *name, int dflt) {
    const char *e = getenv(name);
    return e ? atoi(e) : dflt;
}

template <typename S>
static int z_rows(long n) {
    static const int max_threads = env_int("DBEC_SWEEP_Z_THREADS", 256);
    static const long tile_max = 1024L * env_int("DBEC_SWEEP_Z_TILE_KB", 64);
    const int M = z_seg_len(n);
    const int P = (int)(n / M);
    int L = 64;
    while (L > 4 && ((long)L * P > max_threads || (long)(n + Pk<S>::n) * L * (long)sizeof(S) > tile_max)) L /= 2;
    return L;
}

static size_t arith_elem(const dbec_ctx *c) { return c->mode == DBEC_REAL ? 16 : 8; }

int sweep_z_rows(const dbec_ctx *c) {
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) return f32 ? z_rows<float>(c->Nz) : z_rows<double>(c->Nz);
    return f32 ? z_rows<cpx<float>>(c->Nz) : z_rows<cpx<double>>(c->Nz);
}

void sweep_z_args(const dbec_ctx *c, FftArgs &a) {
    const CoefLayout lay(c->Nx, c->Ny, c->Nz);
    const bool cplx = c->mode == DBEC_REAL;
    const size_t el = sizeof(double) * (cplx ? 2 : 1);  // FP64 tables in every precision
    const char *base = reinterpret_cast<const char *>(c->b.coef);
    a.ztab[0] = base + el * lay.ca(2);
    a.ztab[1] = base + el * lay.cg(2);
    a.ztab[2] = base + el * lay.gb(2, 0);
    a.ztab[3] = base + el * lay.fa(2, 0);
    for (int q = 0; q < 2; q++) {
        a.zk[0][q] = c->A0r[2][q];
        a.zk[1][q] = c->A[2][q];
        a.zk[2][q] = c->mA[2][q];
    }
    a.wz = c->b.wgt + c->Nx + c->Ny;
    a.wy = c->b.wgt + c->Nx;
    a.zrows = sweep_z_rows(c);
}

long sweep_z_tiles(const dbec_ctx *c) {
    const bool f32 = c->prec == DBEC_MIXED;
    int L;
    if (c->mode == DBEC_IMAG) L = f32 ? z_rows<float>(c->Nz) : z_rows<double>(c->Nz);
    else L = f32 ? z_rows<cpx<float>>(c->Nz) : z_rows<cpx<double>>(c->Nz);
    return (c->Ny + L - 1) / L;
}

// Solves of one launch: the by-value `one` or nb entries of `many`.
struct Launch {
    hipStream_t s;
    Solves v;
};

// The tables of one axis for both segment lengths (16 and 32).
template <typename A>
struct SweepTabs {
    const typename TabOf<A>::T *ca, *cg, *gb[2], *fa[2];
    SweepTab<A> at(int M) const {
        const int m = M == 32 ? 1 : 0;
        return {ca, cg, gb[m], fa[m]};
    }
};

template <typename S, int AXIS, int L, bool PEEL>
static hipError_t sweep_xy_launch(const Launch &q, int Nx, int Ny, int Nz, const SweepTabs<typename Arith<S>::T> &t,
                                  const SweepConst<typename Arith<S>::T> &k, int M) {
    using A = typename Arith<S>::T;
    const int n = AXIS == 0 ? Nx : Ny;
    const int P = n / M;
    if (n % M != 0 || L * P > 512) return hipErrorInvalidValue;
    const long nblocks = (long)((Nz + L - 1) / L) * (AXIS == 0 ? Ny : Nx);
    dim3 grid((unsigned)nblocks, (unsigned)q.v.nb), block(L * P);
    const size_t lds =
        M == 32 ? SweepShared<A, 32>::bytes(n, L * P, L, false) : SweepShared<A, 16>::bytes(n, L * P, L, false);
    if constexpr (sizeof(A) <= 8) {
        if (M == 32) {
            k_sweep_xy<S, AXIS, L, 32, PEEL><<<grid, block, lds, q.s>>>(q.v, Nx, Ny, Nz, t.at(32), k);
            return hipGetLastError();
        }
    } else {
        if (M != 16) return hipErrorInvalidValue;
    }
    k_sweep_xy<S, AXIS, L, 16, PEEL><<<grid, block, lds, q.s>>>(q.v, Nx, Ny, Nz, t.at(16), k);
    return hipGetLastError();
}

// x / y sweeps: segments of seg_len elements; per axis the row width (bytes
// of z per load: 256, or 128 for FP32 real lines) and the boundary peeling
// (real lines) by rule, or as tuned for the context.  DBEC_SWEEP_XY_{ROWB,M}
// (FP64 and complex lines), DBEC_SWEEP_XY_{ROWB,M}32 (FP32 real lines) and
// DBEC_SWEEP_XY_PEEL (0 or 1) override (tuning runs).
template <typename S, int AXIS, bool PEEL>
static hipError_t sweep_xy_rows(const Launch &q, int Nx, int Ny, int Nz, const SweepTabs<typename Arith<S>::T> &t,
                                const SweepConst<typename Arith<S>::T> &k, int rowb, int M) {
    const int n = AXIS == 0 ? Nx : Ny;
    // 256-byte rows only where their workgroup stays within 512 threads
    if (rowb == 128 || (long)(256 / sizeof(S)) * (n / M) > 512)
        return sweep_xy_launch<S, AXIS, 128 / sizeof(S), PEEL>(q, Nx, Ny, Nz, t, k, M);
    return sweep_xy_launch<S, AXIS, 256 / sizeof(S), PEEL>(q, Nx, Ny, Nz, t, k, M);
}
template <typename S, int AXIS>
static hipError_t sweep_xy_dispatch(const Launch &q, int Nx, int Ny, int Nz, const SweepTabs<typename Arith<S>::T> &t,
                                    const SweepConst<typename Arith<S>::T> &k, const Tuning &tu) {
    using A = typename Arith<S>::T;
    static const int rowb_env = sizeof(S) == 4 ? env_int("DBEC_SWEEP_XY_ROWB32", 0) : env_int("DBEC_SWEEP_XY_ROWB", 0);
    static const int mset = sizeof(S) == 4 ? env_int("DBEC_SWEEP_XY_M32", 0) : env_int("DBEC_SWEEP_XY_M", 0);
    static const int peel_env = env_int("DBEC_SWEEP_XY_PEEL", -1);
    const int n = AXIS == 0 ? Nx : Ny;
    int M = (mset == 16 || mset == 32) ? mset : seg_len<S>(n);
    if (sizeof(A) > 8) M = 16;
    const int rowb = rowb_env ? rowb_env : (tu.xy_rowb[AXIS] ? tu.xy_rowb[AXIS] : (sizeof(S) == 4 ? 128 : 256));
    const int peel_t = peel_env >= 0 ? peel_env : tu.xy_peel[AXIS];
    const bool peel = peel_t >= 0 ? peel_t != 0 : std::is_floating_point<A>::value;
    if (peel) return sweep_xy_rows<S, AXIS, true>(q, Nx, Ny, Nz, t, k, rowb, M);
    return sweep_xy_rows<S, AXIS, false>(q, Nx, Ny, Nz, t, k, rowb, M);
}

template <typename S, int L, bool COLS, bool DIRECT>
static hipError_t sweep_z_launch(const Launch &q, int Nx, int Ny, int Nz, const SweepTab<typename Arith<S>::T> &t,
                                 const SweepConst<typename Arith<S>::T> &k, const double *wz, const double *wy) {
    using A = typename Arith<S>::T;
    const int P = Nz / 16;
    const long nblocks = (long)((Ny + L - 1) / L) * Nx;
    dim3 grid((unsigned)nblocks, (unsigned)q.v.nb), block(L * P);
    size_t lds = (SweepShared<A, 16>::bytes(Nz, L * P, L, COLS) + 15) & ~size_t(15);
    if (!DIRECT) lds += (size_t)(Nz + Pk<S>::n) * L * sizeof(S);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sweep_z<S, L, 16, COLS, DIRECT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    k_sweep_z<S, L, 16, COLS, DIRECT><<<grid, block, lds, q.s>>>(q.v, Nx, Ny, Nz, t, k, wz, wy);
    return hipGetLastError();
}

template <typename S, bool COLS, bool DIRECT>
static hipError_t sweep_z_rows(const Launch &q, int Nx, int Ny, int Nz, const SweepTab<typename Arith<S>::T> &t,
                               const SweepConst<typename Arith<S>::T> &k, const double *wz, const double *wy) {
    switch (z_rows<S>(Nz)) {
    case 64: return sweep_z_launch<S, 64, COLS, DIRECT>(q, Nx, Ny, Nz, t, k, wz, wy);
    case 32: return sweep_z_launch<S, 32, COLS, DIRECT>(q, Nx, Ny, Nz, t, k, wz, wy);
    case 16: return sweep_z_launch<S, 16, COLS, DIRECT>(q, Nx, Ny, Nz, t, k, wz, wy);
    case 8: return sweep_z_launch<S, 8, COLS, DIRECT>(q, Nx, Ny, Nz, t, k, wz, wy);
    default: return sweep_z_launch<S, 4, COLS, DIRECT>(q, Nx, Ny, Nz, t, k, wz, wy);
    }
}

// The z sweep staged through LDS or direct (DBEC_SWEEP_Z_DIRECT, else the
// tuned choice, else staged).
template <typename S, bool COLS>
static hipError_t sweep_z_dispatch(const Launch &q, int Nx, int Ny, int Nz, const SweepTabs<typename Arith<S>::T> &ts,
                                   const SweepConst<typename Arith<S>::T> &k, const double *wz, const double *wy,
                                   const Tuning &tu) {
    if (Nz % 16 != 0) return hipErrorInvalidValue;
    static const int direct_env = env_int("DBEC_SWEEP_Z_DIRECT", -1);
    const int d = direct_env >= 0 ? direct_env : tu.z_direct;
    const SweepTab<typename Arith<S>::T> t = ts.at(16);
    if (d > 0) return sweep_z_rows<S, COLS, true>(q, Nx, Ny, Nz, t, k, wz, wy);
    return sweep_z_rows<S, COLS, false>(q, Nx, Ny, Nz, t, k, wz, wy);
}

template <typename S, typename A>
static hipError_t sweep_axis(const Launch &q, const dbec_ctx *c, int axis, const typename TabOf<A>::T *base,
                             const SweepConst<A> &k, bool cols) {
    const int Nx = (int)c->Nx, Ny = (int)c->Ny, Nz = (int)c->Nz;
    const CoefLayout lay(c->Nx, c->Ny, c->Nz);
    const SweepTabs<A> t{base + lay.ca(axis), base + lay.cg(axis),
                         {base + lay.gb(axis, 0), base + lay.gb(axis, 1)},
                         {base + lay.fa(axis, 0), base + lay.fa(axis, 1)}};
    const double *wy = c->b.wgt + c->Nx, *wz = c->b.wgt + c->Nx + c->Ny;
    if (axis == 0) return sweep_xy_dispatch<S, 0>(q, Nx, Ny, Nz, t, k, c->tune);
    if (axis == 1) return sweep_xy_dispatch<S, 1>(q, Nx, Ny, Nz, t, k, c->tune);
    if (cols) return sweep_z_dispatch<S, true>(q, Nx, Ny, Nz, t, k, wz, wy, c->tune);
    return sweep_z_dispatch<S, false>(q, Nx, Ny, Nz, t, k, wz, wy, c->tune);
}

hipError_t launch_sweep(hipStream_t s, const dbec_ctx *c, int axis, const Solves &v, bool cols) {
    const auto &b = c->b;
    const Launch q{s, v};
    const bool f32 = c->prec == DBEC_MIXED;
    if (c->mode == DBEC_IMAG) {
        if (f32) {
            SweepConst<float> k;
            k.A0r = (float)c->A0r[axis][0];
            k.A = (float)c->A[axis][0];
            k.mA = -k.A;
            return sweep_axis<float>(q, c, axis, b.coef, k, cols);
        }
        SweepConst<double> k;
        k.A0r = c->A0r[axis][0];
        k.A = c->A[axis][0];
        k.mA = -c->A[axis][0];
        return sweep_axis<double>(q, c, axis, b.coef, k, cols);
    }
    // complex lines: the sweep kernels use A = i a, mA = -A, A0r = 1 + i b
    if (c->A[axis][0] != 0.0 || c->A0r[axis][0] != 1.0 || c->mA[axis][0] != 0.0 || c->mA[axis][1] != -c->A[axis][1])
        return hipErrorInvalidValue;
    if (f32) {
        SweepConst<cpx<float>> k;
        k.A0r = {(float)c->A0r[axis][0], (float)c->A0r[axis][1]};
        k.A = {(float)c->A[axis][0], (float)c->A[axis][1]};
        k.mA = {(float)c->mA[axis][0], (float)c->mA[axis][1]};
        return sweep_axis<cpx<float>>(q, c, axis, reinterpret_cast<const cpx<double> *>(b.coef), k, cols);
    }
    SweepConst<cpx<double>> k;
    k.A0r = {c->A0r[axis][0], c->A0r[axis][1]};
    k.A = {c->A[axis][0], c->A[axis][1]};
    k.mA = {c->mA[axis][0], c->mA[axis][1]};
    return sweep_axis<cpx<double>>(q, c, axis, reinterpret_cast<const cpx<double> *>(b.coef), k, cols);
}

// norm^2 = sum_x wx[x] sum_t colpart[t][x]; scal[NORM] = norm, scal[SCALE] = 1/norm.
__global__ void __launch_bounds__(256)
k_norm_cols(Solves v, int tiles, const double *__restrict__ wx, int Nx) {
    __shared__ double sh[256];
    const SolveAt sr = solve_at(v);
    norm_from_cols<256>(sr.rows, tiles, wx, Nx, sr.scal, v.shift, sh);
}

hipError_t launch_norm_from_cols(hipStream_t s, const dbec_ctx *c, const Solves &v) {
    k_norm_cols<<<dim3(1, (unsigned)v.nb), 256, 0, s>>>(v, (int)sweep_z_tiles(c), c->b.wgt, (int)c->Nx);
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

// Components a derivative is taken of: |psi| (the reference's calcmuen), or
// the real or imaginary part (the phase-gradient energy of real time).
enum { COMP_AMP = 0, COMP_RE = 1, COMP_IM = 2 };
template <int C> __device__ __forceinline__ double comp_of(double v) { return v; }
template <int C> __device__ __forceinline__ double comp_of(float v) { return (double)v; }
template <int C> __device__ __forceinline__ double comp_of(cpx<double> v) {
    return C == COMP_RE ? v.x : (C == COMP_IM ? v.y : amp_of(v));
}
template <int C> __device__ __forceinline__ double comp_of(cpx<float> v) {
    return C == COMP_RE ? (double)v.x : (C == COMP_IM ? (double)v.y : amp_of(v));
}

// Five-point derivative with the boundary rule of calcmuen_kernel1x
// (imag3d.cu:1303-1332): 0 at both ends, centred 3-point at 1 and N-2.
template <int C = COMP_AMP, typename P>
__device__ __forceinline__ double deriv5(const P *__restrict__ f, int i, int n, long st, double h) {
    if (i == 0 || i == n - 1) return 0.0;
    if (i == 1) return (comp_of<C>(f[2 * st]) - comp_of<C>(f[0])) / (2.0 * h);
    if (i == n - 2) return (comp_of<C>(f[(long)(n - 1) * st]) - comp_of<C>(f[(long)(n - 3) * st])) / (2.0 * h);
    return (comp_of<C>(f[(long)(i - 2) * st]) - 8.0 * comp_of<C>(f[(long)(i - 1) * st]) +
            8.0 * comp_of<C>(f[(long)(i + 1) * st]) - comp_of<C>(f[(long)(i + 2) * st])) / (12.0 * h);
}

// Kinetic density sum over the axes of (d psi)^2 / (3 - par) at (i, j, k) of
// the row `row`: of |psi|, or with PHASE of the real and imaginary parts.
template <int C, typename P>
__device__ __forceinline__ double kin_comp(const P *__restrict__ psi, int i, int j, int k, long row, int Nx, int Ny,
                                           int Nz, const MuenConst &mc) {
    const long NyNz = (long)Ny * Nz;
    const double dxv = deriv5<C>(psi + (long)j * Nz + k, i, Nx, NyNz, mc.dx);
    double d = (dxv * dxv) / mc.kin;
    const double dyv = deriv5<C>(psi + (long)i * NyNz + k, j, Ny, (long)Nz, mc.dy);
    d += (dyv * dyv) / mc.kin;
    const double dzv = deriv5<C>(psi + row * Nz, k, Nz, 1L, mc.dz);
    d += (dzv * dzv) / mc.kin;
    return d;
}

// Per-row z integrals of the chemical potential and energy densities
// (calcmuen_kernel1x/1y/1z and kernel2, imag3d.cu:1293-1450; real3d.cu
// 1382-1543, where real time differentiates |psi|).
// PHASE (complex psi): the kinetic density of the real and imaginary parts
// (|grad psi|^2) instead of that of |psi|.
template <typename P, typename R, bool PHASE = false>
__global__ void __launch_bounds__(256)
k_rows_muen(const P *__restrict__ psi, const R *__restrict__ phi, const double *__restrict__ sq,
            const double *__restrict__ wz, TrapConst tc, MuenConst mc, double *__restrict__ rmu,
            double *__restrict__ ren, int Nx, int Ny, int Nz) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * 4 + wave;
    if (row >= (long)Nx * Ny) return;
    const int i = (int)(row / Ny), j = (int)(row % Ny);
    double amu = 0.0, aen = 0.0;
    for (int k = lane; k < Nz; k += 64) {
        const long n = row * Nz + k;
        double d;
        if constexpr (PHASE)
            d = kin_comp<COMP_RE>(psi, i, j, k, row, Nx, Ny, Nz, mc) + kin_comp<COMP_IM>(psi, i, j, k, row, Nx, Ny, Nz, mc);
        else
            d = kin_comp<COMP_AMP>(psi, i, j, k, row, Nx, Ny, Nz, mc);
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

hipError_t launch_rows_muen(hipStream_t s, const dbec_ctx *c, double *rmu, double *ren, bool phase) {
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
    } else if (phase) {
        if (f32)
            k_rows_muen<cpx<float>, float, true><<<grid, block, 0, s>>>(static_cast<const cpx<float> *>(b.psi),
                                                                        static_cast<const float *>(b.phi), b.sq, wz,
                                                                        tc, mc, rmu, ren, Nx, Ny, Nz);
        else
            k_rows_muen<cpx<double>, double, true><<<grid, block, 0, s>>>(static_cast<const cpx<double> *>(b.psi),
                                                                          static_cast<const double *>(b.phi), b.sq,
                                                                          wz, tc, mc, rmu, ren, Nx, Ny, Nz);
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


// Sampled extracts on the device (the allocation-free density entry point):
// dst[a][b][k] = |psi|^2 at (x0 + a ox, y0 + b oy, z0 + k oz), the same
// mod2_of as k_mod2, so the values equal those of the full extract bit for bit.
template <typename P>
__global__ void k_pick_mod2(const P *__restrict__ psi, long x0, long y0, long z0, long ox, long oy, long oz, long cy,
                            long cz, long n, double *__restrict__ dst, long Ny, long Nz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long a = t / (cy * cz), b = (t / cz) % cy, k = t % cz;
    dst[t] = mod2_of(psi[((x0 + a * ox) * Ny + y0 + b * oy) * Nz + z0 + k * oz]);
}

hipError_t launch_pick_mod2(hipStream_t s, const dbec_ctx *c, const long start[3], const long step[3],
                            const long cnt[3], double *dst) {
    const long n = cnt[0] * cnt[1] * cnt[2];
    if (n <= 0) return hipSuccess;
    const unsigned g = (unsigned)((n + 255) / 256);
    const bool f32 = c->prec == DBEC_MIXED;
#define DBEC_PICK_ARGS start[0], start[1], start[2], step[0], step[1], step[2], cnt[1], cnt[2], n, dst, c->Ny, c->Nz
    if (c->mode == DBEC_IMAG) {
        if (f32) k_pick_mod2<float><<<g, 256, 0, s>>>(static_cast<const float *>(c->b.psi), DBEC_PICK_ARGS);
        else k_pick_mod2<double><<<g, 256, 0, s>>>(static_cast<const double *>(c->b.psi), DBEC_PICK_ARGS);
    } else {
        if (f32) k_pick_mod2<cpx<float>><<<g, 256, 0, s>>>(static_cast<const cpx<float> *>(c->b.psi), DBEC_PICK_ARGS);
        else k_pick_mod2<cpx<double>><<<g, 256, 0, s>>>(static_cast<const cpx<double> *>(c->b.psi), DBEC_PICK_ARGS);
    }
#undef DBEC_PICK_ARGS
    return hipGetLastError();
}

// dst[a][b] = src[(a o0) n1 + b o1], a < c0, b < c1
__global__ void k_pick2(const double *__restrict__ src, long n1, long o0, long o1, long c1, long n,
                        double *__restrict__ dst) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    dst[t] = src[(t / c1) * o0 * n1 + (t % c1) * o1];
}

hipError_t launch_pick2(hipStream_t s, const double *src, long n1, long o0, long o1, long c0, long c1, double *dst) {
    const long n = c0 * c1;
    if (n <= 0) return hipSuccess;
    k_pick2<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, n1, o0, o1, c1, n, dst);
    return hipGetLastError();
}

}  // namespace dbec
