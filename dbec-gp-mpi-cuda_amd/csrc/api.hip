// api.hip - the C ABI of include/dbec.h on the HIP backend.
//
// Reference aliases: imag3d.cu = src/imag3d-mpicuda/imag3d-mpicuda.cu,
// real3d.cu = src/real3d-mpicuda/real3d-mpicuda.cu.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "core.hpp"
#include "fft_device.hpp"
#include "fft_launch.hpp"
#include "ops.hpp"

namespace dbec {

hipError_t launch_zfwd(long n, const FftArgs &a) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::zfwd(a);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return hipErrorInvalidValue;
    }
}
hipError_t launch_ypass(long n, int dir, const FftArgs &a) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::ypass(dir, a);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return hipErrorInvalidValue;
    }
}
hipError_t launch_xconv(long n, const FftArgs &a) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::xconv(a);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return hipErrorInvalidValue;
    }
}
hipError_t launch_zinv(long n, const FftArgs &a) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::zinv(a);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return hipErrorInvalidValue;
    }
}
hipError_t launch_sweepz_zfwd(long n, const FftArgs &a) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::sweepz_zfwd(a);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return hipErrorInvalidValue;
    }
}
bool sweepz_zfwd_fusable(long n, bool fp32, bool cpsi, int zrows) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::fusable(fp32, cpsi, zrows);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return false;
    }
}
hipError_t launch_zinv_nu(long n, const FftArgs &a) {
    switch (n) {
#define DBEC_X(NN) case NN: return FftLaunch<NN>::fused(a);
        DBEC_FOR_EACH_FFT_LEN(DBEC_X)
#undef DBEC_X
    default: return hipErrorInvalidValue;
    }
}

// Per-kernel event timing used by dbec_profile_step.
struct KernelTimer {
    struct Rec {
        std::string name;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    int burst = 1;  // back-to-back launches of each kernel between its two events
};

}  // namespace dbec

// Device scratch of dbec_density_noalloc: DBEC_NOALLOC_SCRATCH_BYTES, one
// copy per GPU, held until the process ends.  In host code this name is not a
// device address; the address of each GPU's copy comes from
// hipGetSymbolAddress (scratch_setup) and only that address is used.
static __device__ __align__(256) double scratch[DBEC_NOALLOC_SCRATCH_BYTES / sizeof(double)];

using namespace dbec;

namespace {

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;  // device ordinal -> address of that device's scratch

// Once per device, with that device current (dbec_create): look up and keep
// the address of the device's scratch.  A failure is not fatal here; the
// allocation-free entry point reports it.
void scratch_setup(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(dev)) return;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();  // leave no stale error for the next launch check
        return;
    }
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
}

double *scratch_of(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}

thread_local std::string g_thread_error;

int fail(dbec_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    g_thread_error = msg;
    return code;
}

int hip_fail(dbec_ctx *c, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, DBEC_EDEVICE, m);
}

#define HIPCHK(c, expr, what)                               \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail((c), e_, (what)); \
    } while (0)

#define CHK(expr)                    \
    do {                             \
        int rc_ = (expr);            \
        if (rc_ != DBEC_OK) return rc_; \
    } while (0)

size_t psi_elem(const dbec_ctx *c) {
    size_t r = (c->prec == DBEC_MIXED) ? sizeof(float) : sizeof(double);
    return c->mode == DBEC_REAL ? 2 * r : r;
}
size_t real_elem(const dbec_ctx *c) { return (c->prec == DBEC_MIXED) ? sizeof(float) : sizeof(double); }

// Timed launch: with a profiler attached, launch `burst` times back to back
// between two events (consecutive launches of the kernel pipeline as in a run
// of steps, and the events' own cost is spread over the burst).
template <typename F>
hipError_t timed(dbec_ctx *c, const char *name, double bytes, F &&launch) {
    if (!c->timer) return launch();
    KernelTimer::Rec r;
    r.name = name;
    r.bytes = bytes;
    (void)hipEventCreate(&r.a);
    (void)hipEventCreate(&r.b);
    (void)hipEventRecord(r.a, c->stream);
    hipError_t e = hipSuccess;
    for (int i = 0; i < c->timer->burst && e == hipSuccess; i++) e = launch();
    (void)hipEventRecord(r.b, c->stream);
    c->timer->recs.push_back(r);
    return e;
}

// The solve held by a single context; use_scale: psi is read as psi / norm of
// the last step (renormalisation pending).
Solves solve_ref(const dbec_ctx *c, bool use_scale) {
    Solves v;
    v.psi = c->b.psi;
    v.spec = c->b.spec;
    v.rows = c->b.rows;
    v.scal = c->b.scal;
    v.psi_st = v.spec_st = v.rows_st = v.scal_st = 0;
    v.slot = nullptr;
    v.par = nullptr;
    v.p0.tc = trap_const(c);
    v.p0.g = c->g;
    v.p0.gd = c->gd;
    v.nb = 1;
    v.use_scale = use_scale ? 1 : 0;
    v.shift = (c->prec == DBEC_MIXED && c->mode == DBEC_IMAG) ? 1 : 0;
    v.fold = 0;
    return v;
}

FftArgs fft_args(const dbec_ctx *c, int axis, const Solves &v) {
    FftArgs a;
    a.stream = c->stream;
    a.Nx = c->Nx;
    a.Ny = c->Ny;
    a.Nz = c->Nz;
    a.Nkz = c->Nkz;
    a.fp32 = c->prec == DBEC_MIXED;
    a.cpsi = c->mode == DBEC_REAL;
    a.v = v;
    a.phi = c->b.phi;
    a.potddq = c->b.potddq;
    a.tw = a.fp32 ? c->b.twf[axis] : c->b.tw[axis];
    a.sq = c->b.sq;
    a.nc = nu_const(c);
    a.wx = c->b.wgt;
    a.tiles = (int)sweep_z_tiles(c);
    a.tune = c->tune;
    return a;
}

// calcpsidd2 of the context's own psi (no pending scale): |psi|^2 -> half
// spectrum -> * potdd -> phi, then the face copies.
int run_calcpsidd2(dbec_ctx *c) {
    const Solves v = solve_ref(c, false);
    const double ce = (double)(2 * real_elem(c));
    const double spec_b = (double)c->Nkz * c->Nx * c->Ny * ce;
    const double nre = (double)c->N * real_elem(c);
    HIPCHK(c, timed(c, "zfwd", (double)c->N * psi_elem(c) + spec_b, [&] { return launch_zfwd(c->Nz, fft_args(c, 2, v)); }), "zfwd");
    HIPCHK(c, timed(c, "ypass_fwd", 2 * spec_b, [&] { return launch_ypass(c->Ny, -1, fft_args(c, 1, v)); }), "ypass");
    const double pq = (double)c->Nkz * (c->Nx / 2 + 1) * (c->Ny / 2 + 1) * real_elem(c);
    HIPCHK(c, timed(c, "xconv", 2 * spec_b + pq, [&] { return launch_xconv(c->Nx, fft_args(c, 0, v)); }), "xconv");
    HIPCHK(c, timed(c, "ypass_inv", 2 * spec_b, [&] { return launch_ypass(c->Ny, +1, fft_args(c, 1, v)); }), "ypass");
    HIPCHK(c, timed(c, "zinv", spec_b + nre, [&] { return launch_zinv(c->Nz, fft_args(c, 2, v)); }), "zinv");
    const double fb = 2.0 * (double)(c->Ny * c->Nz + c->Nx * c->Nz + c->Nx * c->Ny) * real_elem(c);
    HIPCHK(c, timed(c, "facefix", fb, [&] {
               return launch_facefix(c->stream, c->b.phi, c->prec == DBEC_MIXED, c->Nx, c->Ny, c->Nz);
           }), "facefix");
    c->phi_valid = true;
    return DBEC_OK;
}

// The composite step's convolution: the forward passes (the z pass unless
// the previous step's fused z sweep did it: z_done), then the inverse z pass
// fused with calcnu (interior rows) and the face rows.
int run_conv_nu(dbec_ctx *c, const Solves &v, bool z_done) {
    const double nb = v.nb;
    const double ce = (double)(2 * real_elem(c));
    const double spec_b = (double)c->Nkz * c->Nx * c->Ny * ce;
    if (!z_done)
        HIPCHK(c, timed(c, "zfwd", nb * ((double)c->N * psi_elem(c) + spec_b), [&] { return launch_zfwd(c->Nz, fft_args(c, 2, v)); }), "zfwd");
    HIPCHK(c, timed(c, "ypass_fwd", nb * 2 * spec_b, [&] { return launch_ypass(c->Ny, -1, fft_args(c, 1, v)); }), "ypass");
    const double pq = (double)c->Nkz * (c->Nx / 2 + 1) * (c->Ny / 2 + 1) * real_elem(c);
    HIPCHK(c, timed(c, "xconv", nb * (2 * spec_b + pq), [&] { return launch_xconv(c->Nx, fft_args(c, 0, v)); }), "xconv");
    HIPCHK(c, timed(c, "ypass_inv", nb * 2 * spec_b, [&] { return launch_ypass(c->Ny, +1, fft_args(c, 1, v)); }), "ypass");
    const double nf = (double)(c->Nx + c->Ny - 1);
    HIPCHK(c, timed(c, "zinv_nu", nb * (spec_b + 2.0 * c->N * psi_elem(c) + nf * c->Nkz * 2 * ce),
                    [&] { return launch_zinv_nu(c->Nz, fft_args(c, 2, v)); }), "zinv_nu");
    c->phi_valid = false;
    return DBEC_OK;
}

int run_sweep(dbec_ctx *c, int axis, const Solves &v, bool cols) {
    static const char *names[3] = {"sweep_x", "sweep_y", "sweep_z"};
    double bytes = v.nb * (2.0 * (double)c->N * psi_elem(c) + (cols ? (double)c->Nx * sweep_z_tiles(c) * 8.0 : 0.0));
    HIPCHK(c, timed(c, names[axis], bytes, [&] { return launch_sweep(c->stream, c, axis, v, cols); }), "sweep");
    return DBEC_OK;
}

// Whether the z sweep of a step can run fused with the next step's forward
// z pass (the z length's plan has T = N / 8 and the kernel fits for the
// precision and the z sweep's rows per workgroup; DBEC_FUSE_Z=0 turns it off).
bool fuse_z(const dbec_ctx *c) {
    static const bool on = [] {
        const char *e = getenv("DBEC_FUSE_Z");
        return !e || std::strcmp(e, "0") != 0;
    }();
    return on && sweepz_zfwd_fusable(c->Nz, c->prec == DBEC_MIXED, c->mode == DBEC_REAL, sweep_z_rows(c));
}

// One split step (calcpsidd2, calcnu, calclux/y/z) of the solves v.  In
// imaginary time the z sweep leaves the Simpson column partials of the new
// psi: the next step folds the norm in (v.fold), or run_norm computes it.
// z_done: this step's forward z pass already ran (fused into the previous
// step's z sweep).  fuse_next: run this step's z sweep fused with the next
// step's forward z pass, whose |psi|^2 it transforms unscaled (that step
// then has v.fold == 2).
int run_step(dbec_ctx *c, const Solves &v, bool z_done, bool fuse_next) {
    const bool imag = c->mode == DBEC_IMAG;
    CHK(run_conv_nu(c, v, z_done));
    CHK(run_sweep(c, 0, v, false));
    CHK(run_sweep(c, 1, v, false));
    if (!fuse_next) return run_sweep(c, 2, v, imag);
    const double ce = (double)(2 * real_elem(c));
    const double spec_b = (double)c->Nkz * c->Nx * c->Ny * ce;
    const double bytes = v.nb * (2.0 * (double)c->N * psi_elem(c) + spec_b +
                                 (imag ? (double)c->Nx * sweep_z_tiles(c) * 8.0 : 0.0));
    FftArgs a = fft_args(c, 2, v);
    sweep_z_args(c, a);
    HIPCHK(c, timed(c, "sweep_z_zfwd", bytes, [&] { return launch_sweepz_zfwd(c->Nz, a); }), "sweep_z_zfwd");
    return DBEC_OK;
}

// calcnorm from the column partials: scal[NORM], scal[SCALE] = 1/norm.
int run_norm(dbec_ctx *c, const Solves &v) {
    HIPCHK(c, timed(c, "norm", v.nb * (double)c->Nx * sweep_z_tiles(c) * 8.0,
                    [&] { return launch_norm_from_cols(c->stream, c, v); }), "norm");
    return DBEC_OK;
}

int run_calcnu(dbec_ctx *c, const double *scale) {
    double bytes = (double)c->N * (2 * psi_elem(c) + real_elem(c));
    HIPCHK(c, timed(c, "calcnu", bytes, [&] { return launch_calcnu(c->stream, c, scale); }), "calcnu");
    return DBEC_OK;
}

int read_scalars(dbec_ctx *c, double *dst, int first, int count) {
    HIPCHK(c, hipMemcpyAsync(dst, c->b.scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream), "copy scalars");
    HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
    return DBEC_OK;
}

int materialize_scale(dbec_ctx *c) {
    if (!c->scale_pending) return DBEC_OK;
    HIPCHK(c, timed(c, "scale", 2.0 * (double)c->N * psi_elem(c),
                    [&] { return launch_scale(c->stream, c, c->b.scal + SC_SCALE); }), "scale");
    c->scale_pending = false;
    return DBEC_OK;
}

// Simpson weights h/3 * (1, 4, 2, ..., 4, 1) plus the even-N end correction
// (5 f[N-1] + 8 f[N-2] - f[N-3]) / 4 of simpint (src/utils/diffint.cu:37-51).
void simpson_weights(double h, long n, double *w) {
    for (long i = 0; i < n; i++) w[i] = 0.0;
    for (long i = 1; i < n - 1; i += 2) {
        w[i - 1] += 1.0;
        w[i] += 4.0;
        w[i + 1] += 1.0;
    }
    if (n % 2 == 0) {
        w[n - 1] += 5.0 / 4.0;
        w[n - 2] += 8.0 / 4.0;
        w[n - 3] -= 1.0 / 4.0;
    }
    for (long i = 0; i < n; i++) w[i] = w[i] * h / 3.0;
}

struct Cx {
    double re, im;
};
Cx cx_mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
Cx cx_add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
// Smith's (1962) complex division.
Cx cx_div(Cx a, Cx b) {
    if (std::fabs(b.re) >= std::fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    double r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

void free_buffers(dbec_ctx *c) {
    Buffers &b = c->b;
    void *ptrs[] = {b.psi, b.spec, b.phi, b.potddq, b.rows, b.red, b.scal, b.coef, b.wgt, b.sq,
                    b.tw[0], b.tw[1], b.tw[2], b.twf[0], b.twf[1], b.twf[2]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    b = Buffers();
}

int dev_alloc(dbec_ctx *c, void **p, size_t bytes, const char *what) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, DBEC_ENOMEM, std::string("device allocation of ") + what + " failed: " + hipGetErrorString(e));
    }
    return DBEC_OK;
}

int setup(dbec_ctx *c) {
    const dbec_params &p = c->p;
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    const double pi = c->pi, par = p.par;
    Buffers &b = c->b;
    const bool f32 = c->prec == DBEC_MIXED;
    const bool cplx = c->mode == DBEC_REAL;

    // Grid squares (initpsi imag3d.cu:842-855) and Simpson weights.
    std::vector<double> sq(Nx + Ny + Nz), wgt(Nx + Ny + Nz);
    for (long i = 0; i < Nx; i++) { double x = (i - Nx / 2) * p.dx; sq[i] = x * x; }
    for (long j = 0; j < Ny; j++) { double y = (j - Ny / 2) * p.dy; sq[Nx + j] = y * y; }
    for (long k = 0; k < Nz; k++) { double z = (k - Nz / 2) * p.dz; sq[Nx + Ny + k] = z * z; }
    simpson_weights(p.dx, Nx, wgt.data());
    simpson_weights(p.dy, Ny, wgt.data() + Nx);
    simpson_weights(p.dz, Nz, wgt.data() + Nx + Ny);

    // Crank-Nicolson coefficients (gencoef imag3d.cu:894-931; real3d.cu:991-1035)
    // and the carry multipliers of the partitioned sweeps (CoefLayout).
    const double h2[3] = {p.dx * p.dx, p.dy * p.dy, p.dz * p.dz};
    const long n3[3] = {Nx, Ny, Nz};
    const CoefLayout lay(Nx, Ny, Nz);
    const int cw = cplx ? 2 : 1;  // doubles per element
    std::vector<double> coef(cw * lay.size(), 0.0);
    auto put = [&](long at, Cx v) {
        coef[cw * at] = v.re;
        if (cplx) coef[cw * at + 1] = v.im;
    };
    for (int a = 0; a < 3; a++) {
        const long n = n3[a];
        const double r = p.dt / h2[a] / (3. - par);
        // imaginary time: real A0, A0r, A; real time: A0 = 1 + i r, A0r = 1 - i r, A = -i r / 2
        const Cx A0 = cplx ? Cx{1., r} : Cx{1. + r, 0.};
        const Cx A0r = cplx ? Cx{1., -r} : Cx{1. - r, 0.};
        const Cx A = cplx ? Cx{0., -0.5 * p.dt / h2[a] / (3. - par)} : Cx{-0.5 * p.dt / h2[a] / (3. - par), 0.};
        const Cx minus1 = {-1., 0.};
        c->A0r[a][0] = A0r.re; c->A0r[a][1] = A0r.im;
        c->A[a][0] = A.re; c->A[a][1] = A.im;
        c->mA[a][0] = -A.re; c->mA[a][1] = -A.im;
        std::vector<Cx> al(n - 1), ga(n - 1);
        al[n - 2] = {0., 0.};
        ga[n - 2] = cplx ? cx_div(minus1, A0) : Cx{-1. / A0.re, 0.};
        for (long i = n - 2; i > 0; i--) {
            if (cplx) {
                al[i - 1] = cx_mul(A, ga[i]);
                ga[i - 1] = cx_div(minus1, cx_add(A0, cx_mul(A, al[i - 1])));
            } else {
                al[i - 1] = {A.re * ga[i].re, 0.};
                ga[i - 1] = {-1. / (A0.re + A.re * al[i - 1].re), 0.};
            }
        }
        c->h_calpha[a].assign(cw * (n - 1), 0.);
        c->h_cgamma[a].assign(cw * (n - 1), 0.);
        for (long i = 0; i < n - 1; i++) {
            put(lay.ca(a) + i, al[i]);
            put(lay.cg(a) + i, ga[i]);
            c->h_calpha[a][cw * i] = al[i].re;
            c->h_cgamma[a][cw * i] = ga[i].re;
            if (cplx) {
                c->h_calpha[a][cw * i + 1] = al[i].im;
                c->h_cgamma[a][cw * i + 1] = ga[i].im;
            }
        }
        // gb[i] = prod_{t=i..top} cgamma[t] A and fa[e] = prod_{t=bot..e} calpha[t-1]
        // within each segment of M elements, over the in-range indices 1..n-2,
        // multiplied in the order of the sequential chains they replace
        auto mul = [&](Cx x, Cx y) { return cplx ? cx_mul(x, y) : Cx{x.re * y.re, 0.}; };
        for (int m = 0; m < 2; m++) {
            const long M = 16L << m;
            if (n % M != 0) continue;
            for (long e0 = 0; e0 < n; e0 += M) {
                const long bot = std::max(e0, 1L), top = std::min(e0 + M - 1, n - 2);
                Cx pr = {1., 0.};
                for (long i = top; i >= bot; i--) {
                    pr = mul(mul(ga[i], A), pr);
                    put(lay.gb(a, m) + i, pr);
                }
                pr = {1., 0.};
                for (long e = bot; e <= top; e++) {
                    pr = mul(al[e - 1], pr);
                    put(lay.fa(a, m) + e, pr);
                }
            }
        }
    }

    // Dipolar kernel (initpotdd imag3d.cu:945-980) on the quarter |k| grid:
    // the reference's k ordering gives k[N-i] = -k[i] exactly, so potdd is the
    // same number at +k and -k in every component.
    const long qx = Nx / 2 + 1, qy = Ny / 2 + 1, qz = Nz / 2 + 1;
    const double dkx = 2. * pi / (Nx * p.dx), dky = 2. * pi / (Ny * p.dy), dkz = 2. * pi / (Nz * p.dz);
    std::vector<double> kx2(qx), ky2(qy), kz2(qz);
    for (long i = 0; i < qx; i++) { double k = i * dkx; kx2[i] = k * k; }
    for (long j = 0; j < qy; j++) { double k = j * dky; ky2[j] = k * k; }
    for (long k = 0; k < qz; k++) { double kk = k * dkz; kz2[k] = kk * kk; }
    std::vector<double> pq(qz * qx * qy);
    const double R = p.cutoff;
    for (long k = 0; k < qz; k++)
        for (long i = 0; i < qx; i++)
            for (long j = 0; j < qy; j++) {
                double xk = std::sqrt(kz2[k] + kx2[i] + ky2[j]);
                double tmp = 1. + 3. * std::cos(xk * R) / (xk * xk * R * R) -
                             3. * std::sin(xk * R) / (xk * xk * xk * R * R * R);
                pq[(k * qx + i) * qy + j] = (4. * pi * (3. * kz2[k] / (kx2[i] + ky2[j] + kz2[k]) - 1.) / 3.) * tmp;
            }
    pq[0] = 0.;
    // device layout [kz][|ky|][|kx|]: rows contiguous in kx for the x convolution
    std::vector<double> pqd(pq.size());
    for (long k = 0; k < qz; k++)
        for (long i = 0; i < qx; i++)
            for (long j = 0; j < qy; j++) pqd[(k * qy + j) * qx + i] = pq[(k * qx + i) * qy + j];
    // host copy of the half table [kx][ky][kz<=Nz/2] for dbec_get_table
    c->h_potdd_half.resize(Nx * Ny * qz);
    for (long i = 0; i < Nx; i++)
        for (long j = 0; j < Ny; j++)
            for (long k = 0; k < qz; k++) {
                long ii = i <= Nx / 2 ? i : Nx - i, jj = j <= Ny / 2 ? j : Ny - j;
                c->h_potdd_half[(i * Ny + j) * qz + k] = pq[(k * qx + ii) * qy + jj];
            }

    // Twiddles exp(-2 pi i t / n) per axis.
    auto twiddles = [](long n, std::vector<double> &d, std::vector<float> &f) {
        d.resize(2 * n);
        f.resize(2 * n);
        const long double twopi = 6.283185307179586476925286766559005768L;
        for (long t = 0; t < n; t++) {
            long double a = twopi * (long double)t / (long double)n;
            d[2 * t] = (double)cosl(a);
            d[2 * t + 1] = (double)-sinl(a);
            f[2 * t] = (float)cosl(a);
            f[2 * t + 1] = (float)-sinl(a);
        }
    };

    // Device buffers.
    const size_t re = f32 ? 4 : 8;
    CHK(dev_alloc(c, &b.psi, (size_t)c->N * psi_elem(c), "psi"));
    CHK(dev_alloc(c, &b.spec, (size_t)c->Nkz * Nx * Ny * 2 * re, "spectrum"));
    CHK(dev_alloc(c, &b.phi, (size_t)c->N * re, "psidd2"));
    CHK(dev_alloc(c, &b.potddq, pq.size() * re, "potdd"));
    CHK(dev_alloc(c, (void **)&b.rows, (size_t)2 * Nx * Ny * sizeof(double), "row sums"));
    CHK(dev_alloc(c, (void **)&b.scal, SC_COUNT * sizeof(double), "scalars"));
    CHK(dev_alloc(c, (void **)&b.coef, coef.size() * sizeof(double), "CN coefficients"));
    CHK(dev_alloc(c, (void **)&b.wgt, wgt.size() * sizeof(double), "Simpson weights"));
    CHK(dev_alloc(c, (void **)&b.sq, sq.size() * sizeof(double), "grid squares"));
    hipStream_t s = c->stream;
    HIPCHK(c, hipMemsetAsync(b.psi, 0, (size_t)c->N * psi_elem(c), s), "memset");
    HIPCHK(c, hipMemsetAsync(b.phi, 0, (size_t)c->N * re, s), "memset");
    HIPCHK(c, hipMemcpyAsync(b.coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
    HIPCHK(c, hipMemcpyAsync(b.wgt, wgt.data(), wgt.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
    HIPCHK(c, hipMemcpyAsync(b.sq, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
    double sc[SC_COUNT] = {0};
    sc[SC_NORM] = 1.0;
    sc[SC_SCALE] = 1.0;
    sc[SC_ONE] = 1.0;
    HIPCHK(c, hipMemcpyAsync(b.scal, sc, sizeof sc, hipMemcpyHostToDevice, s), "copy");
    if (!f32) {
        HIPCHK(c, hipMemcpyAsync(b.potddq, pqd.data(), pqd.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
    } else {
        std::vector<float> pf(pqd.begin(), pqd.end());
        HIPCHK(c, hipMemcpyAsync(b.potddq, pf.data(), pf.size() * sizeof(float), hipMemcpyHostToDevice, s), "copy");
        HIPCHK(c, hipStreamSynchronize(s), "synchronize");
    }
    for (int a = 0; a < 3; a++) {
        std::vector<double> td;
        std::vector<float> tf;
        twiddles(n3[a], td, tf);
        CHK(dev_alloc(c, &b.tw[a], td.size() * sizeof(double), "twiddles"));
        CHK(dev_alloc(c, &b.twf[a], tf.size() * sizeof(float), "twiddles"));
        HIPCHK(c, hipMemcpyAsync(b.tw[a], td.data(), td.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        HIPCHK(c, hipMemcpyAsync(b.twf[a], tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice, s), "copy");
        HIPCHK(c, hipStreamSynchronize(s), "synchronize");
    }
    HIPCHK(c, hipStreamSynchronize(s), "synchronize");
    return DBEC_OK;
}

// Autotuning of the launch choices in Tuning: each candidate runs a few
// times on the context's stream (psi and the spectrum are zero: the kernels
// move the same bytes and do the same arithmetic as on a state) and the
// fastest is kept, pass by pass.  The choices never change a result, only
// its speed.  The built-in rule (Tuning's 0 / -1) is always a candidate and
// is kept unless another value is faster by more than `margin`, so timing
// noise cannot move a pass off the rule.  Candidates are timed in
// interleaved rounds (rule, a, b, rule, a, b, ...) and compared by their
// median round, so a clock ramp at context creation affects all of them
// alike.  Results are kept per (grid, mode, precision, device) for the
// process, so contexts of one shape share them.  DBEC_AUTOTUNE=0 keeps the
// built-in rules; DBEC_AUTOTUNE_LOG=1 prints the choices to stderr;
// dbec_tuning() returns them as text.
struct TuneKey {
    long nx, ny, nz;
    int mode, prec, dev;
    bool operator<(const TuneKey &o) const {
        return std::tie(nx, ny, nz, mode, prec, dev) < std::tie(o.nx, o.ny, o.nz, o.mode, o.prec, o.dev);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, Tuning> g_tune;

constexpr int kTuneRounds = 7, kTuneReps = 3;
constexpr double kTuneMargin = 0.02;

// Median over kTuneRounds interleaved rounds of kTuneReps back-to-back calls
// of set(i) + launch(), for every candidate i < ncand; ms[i] in ms per call.
template <typename S, typename F>
int time_candidates(dbec_ctx *c, int ncand, S &&set, F &&launch, std::vector<double> &ms) {
    std::vector<hipEvent_t> ev(2 * ncand * kTuneRounds, nullptr);
    hipError_t err = hipSuccess;
    for (auto &e : ev)
        if (err == hipSuccess) err = hipEventCreate(&e);
    for (int i = 0; i < ncand && err == hipSuccess; i++) {  // warm-up: first launch of each variant
        set(i);
        err = launch();
    }
    for (int r = 0; r < kTuneRounds && err == hipSuccess; r++)
        for (int i = 0; i < ncand && err == hipSuccess; i++) {
            set(i);
            hipEvent_t a = ev[2 * (r * ncand + i)], b = ev[2 * (r * ncand + i) + 1];
            err = hipEventRecord(a, c->stream);
            for (int k = 0; k < kTuneReps && err == hipSuccess; k++) err = launch();
            if (err == hipSuccess) err = hipEventRecord(b, c->stream);
        }
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    ms.assign(ncand, 0.);
    for (int i = 0; i < ncand && err == hipSuccess; i++) {
        std::vector<double> t(kTuneRounds);
        for (int r = 0; r < kTuneRounds && err == hipSuccess; r++) {
            float v = 0.f;
            err = hipEventElapsedTime(&v, ev[2 * (r * ncand + i)], ev[2 * (r * ncand + i) + 1]);
            t[r] = (double)v / kTuneReps;
        }
        std::sort(t.begin(), t.end());
        ms[i] = t[kTuneRounds / 2];
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    HIPCHK(c, err, "autotune launch");
    return DBEC_OK;
}

// Index of the candidate to keep: 0 is the built-in rule, kept unless another
// is faster by more than kTuneMargin.
int pick_candidate(const std::vector<double> &ms) {
    int best = 0;
    for (int i = 1; i < (int)ms.size(); i++)
        if (ms[i] < ms[best]) best = i;
    return (best != 0 && ms[best] < (1.0 - kTuneMargin) * ms[0]) ? best : 0;
}

void describe_tuning(const dbec_ctx *c, char *buf, size_t len) {
    const Tuning &t = c->tune;
    std::snprintf(buf, len,
                  "%ldx%ldx%ld %s %s: lm zfwd %d ypass %d xconv %d fused %d; sweep x rowb %d peel %d, "
                  "y rowb %d peel %d, z direct %d; fused layout %d (0 / -1 = built-in rule)",
                  c->Nx, c->Ny, c->Nz, c->mode == DBEC_IMAG ? "imag" : "real",
                  c->prec == DBEC_MIXED ? "mixed" : "fp64", t.lm[0], t.lm[1], t.lm[2], t.lm[3], t.xy_rowb[0],
                  t.xy_peel[0], t.xy_rowb[1], t.xy_peel[1], t.z_direct, t.zinv_swz);
}

int autotune(dbec_ctx *c) {
    const char *env = getenv("DBEC_AUTOTUNE");
    if (env && std::strcmp(env, "0") == 0) return DBEC_OK;
    const TuneKey key{c->Nx, c->Ny, c->Nz, (int)c->mode, (int)c->prec, c->device};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            c->tune = it->second;
            return DBEC_OK;
        }
    }
    const Solves v = solve_ref(c, false);
    // one field: candidate 0 is its rule value
    auto choose = [&](int *field, std::vector<int> cands, auto &&launch) -> int {
        std::vector<double> ms;
        CHK(time_candidates(c, (int)cands.size(), [&](int i) { *field = cands[i]; }, launch, ms));
        *field = cands[pick_candidate(ms)];
        return DBEC_OK;
    };
    Tuning &t = c->tune;
    CHK(choose(&t.lm[0], {0, 1, 2}, [&] { return launch_zfwd(c->Nz, fft_args(c, 2, v)); }));
    CHK(choose(&t.lm[1], {0, 1, 2}, [&] { return launch_ypass(c->Ny, -1, fft_args(c, 1, v)); }));
    CHK(choose(&t.lm[2], {0, 1, 2}, [&] { return launch_xconv(c->Nx, fft_args(c, 0, v)); }));
    if (c->prec != DBEC_MIXED) {
        // the FP64 fused inverse pass has two buffer layouts: its tile
        // multiplier and layout are chosen together
        const int lms[5] = {0, 1, 2, 1, 2}, sws[5] = {-1, 0, 0, 1, 1};
        std::vector<double> ms;
        CHK(time_candidates(
            c, 5, [&](int i) { t.lm[3] = lms[i]; t.zinv_swz = sws[i]; },
            [&] { return launch_zinv_nu(c->Nz, fft_args(c, 2, v)); }, ms));
        const int i = pick_candidate(ms);
        t.lm[3] = lms[i];
        t.zinv_swz = sws[i];
    } else {
        CHK(choose(&t.lm[3], {0, 1, 2}, [&] { return launch_zinv_nu(c->Nz, fft_args(c, 2, v)); }));
    }
    for (int axis = 0; axis < 2; axis++) {
        auto sweep = [&] { return launch_sweep(c->stream, c, axis, v, false); };
        CHK(choose(&t.xy_peel[axis], {-1, 0, 1}, sweep));
        CHK(choose(&t.xy_rowb[axis], {0, 128, 256}, sweep));
    }
    CHK(choose(&t.z_direct, {-1, 0, 1}, [&] { return launch_sweep(c->stream, c, 2, v, c->mode == DBEC_IMAG); }));
    // the candidates ran on zero data; psi and the spectrum stay zero
    if (const char *lg = getenv("DBEC_AUTOTUNE_LOG"); lg && std::strcmp(lg, "0") != 0) {
        char buf[256];
        describe_tuning(c, buf, sizeof buf);
        std::fprintf(stderr, "[dbec autotune] %s\n", buf);
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = t;
    return DBEC_OK;
}

// Host-side selection of every o-th entry of a row-major 2-D array.
void pick2(const std::vector<double> &m, long n0, long n1, long o0, long o1, double *out) {
    long w = 0;
    for (long a = 0; a < n0; a += o0)
        for (long b2 = 0; b2 < n1; b2 += o1) out[w++] = m[a * n1 + b2];
}

}  // namespace

extern "C" {

const char *dbec_last_error(const dbec_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    return g_thread_error.c_str();
}

int dbec_create(const dbec_params *p, dbec_mode mode, dbec_prec prec, void *stream, dbec_ctx **out) {
    if (!p || !out) return fail(nullptr, DBEC_EINVAL, "null argument");
    *out = nullptr;
    if (mode != DBEC_IMAG && mode != DBEC_REAL) return fail(nullptr, DBEC_EINVAL, "bad mode");
    if (prec != DBEC_FP64 && prec != DBEC_MIXED) return fail(nullptr, DBEC_EINVAL, "bad precision");
    if (!fft_len_supported(p->nx) || !fft_len_supported(p->ny) || !fft_len_supported(p->nz))
        return fail(nullptr, DBEC_EUNSUPPORTED,
                    "grid sizes must each be one of 16 32 48 64 80 96 128 160 192 256 320 384 512");
    if (!(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0) || !(p->dt > 0))
        return fail(nullptr, DBEC_EINVAL, "dx, dy, dz, dt must be positive");
    if (!(p->par == 1.0 || p->par == 2.0)) return fail(nullptr, DBEC_EINVAL, "par must be 1 or 2");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) return fail(nullptr, DBEC_EDEVICE, "no HIP device available");
    int dev = p->device;
    if (dev < 0) {
        e = hipGetDevice(&dev);
        if (e != hipSuccess) return fail(nullptr, DBEC_EDEVICE, "hipGetDevice failed");
    }
    if (dev >= ndev) return fail(nullptr, DBEC_EINVAL, "DEVICE out of range");
    e = hipSetDevice(dev);
    if (e != hipSuccess) return fail(nullptr, DBEC_EDEVICE, "hipSetDevice failed");
    scratch_setup(dev);
    dbec_ctx *c = new (std::nothrow) dbec_ctx();
    if (!c) return fail(nullptr, DBEC_ENOMEM, "out of host memory");
    c->p = *p;
    c->mode = mode;
    c->prec = prec;
    c->device = dev;
    c->Nx = p->nx;
    c->Ny = p->ny;
    c->Nz = p->nz;
    c->N = c->Nx * c->Ny * c->Nz;
    c->Nkz = c->Nz / 2 + 1;
    c->pi = (p->pi != 0.) ? p->pi : 3.14159265358979;
    if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
    } else {
        e = hipStreamCreate(&c->stream);
        if (e != hipSuccess) {
            delete c;
            return fail(nullptr, DBEC_EDEVICE, "hipStreamCreate failed");
        }
        c->own_stream = true;
    }
    int rc = setup(c);
    if (rc == DBEC_OK) rc = autotune(c);
    if (rc != DBEC_OK) {
        std::string m = c->err;
        (void)hipStreamSynchronize(c->stream);
        free_buffers(c);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, rc, m);
    }
    *out = c;
    return DBEC_OK;
}

int dbec_destroy(dbec_ctx *c) {
    if (!c) return DBEC_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (!c->view) free_buffers(c);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DBEC_OK;
}

int dbec_set_psi(dbec_ctx *c, const double *psi) {
    if (!c || !psi) return fail(c, DBEC_EINVAL, "null argument");
    (void)hipSetDevice(c->device);
    const size_t nd = (size_t)c->N * (c->mode == DBEC_REAL ? 2 : 1);
    if (c->prec == DBEC_FP64) {
        HIPCHK(c, hipMemcpyAsync(c->b.psi, psi, nd * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy psi");
    } else {
        std::vector<float> f(psi, psi + nd);
        HIPCHK(c, hipMemcpyAsync(c->b.psi, f.data(), nd * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy psi");
    }
    const double zero = 0.0;  // a new psi starts without exponent shift
    HIPCHK(c, hipMemcpyAsync(c->b.scal + SC_SHIFT, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream), "copy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
    c->scale_pending = false;
    return DBEC_OK;
}

int dbec_get_psi(dbec_ctx *c, double *psi) {
    if (!c || !psi) return fail(c, DBEC_EINVAL, "null argument");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    const size_t nd = (size_t)c->N * (c->mode == DBEC_REAL ? 2 : 1);
    if (c->prec == DBEC_FP64) {
        HIPCHK(c, hipMemcpyAsync(psi, c->b.psi, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy psi");
        HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
    } else {
        std::vector<float> f(nd);
        HIPCHK(c, hipMemcpyAsync(f.data(), c->b.psi, nd * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy psi");
        HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
        for (size_t i = 0; i < nd; i++) psi[i] = f[i];
    }
    return DBEC_OK;
}

int dbec_set_coupling(dbec_ctx *c, double g, double gd) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    c->g = g;
    c->gd = gd;
    return DBEC_OK;
}

int dbec_get_coupling(const dbec_ctx *c, double *g, double *gd) {
    if (!c) return DBEC_EINVAL;
    if (g) *g = c->g;
    if (gd) *gd = c->gd;
    return DBEC_OK;
}

int dbec_calcpsidd2(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    return run_calcpsidd2(c);
}

int dbec_get_psidd2(dbec_ctx *c, double *phi) {
    if (!c || !phi) return fail(c, DBEC_EINVAL, "null argument");
    (void)hipSetDevice(c->device);
    if (c->prec == DBEC_FP64) {
        HIPCHK(c, hipMemcpyAsync(phi, c->b.phi, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
        HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
    } else {
        std::vector<float> f(c->N);
        HIPCHK(c, hipMemcpyAsync(f.data(), c->b.phi, c->N * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
        HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
        for (long i = 0; i < c->N; i++) phi[i] = f[i];
    }
    return DBEC_OK;
}

int dbec_calcnu(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    return run_calcnu(c, nullptr);
}

static int sweep_api(dbec_ctx *c, int axis) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    return run_sweep(c, axis, solve_ref(c, false), false);
}
int dbec_calclux(dbec_ctx *c) { return sweep_api(c, 0); }
int dbec_calcluy(dbec_ctx *c) { return sweep_api(c, 1); }
int dbec_calcluz(dbec_ctx *c) { return sweep_api(c, 2); }

int dbec_calcnorm(dbec_ctx *c, double *norm, int rescale) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    HIPCHK(c, launch_rows_mod2(c->stream, c, c->b.rows, nullptr), "rows");
    HIPCHK(c, launch_finalize(c->stream, c, 0, c->b.rows, nullptr), "finalize");
    if (rescale) HIPCHK(c, launch_scale(c->stream, c, c->b.scal + SC_SCALE), "scale");
    double v;
    CHK(read_scalars(c, &v, SC_NORM, 1));
    if (norm) *norm = v;
    return DBEC_OK;
}

static int calcmuen_api(dbec_ctx *c, double *mu, double *en, bool phase) {
    if (!c || !mu || !en) return fail(c, DBEC_EINVAL, "null argument");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    CHK(run_calcpsidd2(c));
    double *rmu = c->b.rows, *ren = c->b.rows + c->Nx * c->Ny;
    HIPCHK(c, launch_rows_muen(c->stream, c, rmu, ren, phase), "muen rows");
    HIPCHK(c, launch_finalize(c->stream, c, 2, rmu, ren), "finalize");
    double v[2];
    CHK(read_scalars(c, v, SC_MU, 2));
    *mu = v[0];
    *en = v[1];
    return DBEC_OK;
}

int dbec_calcmuen(dbec_ctx *c, double *mu, double *en) { return calcmuen_api(c, mu, en, false); }
int dbec_calcmuen_phase(dbec_ctx *c, double *mu, double *en) { return calcmuen_api(c, mu, en, true); }

int dbec_calcrms(dbec_ctx *c, double rms[4]) {
    if (!c || !rms) return fail(c, DBEC_EINVAL, "null argument");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    double *r0 = c->b.rows, *rz = c->b.rows + c->Nx * c->Ny;
    HIPCHK(c, launch_rows_mod2(c->stream, c, r0, rz), "rms rows");
    HIPCHK(c, launch_finalize(c->stream, c, 1, r0, rz), "finalize");
    double v[4];
    CHK(read_scalars(c, v, SC_R0, 4));
    for (int i = 0; i < 4; i++) rms[i] = v[i];
    return DBEC_OK;
}

int dbec_psi_center(dbec_ctx *c, double out[2]) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    const long n = ((c->Nx / 2) * c->Ny + c->Ny / 2) * c->Nz + c->Nz / 2;
    const size_t el = psi_elem(c);
    unsigned char buf[16];
    HIPCHK(c, hipMemcpyAsync(buf, static_cast<char *>(c->b.psi) + n * el, el, hipMemcpyDeviceToHost, c->stream), "copy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
    double v[2] = {0., 0.};
    if (c->prec == DBEC_FP64) std::memcpy(v, buf, el);
    else {
        float f[2] = {0.f, 0.f};
        std::memcpy(f, buf, el);
        v[0] = f[0];
        v[1] = f[1];
    }
    out[0] = v[0];
    out[1] = (c->mode == DBEC_REAL) ? v[1] : 0.;
    return DBEC_OK;
}

int dbec_step(dbec_ctx *c, long nsteps, double g_stp, double gd_stp, double *norm) {
    if (!c || nsteps < 0) return fail(c, DBEC_EINVAL, "bad argument");
    (void)hipSetDevice(c->device);
    const bool imag = c->mode == DBEC_IMAG;
    const bool fuse = fuse_z(c);
    bool z_done = false;
    for (long s = 0; s < nsteps; s++) {
        if (g_stp != 0. || gd_stp != 0.) {
            c->g += g_stp;
            c->gd += gd_stp;
        }
        // after the first step the previous norm is folded into this one
        Solves v = solve_ref(c, imag && (c->scale_pending || s > 0));
        v.fold = (imag && s > 0) ? (z_done ? 2 : 1) : 0;
        const bool fuse_next = fuse && s + 1 < nsteps;
        CHK(run_step(c, v, z_done, fuse_next));
        z_done = fuse_next;
    }
    if (imag && nsteps > 0) {
        CHK(run_norm(c, solve_ref(c, false)));
        c->scale_pending = true;
    }
    if (imag && nsteps > 0 && norm) {
        CHK(materialize_scale(c));
        CHK(read_scalars(c, norm, SC_NORM, 1));
    }
    return DBEC_OK;
}

long dbec_density_count(const dbec_ctx *c, int kind, long ox, long oy, long oz) {
    if (!c || ox < 1 || oy < 1 || oz < 1) return -1;
    long sx = (c->Nx + ox - 1) / ox, sy = (c->Ny + oy - 1) / oy, sz = (c->Nz + oz - 1) / oz;
    switch (kind) {
    case DBEC_DEN_X: return sx;
    case DBEC_DEN_Y: return sy;
    case DBEC_DEN_Z: return sz;
    case DBEC_DEN_XY: case DBEC_CUT_XY0: return sx * sy;
    case DBEC_DEN_XZ: case DBEC_CUT_X0Z: return sx * sz;
    case DBEC_DEN_YZ: case DBEC_CUT_0YZ: return sy * sz;
    case DBEC_DEN_XYZ: return sx * sy * sz;
    default: return -1;
    }
}

int dbec_density(dbec_ctx *c, int kind, long ox, long oy, long oz, double *out) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    if (dbec_density_count(c, kind, ox, oy, oz) < 0) return fail(c, DBEC_EINVAL, "bad density request");
    (void)hipSetDevice(c->device);
    CHK(materialize_scale(c));
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    const double *wx = c->b.wgt, *wy = c->b.wgt + Nx;
    size_t need;
    switch (kind) {
    case DBEC_DEN_X: case DBEC_DEN_Y: case DBEC_DEN_XY: need = Nx * Ny + std::max(Nx, Ny); break;
    case DBEC_DEN_Z: case DBEC_DEN_XZ: need = Nx * Nz + Nz; break;
    case DBEC_DEN_YZ: need = Ny * Nz; break;
    default: need = c->N; break;
    }
    double *d = nullptr;
    HIPCHK(c, hipMalloc(&d, need * sizeof(double)), "density scratch");
    std::vector<double> h;
    hipError_t e = hipSuccess;
    auto fetch = [&](const double *src, size_t n) {
        h.resize(n);
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    };
    switch (kind) {
    case DBEC_DEN_X:
    case DBEC_DEN_Y:
    case DBEC_DEN_XY:
        e = launch_reduce3(c->stream, c, 2, d);  // [x][y]
        if (kind == DBEC_DEN_XY) {
            fetch(d, Nx * Ny);
            if (e == hipSuccess) pick2(h, Nx, Ny, ox, oy, out);
        } else if (kind == DBEC_DEN_X) {
            if (e == hipSuccess) e = launch_reduce2(c->stream, d, Nx, Ny, 1, wy, d + Nx * Ny);
            fetch(d + Nx * Ny, Nx);
            if (e == hipSuccess) pick2(h, Nx, 1, ox, 1, out);
        } else {
            if (e == hipSuccess) e = launch_reduce2(c->stream, d, Nx, Ny, 0, wx, d + Nx * Ny);
            fetch(d + Nx * Ny, Ny);
            if (e == hipSuccess) pick2(h, Ny, 1, oy, 1, out);
        }
        break;
    case DBEC_DEN_Z:
    case DBEC_DEN_XZ:
        e = launch_reduce3(c->stream, c, 1, d);  // [x][z]
        if (kind == DBEC_DEN_XZ) {
            fetch(d, Nx * Nz);
            if (e == hipSuccess) pick2(h, Nx, Nz, ox, oz, out);
        } else {
            if (e == hipSuccess) e = launch_reduce2(c->stream, d, Nx, Nz, 0, wx, d + Nx * Nz);
            fetch(d + Nx * Nz, Nz);
            if (e == hipSuccess) pick2(h, Nz, 1, oz, 1, out);
        }
        break;
    case DBEC_DEN_YZ:
        e = launch_reduce3(c->stream, c, 0, d);  // [y][z]
        fetch(d, Ny * Nz);
        if (e == hipSuccess) pick2(h, Ny, Nz, oy, oz, out);
        break;
    default: {
        e = launch_mod2(c->stream, c, d);
        fetch(d, c->N);
        if (e != hipSuccess) break;
        long w = 0;
        if (kind == DBEC_CUT_XY0) {
            for (long i = 0; i < Nx; i += ox)
                for (long j = 0; j < Ny; j += oy) out[w++] = h[(i * Ny + j) * Nz + Nz / 2];
        } else if (kind == DBEC_CUT_X0Z) {
            for (long i = 0; i < Nx; i += ox)
                for (long k = 0; k < Nz; k += oz) out[w++] = h[(i * Ny + Ny / 2) * Nz + k];
        } else if (kind == DBEC_CUT_0YZ) {
            for (long j = 0; j < Ny; j += oy)
                for (long k = 0; k < Nz; k += oz) out[w++] = h[((Nx / 2) * Ny + j) * Nz + k];
        } else {
            for (long i = 0; i < Nx; i += ox)
                for (long j = 0; j < Ny; j += oy)
                    for (long k = 0; k < Nz; k += oz) out[w++] = h[(i * Ny + j) * Nz + k];
        }
    } break;
    }
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(c, e, "density");
    return DBEC_OK;
}

int dbec_density_noalloc(dbec_ctx *c, int kind, long ox, long oy, long oz, double *out) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    const long cnt = dbec_density_count(c, kind, ox, oy, oz);
    if (cnt < 0) return fail(c, DBEC_EINVAL, "bad density request");
    (void)hipSetDevice(c->device);
    double *d = scratch_of(c->device);
    if (!d) return fail(c, DBEC_EDEVICE, "dbec_density_noalloc: the device scratch of this GPU has no address");
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    const long sx = (Nx + ox - 1) / ox, sy = (Ny + oy - 1) / oy, sz = (Nz + oz - 1) / oz;
    // doubles of scratch: the weighted sums of the extract, then its sampled values
    long sums;
    switch (kind) {
    case DBEC_DEN_XY: sums = Nx * Ny; break;
    case DBEC_DEN_X: sums = Nx * Ny + Nx; break;
    case DBEC_DEN_Y: sums = Nx * Ny + Ny; break;
    case DBEC_DEN_XZ: sums = Nx * Nz; break;
    case DBEC_DEN_Z: sums = Nx * Nz + Nz; break;
    case DBEC_DEN_YZ: sums = Ny * Nz; break;
    default: sums = 0; break;
    }
    const long cap = (long)(DBEC_NOALLOC_SCRATCH_BYTES / sizeof(double));
    if (sums + cnt > cap)
        return fail(c, DBEC_EINVAL,
                    "dbec_density_noalloc: this extract needs " + std::to_string((sums + cnt) * (long)sizeof(double)) +
                        " bytes of device scratch, more than its " + std::to_string((long)DBEC_NOALLOC_SCRATCH_BYTES) +
                        " bytes (use dbec_density or coarser output steps)");
    CHK(materialize_scale(c));
    const double *wx = c->b.wgt, *wy = c->b.wgt + Nx;
    double *dst = d + sums;
    hipStream_t s = c->stream;
    hipError_t e = hipSuccess;
    switch (kind) {
    case DBEC_DEN_XY:
        e = launch_reduce3(s, c, 2, d);  // [x][y]
        if (e == hipSuccess) e = launch_pick2(s, d, Ny, ox, oy, sx, sy, dst);
        break;
    case DBEC_DEN_X:
        e = launch_reduce3(s, c, 2, d);
        if (e == hipSuccess) e = launch_reduce2(s, d, Nx, Ny, 1, wy, d + Nx * Ny);
        if (e == hipSuccess) e = launch_pick2(s, d + Nx * Ny, 1, ox, 1, sx, 1, dst);
        break;
    case DBEC_DEN_Y:
        e = launch_reduce3(s, c, 2, d);
        if (e == hipSuccess) e = launch_reduce2(s, d, Nx, Ny, 0, wx, d + Nx * Ny);
        if (e == hipSuccess) e = launch_pick2(s, d + Nx * Ny, 1, oy, 1, sy, 1, dst);
        break;
    case DBEC_DEN_XZ:
        e = launch_reduce3(s, c, 1, d);  // [x][z]
        if (e == hipSuccess) e = launch_pick2(s, d, Nz, ox, oz, sx, sz, dst);
        break;
    case DBEC_DEN_Z:
        e = launch_reduce3(s, c, 1, d);
        if (e == hipSuccess) e = launch_reduce2(s, d, Nx, Nz, 0, wx, d + Nx * Nz);
        if (e == hipSuccess) e = launch_pick2(s, d + Nx * Nz, 1, oz, 1, sz, 1, dst);
        break;
    case DBEC_DEN_YZ:
        e = launch_reduce3(s, c, 0, d);  // [y][z]
        if (e == hipSuccess) e = launch_pick2(s, d, Nz, oy, oz, sy, sz, dst);
        break;
    default: {
        long start[3] = {0, 0, 0}, step[3] = {ox, oy, oz}, n[3] = {sx, sy, sz};
        if (kind == DBEC_CUT_XY0) {
            start[2] = Nz / 2;
            step[2] = 1;
            n[2] = 1;
        } else if (kind == DBEC_CUT_X0Z) {
            start[1] = Ny / 2;
            step[1] = 1;
            n[1] = 1;
        } else if (kind == DBEC_CUT_0YZ) {
            start[0] = Nx / 2;
            step[0] = 1;
            n[0] = 1;
        }
        e = launch_pick_mod2(s, c, start, step, n, dst);
    } break;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dst, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "density (no allocation)");
    return DBEC_OK;
}

int dbec_get_table(dbec_ctx *c, int which, int axis, double *out) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    switch (which) {
    case DBEC_TABLE_POT: {
        const dbec_params &p = c->p;
        const double vg2 = p.vgamma * p.vgamma, vn2 = p.vnu * p.vnu, vl2 = p.vlambda * p.vlambda;
        for (long i = 0; i < c->Nx; i++) {
            double x = (i - c->Nx / 2) * p.dx, x2 = x * x;
            for (long j = 0; j < c->Ny; j++) {
                double y = (j - c->Ny / 2) * p.dy, y2 = y * y;
                for (long k = 0; k < c->Nz; k++) {
                    double z = (k - c->Nz / 2) * p.dz, z2 = z * z;
                    out[(i * c->Ny + j) * c->Nz + k] = 0.5 * p.par * (vg2 * x2 + vn2 * y2 + vl2 * z2);
                }
            }
        }
        return DBEC_OK;
    }
    case DBEC_TABLE_POTDD:
        std::memcpy(out, c->h_potdd_half.data(), c->h_potdd_half.size() * sizeof(double));
        return DBEC_OK;
    case DBEC_TABLE_CALPHA:
    case DBEC_TABLE_CGAMMA: {
        if (axis < 0 || axis > 2) return fail(c, DBEC_EINVAL, "bad axis");
        const std::vector<double> &v = (which == DBEC_TABLE_CALPHA) ? c->h_calpha[axis] : c->h_cgamma[axis];
        std::memcpy(out, v.data(), v.size() * sizeof(double));
        return DBEC_OK;
    }
    default:
        return fail(c, DBEC_EINVAL, "bad table");
    }
}

int dbec_tuning(const dbec_ctx *c, char *buf, long len) {
    if (!c || !buf || len < 1) return DBEC_EINVAL;
    describe_tuning(c, buf, (size_t)len);
    return DBEC_OK;
}

int dbec_synchronize(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
    return DBEC_OK;
}

int dbec_profile_step(dbec_ctx *c, int reps, dbec_kernel_stat *stats, int max, int *n) {
    if (!c || reps < 1 || !stats || !n) return fail(c, DBEC_EINVAL, "bad argument");
    (void)hipSetDevice(c->device);
    // calls of 2 and of 3 steps, every launch a burst of reps: the difference
    // of their launch counts is the launches per step of a long run (the
    // first forward z pass, the last z sweep and the norm run once per call);
    // the times are those of the 3-step call
    struct Agg {
        std::vector<std::string> names;
        std::vector<double> ms, bytes, cnt;
    } agg[2];
    int rc = DBEC_OK;
    hipError_t e = hipSuccess;
    for (int pass = 0; pass < 2 && rc == DBEC_OK; pass++) {
        KernelTimer timer;
        timer.burst = reps;
        c->timer = &timer;
        rc = dbec_step(c, 2 + pass, 0., 0., nullptr);
        if (rc == DBEC_OK) rc = materialize_scale(c);
        c->timer = nullptr;
        e = hipStreamSynchronize(c->stream);
        Agg &g = agg[pass];
        for (auto &r : timer.recs) {  // aggregate by name, in first-seen order
            float t = 0.f;
            if (e == hipSuccess) (void)hipEventElapsedTime(&t, r.a, r.b);
            size_t k = 0;
            while (k < g.names.size() && g.names[k] != r.name) k++;
            if (k == g.names.size()) {
                g.names.push_back(r.name);
                g.ms.push_back(0.);
                g.bytes.push_back(r.bytes);
                g.cnt.push_back(0.);
            }
            g.ms[k] += t;
            g.cnt[k] += 1.;
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        if (e != hipSuccess) break;
    }
    if (rc != DBEC_OK) return rc;
    if (e != hipSuccess) return hip_fail(c, e, "synchronize");
    const Agg &g3 = agg[1], &g2 = agg[0];
    int m = 0;
    for (size_t k = 0; k < g3.names.size() && m < max; k++, m++) {
        double c2 = 0.;
        for (size_t j = 0; j < g2.names.size(); j++)
            if (g2.names[j] == g3.names[k]) c2 = g2.cnt[j];
        std::memset(&stats[m], 0, sizeof(dbec_kernel_stat));
        std::snprintf(stats[m].name, sizeof stats[m].name, "%s", g3.names[k].c_str());
        stats[m].avg_ms = g3.ms[k] / (g3.cnt[k] * reps);
        stats[m].alg_bytes = g3.bytes[k];
        stats[m].launches = g3.cnt[k] - c2;
    }
    *n = (int)g3.names.size();
    return DBEC_OK;
}

}  // extern "C"

namespace dbec {

Solves core_solves(const dbec_ctx *c, bool use_scale) { return solve_ref(c, use_scale); }

int core_step(dbec_ctx *c, const Solves &v, bool z_done, bool fuse_next) {
    return run_step(c, v, z_done, fuse_next);
}
bool core_fuse_z(const dbec_ctx *c) { return fuse_z(c); }

int core_norm(dbec_ctx *c, const Solves &v) { return run_norm(c, v); }

int core_observe_async(dbec_ctx *c, int what) {
    if (c->mode == DBEC_IMAG) {
        HIPCHK(c, launch_scale(c->stream, c, c->b.scal + SC_SCALE), "scale");
        HIPCHK(c, launch_set_scalar(c->stream, c->b.scal + SC_SCALE, 1.0), "scalar");
    }
    double *r0 = c->b.rows, *r1 = c->b.rows + c->Nx * c->Ny;
    if (what & OBS_MUEN) {
        CHK(run_calcpsidd2(c));
        HIPCHK(c, launch_rows_muen(c->stream, c, r0, r1), "muen rows");
        HIPCHK(c, launch_finalize(c->stream, c, 2, r0, r1), "finalize");
    }
    if (what & OBS_RMS) {
        HIPCHK(c, launch_rows_mod2(c->stream, c, r0, r1), "rms rows");
        HIPCHK(c, launch_finalize(c->stream, c, 1, r0, r1), "finalize");
    }
    return DBEC_OK;
}

int core_fail(dbec_ctx *c, int code, const std::string &msg) { return fail(c, code, msg); }
size_t core_psi_elem(const dbec_ctx *c) { return psi_elem(c); }
size_t core_real_elem(const dbec_ctx *c) { return real_elem(c); }

}  // namespace dbec
