/*
 * oracle.cpp - TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * CPU restatement, in plain FP64 C++ with OpenMP, of the per-step algorithm of
 * the reference programs imag3d-mpicuda and real3d-mpicuda at one MPI rank.
 * It implements the same C ABI as the HIP library (include/dbec.h) so that
 * tests drive both backends with identical calls.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load it.
 *
 * Reference aliases: imag3d.cu = src/imag3d-mpicuda/imag3d-mpicuda.cu,
 * real3d.cu = src/real3d-mpicuda/real3d-mpicuda.cu, diffint.cu =
 * src/utils/diffint.cu.  Each function names the lines it restates.
 *
 * Deliberate differences from the reference, all at round-off level:
 *  - FFTs: the reference calls cuFFT (R2C over y,z per x slab, C2C along x,
 *    C2R back).  Here the same unnormalised convolution is computed with a
 *    full complex 3-D DFT of |psi|^2 (own Stockham mixed-radix FFT, twiddles
 *    in long double), multiplied by potdd over all k, and the real part of
 *    the inverse is taken.  potdd is even in every k component, so this is
 *    the same mathematical quantity.
 *  - The last Simpson level on the GPU is a block tree reduction
 *    (diffint.cu:141-200); here every Simpson sum runs sequentially in the
 *    order of simpint (diffint.cu:37-51).
 *  - Complex helpers (|z|, division) are written from the mathematics:
 *    hypot for the modulus, Smith's (1962) algorithm for division.
 * Results do not depend on the number of OpenMP threads: parallel loops only
 * distribute independent lines, and every reduction is sequential.
 */
#include "dbec.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include <omp.h>

namespace {

struct cplx {
    double re, im;
};
inline cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
inline cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cplx cconj(cplx a) { return {a.re, -a.im}; }
inline double cabs_(cplx a) { return std::hypot(a.re, a.im); }
/* a / b by Smith's algorithm. */
inline cplx cdiv(cplx a, cplx b) {
    if (std::fabs(b.re) >= std::fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    } else {
        double r = b.re / b.im, d = b.im + b.re * r;
        return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
    }
}
/* x * y + d with fused multiply-adds (the reference's cuCfma, real3d.cu:1936). */
inline cplx cfma(cplx x, cplx y, cplx d) {
    double re = std::fma(x.re, y.re, d.re);
    re = std::fma(-x.im, y.im, re);
    double im = std::fma(x.re, y.im, d.im);
    im = std::fma(x.im, y.re, im);
    return {re, im};
}

thread_local std::string g_thread_error;

/* ---------------------------------------------------------------------------
 * 1-D complex FFT: Stockham autosort, mixed radix (4, 2, 3, 5, then any prime
 * by a direct DFT).  Unnormalised; sign -1 forward, +1 inverse.
 * ------------------------------------------------------------------------ */
struct Fft1d {
    long n = 0;
    std::vector<int> radix;
    std::vector<cplx> tw;  /* tw[k] = exp(-2 pi i k / n) */

    void init(long n_) {
        n = n_;
        radix.clear();
        long m = n;
        while (m % 4 == 0) { radix.push_back(4); m /= 4; }
        while (m % 2 == 0) { radix.push_back(2); m /= 2; }
        for (long p = 3; m > 1; p += 2) {
            while (m % p == 0) { radix.push_back((int)p); m /= p; }
        }
        tw.resize(n);
        const long double twopi = 6.283185307179586476925286766559005768L;
        for (long k = 0; k < n; k++) {
            long double a = twopi * (long double)k / (long double)n;
            tw[k] = {(double)cosl(a), (double)-sinl(a)};
        }
    }

    /* data: n values, scratch: n values. */
    void run(cplx *data, cplx *scratch, int sign) const {
        cplx *a = data, *b = scratch;
        long ns = 1;
        cplx v[64], w[64];
        for (int R : radix) {
            long m = n / R;
            long tstep = n / (ns * R);
            for (long j = 0; j < m; j++) {
                long k = j % ns;
                for (int r = 0; r < R; r++) {
                    cplx x = a[j + r * m];
                    long ti = (long)r * k * tstep;
                    cplx t = tw[ti];
                    if (sign > 0) t = cconj(t);
                    v[r] = (ti == 0) ? x : cmul(x, t);
                }
                if (R == 2) {
                    w[0] = cadd(v[0], v[1]);
                    w[1] = csub(v[0], v[1]);
                } else if (R == 4) {
                    cplx t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
                    cplx t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
                    /* t3 = sign * i * (v1 - v3) */
                    cplx t3 = (sign < 0) ? cplx{d.im, -d.re} : cplx{-d.im, d.re};
                    w[0] = cadd(t0, t2);
                    w[2] = csub(t0, t2);
                    w[1] = cadd(t1, t3);
                    w[3] = csub(t1, t3);
                } else {
                    /* direct DFT of size R with roots exp(sign 2 pi i q r / R) */
                    long step = n / R;
                    for (int q = 0; q < R; q++) {
                        cplx s = {0.0, 0.0};
                        for (int r = 0; r < R; r++) {
                            long ti = ((long)q * r % R) * step;
                            cplx t = tw[ti];
                            if (sign > 0) t = cconj(t);
                            s = cadd(s, cmul(v[r], t));
                        }
                        w[q] = s;
                    }
                }
                long base = (j / ns) * ns * R + k;
                for (int r = 0; r < R; r++) b[base + r * ns] = w[r];
            }
            cplx *tmp = a; a = b; b = tmp;
            ns *= R;
        }
        if (a != data) std::memcpy(data, a, sizeof(cplx) * n);
    }
};

/* Simpson rule of the reference (diffint.cu:37-51): pairwise over odd centres
 * plus, for even N, the quadratic end correction (5f[N-1] + 8f[N-2] - f[N-3])/4. */
double simpint(double h, const double *f, long N) {
    double sumi = 0., sumj = 0., sumk = 0.;
    for (long i = 1; i < N - 1; i += 2) {
        sumi += f[i];
        sumj += f[i - 1];
        sumk += f[i + 1];
    }
    double sum = sumj + 4. * sumi + sumk;
    if (N % 2 == 0) sum += (5. * f[N - 1] + 8. * f[N - 2] - f[N - 3]) / 4.;
    return sum * h / 3.;
}

/* Five-point derivative with the reference's boundary rule
 * (imag3d.cu:1303-1332): 0 at both ends, centred 3-point at 1 and N-2. */
inline double deriv5(const double *f, long i, long N, long stride, double h) {
    if (i == 0 || i == N - 1) return 0.;
    if (i == 1) return (f[2 * stride] - f[0]) / (2. * h);
    if (i == N - 2) return (f[(N - 1) * stride] - f[(N - 3) * stride]) / (2. * h);
    return (f[(i - 2) * stride] - 8. * f[(i - 1) * stride] + 8. * f[(i + 1) * stride] -
            f[(i + 2) * stride]) / (12. * h);
}

}  // namespace

struct dbec_ctx {
    dbec_params p;
    dbec_mode mode;
    dbec_prec prec;
    long Nx, Ny, Nz, Nx2, Ny2, Nz2, N;
    double pi;
    std::vector<double> x, y, z, x2, y2, z2;
    std::vector<double> pot;    /* [x][y][z] */
    std::vector<double> potdd;  /* [kx][ky][kz], all kz */
    /* CN coefficients, per axis 0=x 1=y 2=z.  Imaginary time: real; real time: complex. */
    double A0[3], A0r[3], A[3];
    std::vector<double> calpha[3], cgamma[3];
    cplx A0c[3], A0rc[3], Ac[3], minusAc[3];
    std::vector<cplx> calphac[3], cgammac[3];
    std::vector<double> psir;   /* imag: [x][y][z] */
    std::vector<cplx> psic;     /* real: [x][y][z] */
    std::vector<double> phi;    /* psidd2: unnormalised dipolar potential [x][y][z] */
    std::vector<cplx> work;     /* 3-D FFT buffer */
    Fft1d fx, fy, fz;
    double g = 0., gd = 0.;
    std::string err;
};

namespace {

int fail(dbec_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    g_thread_error = msg;
    return code;
}

inline long idx(const dbec_ctx *c, long i, long j, long k) { return (i * c->Ny + j) * c->Nz + k; }

/* |psi|^2 at a flat index. */
inline double psi2_at(const dbec_ctx *c, long n) {
    if (c->mode == DBEC_IMAG) {
        double t = c->psir[n];
        return t * t;
    }
    double t = cabs_(c->psic[n]);
    return t * t;
}

/* FFT of every line of the 3-D complex buffer along one axis. */
void fft_axis(dbec_ctx *c, int axis, int sign) {
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    cplx *w = c->work.data();
    if (axis == 2) {
#pragma omp parallel
        {
            std::vector<cplx> s(Nz);
#pragma omp for schedule(static)
            for (long r = 0; r < Nx * Ny; r++) c->fz.run(w + r * Nz, s.data(), sign);
        }
    } else if (axis == 1) {
#pragma omp parallel
        {
            std::vector<cplx> buf(Ny * Nz), s(Ny);
#pragma omp for schedule(static)
            for (long i = 0; i < Nx; i++) {
                cplx *pl = w + i * Ny * Nz;
                for (long j = 0; j < Ny; j++)
                    for (long k = 0; k < Nz; k++) buf[k * Ny + j] = pl[j * Nz + k];
                for (long k = 0; k < Nz; k++) c->fy.run(buf.data() + k * Ny, s.data(), sign);
                for (long j = 0; j < Ny; j++)
                    for (long k = 0; k < Nz; k++) pl[j * Nz + k] = buf[k * Ny + j];
            }
        }
    } else {
#pragma omp parallel
        {
            std::vector<cplx> buf(Nx * Nz), s(Nx);
#pragma omp for schedule(static)
            for (long j = 0; j < Ny; j++) {
                for (long i = 0; i < Nx; i++)
                    for (long k = 0; k < Nz; k++) buf[k * Nx + i] = w[idx(c, i, j, k)];
                for (long k = 0; k < Nz; k++) c->fx.run(buf.data() + k * Nx, s.data(), sign);
                for (long i = 0; i < Nx; i++)
                    for (long k = 0; k < Nz; k++) w[idx(c, i, j, k)] = buf[k * Nx + i];
            }
        }
    }
}

/* calcpsidd2_mem (imag3d.cu:1460-1510; real3d.cu:1554-1604):
 * phi = IDFT(potdd * DFT(|psi|^2)) unnormalised, then the periodic face
 * copies of calcpsidd2_kernel2a/2b (imag3d.cu:1491-1505, 1591-1625) at one rank:
 *   x = Nx-1, y < Ny-1, z < Nz-1  <- x = 0
 *   x < Nx-1, y = Ny-1, z < Nz-1  <- y = 0
 *   x < Nx-1, y < Ny-1, z = Nz-1  <- z = 0
 * and the edges (Nx-1, Ny-1, .), (Nx-1, ., Nz-1), (x<Nx-1, Ny-1, Nz-1) keep
 * their transform values. */
void calcpsidd2_impl(dbec_ctx *c) {
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz, N = c->N;
    cplx *w = c->work.data();
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) w[n] = {psi2_at(c, n), 0.};
    fft_axis(c, 2, -1);
    fft_axis(c, 1, -1);
    fft_axis(c, 0, -1);
    const double *pd = c->potdd.data();
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) {
        w[n].re *= pd[n];
        w[n].im *= pd[n];
    }
    fft_axis(c, 0, +1);
    fft_axis(c, 1, +1);
    fft_axis(c, 2, +1);
    double *phi = c->phi.data();
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) phi[n] = w[n].re;
    /* kernel2a: last x plane <- first (cudaMemcpy of plane 0 happens first). */
    for (long j = 0; j < Ny - 1; j++)
        for (long k = 0; k < Nz - 1; k++) phi[idx(c, Nx - 1, j, k)] = phi[idx(c, 0, j, k)];
    /* kernel2b with last = 1. */
    for (long i = 0; i < Nx - 1; i++) {
        for (long k = 0; k < Nz - 1; k++) phi[idx(c, i, Ny - 1, k)] = phi[idx(c, i, 0, k)];
        for (long j = 0; j < Ny - 1; j++) phi[idx(c, i, j, Nz - 1)] = phi[idx(c, i, j, 0)];
    }
}

/* calcnu (imag3d.cu:1759-1783; real3d.cu:1856-1883). */
void calcnu_impl(dbec_ctx *c) {
    const long N = c->N;
    const double Nd = (double)(c->Nx * c->Ny * c->Nz);
    const double dt = c->p.dt, g = c->g, gd = c->gd;
    const double *phi = c->phi.data(), *pot = c->pot.data();
    if (c->mode == DBEC_IMAG) {
        double *psi = c->psir.data();
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; n++) {
            double psitmp = psi[n];
            double psi2lin = psitmp * psitmp * g;
            double psidd2lin = (phi[n] / Nd) * gd;
            double pottmp = dt * (pot[n] + psi2lin + psidd2lin);
            psi[n] = psitmp * std::exp(-pottmp);
        }
    } else {
        cplx *psi = c->psic.data();
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; n++) {
            cplx psitmp = psi[n];
            double a = cabs_(psitmp);
            double psi2lin = a * a * g;
            double psidd2lin = (phi[n] / Nd) * gd;
            double pottmp = dt * (pot[n] + psi2lin + psidd2lin);
            cplx e = {std::cos(-pottmp), std::sin(-pottmp)};
            psi[n] = cmul(psitmp, e);
        }
    }
}

/* Thomas sweep of one line (imag3d.cu:1802-1844 / 1858-1900 / 1914-1942):
 * backward cbeta recurrence, forward substitution with fma, Dirichlet zeros. */
void sweep_line_real(double *f, long n, long stride, const double *calpha, const double *cgamma,
                     double A0r, double A, double *cbeta) {
    cbeta[n - 2] = f[(n - 1) * stride];
    for (long i = n - 2; i > 0; i--) {
        double c = -A * f[(i + 1) * stride] + A0r * f[i * stride] - A * f[(i - 1) * stride];
        cbeta[i - 1] = cgamma[i] * (A * cbeta[i] - c);
    }
    f[0] = 0.;
    for (long i = 0; i < n - 2; i++) f[(i + 1) * stride] = std::fma(calpha[i], f[i * stride], cbeta[i]);
    f[(n - 1) * stride] = 0.;
}

/* Complex form (real3d.cu:1902-1944): c = minusA psi[i+1] + A0r psi[i] - A psi[i-1]. */
void sweep_line_cplx(cplx *f, long n, long stride, const cplx *calpha, const cplx *cgamma,
                     cplx A0r, cplx A, cplx minusA, cplx *cbeta) {
    cbeta[n - 2] = f[(n - 1) * stride];
    for (long i = n - 2; i > 0; i--) {
        cplx c = csub(cadd(cmul(minusA, f[(i + 1) * stride]), cmul(A0r, f[i * stride])),
                      cmul(A, f[(i - 1) * stride]));
        cbeta[i - 1] = cmul(cgamma[i], csub(cmul(A, cbeta[i]), c));
    }
    f[0] = {0., 0.};
    for (long i = 0; i < n - 2; i++) f[(i + 1) * stride] = cfma(calpha[i], f[i * stride], cbeta[i]);
    f[(n - 1) * stride] = {0., 0.};
}

void sweep_axis(dbec_ctx *c, int axis) {
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    long n = axis == 0 ? Nx : axis == 1 ? Ny : Nz;
    long stride = axis == 0 ? Ny * Nz : axis == 1 ? Nz : 1;
    long nlines = c->N / n;
#pragma omp parallel
    {
        std::vector<double> cb(c->mode == DBEC_IMAG ? n : 0);
        std::vector<cplx> cbc(c->mode == DBEC_REAL ? n : 0);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; l++) {
            long base;
            if (axis == 0) base = l;                                /* l = j*Nz + k */
            else if (axis == 1) base = (l / Nz) * Ny * Nz + l % Nz; /* l = i*Nz + k */
            else base = l * Nz;                                     /* l = i*Ny + j */
            if (c->mode == DBEC_IMAG)
                sweep_line_real(c->psir.data() + base, n, stride, c->calpha[axis].data(),
                                c->cgamma[axis].data(), c->A0r[axis], c->A[axis], cb.data());
            else
                sweep_line_cplx(c->psic.data() + base, n, stride, c->calphac[axis].data(),
                                c->cgammac[axis].data(), c->A0rc[axis], c->Ac[axis],
                                c->minusAc[axis], cbc.data());
        }
    }
}

/* Integral of f[x][y][z] in the order z, y, x (simpint3d, simpint2d,
 * simpint_gpu as in calcnorm imag3d.cu:1198-1209). */
double integrate_zyx(const dbec_ctx *c, const double *f) {
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    std::vector<double> tx(Nx);
#pragma omp parallel
    {
        std::vector<double> ty(Ny);
#pragma omp for schedule(static)
        for (long i = 0; i < Nx; i++) {
            for (long j = 0; j < Ny; j++) ty[j] = simpint(c->p.dz, f + idx(c, i, j, 0), Nz);
            tx[i] = simpint(c->p.dy, ty.data(), Ny);
        }
    }
    return simpint(c->p.dx, tx.data(), Nx);
}

/* calcnorm (imag3d.cu:1194-1218; real3d.cu:1304-1322). */
double calcnorm_impl(dbec_ctx *c, int rescale) {
    const long N = c->N;
    std::vector<double> t(N);
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) t[n] = psi2_at(c, n);
    double norm = std::sqrt(integrate_zyx(c, t.data()));
    if (rescale) {
        double s = 1. / norm;
        if (c->mode == DBEC_IMAG) {
#pragma omp parallel for schedule(static)
            for (long n = 0; n < N; n++) c->psir[n] *= s;
        } else {
#pragma omp parallel for schedule(static)
            for (long n = 0; n < N; n++) {
                c->psic[n].re *= s;
                c->psic[n].im *= s;
            }
        }
    }
    return norm;
}

/* calcmuen (imag3d.cu:1250-1450; real3d.cu:1339-1543).  Real time
 * differentiates |psi| (real3d.cu:1401-1510).  phase (an extension, real
 * time only): the kinetic density is |grad psi|^2 of the complex psi, the
 * same stencil applied to the real and imaginary parts, squares added. */
void calcmuen_impl(dbec_ctx *c, double *mu, double *en, bool phase = false) {
    calcpsidd2_impl(c);
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz, N = c->N;
    const double par = c->p.par, Nd = (double)(Nx * Ny * Nz);
    const bool cplx = phase && c->mode == DBEC_REAL;
    std::vector<double> f(N), dpsi(N), tmu(N), ten(N), fi(cplx ? N : 0);
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) {
        if (cplx) {
            f[n] = c->psic[n].re;
            fi[n] = c->psic[n].im;
        } else {
            f[n] = (c->mode == DBEC_IMAG) ? c->psir[n] : cabs_(c->psic[n]);
        }
    }
    auto kin = [&](const std::vector<double> &g, long i, long j, long k) {
        double dxv = deriv5(g.data() + idx(c, 0, j, k), i, Nx, Ny * Nz, c->p.dx);
        double d = (dxv * dxv) / (3. - par);
        double dyv = deriv5(g.data() + idx(c, i, 0, k), j, Ny, Nz, c->p.dy);
        d += (dyv * dyv) / (3. - par);
        double dzv = deriv5(g.data() + idx(c, i, j, 0), k, Nz, 1, c->p.dz);
        d += (dzv * dzv) / (3. - par);
        return d;
    };
#pragma omp parallel for schedule(static)
    for (long i = 0; i < Nx; i++)
        for (long j = 0; j < Ny; j++)
            for (long k = 0; k < Nz; k++) {
                long n = idx(c, i, j, k);
                dpsi[n] = cplx ? kin(f, i, j, k) + kin(fi, i, j, k) : kin(f, i, j, k);
            }
    if (cplx) {
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; n++) f[n] = cabs_(c->psic[n]);
    }
    const double g = c->g, gd = c->gd;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) {
        double psi2 = f[n];
        psi2 *= psi2;
        double psi2lin = psi2 * g;
        double psidd2lin = (c->phi[n] / Nd) * gd;
        double tmppot = c->pot[n];
        tmu[n] = (tmppot + psi2lin + psidd2lin) * psi2 + dpsi[n];
        ten[n] = (tmppot + 0.5 * psi2lin + 0.5 * psidd2lin) * psi2 + dpsi[n];
    }
    *mu = integrate_zyx(c, tmu.data());
    *en = integrate_zyx(c, ten.data());
}

/* calcrms (imag3d.cu:1639-1744; real3d.cu:1733-1848): integration orders
 * x,z,y for <x^2>; y,z,x for <y^2>; z,y,x for <z^2>; no centring. */
void calcrms_impl(dbec_ctx *c, double rms[4]) {
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz, N = c->N;
    std::vector<double> p2(N);
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) p2[n] = psi2_at(c, n);
    /* <x^2>: over x, then z, then y. */
    std::vector<double> ty(Ny);
#pragma omp parallel
    {
        std::vector<double> lx(Nx), lz(Nz);
#pragma omp for schedule(static)
        for (long j = 0; j < Ny; j++) {
            for (long k = 0; k < Nz; k++) {
                for (long i = 0; i < Nx; i++) lx[i] = c->x2[i] * p2[idx(c, i, j, k)];
                lz[k] = simpint(c->p.dx, lx.data(), Nx);
            }
            ty[j] = simpint(c->p.dz, lz.data(), Nz);
        }
    }
    rms[1] = std::sqrt(simpint(c->p.dy, ty.data(), Ny));
    /* <y^2>: over y, then z, then x. */
    std::vector<double> tx(Nx);
#pragma omp parallel
    {
        std::vector<double> ly(Ny), lz(Nz);
#pragma omp for schedule(static)
        for (long i = 0; i < Nx; i++) {
            for (long k = 0; k < Nz; k++) {
                for (long j = 0; j < Ny; j++) ly[j] = c->y2[j] * p2[idx(c, i, j, k)];
                lz[k] = simpint(c->p.dy, ly.data(), Ny);
            }
            tx[i] = simpint(c->p.dz, lz.data(), Nz);
        }
    }
    rms[2] = std::sqrt(simpint(c->p.dx, tx.data(), Nx));
    /* <z^2>: over z, then y, then x. */
#pragma omp parallel
    {
        std::vector<double> lz(Nz), ly(Ny);
#pragma omp for schedule(static)
        for (long i = 0; i < Nx; i++) {
            for (long j = 0; j < Ny; j++) {
                for (long k = 0; k < Nz; k++) lz[k] = c->z2[k] * p2[idx(c, i, j, k)];
                ly[j] = simpint(c->p.dz, lz.data(), Nz);
            }
            tx[i] = simpint(c->p.dy, ly.data(), Ny);
        }
    }
    rms[3] = std::sqrt(simpint(c->p.dx, tx.data(), Nx));
    rms[0] = std::sqrt(rms[1] * rms[1] + rms[2] * rms[2] + rms[3] * rms[3]);
}

bool supported_len(long n) {
    if (n < 4 || n % 2 != 0) return false;
    return true;
}

}  // namespace

extern "C" {

const char *dbec_last_error(const dbec_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    return g_thread_error.c_str();
}

int dbec_create(const dbec_params *p, dbec_mode mode, dbec_prec prec, void *stream, dbec_ctx **out) {
    (void)stream;
    if (!p || !out) return fail(nullptr, DBEC_EINVAL, "null argument");
    *out = nullptr;
    if (mode != DBEC_IMAG && mode != DBEC_REAL) return fail(nullptr, DBEC_EINVAL, "bad mode");
    if (prec != DBEC_FP64) return fail(nullptr, DBEC_EUNSUPPORTED, "the oracle computes in FP64 only");
    if (!supported_len(p->nx) || !supported_len(p->ny) || !supported_len(p->nz))
        return fail(nullptr, DBEC_EUNSUPPORTED, "grid sizes must be even and >= 4");
    if (!(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0) || !(p->dt > 0))
        return fail(nullptr, DBEC_EINVAL, "dx, dy, dz, dt must be positive");
    dbec_ctx *c = new (std::nothrow) dbec_ctx();
    if (!c) return fail(nullptr, DBEC_ENOMEM, "out of host memory");
    c->p = *p;
    c->mode = mode;
    c->prec = prec;
    c->Nx = p->nx; c->Ny = p->ny; c->Nz = p->nz;
    c->Nx2 = c->Nx / 2; c->Ny2 = c->Ny / 2; c->Nz2 = c->Nz / 2;
    c->N = c->Nx * c->Ny * c->Nz;
    const double pi = (p->pi != 0.) ? p->pi : 3.14159265358979;
    c->pi = pi;
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    const double dx = p->dx, dy = p->dy, dz = p->dz, par = p->par;
    try {
        /* Grid (initpsi, imag3d.cu:842-855). */
        c->x.resize(Nx); c->y.resize(Ny); c->z.resize(Nz);
        c->x2.resize(Nx); c->y2.resize(Ny); c->z2.resize(Nz);
        for (long i = 0; i < Nx; i++) { c->x[i] = (i - c->Nx2) * dx; c->x2[i] = c->x[i] * c->x[i]; }
        for (long j = 0; j < Ny; j++) { c->y[j] = (j - c->Ny2) * dy; c->y2[j] = c->y[j] * c->y[j]; }
        for (long k = 0; k < Nz; k++) { c->z[k] = (k - c->Nz2) * dz; c->z2[k] = c->z[k] * c->z[k]; }
        /* Trap (initpot, imag3d.cu:872-889). */
        c->pot.resize(c->N);
        const double vg2 = p->vgamma * p->vgamma, vn2 = p->vnu * p->vnu, vl2 = p->vlambda * p->vlambda;
        for (long i = 0; i < Nx; i++)
            for (long j = 0; j < Ny; j++)
                for (long k = 0; k < Nz; k++)
                    c->pot[idx(c, i, j, k)] = 0.5 * par * (vg2 * c->x2[i] + vn2 * c->y2[j] + vl2 * c->z2[k]);
        /* CN coefficients (gencoef, imag3d.cu:894-931; real3d.cu:991-1035). */
        const double h2[3] = {dx * dx, dy * dy, dz * dz};
        const long n3[3] = {Nx, Ny, Nz};
        for (int a = 0; a < 3; a++) {
            long n = n3[a];
            if (mode == DBEC_IMAG) {
                c->A0[a] = 1. + p->dt / h2[a] / (3. - par);
                c->A0r[a] = 1. - p->dt / h2[a] / (3. - par);
                c->A[a] = -0.5 * p->dt / h2[a] / (3. - par);
                c->calpha[a].assign(n - 1, 0.);
                c->cgamma[a].assign(n - 1, 0.);
                c->calpha[a][n - 2] = 0.;
                c->cgamma[a][n - 2] = -1. / c->A0[a];
                for (long i = n - 2; i > 0; i--) {
                    c->calpha[a][i - 1] = c->A[a] * c->cgamma[a][i];
                    c->cgamma[a][i - 1] = -1. / (c->A0[a] + c->A[a] * c->calpha[a][i - 1]);
                }
            } else {
                c->A0c[a] = {1., p->dt / h2[a] / (3. - par)};
                c->A0rc[a] = cconj(c->A0c[a]);
                c->Ac[a] = {0., -0.5 * p->dt / h2[a] / (3. - par)};
                c->minusAc[a] = {0., -c->Ac[a].im};
                const cplx minus1 = {-1., 0.};
                c->calphac[a].assign(n - 1, {0., 0.});
                c->cgammac[a].assign(n - 1, {0., 0.});
                c->cgammac[a][n - 2] = cdiv(minus1, c->A0c[a]);
                for (long i = n - 2; i > 0; i--) {
                    c->calphac[a][i - 1] = cmul(c->Ac[a], c->cgammac[a][i]);
                    c->cgammac[a][i - 1] = cdiv(minus1, cadd(c->A0c[a], cmul(c->Ac[a], c->calphac[a][i - 1])));
                }
            }
        }
        /* Dipolar kernel (initpotdd, imag3d.cu:945-980). */
        std::vector<double> kx(Nx), ky(Ny), kz(Nz), kx2(Nx), ky2(Ny), kz2(Nz);
        const double dkx = 2. * pi / (Nx * dx), dky = 2. * pi / (Ny * dy), dkz = 2. * pi / (Nz * dz);
        for (long i = 0; i < c->Nx2; i++) kx[i] = i * dkx;
        for (long i = 0; i < c->Nx2; i++) kx[i + c->Nx2] = (i - c->Nx2) * dkx;
        for (long j = 0; j < c->Ny2; j++) ky[j] = j * dky;
        for (long j = 0; j < c->Ny2; j++) ky[j + c->Ny2] = (j - c->Ny2) * dky;
        for (long k = 0; k < c->Nz2; k++) kz[k] = k * dkz;
        for (long k = 0; k < c->Nz2; k++) kz[k + c->Nz2] = (k - c->Nz2) * dkz;
        for (long i = 0; i < Nx; i++) kx2[i] = kx[i] * kx[i];
        for (long j = 0; j < Ny; j++) ky2[j] = ky[j] * ky[j];
        for (long k = 0; k < Nz; k++) kz2[k] = kz[k] * kz[k];
        c->potdd.resize(c->N);
        const double R = p->cutoff;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < Nx; i++)
            for (long j = 0; j < Ny; j++)
                for (long k = 0; k < Nz; k++) {
                    double xk = std::sqrt(kz2[k] + kx2[i] + ky2[j]);
                    double tmp = 1. + 3. * std::cos(xk * R) / (xk * xk * R * R) -
                                 3. * std::sin(xk * R) / (xk * xk * xk * R * R * R);
                    c->potdd[idx(c, i, j, k)] =
                        (4. * pi * (3. * kz2[k] / (kx2[i] + ky2[j] + kz2[k]) - 1.) / 3.) * tmp;
                }
        c->potdd[0] = 0.;
        if (mode == DBEC_IMAG) c->psir.assign(c->N, 0.);
        else c->psic.assign(c->N, {0., 0.});
        c->phi.assign(c->N, 0.);
        c->work.resize(c->N);
        c->fx.init(Nx); c->fy.init(Ny); c->fz.init(Nz);
    } catch (const std::bad_alloc &) {
        delete c;
        return fail(nullptr, DBEC_ENOMEM, "out of host memory");
    }
    *out = c;
    return DBEC_OK;
}

int dbec_destroy(dbec_ctx *ctx) {
    delete ctx;
    return DBEC_OK;
}

int dbec_set_psi(dbec_ctx *c, const double *psi) {
    if (!c || !psi) return fail(c, DBEC_EINVAL, "null argument");
    if (c->mode == DBEC_IMAG) std::memcpy(c->psir.data(), psi, sizeof(double) * c->N);
    else std::memcpy(c->psic.data(), psi, sizeof(cplx) * c->N);
    return DBEC_OK;
}

int dbec_get_psi(dbec_ctx *c, double *psi) {
    if (!c || !psi) return fail(c, DBEC_EINVAL, "null argument");
    if (c->mode == DBEC_IMAG) std::memcpy(psi, c->psir.data(), sizeof(double) * c->N);
    else std::memcpy(psi, c->psic.data(), sizeof(cplx) * c->N);
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
    calcpsidd2_impl(c);
    return DBEC_OK;
}

int dbec_get_psidd2(dbec_ctx *c, double *phi) {
    if (!c || !phi) return fail(c, DBEC_EINVAL, "null argument");
    std::memcpy(phi, c->phi.data(), sizeof(double) * c->N);
    return DBEC_OK;
}

int dbec_calcnu(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    calcnu_impl(c);
    return DBEC_OK;
}

int dbec_calclux(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    sweep_axis(c, 0);
    return DBEC_OK;
}

int dbec_calcluy(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    sweep_axis(c, 1);
    return DBEC_OK;
}

int dbec_calcluz(dbec_ctx *c) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    sweep_axis(c, 2);
    return DBEC_OK;
}

int dbec_calcnorm(dbec_ctx *c, double *norm, int rescale) {
    if (!c) return fail(c, DBEC_EINVAL, "null context");
    double nv = calcnorm_impl(c, rescale);
    if (norm) *norm = nv;
    return DBEC_OK;
}

int dbec_calcmuen(dbec_ctx *c, double *mu, double *en) {
    if (!c || !mu || !en) return fail(c, DBEC_EINVAL, "null argument");
    calcmuen_impl(c, mu, en);
    return DBEC_OK;
}

int dbec_calcmuen_phase(dbec_ctx *c, double *mu, double *en) {
    if (!c || !mu || !en) return fail(c, DBEC_EINVAL, "null argument");
    calcmuen_impl(c, mu, en, true);
    return DBEC_OK;
}

int dbec_calcrms(dbec_ctx *c, double rms[4]) {
    if (!c || !rms) return fail(c, DBEC_EINVAL, "null argument");
    calcrms_impl(c, rms);
    return DBEC_OK;
}

int dbec_psi_center(dbec_ctx *c, double out[2]) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    long n = idx(c, c->Nx2, c->Ny2, c->Nz2);
    if (c->mode == DBEC_IMAG) { out[0] = c->psir[n]; out[1] = 0.; }
    else { out[0] = c->psic[n].re; out[1] = c->psic[n].im; }
    return DBEC_OK;
}

int dbec_step(dbec_ctx *c, long nsteps, double g_stp, double gd_stp, double *norm) {
    if (!c || nsteps < 0) return fail(c, DBEC_EINVAL, "bad argument");
    double nv = 0.;
    for (long s = 0; s < nsteps; s++) {
        if (g_stp != 0. || gd_stp != 0.) {
            c->g += g_stp;
            c->gd += gd_stp;
        }
        calcpsidd2_impl(c);
        calcnu_impl(c);
        sweep_axis(c, 0);
        sweep_axis(c, 1);
        sweep_axis(c, 2);
        if (c->mode == DBEC_IMAG) nv = calcnorm_impl(c, 1);
    }
    if (norm) *norm = nv;
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

/* Output writers' quantities (imag3d.cu:1944-2399). */
int dbec_density(dbec_ctx *c, int kind, long ox, long oy, long oz, double *out) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    if (dbec_density_count(c, kind, ox, oy, oz) < 0) return fail(c, DBEC_EINVAL, "bad density request");
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    std::vector<double> p2(c->N);
    for (long n = 0; n < c->N; n++) p2[n] = psi2_at(c, n);
    std::vector<double> lx(Nx), ly(Ny), lz(Nz);
    long o = 0;
    switch (kind) {
    case DBEC_DEN_X:
        for (long i = 0; i < Nx; i += ox) {
            for (long j = 0; j < Ny; j++) ly[j] = simpint(c->p.dz, &p2[idx(c, i, j, 0)], Nz);
            out[o++] = simpint(c->p.dy, ly.data(), Ny);
        }
        break;
    case DBEC_DEN_Y:
        for (long j = 0; j < Ny; j += oy) {
            for (long i = 0; i < Nx; i++) lx[i] = simpint(c->p.dz, &p2[idx(c, i, j, 0)], Nz);
            out[o++] = simpint(c->p.dx, lx.data(), Nx);
        }
        break;
    case DBEC_DEN_Z:
        for (long k = 0; k < Nz; k += oz) {
            for (long i = 0; i < Nx; i++) {
                for (long j = 0; j < Ny; j++) ly[j] = p2[idx(c, i, j, k)];
                lx[i] = simpint(c->p.dy, ly.data(), Ny);
            }
            out[o++] = simpint(c->p.dx, lx.data(), Nx);
        }
        break;
    case DBEC_DEN_XY:
        for (long i = 0; i < Nx; i += ox)
            for (long j = 0; j < Ny; j += oy) out[o++] = simpint(c->p.dz, &p2[idx(c, i, j, 0)], Nz);
        break;
    case DBEC_DEN_XZ:
        for (long i = 0; i < Nx; i += ox)
            for (long k = 0; k < Nz; k += oz) {
                for (long j = 0; j < Ny; j++) ly[j] = p2[idx(c, i, j, k)];
                out[o++] = simpint(c->p.dy, ly.data(), Ny);
            }
        break;
    case DBEC_DEN_YZ:
        for (long j = 0; j < Ny; j += oy)
            for (long k = 0; k < Nz; k += oz) {
                for (long i = 0; i < Nx; i++) lx[i] = p2[idx(c, i, j, k)];
                out[o++] = simpint(c->p.dx, lx.data(), Nx);
            }
        break;
    case DBEC_CUT_XY0:
        for (long i = 0; i < Nx; i += ox)
            for (long j = 0; j < Ny; j += oy) out[o++] = p2[idx(c, i, j, c->Nz2)];
        break;
    case DBEC_CUT_X0Z:
        for (long i = 0; i < Nx; i += ox)
            for (long k = 0; k < Nz; k += oz) out[o++] = p2[idx(c, i, c->Ny2, k)];
        break;
    case DBEC_CUT_0YZ:
        for (long j = 0; j < Ny; j += oy)
            for (long k = 0; k < Nz; k += oz) out[o++] = p2[idx(c, c->Nx2, j, k)];
        break;
    case DBEC_DEN_XYZ:
        for (long i = 0; i < Nx; i += ox)
            for (long j = 0; j < Ny; j += oy)
                for (long k = 0; k < Nz; k += oz) out[o++] = p2[idx(c, i, j, k)];
        break;
    }
    return DBEC_OK;
}

int dbec_get_table(dbec_ctx *c, int which, int axis, double *out) {
    if (!c || !out) return fail(c, DBEC_EINVAL, "null argument");
    const long Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    switch (which) {
    case DBEC_TABLE_POT:
        std::memcpy(out, c->pot.data(), sizeof(double) * c->N);
        return DBEC_OK;
    case DBEC_TABLE_POTDD: {
        long o = 0;
        for (long i = 0; i < Nx; i++)
            for (long j = 0; j < Ny; j++)
                for (long k = 0; k <= c->Nz2; k++) out[o++] = c->potdd[idx(c, i, j, k)];
        return DBEC_OK;
    }
    case DBEC_TABLE_CALPHA:
    case DBEC_TABLE_CGAMMA: {
        if (axis < 0 || axis > 2) return fail(c, DBEC_EINVAL, "bad axis");
        long n = axis == 0 ? Nx : axis == 1 ? Ny : Nz;
        if (c->mode == DBEC_IMAG) {
            const std::vector<double> &v = (which == DBEC_TABLE_CALPHA) ? c->calpha[axis] : c->cgamma[axis];
            std::memcpy(out, v.data(), sizeof(double) * (n - 1));
        } else {
            const std::vector<cplx> &v = (which == DBEC_TABLE_CALPHA) ? c->calphac[axis] : c->cgammac[axis];
            std::memcpy(out, v.data(), sizeof(cplx) * (n - 1));
        }
        return DBEC_OK;
    }
    default:
        return fail(c, DBEC_EINVAL, "bad table");
    }
}

/* The allocation-free variant computes the same extracts; the oracle holds
 * no device memory, so it is dbec_density itself. */
int dbec_density_noalloc(dbec_ctx *c, int kind, long ox, long oy, long oz, double *out) {
    return dbec_density(c, kind, ox, oy, oz, out);
}

int dbec_synchronize(dbec_ctx *c) { return c ? DBEC_OK : DBEC_EINVAL; }

int dbec_tuning(const dbec_ctx *c, char *buf, long len) {
    if (!c || !buf || len < 1) return DBEC_EINVAL;
    std::snprintf(buf, (size_t)len, "oracle: no launch choices");
    return DBEC_OK;
}

int dbec_profile_step(dbec_ctx *c, int reps, dbec_kernel_stat *stats, int max, int *n) {
    (void)reps; (void)stats; (void)max;
    if (n) *n = 0;
    return fail(c, DBEC_EUNSUPPORTED, "the oracle has no kernels to profile");
}

}  /* extern "C" */
