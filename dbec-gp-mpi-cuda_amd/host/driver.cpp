/*
 * driver.cpp - the imag3d / real3d programs: command line, parameter file,
 * phase loops, text outputs and density files, driving a solver through the
 * C ABI of include/dbec.h.  Compiled once per program (DBEC_PROGRAM_REAL
 * selects real3d) and linked against the HIP library.
 *
 * Drop-in for src/imag3d-mpicuda/imag3d-mpicuda.cu:29-645 (alias imag3d.cu)
 * and src/real3d-mpicuda/real3d-mpicuda.cu:29-670 (alias real3d.cu) at one
 * process: same argv rules, messages, exit codes, keys, per-step sequence and
 * output bytes; the header reports MPI_NUM_PROCS = 1.
 *
 * Extension keys (absent by default, so default runs behave like the
 * reference):
 *   PRECISION = FP64 | MIXED      arithmetic of the bulk operators
 *   POLISH    = K                 after a MIXED imaginary-time NRUN, K more
 *                                 steps in FP64 before the NRUN report
 *   OBSSTP    = K                 observables every K steps of NPAS and NRUN
 *                                 (outside the timed region; psi unchanged)
 *   OBSOUT    = name              write them to name.txt (needs OBSSTP)
 *   MUTOL     = eps               imaginary time: end NPAS / NRUN early once
 *                                 |mu - mu_prev| <= eps |mu| between two
 *                                 OBSSTP points (needs OBSSTP)
 */
#include "dbec.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

const double BOHR_RADIUS = 5.2917720859e-11; /* imag3d.cuh:46 */

/* Parameter file (src/utils/cfg.c:34-63): one "KEY = VALUE" per line, read
 * with the same sscanf pattern; the first occurrence of a key wins; at most
 * 256 keys. */
struct Cfg {
    std::vector<std::string> keys, vals;
    int init(const char *path) {
        FILE *f = std::fopen(path, "r");
        if (!f) return 0;
        char buf[256], k[256], v[256];
        while (std::fgets(buf, sizeof buf, f) != nullptr) {
            if (keys.size() >= 256) break;
            if (std::sscanf(buf, "%255s = %255s", k, v) == 2) {
                keys.emplace_back(k);
                vals.emplace_back(v);
            }
        }
        std::fclose(f);
        return (int)keys.size();
    }
    const char *read(const char *key) const {
        for (size_t i = 0; i < keys.size(); i++)
            if (keys[i] == key) return vals[i].c_str();
        return nullptr;
    }
};

struct Params {
    int opt = 0;
    long Na = 0;
    double g0 = 0., gd0 = 0., aho = 0., as = 0., add = 0.;
    double gpar = 0., gdpar = 0.;
    long Nx = 0, Ny = 0, Nz = 0;
    double dx = 0., dy = 0., dz = 0., dt = 0.;
    double vgamma = 0., vnu = 0., vlambda = 0.;
    long Nstp = 0, Npas = 0, Nrun = 0;
    double cutoff = 0.;
    const char *output = nullptr, *rmsout = nullptr, *initout = nullptr;
    const char *Nstpout = nullptr, *Npasout = nullptr, *Nrunout = nullptr, *dynaout = nullptr;
    long outstpx = 0, outstpy = 0, outstpz = 0, outstpt = 0;
    int device = -1;
    int potmem = 2;
    /* extensions */
    dbec_prec prec = DBEC_FP64;
    long polish = 0;
    const char *obsout = nullptr;
    long obsstp = 0;
    double mutol = 0.;
};

[[noreturn]] void die(const char *msg, bool to_stdout = false) {
    std::fputs(msg, to_stdout ? stdout : stderr);
    std::exit(EXIT_FAILURE);
}

/* readpar (imag3d.cu:650-829; real3d.cu:675-883), same messages and order. */
void readpar(const Cfg &cfg, Params &P, bool real) {
    const char *t;
    if ((t = cfg.read("OPTION")) == nullptr) die("OPTION is not defined in the configuration file\n");
    P.opt = (int)std::atol(t);
    if ((t = cfg.read("NATOMS")) == nullptr) die("NATOMS is not defined in the configuration file.\n");
    P.Na = std::atol(t);
    const double pi = 3.14159265358979;
    if ((t = cfg.read("G0")) == nullptr) {
        if ((t = cfg.read("AHO")) == nullptr) die("AHO is not defined in the configuration file.\n");
        P.aho = std::atof(t);
        if ((t = cfg.read("AS")) == nullptr) die("AS is not defined in the configuration file.\n");
        P.as = std::atof(t);
        P.g0 = 4. * pi * P.as * P.Na * BOHR_RADIUS / P.aho;
    } else {
        P.g0 = std::atof(t);
    }
    if (real) {
        if ((t = cfg.read("GPAR")) == nullptr) die("GPAR is not defined in the configuration file.\n");
        P.gpar = std::atof(t);
    }
    if ((t = cfg.read("GDD0")) == nullptr) {
        if ((t = cfg.read("AHO")) == nullptr) die("AHO is not defined in the configuration file.\n");
        P.aho = std::atof(t);
        if ((t = cfg.read("ADD")) == nullptr) die("ADD is not defined in the configuration file.\n");
        P.add = std::atof(t);
        P.gd0 = 3. * P.add * P.Na * BOHR_RADIUS / P.aho;
    } else {
        P.gd0 = std::atof(t);
    }
    if (real) {
        if ((t = cfg.read("GDDPAR")) == nullptr) die("GDDPAR is not defined in the configuration file.\n");
        P.gdpar = std::atof(t);
    }
    if ((t = cfg.read("NX")) == nullptr) die("NX is not defined in the configuration file.\n");
    P.Nx = std::atol(t);
    if ((t = cfg.read("NY")) == nullptr) die("NY is not defined in the configuration file.\n");
    P.Ny = std::atol(t);
    if ((t = cfg.read("NZ")) == nullptr) die("Nz is not defined in the configuration file.\n");
    P.Nz = std::atol(t);
    if ((t = cfg.read("DX")) == nullptr) die("DX is not defined in the configuration file.\n");
    P.dx = std::atof(t);
    if ((t = cfg.read("DY")) == nullptr) die("DY is not defined in the configuration file.\n");
    P.dy = std::atof(t);
    if ((t = cfg.read("DZ")) == nullptr) die("DZ is not defined in the configuration file.\n");
    P.dz = std::atof(t);
    if ((t = cfg.read("DT")) == nullptr) die("DT is not defined in the configuration file.\n");
    P.dt = std::atof(t);
    if ((t = cfg.read("GAMMA")) == nullptr) die("GAMMA is not defined in the configuration file.\n");
    P.vgamma = std::atof(t);
    if ((t = cfg.read("NU")) == nullptr) die("NU is not defined in the configuration file.\n");
    P.vnu = std::atof(t);
    if ((t = cfg.read("LAMBDA")) == nullptr) die("LAMBDA is not defined in the configuration file.\n");
    P.vlambda = std::atof(t);
    if ((t = cfg.read("NSTP")) == nullptr) die("NSTP is not defined in the configuration file.\n");
    P.Nstp = std::atol(t);
    if ((t = cfg.read("NPAS")) == nullptr) die("NPAS is not defined in the configuration file.\n");
    P.Npas = std::atol(t);
    if ((t = cfg.read("NRUN")) == nullptr) die("NRUN is not defined in the configuration file.\n");
    P.Nrun = std::atol(t);
    if ((t = cfg.read("CUTOFF")) == nullptr) die("CUTOFF is not defined in the configuration file.\n");
    P.cutoff = std::atof(t);
    P.output = cfg.read("OUTPUT");
    P.rmsout = cfg.read("RMSOUT");
    P.initout = cfg.read("INITOUT");
    P.Nstpout = cfg.read("NSTPOUT");
    P.Npasout = cfg.read("NPASOUT");
    P.Nrunout = cfg.read("NRUNOUT");
    if (real) P.dynaout = cfg.read("DYNAOUT");
    if (P.initout || P.Nstpout || P.Npasout || P.Nrunout) {
        if ((t = cfg.read("OUTSTPX")) == nullptr) die("OUTSTPX is not defined in the configuration file.\n");
        P.outstpx = std::atol(t);
        if ((t = cfg.read("OUTSTPY")) == nullptr) die("OUTSTPY is not defined in the configuration file.\n");
        P.outstpy = std::atol(t);
        if ((t = cfg.read("OUTSTPZ")) == nullptr) die("OUTSTPZ is not defined in the configuration file.\n");
        P.outstpz = std::atol(t);
    }
    if (real && (P.rmsout || P.dynaout)) {
        if ((t = cfg.read("OUTSTPT")) == nullptr) die("OUTSTPT is not defined in the configuration file.\n");
        P.outstpt = std::atol(t);
    }
    if ((t = cfg.read("DEVICE")) == nullptr) P.device = -1;
    else P.device = std::atoi(t);
    if ((t = cfg.read("POTMEM")) == nullptr) {
        P.potmem = 2;
    } else {
        P.potmem = std::atoi(t);
        if (P.potmem < 0 || P.potmem > 2) die("POTMEM can be 0, 1 or 2.\n");
    }
    /* Extension keys. */
    if ((t = cfg.read("PRECISION")) != nullptr) {
        if (!std::strcmp(t, "FP64")) P.prec = DBEC_FP64;
        else if (!std::strcmp(t, "MIXED")) P.prec = DBEC_MIXED;
        else die("PRECISION can be FP64 or MIXED.\n");
    }
    if ((t = cfg.read("POLISH")) != nullptr) {
        P.polish = std::atol(t);
        if (P.polish < 0) die("POLISH must not be negative.\n");
        if (P.polish > 0 && (real || P.prec != DBEC_MIXED))
            die("POLISH applies to an imaginary-time run with PRECISION = MIXED.\n");
    }
    P.obsout = cfg.read("OBSOUT");
    if ((t = cfg.read("MUTOL")) != nullptr) {
        P.mutol = std::atof(t);
        if (real) die("MUTOL applies to imaginary time only.\n");
        if (!(P.mutol > 0.)) die("MUTOL must be positive.\n");
    }
    if (P.obsout || P.mutol > 0.) {
        if ((t = cfg.read("OBSSTP")) == nullptr) die("OBSSTP is not defined in the configuration file.\n");
        P.obsstp = std::atol(t);
        if (P.obsstp < 1) die("OBSSTP must be positive.\n");
    }
    if (real && P.dynaout && P.outstpt < 1) die("OUTSTPT must be positive.\n");
}

void check(int rc, dbec_ctx *ctx) {
    if (rc != DBEC_OK) {
        std::fprintf(stderr, "dbec error %d: %s\n", rc, dbec_last_error(ctx));
        std::exit(EXIT_FAILURE);
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Grid {
    std::vector<double> x, y, z;
};

/* Density files of the writers (imag3d.cu:1944-2399): raw little-endian
 * float64; 1-D files hold (coordinate, value) pairs, 2-D files (c1, c2, value)
 * triples, the 3-D file bare values. */
void write_density(dbec_ctx *ctx, const Params &P, const Grid &G, const char *pfx, bool npas_quirk) {
    char fn[512];
    const long ox = P.outstpx, oy = P.outstpy, oz = P.outstpz;
    std::vector<double> v;
    auto fetch = [&](int kind) {
        long n = dbec_density_count(ctx, kind, ox, oy, oz);
        if (n < 0) {
            std::fprintf(stderr, "bad OUTSTP values\n");
            std::exit(EXIT_FAILURE);
        }
        v.assign(n, 0.);
        check(dbec_density(ctx, kind, ox, oy, oz, v.data()), ctx);
    };
    auto open = [&](const char *suffix) {
        std::snprintf(fn, sizeof fn, "%s%s", pfx, suffix);
        FILE *f = std::fopen(fn, "wb");
        if (!f) {
            std::fprintf(stderr, "cannot open %s\n", fn);
            std::exit(EXIT_FAILURE);
        }
        return f;
    };
    auto put = [](FILE *f, std::initializer_list<double> vals) {
        for (double d : vals) std::fwrite(&d, sizeof d, 1, f);
    };
    FILE *f;
    /* <PFX>.bin: |psi|^2 [x][y][z] (outdenxyz 2355-2381). */
    fetch(DBEC_DEN_XYZ);
    f = open(".bin");
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
    /* 1-D (outdenx/outdeny/outdenz). */
    fetch(DBEC_DEN_X);
    f = open("1d_x.bin");
    for (long i = 0, o = 0; i < P.Nx; i += ox) put(f, {G.x[i], v[o++]});
    std::fclose(f);
    fetch(DBEC_DEN_Y);
    f = open("1d_y.bin");
    for (long j = 0, o = 0; j < P.Ny; j += oy) put(f, {G.y[j], v[o++]});
    std::fclose(f);
    fetch(DBEC_DEN_Z);
    /* imag3d names the NPAS 1d_z file .txt although it is binary (imag3d.cu:431). */
    f = open(npas_quirk ? "1d_z.txt" : "1d_z.bin");
    for (long k = 0, o = 0; k < P.Nz; k += oz) put(f, {G.z[k], v[o++]});
    std::fclose(f);
    /* 2-D integrals (outdenxy/xz/yz). */
    fetch(DBEC_DEN_XY);
    f = open("2d_xy.bin");
    for (long i = 0, o = 0; i < P.Nx; i += ox)
        for (long j = 0; j < P.Ny; j += oy) put(f, {G.x[i], G.y[j], v[o++]});
    std::fclose(f);
    fetch(DBEC_DEN_XZ);
    f = open("2d_xz.bin");
    for (long i = 0, o = 0; i < P.Nx; i += ox)
        for (long k = 0; k < P.Nz; k += oz) put(f, {G.x[i], G.z[k], v[o++]});
    std::fclose(f);
    fetch(DBEC_DEN_YZ);
    f = open("2d_yz.bin");
    for (long j = 0, o = 0; j < P.Ny; j += oy)
        for (long k = 0; k < P.Nz; k += oz) put(f, {G.y[j], G.z[k], v[o++]});
    std::fclose(f);
    /* 3-D cuts through the centre (outpsi2xy/xz/yz). */
    fetch(DBEC_CUT_XY0);
    f = open("3d_xy0.bin");
    for (long i = 0, o = 0; i < P.Nx; i += ox)
        for (long j = 0; j < P.Ny; j += oy) put(f, {G.x[i], G.y[j], v[o++]});
    std::fclose(f);
    fetch(DBEC_CUT_X0Z);
    f = open("3d_x0z.bin");
    for (long i = 0, o = 0; i < P.Nx; i += ox)
        for (long k = 0; k < P.Nz; k += oz) put(f, {G.x[i], G.z[k], v[o++]});
    std::fclose(f);
    fetch(DBEC_CUT_0YZ);
    f = open("3d_0yz.bin");
    for (long j = 0, o = 0; j < P.Ny; j += oy)
        for (long k = 0; k < P.Nz; k += oz) put(f, {G.y[j], G.z[k], v[o++]});
    std::fclose(f);
}

struct Obs {
    double norm = 0., mu = 0., en = 0., rms[4] = {0, 0, 0, 0}, c2 = 0.;
};

/* The observables printed at a phase end: norm (pre-rescale value of the last
 * step in imaginary time, a measurement in real time), mu, E, rms and
 * |psi(Nx/2,Ny/2,Nz/2)|^2. */
void observe(dbec_ctx *ctx, Obs &o, bool measure_norm) {
    if (measure_norm) check(dbec_calcnorm(ctx, &o.norm, 0), ctx);
    check(dbec_calcmuen(ctx, &o.mu, &o.en), ctx);
    check(dbec_calcrms(ctx, o.rms), ctx);
    double c[2];
    check(dbec_psi_center(ctx, c), ctx);
    double a = std::hypot(c[0], c[1]);
    o.c2 = a * a;
}

/* Extension: observables every OBSSTP steps.  Real time adds the energy
 * with the phase gradient in the kinetic term (dbec_calcmuen_phase), the one
 * the dynamics conserves; E/par is the reference's, of |psi|. */
struct ObsLog {
    FILE *f = nullptr;
    bool real = false;
    void open(const Params &P, bool real_) {
        real = real_;
        if (!P.obsout) return;
        char fn[512];
        std::snprintf(fn, sizeof fn, "%s.txt", P.obsout);
        f = std::fopen(fn, "w");
        if (!f) die("cannot open OBSOUT file\n");
        std::fprintf(f, "# %s: step  t  norm  mu/par  E/par  <r>  <x>  <y>  <z>%s\n",
                     real ? "real time" : "imaginary time", real ? "  Ephase/par" : "");
    }
    /* Measures the observables (norm measured, no rescale) and logs them;
     * returns mu. */
    double line(dbec_ctx *ctx, long step, double t, double par) {
        Obs o;
        observe(ctx, o, true);
        if (f) {
            std::fprintf(f, "%ld %.10e %.16e %.16e %.16e %.16e %.16e %.16e %.16e", step, t, o.norm, o.mu / par,
                         o.en / par, o.rms[0], o.rms[1], o.rms[2], o.rms[3]);
            if (real) {
                double mup, enp;
                check(dbec_calcmuen_phase(ctx, &mup, &enp), ctx);
                std::fprintf(f, " %.16e", enp / par);
            }
            std::fprintf(f, "\n");
            std::fflush(f);
        }
        return o.mu;
    }
    ~ObsLog() {
        if (f) std::fclose(f);
    }
};

int run(int argc, char **argv, bool real) {
    double t_total0 = now_s();
    double iter_time = 0.;
    if (argc != 3 || std::strcmp(argv[1], "-i") != 0) {
        std::fprintf(stderr, "Usage: %s -i <input parameter file> \n", argv[0]);
        return EXIT_FAILURE;
    }
    Cfg cfg;
    if (!cfg.init(argv[2])) {
        std::fprintf(stderr, "Wrong input parameter file.\n");
        return EXIT_FAILURE;
    }
    Params P;
    readpar(cfg, P, real);

    const double pi = 3.14159265358979;
    const double par = (P.opt == 2) ? 2. : 1.;
    const long Nx = P.Nx, Ny = P.Ny, Nz = P.Nz;
    const long Nx2 = Nx / 2, Ny2 = Ny / 2, Nz2 = Nz / 2;

    char filename[512];
    FILE *out = stdout;
    if (P.output) {
        std::snprintf(filename, sizeof filename, "%s.txt", P.output);
        out = std::fopen(filename, "w");
        if (!out) die("cannot open OUTPUT file\n");
    }
    FILE *filerms = nullptr;
    if (P.rmsout) {
        std::snprintf(filename, sizeof filename, "%s.txt", P.rmsout);
        filerms = std::fopen(filename, "w");
        if (!filerms) die("cannot open RMSOUT file\n");
    }
    FILE *dyna = nullptr;
    if (real && P.dynaout) {
        std::snprintf(filename, sizeof filename, "%s.txt", P.dynaout);
        dyna = std::fopen(filename, "w");
        if (!dyna) die("cannot open DYNAOUT file\n");
    }

    /* Header (imag3d.cu:204-216; real3d.cu:213-227). */
    std::fprintf(out, " %s time propagation 3D,   OPTION = %d, MPI_NUM_PROCS = %d\n\n",
                 real ? "Real" : "Imaginary", P.opt, 1);
    std::fprintf(out, "  Number of Atoms N = %li, Unit of length AHO = %.8f m\n", P.Na, P.aho);
    std::fprintf(out, "  Scattering length a = %.2f*a0, Dipolar ADD = %.2f*a0\n", P.as, P.add);
    std::fprintf(out, "  Nonlinearity G_3D = %.4f, Strength of DDI GD_3D = %.5f\n", P.g0, P.gd0);
    std::fprintf(out, "  Parameters of trap: GAMMA = %.2f, NU = %.2f, LAMBDA = %.2f\n\n", P.vgamma, P.vnu,
                 P.vlambda);
    std::fprintf(out, " # Space Stp: NX = %li, NY = %li, NZ = %li\n", Nx, Ny, Nz);
    std::fprintf(out, "  Space Step: DX = %.6f, DY = %.6f, DZ = %.6f\n", P.dx, P.dy, P.dz);
    std::fprintf(out, " # Time Stp : NSTP = %li, NPAS = %li, NRUN = %li\n", P.Nstp, P.Npas, P.Nrun);
    std::fprintf(out, "   Time Step:   DT = %.6f\n", P.dt);
    std::fprintf(out, "   Dipolar Cut off:   R = %.3f\n\n", P.cutoff);
    if (real) std::fprintf(out, " * Change for dynamics: GPAR = %.3f, GDPAR = %.3f *\n\n", P.gpar, P.gdpar);
    std::fprintf(out, "                  --------------------------------------------------------\n");
    std::fprintf(out, "                    Norm      Chem        Ener/N     <r>     |Psi(0,0,0)|^2\n");
    std::fprintf(out, "                  --------------------------------------------------------\n");
    std::fflush(out);

    if (Nx < 8 || Ny < 8 || Nz < 8 || Nx % 2 || Ny % 2 || Nz % 2)
        die("NX, NY and NZ must be even and at least 8.\n");

    /* Grid (initpsi imag3d.cu:842-855). */
    Grid G;
    G.x.resize(Nx);
    G.y.resize(Ny);
    G.z.resize(Nz);
    std::vector<double> x2(Nx), y2(Ny), z2(Nz);
    for (long i = 0; i < Nx; i++) { G.x[i] = (i - Nx2) * P.dx; x2[i] = G.x[i] * G.x[i]; }
    for (long j = 0; j < Ny; j++) { G.y[j] = (j - Ny2) * P.dy; y2[j] = G.y[j] * G.y[j]; }
    for (long k = 0; k < Nz; k++) { G.z[k] = (k - Nz2) * P.dz; z2[k] = G.z[k] * G.z[k]; }

    /* Initial wave function (imag3d.cu:835-867; real3d.cu:890-950). */
    const long N = Nx * Ny * Nz;
    std::vector<double> psi0(real ? 2 * N : N, 0.);
    if (!real || P.Nstp != 0) {
        const double cpsi = std::sqrt(pi * std::sqrt(pi / (P.vgamma * P.vnu * P.vlambda)));
        for (long i = 0; i < Nx; i++)
            for (long j = 0; j < Ny; j++)
                for (long k = 0; k < Nz; k++) {
                    double tmp = std::exp(-0.5 * (P.vgamma * x2[i] + P.vnu * y2[j] + P.vlambda * z2[k]));
                    long n = (i * Ny + j) * Nz + k;
                    if (real) { psi0[2 * n] = tmp / cpsi; psi0[2 * n + 1] = 0.; }
                    else psi0[n] = tmp / cpsi;
                }
    } else {
        FILE *f = std::fopen("imag3d-den.bin", "rb");
        if (!f) die("Run the program using the input file to read. e.g.: imag3d-den.bin\n", true);
        std::vector<double> row(Nz);
        for (long r = 0; r < Nx * Ny; r++) {
            size_t got = std::fread(row.data(), sizeof(double), Nz, f);
            for (long k = 0; k < Nz; k++) {
                /* a short file leaves the remainder zero, like a short MPI read */
                double d = (k < (long)got) ? row[k] : 0.;
                psi0[2 * (r * Nz + k)] = std::sqrt(d);
                psi0[2 * (r * Nz + k) + 1] = 0.;
            }
        }
        std::fclose(f);
    }

    dbec_params dp;
    std::memset(&dp, 0, sizeof dp);
    dp.nx = Nx; dp.ny = Ny; dp.nz = Nz;
    dp.dx = P.dx; dp.dy = P.dy; dp.dz = P.dz; dp.dt = P.dt;
    dp.vgamma = P.vgamma; dp.vnu = P.vnu; dp.vlambda = P.vlambda;
    dp.cutoff = P.cutoff; dp.par = par; dp.pi = pi; dp.device = P.device;
    dbec_ctx *ctx = nullptr;
    check(dbec_create(&dp, real ? DBEC_REAL : DBEC_IMAG, P.prec, nullptr, &ctx), nullptr);
    check(dbec_set_psi(ctx, psi0.data()), ctx);
    psi0.clear();
    psi0.shrink_to_fit();

    double g = 0., gd = 0.;
    if (real) {
        /* real3d sets the couplings before the initial observables (real3d.cu:229-235). */
        if (P.Nstp == 0) { g = par * P.g0; gd = par * P.gd0; }
    }
    check(dbec_set_coupling(ctx, g, gd), ctx);

    Obs o;
    /* Initial observables: imag3d normalises psi0 first (imag3d.cu:239) with
     * g = gd = 0; real3d only measures the norm (real3d.cu:257). */
    check(dbec_calcnorm(ctx, &o.norm, real ? 0 : 1), ctx);
    observe(ctx, o, false);
    double last_norm = o.norm;
    std::fprintf(out, "Initial : %15.4f %11.5f %11.5f %10.5f %10.5f\n", o.norm, o.mu / par, o.en / par, o.rms[0],
                 o.c2);
    std::fflush(out);

    auto rms_header = [&]() {
        std::fprintf(filerms, " %s time propagation 3D,   OPTION = %d\n\n", real ? "Real" : "Imaginary", P.opt);
        std::fprintf(filerms, "                  --------------------------------------------------------\n");
        std::fprintf(filerms, "Values of rms size:     <r>          <x>          <y>          <z>\n");
        std::fprintf(filerms, "                  --------------------------------------------------------\n");
        std::fprintf(filerms, "           Initial:%12.5f %12.5f %12.5f %12.5f\n", o.rms[0], o.rms[1], o.rms[2],
                     o.rms[3]);
        std::fflush(filerms);
    };
    if (!real) {
        if (P.initout) write_density(ctx, P, G, P.initout, false);
        if (P.rmsout) rms_header();
    } else {
        if (P.rmsout) rms_header();
        if (P.initout) write_density(ctx, P, G, P.initout, false);
    }

    ObsLog olog;
    olog.open(P, real);

    /* Timed loop of `n` steps; the time brackets the steps only. */
    auto steps = [&](long n, double g_stp, double gd_stp) {
        if (n <= 0) return;
        check(dbec_synchronize(ctx), ctx);
        double t0 = now_s();
        check(dbec_step(ctx, n, g_stp, gd_stp, real ? nullptr : &last_norm), ctx);
        check(dbec_synchronize(ctx), ctx);
        iter_time += now_s() - t0;
    };
    /* A phase of `n` steps with observables every OBSSTP steps (extension) and,
     * in real time, DYNAOUT rms every OUTSTPT steps inside the timed loop
     * (real3d.cu:419-423, 504-507). */
    const bool cadence = P.obsstp > 0;
    auto phase = [&](long n, long label_base, bool dyn, const char *name) {
        long done = 0;
        double mu_prev = 0.;
        bool have_prev = false;
        while (done < n) {
            long chunk = n - done;
            if (cadence) chunk = std::min(chunk, P.obsstp - (done % P.obsstp));
            if (dyn) {
                /* next step count d with (d - 1) % OUTSTPT == 0 */
                long nxt = ((done + P.outstpt - 1) / P.outstpt) * P.outstpt + 1;
                chunk = std::min(chunk, nxt - done);
            }
            steps(chunk, 0., 0.);
            done += chunk;
            if (dyn && ((done - 1) % P.outstpt == 0)) {
                long cnti = done - 1;
                double rms[4];
                double t0 = now_s();
                check(dbec_calcrms(ctx, rms), ctx);
                iter_time += now_s() - t0;
                double tl = (label_base == 0) ? cnti * P.dt * par : (cnti * P.dt + P.dt * label_base) * par;
                std::fprintf(dyna, "%8le   %8le   %8le   %8le   %8le\n", tl, rms[0], rms[1], rms[2], rms[3]);
                std::fflush(dyna);
            }
            if (cadence && done % P.obsstp == 0) {
                const double mu = olog.line(ctx, done + label_base, (done + label_base) * P.dt * par, par);
                if (P.mutol > 0. && have_prev && std::fabs(mu - mu_prev) <= P.mutol * std::fabs(mu)) {
                    std::fprintf(out, " * %s converged after %ld steps: |dmu/mu| = %.3e *\n", name, done,
                                 std::fabs(mu - mu_prev) / std::fabs(mu));
                    std::fflush(out);
                    break;
                }
                mu_prev = mu;
                have_prev = true;
            }
        }
        return done;
    };

    const char *phase_line[3] = {"After NSTP iter.:", "After NPAS iter.:", "After NRUN iter.:"};
    auto report = [&](int ph, const char *pfx) {
        if (!real) {
            o.norm = last_norm;
            observe(ctx, o, false);
        } else {
            observe(ctx, o, true);
        }
        std::fprintf(out, "%s%8.4f %11.5f %11.5f %10.5f %10.5f\n", phase_line[ph], o.norm, o.mu / par, o.en / par,
                     o.rms[0], o.c2);
        std::fflush(out);
        if (real) {
            if (P.rmsout) {
                std::fprintf(filerms, "  %s%12.5f %12.5f %12.5f %12.5f\n", phase_line[ph], o.rms[0], o.rms[1],
                             o.rms[2], o.rms[3]);
                if (ph == 2)
                    std::fprintf(filerms, "                  --------------------------------------------------------\n");
                std::fflush(filerms);
            }
            if (pfx) write_density(ctx, P, G, pfx, false);
        } else {
            if (pfx) write_density(ctx, P, G, pfx, ph == 1);
            if (P.rmsout) {
                std::fprintf(filerms, "  %s%12.5f %12.5f %12.5f %12.5f\n", phase_line[ph], o.rms[0], o.rms[1],
                             o.rms[2], o.rms[3]);
                if (ph == 2)
                    std::fprintf(filerms, "                  --------------------------------------------------------\n");
                std::fflush(filerms);
            }
        }
    };

    /* NSTP: couplings ramp, g += par*g0/NSTP before every step (imag3d.cu:309-317). */
    if (P.Nstp != 0) {
        double g_stp = par * P.g0 / (double)P.Nstp;
        double gd_stp = par * P.gd0 / (double)P.Nstp;
        check(dbec_set_coupling(ctx, 0., 0.), ctx);
        steps(P.Nstp, g_stp, gd_stp);
        report(0, P.Nstpout);
    } else {
        check(dbec_set_coupling(ctx, par * P.g0, par * P.gd0), ctx);
    }

    /* NPAS. */
    const long npas_done = phase(P.Npas, 0, real && dyna != nullptr, "NPAS");
    report(1, P.Npasout);

    /* real3d quench before NRUN (real3d.cu:493-494). */
    if (real) {
        check(dbec_get_coupling(ctx, &g, &gd), ctx);
        check(dbec_set_coupling(ctx, P.gpar * g, P.gdpar * gd), ctx);
    }

    /* NRUN. */
    /* MUTOL ends NPAS early in imaginary time only, so real time labels
     * with the full NPAS as the reference does */
    phase(P.Nrun, npas_done, real && dyna != nullptr, "NRUN");

    /* Extension: FP64 polish of a mixed-precision ground state.  psi moves
     * to an FP64 solver with the same grid and couplings, which takes the
     * remaining observables and outputs. */
    if (P.polish > 0) {
        check(dbec_get_coupling(ctx, &g, &gd), ctx);
        std::vector<double> host(N);
        check(dbec_get_psi(ctx, host.data()), ctx);
        dbec_ctx *fine = nullptr;
        check(dbec_create(&dp, DBEC_IMAG, DBEC_FP64, nullptr, &fine), nullptr);
        check(dbec_destroy(ctx), nullptr);
        ctx = fine;
        check(dbec_set_psi(ctx, host.data()), ctx);
        check(dbec_set_coupling(ctx, g, gd), ctx);
        steps(P.polish, 0., 0.);
        std::fprintf(out, " * POLISH: %ld FP64 steps *\n", P.polish);
        std::fflush(out);
    }
    report(2, P.Nrunout);

    if (filerms) std::fclose(filerms);
    if (dyna) std::fclose(dyna);
    std::fprintf(out, "                  --------------------------------------------------------\n\n");

    check(dbec_destroy(ctx), nullptr);
    double wall = now_s() - t_total0;
    std::fprintf(out, " Initialization/allocation wall-clock time: %.3f seconds\n", wall - iter_time);
    std::fprintf(out, " Calculation (iterations) wall-clock time: %.3f seconds\n", iter_time);
    if (P.output) std::fclose(out);
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char **argv) {
#ifdef DBEC_PROGRAM_REAL
    return run(argc, argv, true);
#else
    return run(argc, argv, false);
#endif
}
