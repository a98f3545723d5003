/*
 * dbec.h - C ABI of the split-step Crank-Nicolson solver for the 3-D dipolar
 * Gross-Pitaevskii equation (imaginary time and real time).
 *
 * One opaque context per solve.  The context owns every device buffer; the
 * caller owns host memory.  All entry points return an int status (0 = OK,
 * see DBEC_E*), never exit the process, and keep no global state (apart from
 * the fixed device scratch of dbec_density_noalloc), so many contexts coexist
 * in one process.  Operators are stream ordered; entry
 * points that return scalars to the host synchronise the context's stream.
 *
 * Host arrays are row-major [x][y][z] (z fastest).  Imaginary-time contexts
 * hold a real wave function (one double per point); real-time contexts hold a
 * complex one (two doubles per point, interleaved re, im).
 *
 * The entry points replace the global-state operator layer of the reference
 * programs imag3d-mpicuda and real3d-mpicuda; each declaration names the
 * reference function it stands for (file:line).  Reference file aliases:
 *   imag3d.cu = src/imag3d-mpicuda/imag3d-mpicuda.cu
 *   real3d.cu = src/real3d-mpicuda/real3d-mpicuda.cu
 */
#ifndef DBEC_H
#define DBEC_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DBEC_OK           0
#define DBEC_EINVAL       1  /* bad argument or parameter */
#define DBEC_ENOMEM       2  /* device or host allocation failed */
#define DBEC_EDEVICE      3  /* HIP runtime error */
#define DBEC_EUNSUPPORTED 4  /* size or mode not supported by this backend */
#define DBEC_ESTATE       5  /* operation not valid in the current state */

typedef struct dbec_ctx dbec_ctx;

typedef enum { DBEC_IMAG = 0, DBEC_REAL = 1 } dbec_mode;

/* DBEC_FP64: every operator in FP64 / complex128.
 * DBEC_MIXED: psi, the spectrum and the dipolar potential are stored in FP32 /
 * complex64, and the FFT convolution, calcnu and the sweeps compute in FP32.
 * FP64 stays for the norm, mu, E and rms reductions (FP64 accumulators), the
 * renormalisation factor, and the Thomas coefficient tables (calpha, cgamma
 * and the partitioned sweeps' carry multipliers), which live in device memory
 * in FP64 and are rounded to FP32 once as a sweep stages them on chip. */
typedef enum { DBEC_FP64 = 0, DBEC_MIXED = 1 } dbec_prec;

/* Physical and numerical parameters of one solve.  Mirrors the keys read by
 * readpar (imag3d.cu:650-829) that the per-step operators consume. */
typedef struct dbec_params {
    long nx, ny, nz;          /* NX, NY, NZ (each even, >= 8) */
    double dx, dy, dz;        /* DX, DY, DZ */
    double dt;                /* DT */
    double vgamma, vnu, vlambda; /* GAMMA, NU, LAMBDA (trap anisotropy) */
    double cutoff;            /* CUTOFF (dipolar spherical cut-off R) */
    double par;               /* 2 when OPTION == 2, else 1 (imag3d.cu:201-202) */
    double pi;                /* value of pi; 0 selects the reference literal 3.14159265358979 (imag3d.cu:69) */
    int device;               /* HIP device ordinal; -1 = current device (imag3d.cu:83-87) */
} dbec_params;

/* Context life cycle: builds grids, trap (initpot imag3d.cu:872), CN
 * coefficients (gencoef imag3d.cu:894; real3d.cu:991), the dipolar kernel
 * (initpotdd imag3d.cu:945) and device buffers (initdevice imag3d.cu:982).
 * `stream` is a hipStream_t (NULL = the context creates its own). */
int dbec_create(const dbec_params *p, dbec_mode mode, dbec_prec prec,
                void *stream, dbec_ctx **out);
int dbec_destroy(dbec_ctx *ctx);

/* Text of the last error of `ctx` (or of the calling thread when ctx is NULL,
 * e.g. after a failed dbec_create).  Replaces cudaCheck's exit (cudautils.cuh:46-52). */
const char *dbec_last_error(const dbec_ctx *ctx);

/* Wave function in / out (host [x][y][z]; real or interleaved complex).
 * Replaces the cudaMemcpy3D of psi (imag3d.cu:226-236). */
int dbec_set_psi(dbec_ctx *ctx, const double *psi);
int dbec_get_psi(dbec_ctx *ctx, double *psi);

/* Contact and dipolar couplings g, gd as the kernels use them (par*G0, par*GDD0,
 * the NSTP ramp and the GPAR/GDDPAR quench; imag3d.cu:309-317, 391-392;
 * real3d.cu:493-494). */
int dbec_set_coupling(dbec_ctx *ctx, double g, double gd);
int dbec_get_coupling(const dbec_ctx *ctx, double *g, double *gd);

/* One export per reference operator. */
int dbec_calcpsidd2(dbec_ctx *ctx);          /* calcpsidd2_mem imag3d.cu:1460; real3d.cu:1554 */
int dbec_get_psidd2(dbec_ctx *ctx, double *phi); /* copy of the unnormalised dipolar potential [x][y][z] */
int dbec_calcnu(dbec_ctx *ctx);              /* calcnu imag3d.cu:1752; real3d.cu:1849 (uses the last calcpsidd2) */
int dbec_calclux(dbec_ctx *ctx);             /* calclux imag3d.cu:1790; real3d.cu:1890 */
int dbec_calcluy(dbec_ctx *ctx);             /* calcluy imag3d.cu:1851; real3d.cu:1951 */
int dbec_calcluz(dbec_ctx *ctx);             /* calcluz imag3d.cu:1907; real3d.cu:2007 */
/* calcnorm imag3d.cu:1194 (rescale=1: psi *= 1/norm) / real3d.cu:1304 (rescale=0).
 * *norm receives the norm before rescaling. */
int dbec_calcnorm(dbec_ctx *ctx, double *norm, int rescale);
int dbec_calcmuen(dbec_ctx *ctx, double *mu, double *en); /* calcmuen imag3d.cu:1250; real3d.cu:1339 (raw, not divided by par) */
/* mu and E with the kinetic density |grad psi|^2 of the complex wave function
 * (the phase gradient included): the energy a real-time run conserves.  The
 * reference's real-time calcmuen differentiates |psi| instead
 * (real3d.cu:1401-1424), so its E ignores the flow; this is an addition with
 * no counterpart there.  Imaginary time: identical to dbec_calcmuen. */
int dbec_calcmuen_phase(dbec_ctx *ctx, double *mu, double *en);
int dbec_calcrms(dbec_ctx *ctx, double rms[4]);          /* calcrms imag3d.cu:1639; real3d.cu:1733: <r>, <x>, <y>, <z> */
/* psi(Nx/2, Ny/2, Nz/2) as {re, im} (imag3d.cu:231, 244; real3d.cu:248). */
int dbec_psi_center(dbec_ctx *ctx, double out[2]);

/* Composite split step, `nsteps` times: [g += g_stp; gd += gd_stp;]
 * calcpsidd2, calcnu, calclux, calcluy, calcluz, and for imaginary time
 * calcnorm with rescale (imag3d.cu:315-325, 397-404; real3d.cu:329-338,
 * 412-418).  `norm` (may be NULL) receives the pre-rescale norm of the last
 * step (imaginary time only).  Semantics are those of the per-operator
 * sequence; the implementation fuses operators. */
int dbec_step(dbec_ctx *ctx, long nsteps, double g_stp, double gd_stp, double *norm);

/* Density extracts for the output writers (outdenx .. outdenxyz,
 * imag3d.cu:1944-2399; real3d.cu:2042-2522).  Values only; the caller adds
 * coordinates.  The number of doubles written is dbec_density_count(). */
#define DBEC_DEN_X    0   /* int dy dz |psi|^2, per sampled x      (outdenx)   */
#define DBEC_DEN_Y    1   /* int dx dz, per sampled y             (outdeny)   */
#define DBEC_DEN_Z    2   /* int dx dy, per sampled z             (outdenz)   */
#define DBEC_DEN_XY   3   /* int dz, [x][y] sampled               (outdenxy)  */
#define DBEC_DEN_XZ   4   /* int dy, [x][z]                       (outdenxz)  */
#define DBEC_DEN_YZ   5   /* int dx, [y][z]                       (outdenyz)  */
#define DBEC_CUT_XY0  6   /* |psi|^2 at z = Nz/2, [x][y]          (outpsi2xy) */
#define DBEC_CUT_X0Z  7   /* |psi|^2 at y = Ny/2, [x][z]          (outpsi2xz) */
#define DBEC_CUT_0YZ  8   /* |psi|^2 at x = Nx/2, [y][z]          (outpsi2yz) */
#define DBEC_DEN_XYZ  9   /* |psi|^2, [x][y][z] sampled           (outdenxyz) */
long dbec_density_count(const dbec_ctx *ctx, int kind, long outstpx, long outstpy, long outstpz);
int dbec_density(dbec_ctx *ctx, int kind, long outstpx, long outstpy, long outstpz, double *out);

/* dbec_density without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  The same extracts, equal to
 * dbec_density's bit for bit; the weighted sums and the sampled values live
 * in a fixed device array of DBEC_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, instead of a buffer allocated per call.  An extract whose
 * sums and samples need more fails with DBEC_EINVAL and a message: every 1-D
 * and 2-D kind fits at every supported grid, the sampled 3-D extracts up to
 * 1048576 values.  All calls on one GPU share its array and the call ends
 * with a synchronisation of the context's stream: calls that overlap on one
 * GPU (from several threads or streams) race, so serialise them. */
#define DBEC_NOALLOC_SCRATCH_BYTES (8L << 20)
int dbec_density_noalloc(dbec_ctx *ctx, int kind, long outstpx, long outstpy, long outstpz, double *out);

/* Set-up tables, for tests (host copies).  POT: [x][y][z]; POTDD: [kx][ky][kz]
 * for kz <= Nz/2; CALPHA/CGAMMA: axis 0,1,2, length N-1 (real: 1 double per
 * entry; real time: 2). */
#define DBEC_TABLE_POT    0
#define DBEC_TABLE_POTDD  1
#define DBEC_TABLE_CALPHA 2
#define DBEC_TABLE_CGAMMA 3
int dbec_get_table(dbec_ctx *ctx, int which, int axis, double *out);

/* The launch choices the context runs with (tile widths, row widths and
 * boundary peeling of its kernels, picked by timing at dbec_create), as one
 * line of text in buf (at most len bytes, NUL-terminated).  The choices never
 * change a result.  No counterpart in the reference (cuFFT planned its own
 * transforms, imag3d.cu:1037-1137). */
int dbec_tuning(const dbec_ctx *ctx, char *buf, long len);

/* Wait for all work queued on the context's stream. */
int dbec_synchronize(dbec_ctx *ctx);

/* Per-kernel timing of the composite step, measured with HIP events on the
 * context's stream: calls of two and of three steps run, and each kernel
 * launch of them is repeated `reps` times back to back between two events,
 * so avg_ms is the pipelined time of one launch; `launches` is per step of a
 * long run (0 for kernels that run once per dbec_step call).  The repeats change psi: the state after the call is
 * not a solver state (profile after reading the results you need).  Fills up
 * to `max` records; *n gets the number of kernels.  No counterpart in the
 * reference (its only timers bracket whole phase loops, imag3d.cu:314/326). */
typedef struct dbec_kernel_stat {
    char name[48];
    double avg_ms;         /* mean duration of one launch */
    double alg_bytes;      /* algorithmic HBM bytes of one launch */
    double launches;       /* launches per step */
} dbec_kernel_stat;
int dbec_profile_step(dbec_ctx *ctx, int reps, dbec_kernel_stat *stats, int max, int *n);

#ifdef __cplusplus
}
#endif

#endif /* DBEC_H */
