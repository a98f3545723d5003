/*
 * dbec_serve.h - a continuous-batching request loop over the solver of
 * dbec.h: many ground-state (imaginary-time) and real-time solves with
 * differing grids, couplings and traps share one GPU.
 *
 * There is no counterpart in the reference, which runs one solve per program
 * invocation (src/imag3d-mpicuda/imag3d-mpicuda.cu:29-645,
 * src/real3d-mpicuda/real3d-mpicuda.cu:29-670).  Each solve follows the same
 * per-step operators as dbec_step; the ground-state stop rule is
 * |mu - mu_prev| <= mutol |mu| between two checks (calcmuen,
 * imag3d-mpicuda.cu:1250-1291).
 *
 * Requests of one grid, dx, dt, cutoff, mode and precision form a size
 * bucket.  A bucket keeps its tables (dipolar kernel, twiddles, coefficients)
 * resident and owns a fixed number of slots; a slot holds one solve's psi and
 * spectrum.  Each bucket advances all of its active solves with one launch per
 * kernel (the step kernels take a list of slots), on its own stream, in chunks
 * of `check_every` steps; between chunks the host admits waiting requests into
 * free slots and retires finished ones (a bucket of real-time solves runs up
 * to 8 chunks back to back when every solve in it has that many left).  Buckets are created while the device
 * memory they need fits `mem_budget`, and idle buckets are released when a new
 * one does not fit.  A solve's result does not depend on which other solves
 * shared its launches.
 *
 * Calls are not thread-safe: one host thread drives a server.
 */
#ifndef DBEC_SERVE_H
#define DBEC_SERVE_H

#include "dbec.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dbec_server dbec_server;

typedef struct {
    int device;          /* HIP device ordinal, -1 = current */
    double mem_budget;   /* bytes of device memory the server may hold; <= 0: half the free memory */
    int max_batch;       /* slots per bucket (solves per launch), 1..64 */
    long check_every;    /* steps per chunk: mu check, admission and retirement points */
    double polish_tol;   /* > 0: a DBEC_MIXED ground state that reaches |dmu/mu| <= polish_tol moves
                            to an FP64 slot of the same grid and converges there to its mutol
                            (FP32 cannot resolve mu to 1e-8); it also moves once |dmu/mu| is below
                            1e-6 and stops falling between checks (FP32's rounding floor);
                            <= 0: mixed solves run to mutol */
} dbec_server_config;

typedef struct {
    long id;             /* caller's tag, returned in the result */
    int mode;            /* DBEC_IMAG (ground state) or DBEC_REAL */
    int prec;            /* DBEC_FP64 or DBEC_MIXED */
    long nx, ny, nz;
    double dx, dy, dz, dt;
    double vgamma, vnu, vlambda, cutoff, par;
    double g, gd;        /* couplings as the operators use them (par * G0, par * GDD0) */
    long max_steps;      /* imaginary time: step cap; real time: steps to run */
    double mutol;        /* imaginary time: stop rule on |dmu/mu| between checks */
} dbec_request;

/* Status of a result.  DIVERGED: psi blew up (NaN or overflow), or the
 * mixed-precision step had to cap its growth factor; both happen when the
 * requested condensate collapses (strong dipolar interaction), and FP64 and
 * mixed solves are stopped at the first check after it. */
enum { DBEC_SOLVE_CONVERGED = 0, DBEC_SOLVE_STEPS_DONE = 1, DBEC_SOLVE_NOT_CONVERGED = 2, DBEC_SOLVE_DIVERGED = 3 };

typedef struct {
    long id;
    int status;
    long steps;          /* split steps taken (mixed and FP64 polish together) */
    long steps_fp64;     /* of which in the FP64 polish */
    double norm;         /* norm of psi (imaginary time: the last step's pre-rescale norm) */
    double mu, en;       /* raw calcmuen values (the programs print them divided by par) */
    double rms[4];       /* <r>, <x>, <y>, <z> */
    double t_submit, t_start, t_done;  /* seconds on the server clock */
} dbec_result;

int dbec_server_create(const dbec_server_config *cfg, dbec_server **out);
int dbec_server_destroy(dbec_server *s);
/* Seconds since the server was created (the clock of the result times). */
double dbec_server_time(const dbec_server *s);
/* Queues a request (copied).  Rejected: bad parameters, or a bucket that
 * could never fit mem_budget. */
int dbec_server_submit(dbec_server *s, const dbec_request *r);
/* Advances the loop without blocking on the device: retires finished
 * chunks, admits queued requests, launches the next chunks.  Up to `max`
 * finished results are written to `out`, their number to *n.  With wait != 0
 * it sleeps briefly when nothing changed. */
int dbec_server_progress(dbec_server *s, int wait, dbec_result *out, int max, int *n);
/* Requests submitted and not yet returned by dbec_server_progress. */
long dbec_server_pending(const dbec_server *s);
/* Device memory held by buckets, in bytes, and the number of buckets. */
int dbec_server_usage(const dbec_server *s, double *bytes, int *buckets);
const char *dbec_server_last_error(const dbec_server *s);

#ifdef __cplusplus
}
#endif
#endif
