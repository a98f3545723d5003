// ops.hpp - launch interfaces of the non-FFT kernels (ops.hip).
#pragma once

#include <algorithm>
#include <type_traits>

#include "internal.hpp"
#include "nu_device.hpp"

namespace dbec {

struct MuenConst {
    double dx, dy, dz, kin, g, gd, nd;  // kin = 3 - par
};
template <typename V>
struct SweepConst {
    V A0r, A, mA;  // diagonal of the explicit half, off-diagonal, and its negative
};

// Layout of the Crank-Nicolson coefficient buffer (b.coef, FP64 real or
// complex elements in both precisions): calpha and cgamma of the axes
// x, y, z back to back (n - 1 each), then the carry multipliers gb and fa of
// the partitioned sweeps (n each per axis) for segments of 16 and of 32:
//   [calpha | cgamma | gb16 | fa16 | gb32 | fa32]
struct CoefLayout {
    long tot, totn, off[3], offn[3];
    CoefLayout(long Nx, long Ny, long Nz)
        : tot(Nx + Ny + Nz - 3), totn(Nx + Ny + Nz), off{0, Nx - 1, Nx - 1 + Ny - 1}, offn{0, Nx, Nx + Ny} {}
    long ca(int a) const { return off[a]; }
    long cg(int a) const { return tot + off[a]; }
    long gb(int a, int m) const { return 2 * tot + 2 * m * totn + offn[a]; }  // m = 0: M = 16, 1: M = 32
    long fa(int a, int m) const { return 2 * tot + (2 * m + 1) * totn + offn[a]; }
    long size() const { return 2 * tot + 4 * totn; }
};

hipError_t launch_facefix(hipStream_t s, void *phi, bool fp32, long Nx, long Ny, long Nz);
hipError_t launch_calcnu(hipStream_t s, const dbec_ctx *c, const double *scale);
hipError_t launch_scale(hipStream_t s, const dbec_ctx *c, const double *scale);
// *p = v on the stream (a one-thread kernel: also valid under graph capture)
hipError_t launch_set_scalar(hipStream_t s, double *p, double v);
// Sweep along `axis` of the solves v.
// The grid, coefficients and weights come from c.  The z sweep with cols also
// writes, per solve and (y tile t, x), rows[t][x] = sum over the tile's rows
// of wy[y] * (Simpson z integral of psi^2).
hipError_t launch_sweep(hipStream_t s, const dbec_ctx *c, int axis, const Solves &v, bool cols);
// number of y tiles of the z sweep (rows of the column partials)
long sweep_z_tiles(const dbec_ctx *c);
int sweep_z_rows(const dbec_ctx *c);
// The z sweep's tables, constants, weights and rows per workgroup for the
// z sweep fused with the forward z pass.
void sweep_z_args(const dbec_ctx *c, FftArgs &a);
// norm of each solve from its column partials: scal[NORM], scal[SCALE]
hipError_t launch_norm_from_cols(hipStream_t s, const dbec_ctx *c, const Solves &v);
TrapConst trap_const(const dbec_ctx *c);
NuConst nu_const(const dbec_ctx *c);
hipError_t launch_rows_mod2(hipStream_t s, const dbec_ctx *c, double *rows0, double *rowsz);
// phase: complex psi's kinetic density |grad psi|^2 instead of (grad |psi|)^2
hipError_t launch_rows_muen(hipStream_t s, const dbec_ctx *c, double *rmu, double *ren, bool phase = false);
// mode 0: norm and 1/norm; 1: norm and rms moments; 2: mu and E
hipError_t launch_finalize(hipStream_t s, const dbec_ctx *c, int mode, const double *A, const double *B);
hipError_t launch_reduce3(hipStream_t s, const dbec_ctx *c, int axis, double *out);
hipError_t launch_reduce2(hipStream_t s, const double *in, long n0, long n1, int axis, const double *w, double *out);
hipError_t launch_mod2(hipStream_t s, const dbec_ctx *c, double *out);
// dst[a][b][k] = |psi|^2 at (start + (a, b, k) * step), (a, b, k) < cnt
hipError_t launch_pick_mod2(hipStream_t s, const dbec_ctx *c, const long start[3], const long step[3],
                            const long cnt[3], double *dst);
// dst[a][b] = src[(a o0) n1 + b o1], a < c0, b < c1
hipError_t launch_pick2(hipStream_t s, const double *src, long n1, long o0, long o1, long c0, long c1, double *dst);

}  // namespace dbec
