// internal.hpp - context and launch interfaces shared by the HIP translation
// units of libdbec_hip (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dbec.h"
#include "fft_device.hpp"
#include "nu_device.hpp"

namespace dbec {

// Device data layout of one solve (all arrays row-major, last index fastest):
//   psi    [x][y][z]        double (imaginary time) or cpx<double> (real time);
//                           float / cpx<float> in the mixed form
//   spec   [kz][x][y]       cpx, kz = 0..Nz/2: half spectrum of |psi|^2, laid
//                           out so that y lines are contiguous and x lines are
//                           strided by Ny (coalesced over y tiles)
//   phi    [x][y][z]        real: unnormalised dipolar potential (psidd2)
//   potddq [kz][ky][kx]     real, kz <= Nz/2, kx <= Nx/2, ky <= Ny/2: the
//                           dipolar kernel is even in every k component, so a
//                           quarter table indexed by |k| holds it exactly
//   rows   [2][Ny][Nx]      double: per-row z integrals for the reductions
struct Buffers {
    void *psi = nullptr;
    void *spec = nullptr;
    void *phi = nullptr;
    void *potddq = nullptr;
    double *rows = nullptr;
    double *red = nullptr;   // reduction scratch (>= 4 * (Nx + Ny + Nz) doubles)
    double *scal = nullptr;  // device scalars, see SC_* below
    double *coef = nullptr;  // CN coefficients and carry multipliers (CoefLayout), real or complex, FP64
                             // in both precisions (the mixed sweeps round them to FP32 in LDS)
    double *wgt = nullptr;   // Simpson weights wx | wy | wz (include h/3)
    double *sq = nullptr;    // x^2 | y^2 | z^2
    void *tw[3] = {nullptr, nullptr, nullptr};  // twiddles for Nx, Ny, Nz (FP64)
    void *twf[3] = {nullptr, nullptr, nullptr}; // twiddles in FP32
};

// Device scalar slots.
// SC_SHIFT: exponent shift of the mixed-precision imaginary-time step (see
// k_norm_cols), 0 elsewhere; SC_CAPPED: set to 1 when that step had to cap
// the growth factor (a collapsing state).
enum { SC_NORM = 0, SC_SCALE = 1, SC_ONE = 2, SC_MU = 3, SC_EN = 4, SC_R0 = 5, SC_RX = 6, SC_RY = 7, SC_RZ = 8,
       SC_SHIFT = 9, SC_LASTNORM = 10, SC_CAPPED = 11, SC_COUNT = 16 };

// Launch choices of the step kernels that leave every result bit-identical
// (which lines share a workgroup, which iterations test their range), picked
// per context by timing (autotune at creation); 0 / -1: the built-in rule.
struct Tuning {
    int lm[4] = {0, 0, 0, 0};   // FFT tile multiplier (1, 2): zfwd, ypass, xconv, fused inverse z
    int xy_rowb[2] = {0, 0};    // x / y sweeps: bytes of z per row load (128, 256)
    int xy_peel[2] = {-1, -1};  // x / y sweeps: boundary iterations peeled (1) or tested (0)
    int z_direct = -1;          // z sweep: segments loaded straight from HBM (1) or staged through LDS (0)
    int zinv_swz = -1;          // fused inverse z pass, FP64: swizzled (1) or padded (0) LDS layout
};

struct KernelTimer;

}  // namespace dbec

struct dbec_ctx {
    dbec_params p;
    dbec_mode mode;
    dbec_prec prec;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    long Nx = 0, Ny = 0, Nz = 0, N = 0, Nkz = 0;
    double pi = 0.;
    double g = 0., gd = 0.;
    // CN constants (imaginary time: .x only; real time: complex)
    double A0r[3][2], A[3][2], mA[3][2];
    std::vector<double> h_calpha[3], h_cgamma[3];  // host copies (1 or 2 doubles per entry)
    std::vector<double> h_potdd_half;              // [kx][ky][kz<=Nz/2], for dbec_get_table
    bool scale_pending = false;  // psi must be multiplied by scal[SC_SCALE]
    bool view = false;           // buffers belong to another owner (request-loop slot)
    bool phi_valid = false;
    dbec::Buffers b;
    dbec::Tuning tune;
    dbec::KernelTimer *timer = nullptr;  // set while dbec_profile_step records
    std::string err;
};

namespace dbec {

// ---- The solves of a launch ----------------------------------------------
// The step kernels advance nb solves of one grid at once: gridDim.y = nb and
// block row y works on solve slot[y] (slot y when `slot` is null).  Slot s
// owns the buffers at base + s * stride, so every buffer address is kernel
// argument arithmetic (global address space, no pointer loads).  Solves of one
// launch share the grid, dx, dt, cutoff and precision, hence every table;
// they differ in their buffers, trap and couplings.  No kernel mixes data of
// two solves, so a solve's result does not depend on the others in its launch.
struct SolveParams {
    TrapConst tc;
    double g, gd;  // couplings of this step
};
struct Solves {
    void *psi;
    void *spec;
    double *rows;  // z-sweep column partials of psi^2 (imaginary time)
    double *scal;  // device scalars, SC_* slots
    long psi_st, spec_st, rows_st, scal_st;  // bytes from one slot to the next
    const int *slot;          // launch row -> slot, or nullptr
    const SolveParams *par;   // per slot, or nullptr: every row uses p0
    SolveParams p0;
    int nb;                   // launch rows (gridDim.y)
    int use_scale;            // psi is read as psi * scal[SC_SCALE] (pending renormalisation)
    int shift;                // calcnu multiplies by exp(scal[SC_SHIFT]) too (mixed imaginary time)
    int fold;                 // the last step's norm is still in the column partials: the forward z
                              // pass computes it (workgroup 0 of each solve), reads psi unscaled, and
                              // the x convolution applies 1/norm^2 (the transform is linear)
};

struct SolveAt {
    char *psi, *spec;
    double *rows, *scal;
    SolveParams p;
};
__device__ __forceinline__ SolveAt solve_at(const Solves &v) {
    const int y = blockIdx.y;
    const long sl = v.slot ? v.slot[y] : y;
    SolveAt r;
    r.psi = static_cast<char *>(v.psi) + sl * v.psi_st;
    r.spec = static_cast<char *>(v.spec) + sl * v.spec_st;
    r.rows = reinterpret_cast<double *>(reinterpret_cast<char *>(v.rows) + sl * v.rows_st);
    r.scal = reinterpret_cast<double *>(reinterpret_cast<char *>(v.scal) + sl * v.scal_st);
    if (v.par) r.p = v.par[sl];
    else r.p = v.p0;
    return r;
}

// ---- FFT passes (fft_kernels.hpp, one instantiation per length) ----------
struct FftArgs {
    hipStream_t stream;
    long Nx, Ny, Nz, Nkz;
    bool fp32;      // spec / phi / psi stored in FP32
    bool cpsi;      // psi complex
    Solves v;             // the solves of the launch
    void *phi;            // standalone calcpsidd2 (single solve): phi output
    const void *potddq;
    const void *tw;       // twiddles of the transformed axis (FP64 or FP32)
    const double *sq;     // x^2 | y^2 | z^2 (fused calcnu)
    NuConst nc;           // dt and grid size; couplings come from the solve
    const double *wx;     // Simpson weights along x (folded norm)
    int tiles;            // rows of the column partials (folded norm)
    Tuning tune;          // the context's launch choices
    // z sweep fused with the forward z pass: tables (calpha, cgamma, gb, fa of
    // the z axis, segments of 16), constants A0r, A, mA (re, im), Simpson
    // weights along z and y, rows per workgroup
    const void *ztab[4];
    double zk[3][2];
    const double *wz, *wy;
    int zrows;
};
// Each returns hipSuccess or an error; `n` is the transformed length.
hipError_t launch_zfwd(long n, const FftArgs &a);    // |psi|^2 -> spec (R2C along z)
hipError_t launch_ypass(long n, int dir, const FftArgs &a);
hipError_t launch_xconv(long n, const FftArgs &a);   // x FFT, * potdd, inverse x FFT
hipError_t launch_zinv(long n, const FftArgs &a);    // spec -> phi (C2R along z)
// C2R along z fused with calcnu: the interior rows with phi kept on chip and
// the face rows x = Nx-1 or y = Ny-1 with the reference's periodic face
// copies of phi, in one launch.
hipError_t launch_zinv_nu(long n, const FftArgs &a);
// The z sweep of one step fused with the forward z pass of the next; whether
// length n has it for the precision and z sweep rows per workgroup.
hipError_t launch_sweepz_zfwd(long n, const FftArgs &a);
bool sweepz_zfwd_fusable(long n, bool fp32, bool cpsi, int zrows);

}  // namespace dbec
