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
//   potddq [kz][kx][ky]     real, kz <= Nz/2, kx <= Nx/2, ky <= Ny/2: the
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
    double *coef = nullptr;  // [axis][calpha | cgamma], real or complex, FP64
    void *coef_lo = nullptr; // same in FP32 for the mixed form
    double *wgt = nullptr;   // Simpson weights wx | wy | wz (include h/3)
    double *sq = nullptr;    // x^2 | y^2 | z^2
    void *tw[3] = {nullptr, nullptr, nullptr};  // twiddles for Nx, Ny, Nz (FP64)
    void *twf[3] = {nullptr, nullptr, nullptr}; // twiddles in FP32
};

// Device scalar slots.
enum { SC_NORM = 0, SC_SCALE = 1, SC_ONE = 2, SC_MU = 3, SC_EN = 4, SC_R0 = 5, SC_RX = 6, SC_RY = 7, SC_RZ = 8, SC_COUNT = 16 };

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
    bool phi_valid = false;
    dbec::Buffers b;
    dbec::KernelTimer *timer = nullptr;  // set while dbec_profile_step records
    std::string err;
};

namespace dbec {

// ---- FFT passes (fft_kernels.hpp, one instantiation per length) ----------
struct FftArgs {
    hipStream_t stream;
    long Nx, Ny, Nz, Nkz;
    bool fp32;      // spec / phi / psi stored in FP32
    bool cpsi;      // psi complex
    void *psi;
    const double *scale;  // multiply psi by *scale while reading (imaginary time), or nullptr
    void *spec;
    void *phi;
    const void *potddq;
    const void *tw;       // twiddles of the transformed axis (FP64 or FP32)
    const double *sq;     // x^2 | y^2 | z^2 (fused calcnu)
    TrapConst tc;
    NuConst nc;
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

}  // namespace dbec
