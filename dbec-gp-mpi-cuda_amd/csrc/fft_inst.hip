// fft_inst.hip - compiled once per transform length (make passes
// -DDBEC_FFT_N=<n>); instantiates the FFT pass launchers of that length so the
// lengths build in parallel.
#include "fft_kernels.hpp"

#ifndef DBEC_FFT_N
#error "DBEC_FFT_N must be defined"
#endif

namespace dbec {
template struct FftLaunch<DBEC_FFT_N>;
}
