// fft_launch.hpp - declarations of the FFT pass launchers of one transform
// length.  Their definitions and the kernels live in fft_kernels.hpp and are
// compiled once per length (fft_inst.hip); callers include only this header,
// so they hold no copy of any kernel or launch geometry.
#pragma once

#include "internal.hpp"

namespace dbec {

template <int N>
struct FftLaunch {
    static hipError_t zfwd(const FftArgs &a);
    static hipError_t ypass(int dir, const FftArgs &a);
    static hipError_t xconv(const FftArgs &a);
    static hipError_t zinv(const FftArgs &a);   // single solve: spectrum -> phi
    static hipError_t fused(const FftArgs &a);  // zinv fused with calcnu
    // the z sweep fused with the next step's forward z pass, for the
    // precisions and z sweep rows per workgroup where fusable() holds
    static hipError_t sweepz_zfwd(const FftArgs &a);
    static bool fusable(bool fp32, bool cpsi, int zrows);
};

}  // namespace dbec
