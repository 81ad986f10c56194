// Batched 3-D complex FFT on n^3 grids, n = 2^a 3^b (48, 64, 128, 216, 256).
//
// One kernel per axis: a workgroup loads 8 lines of the axis into LDS
// (coalesced along x for every axis), runs a Stockham autosort mixed-radix
// (4, 2, 3) transform on them, and stores them back. Twiddles come from a
// per-size table exp(-2 pi i k / n) computed in double on the host.
#pragma once
#include "common.hpp"

namespace pyfof {

struct FftPlan {
    int n = 0;
    int nrad = 0;
    int radix[16] = {};
    float2* tw = nullptr;  // device table, n entries
};

// Builds the plan (radix sequence and twiddle table in `tw_dev`, n float2).
void fft_make_plan(FftPlan& p, int n, float2* tw_dev, hipStream_t s);

// In-place forward (inverse = false: exp(-i)) or inverse (exp(+i),
// unnormalised) transform of the grids whose slot numbers are listed in
// `slots` (nslots entries); grid g starts at data + slot * n^3.
void fft3d(const FftPlan& p, float2* data, const int32_t* slots, int32_t nslots, bool inverse,
           hipStream_t s);

// Potential of zero-padded sources by circular convolution with the real
// spectrum G (n^3 f32): grid slot k <- IFFT(G x FFT(acc_k / fscale)) on the
// nodes [0, n/2 + 1]^3, unnormalised (divide by n^3). acc holds each slot's
// int64 fixed-point deposit, non-zero only on [0, n/2 + 1]^3; slot_halo maps a
// slot to the halo whose scale fscale[] it uses. Other nodes of grid are
// left with intermediate values. lines: grid lines per workgroup (4, 8 or
// 16), a launch shape only: the result does not depend on it.
void fft_conv3d(const FftPlan& p, int lines, float2* grid, const unsigned long long* acc,
                const int32_t* slots, int32_t nslots, const int32_t* slot_halo, const double* fscale,
                const float* G, hipStream_t s);

// True for the grid sizes fft_conv3d has kernels for (48, 64, 128, 216, 256).
bool fft_conv_size_ok(int n);

}  // namespace pyfof
