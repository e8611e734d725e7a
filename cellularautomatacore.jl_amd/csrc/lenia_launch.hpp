// lenia_launch.hpp -- host-side launchers of the Lenia field kernels (lenia.hip),
// kept apart from the band-walk launchers so that the units of those kernels do
// not depend on the Lenia kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels_lenia_mfma.hpp"

// MFMA implicit GEMM: mode LM_F16 / LM_BF16 / LM_F16X2, one block for each of the
// ntiles 64 x 64 tiles of one lattice (a.nbx a row).
void launch_lenia_mfma(int mode, int32_t ntiles, const LeniaMfmaArgs& a, hipStream_t st);
// One step of several lattices: nblocks 64 x 64 output blocks over the items,
// block b of item block_item[b] being its block (b - block_base[item]).
void launch_lenia_mfma_batch(int mode, int nblocks, const LeniaMfmaArgs* items,
                             const int32_t* block_item, const int32_t* block_base,
                             int32_t parity, int64_t dt, hipStream_t st);
// The first row offset and the row-offset pairs the MFMA loop visits for a dense
// (2H+1)^2 FP64 weight grid wd[(dy + H) * (2H + 1) + (dx + H)].
void lenia_mfma_pairs(int H, const double* wd, int32_t& dy0, int32_t& npairs);
// The A fragments of the implicit GEMM for that grid, as the bytes of the device
// buffer (two planes: the values, then their residuals for split FP16).
void lenia_mfma_fragments(int mode, int H, const double* wd, int32_t dy0, int32_t npairs, std::vector<char>& out);

// FP32 FFT field on a periodic lattice (the FFTW extension's cross-correlation).
struct LeniaFFT;
int lenia_fft_create(LeniaFFT** out, int n1, int n2, int64_t pitch, const int32_t* taps,
                     const double* w, int ntaps, hipStream_t st);
// field = irfft(rfft(cur) .* conj(rfft(kpad))) into the plan's dense FP32 buffer
int lenia_fft_field(LeniaFFT* f, const float* cur, hipStream_t st);
float* lenia_fft_buffer(LeniaFFT* f);
void lenia_fft_destroy(LeniaFFT* f);
// A plan-less stand-in with the given (fake) buffers, for launch checks without
// a device: its field is never computed.
LeniaFFT* lenia_fft_make_dry(int n1, int n2, void* spec, void* kspec, float* field);
