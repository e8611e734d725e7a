// launch.hpp -- host-side launchers of the band-walk kernels.  The kernels of
// each family (and each Life radius) are instantiated in their own translation
// unit (life_r1.hip .. life_r5.hip, forest.hip) so that they compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels_forest.hpp"
#include "kernels_life.hpp"

// Life, Moore(R): vb in {1, 2} (R = 5 uses 2); niv = intervals per mask (1 or 2).
template <int R>
void launch_life_r(int vb, bool bin, int niv, int blocks, const LifeArgs& a, hipStream_t st);
template <int R>
void launch_life_batch_r(int vb, bool bin, int niv, int blocks, const LifeArgs* items,
                         const int32_t* wave_item, const int32_t* wave_base, int32_t nwaves,
                         int32_t parity, hipStream_t st);

// forest fire, VonNeumann(1): vb in {1, 2}; small = fewer than 2^32 cells; tri =
// the lattice and its Constant edges hold only 0, 1 and 2.
void launch_forest(int vb, bool small, bool tri, int blocks, const ForestArgs& a, hipStream_t st);
// The batch items carry p, f and their time base; prec = p_f32 | f_f32 << 1;
// vb in {1, 2}; tri: every item's lattice and Constant edges hold only 0, 1 and 2.
void launch_forest_batch(int vb, bool small, bool tri, int prec, int blocks, const ForestArgs* items,
                         const int32_t* wave_item, const int32_t* wave_base, int32_t nwaves,
                         int32_t parity, int64_t dt, hipStream_t st);

// ---------------------------------------------------------------- Lenia (lenia.hip)
#include "kernels_lenia_mfma.hpp"

// MFMA implicit GEMM: mode LM_F16 / LM_BF16 / LM_F16X2.
void launch_lenia_mfma(int mode, dim3 grid, const LeniaMfmaArgs& a, hipStream_t st);
// One step of several lattices: nblocks 64 x 64 output blocks over the items,
// block b of item block_item[b] being its block (b - block_base[item]).
void launch_lenia_mfma_batch(int mode, int nblocks, const LeniaMfmaArgs* items,
                             const int32_t* block_item, const int32_t* block_base,
                             int32_t parity, int64_t dt, hipStream_t st);
// The A fragments of the implicit GEMM for a dense (2H+1)^2 FP64 weight grid
// wd[(dy + H) * (2H + 1) + (dx + H)], as the bytes of the device buffer.
void lenia_mfma_fragments(int mode, int H, const double* wd, std::vector<char>& out);

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
