// launch.hpp -- host-side launchers of the band-walk kernels.  The kernels of
// each family (and each Life radius) are instantiated in their own translation
// unit (life_r1.hip .. life_r5.hip, forest.hip, forest_batch.hip,
// forest_gated.hip) so that they compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels_forest.hpp"
#include "kernels_life.hpp"

// Life, Moore(R): vb in {1, 2} (R = 5 uses 2); niv = intervals per mask (1 or 2).
template <int R>
void launch_life_r(int vb, bool bin, int niv, int blocks, const LifeArgs& a, hipStream_t st);
// Life or Generations under a gate and/or a mask (k_life_gated): vb in {1, 2};
// cells: GATED_ANY, GATED_BIN (every value read is 0 or 1) or GATED_SMALL (every
// value read is 0, 1 or 2, and a Generations rule has at most 3 states);
// niv = intervals per mask (1 or 2); masked = a.g.mask is set.
enum { GATED_ANY = 0, GATED_BIN = 1, GATED_SMALL = 2 };
template <int R>
void launch_life_gated_r(int vb, int cells, int niv, bool masked, int blocks, const LifeArgs& a, hipStream_t st);
// Blocks of that kernel a CU holds at once (0: unknown).
template <int R>
int life_gated_blocks_per_cu_r(int vb, int cells, int niv, bool masked);
template <int R>
void launch_life_batch_r(int vb, bool bin, int niv, int blocks, const LifeArgs* items,
                         const int32_t* wave_item, const int32_t* wave_base, int32_t nwaves,
                         int32_t parity, hipStream_t st);

// forest fire, VonNeumann(1): vb in {1, 2}; small = fewer than 2^32 cells; tri =
// the lattice and its Constant edges hold only 0, 1 and 2.
void launch_forest(int vb, bool small, bool tri, int blocks, const ForestArgs& a, hipStream_t st);
// The same under a gate and/or a mask (k_forest_gated, 4 cells a lane).
void launch_forest_gated(bool small, bool tri, int blocks, const ForestArgs& a, hipStream_t st);
// The batch items carry p, f and their time base; prec = p_f32 | f_f32 << 1;
// vb in {1, 2}; tri: every item's lattice and Constant edges hold only 0, 1 and 2;
// each item's step time is its t + dt (+ *tbase, read on the device, when tbase
// is not null).
void launch_forest_batch(int vb, bool small, bool tri, int prec, int blocks, const ForestArgs* items,
                         const int32_t* wave_item, const int32_t* wave_base, int32_t nwaves,
                         int32_t parity, int64_t dt, const int64_t* tbase, hipStream_t st);
