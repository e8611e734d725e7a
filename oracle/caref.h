/*
 * caref.h -- CPU oracle for the synchronous stencil sweep of CellularAutomataCore.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * Julia algorithm, used to check the HIP backend (tests/, the smoke test and the
 * CPU-baseline leg of bench.py).  Nothing in the product links or calls it.
 *
 * Each function cites the reference file:line it restates.  Indices follow the
 * reference: cellid is 1-based column-major on the unpadded lattice, `t` is the
 * time of the state being read.
 */
#ifndef CAREF_H
#define CAREF_H

#include <stdint.h>
#include "../include/cahip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* philox4x32 (src/random.jl:24-37) */
void ref_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* _uniform / _uniform32 (src/random.jl:102-120) */
double ref_uniform(uint64_t seed, int64_t cellid, int64_t t, int64_t k);
float  ref_uniform32(uint64_t seed, int64_t cellid, int64_t t, int64_t k);
/* _map_axis (src/spaces.jl:46-51); j is 1-based, returns the 1-based index */
int64_t ref_map_axis(int32_t kind, int64_t j, int64_t n, int32_t* outside);
/* _offsets_where (src/neighbourhoods.jl:124-153): kind 0 = Moore, 1 = VonNeumann.
 * Writes pairs (o1, o2) (o2 = 0 in 1-D), returns the count. */
int32_t ref_offsets(int32_t kind, int32_t ndims, int32_t r, int32_t* out);
/* kernel_weights(LeniaKernel(R; peaks)) (src/kernels.jl:87-104, :154, :174-189):
 * kernel offsets (not negated) and normalised weights; returns the count. */
int32_t ref_lenia_kernel(int32_t R, const double* peaks, int32_t npeaks, int32_t* offs,
                         double* w);

/* One ReferenceBackend step per call, nsteps times, in place on the host array u
 * (prod(dims) Int8 or Float32): _advance_plain! (src/backends.jl:314-322) over
 * _plain_sweep! (src/backends.jl:264-278) with boundary reads on every access. */
int ref_advance_plain(const cah_desc* d, void* u, int64_t t0, int64_t nsteps,
                      const cah_params* p);
/* update_cell(prob, u, I, t, p) (src/backends.jl:294-307) for an Int8 rule:
 * the new value of cell (i1, i2) (0-based). */
int ref_update_cell(const cah_desc* d, const void* u, int64_t i1, int64_t i2, int64_t t,
                    const cah_params* p);
/* The ArrayBackend engine (src/backends.jl:392-506): halo-padded state, serial
 * halo fill, sweeps threaded over min(n_last, 4 * nthreads) slabs of the last
 * axis.  The CPU baseline. */
int ref_advance_halo(const cah_desc* d, void* u, int64_t t0, int64_t nsteps,
                     const cah_params* p, int32_t nthreads);
/* The Lenia potential alone, FP64 sequential sum with boundary reads
 * (_compute_field_reference!, src/environment.jl:151-160), from a Float32 state. */
int ref_lenia_field(const cah_desc* d, const float* u, double* out);
/* The field epidemic's KernelField (source ==(1)) over an Int8 state, in the
 * field's precision and widened to double (src/environment.jl:151-160). */
int ref_epi_field(const cah_desc* d, const int8_t* u, double* out);

/* The build's synthetic-input generator (no reference counterpart; documented in
 * include/cahip.h at cah_state_fill_bernoulli). */
void ref_fill_bernoulli(uint64_t seed, int64_t n, double p, int8_t* out);
void ref_fill_uniform_box(uint64_t seed, int64_t n1, int64_t n2, int64_t lo1, int64_t len1,
                          int64_t lo2, int64_t len2, float* out);

/* sizeof and field offsets of cah_desc and cah_params (layout check). */
int ref_desc_layout(int64_t* out);

#ifdef __cplusplus
}
#endif
#endif
