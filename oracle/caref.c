/*
 * caref.c -- CPU oracle for the synchronous stencil sweep of CellularAutomataCore.jl.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's Julia code,
 * linked by tests/, the smoke test and the CPU-baseline leg of bench.py, never by
 * the product.  Built with -ffp-contract=off so that FP64 sums and products round
 * as Julia's (which never contracts a*b+c into an FMA unless asked to).
 *
 * Pinning (what checks each function, so that a wrong index, sign or term fails):
 *   ref_philox4x32      Random123 known answers held by test/test_random.jl:4-14.
 *   ref_uniform(32)     closed form of src/random.jl:102-120 checked bit by bit
 *                       against an independent big-integer Python evaluation of the
 *                       same formula, plus the invariants of test/test_random.jl:18-56;
 *                       the counter / key word order by the two draws the reference
 *                       prints (vignettes/05_stochastic_rules/05_stochastic_rules.md:46-52).
 *   ref_map_axis        the vectors of test/test_spaces.jl:88-107.
 *   ref_offsets         counts and orders of test/test_spaces.jl:164-197 and the
 *                       doctest of src/neighbourhoods.jl:143-146.
 *   Life / Generations  naive definitions (test/test_rules.jl:23-32, :75-87),
 *                       blinker / block / glider closed forms (:109-122), and a
 *                       brute-force Larger-than-Life count on small tori.
 *   forest fire         invariants (fire -> empty, p = 0 keeps an empty lattice
 *                       empty, f = 0 never ignites a fire-free forest) and the
 *                       empty->tree rate p.
 *   ref_lenia_kernel    weights sum to one, none at the centre, 516 taps at R = 13
 *                       (test/test_environment.jl:3-21), radial symmetry.
 *   Lenia step          FFT cross-correlation in numpy (oracle/lenia_fft.py) and the
 *                       uniform-field closed form (test/test_environment.jl:46-53).
 *   ref_advance_halo    equal to ref_advance_plain state by state
 *                       (test/backend_conformance.jl:207-231).
 *   ref_update_cell     equal to one synchronous step over all cells
 *                       (test/test_backend_interface.jl:40-68).
 *   field epidemic      the field: the kernel-field values of
 *                       test/test_environment.jl:46-80 (a uniform radius-1 field is
 *                       1/5 on the cross, 2/5 in a corner beyond Constant(1) edges,
 *                       an eastward kernel carries a source east); the rule: beta = 0
 *                       infects nobody, a huge beta infects every susceptible cell
 *                       with an infected cell in reach and no other, infection is
 *                       permanent without recovery, masked cells never change, the
 *                       stated form equals the Bernoulli form with u -> 1 - u
 *                       swapped in (checked against the draws), infection rates
 *                       match 1 - exp(-beta * field) over many cells.
 */
#include "caref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* The largest neighbourhood a rule's neighbour values are gathered for: Moore(32),
 * the widest the device kernels take (halo 32), has 65 * 65 - 1 cells. */
#define REF_MAX_NB (65 * 65)
#endif

/* ------------------------------------------------------------------ Philox4x32-10 */

/* src/random.jl:6-37 */
void ref_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 1; round <= 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        if (round < 10) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* _philox_draw (src/random.jl:102-107) */
static void draw(uint64_t seed, int64_t cellid, int64_t t, int64_t k, uint32_t r[4]) {
    uint64_t id = (uint64_t)cellid;
    uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), (uint32_t)(uint64_t)t,
                       (uint32_t)k};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ref_philox4x32(ctr, key, r);
}

/* _uniform (src/random.jl:116-120) */
double ref_uniform(uint64_t seed, int64_t cellid, int64_t t, int64_t k) {
    uint32_t r[4];
    draw(seed, cellid, t, k, r);
    uint64_t bits = ((uint64_t)r[0] << 32) | (uint64_t)r[1];
    return (double)(bits >> 11) * 0x1.0p-53;
}

/* _uniform32 (src/random.jl:109-114) */
float ref_uniform32(uint64_t seed, int64_t cellid, int64_t t, int64_t k) {
    uint32_t r[4];
    draw(seed, cellid, t, k, r);
    return (float)(r[0] >> 8) * (1.0f / 16777216.0f);
}

/* bernoulli(ctx, p, k) (src/random.jl:129-130): the draw has the precision of p. */
static int bern(uint64_t seed, int64_t cellid, int64_t t, int64_t k, double p, int p_f32) {
    if (p_f32) return ref_uniform32(seed, cellid, t, k) < (float)p;
    return ref_uniform(seed, cellid, t, k) < p;
}

/* ------------------------------------------------------------------ spaces */

static int64_t fmod64(int64_t a, int64_t n) {
    int64_t m = a % n;
    return m < 0 ? m + n : m;
}

/* _map_axis (src/spaces.jl:46-51) */
int64_t ref_map_axis(int32_t kind, int64_t j, int64_t n, int32_t* outside) {
    *outside = 0;
    if (kind == CAH_BOUNDARY_PERIODIC) return fmod64(j - 1, n) + 1;
    if (kind == CAH_BOUNDARY_REFLECTING) {
        int64_t m = fmod64(j - 1, 2 * n);
        return m < n ? m + 1 : 2 * n - m;
    }
    *outside = !(1 <= j && j <= n);
    return j;
}

/* _offsets_where (src/neighbourhoods.jl:124-153): CartesianIndices((-r:r)^N) in
 * column-major order (first axis fastest), the centre left out. */
int32_t ref_offsets(int32_t kind, int32_t ndims, int32_t r, int32_t* out) {
    int32_t m = 0;
    int32_t r2 = ndims == 2 ? r : 0;
    for (int32_t o2 = -r2; o2 <= r2; ++o2)
        for (int32_t o1 = -r; o1 <= r; ++o1) {
            if (o1 == 0 && o2 == 0) continue;
            if (kind == 1 && abs(o1) + abs(o2) > r) continue;
            out[2 * m] = o1;
            out[2 * m + 1] = o2;
            ++m;
        }
    return m;
}

/* LeniaKernel _kernel_value (src/kernels.jl:96-104) */
static double lenia_value(int32_t R, const double* peaks, int32_t B, double d) {
    double r = d / (double)R;
    if (r <= 0.0 || r >= 1.0) return 0.0;
    double x = (double)B * r;
    int64_t i = (int64_t)floor(x);
    if (i > B - 1) i = B - 1;
    double f = x - (double)i;
    return (0.0 < f && f < 1.0) ? peaks[i] * exp(4.0 - 1.0 / (f * (1.0 - f))) : 0.0;
}

/* kernel_weights (src/kernels.jl:174-189) with _distance (src/kernels.jl:154) */
int32_t ref_lenia_kernel(int32_t R, const double* peaks, int32_t npeaks, int32_t* offs,
                         double* w) {
    int32_t m = 0;
    for (int32_t o2 = -R; o2 <= R; ++o2)
        for (int32_t o1 = -R; o1 <= R; ++o1) {
            double d = sqrt((double)(o1 * o1 + o2 * o2));
            if (!(d <= (double)R + 1e-9)) continue;
            double v = lenia_value(R, peaks, npeaks, d);
            if (v > 0.0) {
                offs[2 * m] = o1;
                offs[2 * m + 1] = o2;
                w[m] = v;
                ++m;
            }
        }
    double s = 0.0;
    for (int32_t k = 0; k < m; ++k) s += w[k];
    for (int32_t k = 0; k < m; ++k) w[k] = w[k] / s;
    return m;
}

/* ------------------------------------------------------------------ cell reads */

typedef struct {
    const cah_desc* d;
    int64_t n1, n2;
    int32_t two_d;
} lat_t;

static lat_t lat_of(const cah_desc* d) {
    lat_t L;
    L.d = d;
    L.n1 = d->dims[0];
    L.n2 = d->ndims == 2 ? d->dims[1] : 1;
    L.two_d = d->ndims == 2;
    return L;
}

/* _boundary_lookup + _boundary_read (src/backends.jl:177-200): 0-based k1, k2;
 * returns the 0-based linear index, or -(axis) when the highest-numbered
 * Constant axis the index lies beyond is `axis`. */
static int64_t lookup(const lat_t* L, int64_t k1, int64_t k2) {
    int32_t out1, out2 = 0;
    int64_t m1 = ref_map_axis(L->d->boundary[0], k1 + 1, L->n1, &out1);
    int64_t m2 = 1;
    if (L->two_d) m2 = ref_map_axis(L->d->boundary[1], k2 + 1, L->n2, &out2);
    if (out2) return -2;
    if (out1) return -1;
    return (m1 - 1) + (m2 - 1) * L->n1;
}

static int8_t read_i8(const lat_t* L, const int8_t* u, int64_t k1, int64_t k2) {
    int64_t j = lookup(L, k1, k2);
    if (j >= 0) return u[j];
    return (int8_t)L->d->boundary_value[-j - 1];
}

static float read_f32(const lat_t* L, const float* u, int64_t k1, int64_t k2) {
    int64_t j = lookup(L, k1, k2);
    if (j >= 0) return u[j];
    return (float)L->d->boundary_value[-j - 1];
}

/* ------------------------------------------------------------------ rules */

typedef struct {
    uint64_t seed;
    int64_t cellid, t;
    const cah_params* p;
    double f64;   /* the field epidemic's field at the cell (precision 64) */
    float f32;    /*   ... (precision 32)                                  */
    double cov;   /*   ... and its covariate (1 without one)               */
} ctx_t;

/* local_rule(::Life) (src/rules.jl:377-389); (mask >> n) & 1 is 0 for n >= 64. */
static int life(const cah_desc* d, int c, const int8_t* nb, int32_t m) {
    int n = 0;
    for (int32_t k = 0; k < m; ++k) n += (nb[k] == 1);
    uint64_t mask = c == 1 ? d->survive : d->birth;
    return n < 64 ? (int)((mask >> n) & 1u) : 0;
}

/* local_rule(::Generations) (src/rules.jl:446-456) */
static int generations(const cah_desc* d, int c, const int8_t* nb, int32_t m) {
    int n = 0;
    for (int32_t k = 0; k < m; ++k) n += (nb[k] == 1);
    int C = d->gen_states;
    if (c == 0) return n < 64 ? (int)((d->birth >> n) & 1u) : 0;
    if (c == 1) {
        if (n < 64 && ((d->survive >> n) & 1u)) return c;
        return C > 2 ? 2 : 0;
    }
    return c + 1 < C ? c + 1 : 0;
}

/* The forest-fire LocalRule (README.md:56-60; vignettes/04_custom_local_rules/
 * 04_custom_local_rules.qmd:129-134) on Int8 states empty=0, tree=1, fire=2, with
 * draw k = 1 for growth and k = 2 for lightning. */
static int forest(const ctx_t* cx, int c, const int8_t* nb, int32_t m) {
    if (c == 2) return 0;
    if (c == 0) return bern(cx->seed, cx->cellid, cx->t, 1, cx->p->ff_p, cx->p->ff_p_is_f32);
    for (int32_t k = 0; k < m; ++k)
        if (nb[k] == 2) return 2;
    return bern(cx->seed, cx->cellid, cx->t, 2, cx->p->ff_f, cx->p->ff_f_is_f32) ? 2 : 1;
}

/* Outcome inversion over (a => 1 - q, b => q) with the draw u (_pick,
 * src/likelihood.jl:74-83): a when 1 - q > 0 and u < 1 - q, otherwise b when
 * q > 0, otherwise a (the last outcome of positive probability, or the first). */
static int pick64(int a, int b, double q, double u) {
    double pa = 1.0 - q;
    if (pa > 0 && u < pa) return a;
    return q > 0 ? b : a;
}
static int pick32(int a, int b, float q, float u) {
    float pa = 1.0f - q;
    if (pa > 0 && u < pa) return a;
    return q > 0 ? b : a;
}

/* The kernel-field epidemic on Int8 states S=0, I=1, R=2, after the kernel-field
 * rules of test/backend_conformance.jl:63-68 (LocalRule, Float64 with a
 * covariate), :82-85 (LocalRule, Float32) and :101-104 (ProbabilisticRule,
 * Float32):
 *   S: q = 1 - exp(-(beta * cov) * field), Julia's promotion deciding the type:
 *      Float32 only when beta and the field are and there is no covariate;
 *      LocalRule form: infected iff bernoulli(ctx, q, 1); stated form: the
 *      inversion of (0 => 1 - q, 1 => q) with uniform(ctx, 1);
 *   I: with recovery, removed iff bernoulli(ctx, gamma, 2) (stated form: the
 *      inversion of (1 => 1 - gamma, 2 => gamma) with uniform(ctx, 1)); else I;
 *   anything else keeps its state. */
static int epidemic(const cah_desc* d, const ctx_t* cx, int c) {
    const cah_params* p = cx->p;
    if (c == 1) {
        if (!d->epi_recover) return 1;
        if (d->epi_form == CAH_EPI_STATED) {
            if (p->ff_f_is_f32)
                return pick32(1, 2, (float)p->ff_f, ref_uniform32(cx->seed, cx->cellid, cx->t, 1));
            return pick64(1, 2, p->ff_f, ref_uniform(cx->seed, cx->cellid, cx->t, 1));
        }
        return bern(cx->seed, cx->cellid, cx->t, 2, p->ff_f, p->ff_f_is_f32) ? 2 : 1;
    }
    if (c != 0) return c;
    int f32 = p->ff_p_is_f32 && d->lenia_precision == 32 && d->covariate == NULL;
    if (f32) {
        float lam = (float)p->ff_p * cx->f32;
        float q = 1.0f - expf(-lam);
        float u = ref_uniform32(cx->seed, cx->cellid, cx->t, 1);
        if (d->epi_form == CAH_EPI_STATED) return pick32(0, 1, q, u);
        return u < q;
    }
    double beta = p->ff_p_is_f32 ? (double)(float)p->ff_p : p->ff_p;
    double f = d->lenia_precision == 32 ? (double)cx->f32 : cx->f64;
    double lam = d->covariate ? (beta * cx->cov) * f : beta * f;
    double q = 1.0 - exp(-lam);
    double u = ref_uniform(cx->seed, cx->cellid, cx->t, 1);
    if (d->epi_form == CAH_EPI_STATED) return pick64(0, 1, q, u);
    return u < q;
}

static int8_t rule_i8(const cah_desc* d, const ctx_t* cx, int8_t c, const int8_t* nb,
                      int32_t m) {
    switch (d->rule) {
    case CAH_RULE_LIFE: return (int8_t)life(d, c, nb, m);
    case CAH_RULE_GENERATIONS: return (int8_t)generations(d, c, nb, m);
    case CAH_RULE_FOREST_FIRE: return (int8_t)forest(cx, c, nb, m);
    case CAH_RULE_FIELD_EPIDEMIC: return (int8_t)epidemic(d, cx, c);
    default: return c;
    }
}

/* local_rule(::Lenia) (src/lenia.jl:42-47) for Lenia() (precision Float64): the
 * growth in FP64 on the Float32 cell, then convert(Float32, .) (round to nearest). */
static float lenia64(const cah_desc* d, float c, double U) {
    double dm = U - d->mu;
    double g = 2.0 * exp(-(dm * dm) / (2.0 * (d->sigma * d->sigma))) - 1.0;
    double x = (double)c + d->dt * g;
    if (x < 0.0) x = 0.0;
    if (x > 1.0) x = 1.0;
    return (float)x;
}

/* The same for Lenia(; precision=Float32): every operation in Float32. */
static float lenia32(const cah_desc* d, float c, float U) {
    float mu = (float)d->mu, sg = (float)d->sigma, dt = (float)d->dt;
    float dm = U - mu;
    float g = 2.0f * expf(-(dm * dm) / (2.0f * (sg * sg))) - 1.0f;
    float x = c + dt * g;
    if (x < 0.0f) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return x;
}

/* _gate_open (src/schemes.jl:84-89); gate 0 or 1 is the Checkerboard half. */
static int gate_open(const cah_desc* d, int gate, uint64_t seed, int64_t cellid, int64_t t,
                     int64_t x, int64_t y) {
    if (d->scheme == CAH_SCHEME_STOCHASTIC) {
        if (d->scheme_p_is_f32) return ref_uniform32(seed, cellid, t, 0) < (float)d->scheme_p;
        return ref_uniform(seed, cellid, t, 0) < d->scheme_p;
    }
    if (d->scheme == CAH_SCHEME_CHECKERBOARD) {
        int64_t s = (x + 1) + (d->ndims == 2 ? (y + 1) : 0);
        return (int)(s & 1) == gate;
    }
    return 1;
}

static int ngates(const cah_desc* d) { return d->scheme == CAH_SCHEME_CHECKERBOARD ? 2 : 1; }

/* ------------------------------------------------------------------ plain engine */

/* _compute_field_reference! (src/environment.jl:151-160) in the weight precision. */
static void field_plain64(const lat_t* L, const float* u, double* out) {
    const cah_desc* d = L->d;
    for (int64_t y = 0; y < L->n2; ++y)
        for (int64_t x = 0; x < L->n1; ++x) {
            double s = 0.0;
            for (int32_t k = 0; k < d->n_taps; ++k)
                s += d->tap_weights[k] *
                     (double)read_f32(L, u, x + d->tap_offsets[2 * k], y + d->tap_offsets[2 * k + 1]);
            out[x + y * L->n1] = s;
        }
}

static void field_plain32(const lat_t* L, const float* u, float* out) {
    const cah_desc* d = L->d;
    for (int64_t y = 0; y < L->n2; ++y)
        for (int64_t x = 0; x < L->n1; ++x) {
            float s = 0.0f;
            for (int32_t k = 0; k < d->n_taps; ++k)
                s += (float)d->tap_weights[k] *
                     read_f32(L, u, x + d->tap_offsets[2 * k], y + d->tap_offsets[2 * k + 1]);
            out[x + y * L->n1] = s;
        }
}

int ref_lenia_field(const cah_desc* d, const float* u, double* out) {
    lat_t L = lat_of(d);
    field_plain64(&L, u, out);
    return 0;
}

/* The epidemic's KernelField at one cell (_compute_field_reference!,
 * src/environment.jl:151-160, source ==(1)): s += w * (x == 1) in tap order, in
 * the field's precision; w * false is +0, which leaves s as it is. */
static double epi_field64(const lat_t* L, const int8_t* u, int64_t x, int64_t y) {
    const cah_desc* d = L->d;
    double s = 0.0;
    for (int32_t k = 0; k < d->n_taps; ++k)
        s += d->tap_weights[k] *
             (double)(read_i8(L, u, x + d->tap_offsets[2 * k], y + d->tap_offsets[2 * k + 1]) == 1);
    return s;
}

static float epi_field32(const lat_t* L, const int8_t* u, int64_t x, int64_t y) {
    const cah_desc* d = L->d;
    float s = 0.0f;
    for (int32_t k = 0; k < d->n_taps; ++k)
        s += (float)d->tap_weights[k] *
             (float)(read_i8(L, u, x + d->tap_offsets[2 * k], y + d->tap_offsets[2 * k + 1]) == 1);
    return s;
}

/* Field values of the epidemic over the lattice, widened to double. */
int ref_epi_field(const cah_desc* d, const int8_t* u, double* out) {
    lat_t L = lat_of(d);
    for (int64_t y = 0; y < L.n2; ++y)
        for (int64_t x = 0; x < L.n1; ++x)
            out[x + y * L.n1] = d->lenia_precision == 32 ? (double)epi_field32(&L, u, x, y)
                                                          : epi_field64(&L, u, x, y);
    return 0;
}

/* The epidemic's per-cell inputs: field (both precisions) and covariate at time t. */
static void epi_inputs(const lat_t* L, const int8_t* u, int64_t x, int64_t y, int64_t t, ctx_t* cx) {
    const cah_desc* d = L->d;
    if (d->lenia_precision == 32) cx->f32 = epi_field32(L, u, x, y);
    else cx->f64 = epi_field64(L, u, x, y);
    cx->cov = 1.0;
    if (d->covariate) {
        int64_t n = L->n1 * L->n2, sl = t % d->covariate_slices;
        cx->cov = d->covariate[sl * n + x + y * L->n1];
    }
}

/* _plain_sweep! (src/backends.jl:264-278).  `fs` (field epidemic) is the state
 * the step's field is computed from: the state at the start of the step, for
 * every gate (src/backends.jl:314-322). */
static void sweep_plain_i8(const lat_t* L, const int8_t* cur, int8_t* next, int64_t t,
                           int gate, const cah_params* p, const int8_t* fs) {
    const cah_desc* d = L->d;
    int8_t nb[REF_MAX_NB];
    for (int64_t y = 0; y < L->n2; ++y)
        for (int64_t x = 0; x < L->n1; ++x) {
            int64_t lin = x + y * L->n1;
            ctx_t cx = {d->seed, lin + 1, t, p, 0.0, 0.0f, 1.0};
            if (d->rule == CAH_RULE_FIELD_EPIDEMIC) epi_inputs(L, fs, x, y, t, &cx);
            int8_t c = cur[lin];
            if ((d->mask && !d->mask[lin]) || !gate_open(d, gate, d->seed, lin + 1, t, x, y)) {
                next[lin] = c;
                continue;
            }
            for (int32_t k = 0; k < d->n_offsets; ++k)
                nb[k] = read_i8(L, cur, x + d->offsets[2 * k], y + d->offsets[2 * k + 1]);
            next[lin] = rule_i8(d, &cx, c, nb, d->n_offsets);
        }
}

static void sweep_plain_lenia(const lat_t* L, const float* cur, float* next, int64_t t,
                              int gate, const double* f64, const float* f32) {
    const cah_desc* d = L->d;
    for (int64_t y = 0; y < L->n2; ++y)
        for (int64_t x = 0; x < L->n1; ++x) {
            int64_t lin = x + y * L->n1;
            float c = cur[lin];
            if ((d->mask && !d->mask[lin]) || !gate_open(d, gate, d->seed, lin + 1, t, x, y)) {
                next[lin] = c;
                continue;
            }
            next[lin] = d->lenia_precision == 32 ? lenia32(d, c, f32[lin]) : lenia64(d, c, f64[lin]);
        }
}

/* _advance_plain! (src/backends.jl:314-322): fields once, then one sweep and swap
 * per gate; nsteps steps from t0. */
int ref_advance_plain(const cah_desc* d, void* u, int64_t t0, int64_t nsteps,
                      const cah_params* p) {
    if (d->n_offsets > REF_MAX_NB) return -1;
    lat_t L = lat_of(d);
    int64_t n = L.n1 * L.n2;
    size_t es = d->dtype == CAH_DTYPE_F32 ? 4 : 1;
    void* a = u;
    void* b = malloc((size_t)n * es);
    double* f64 = NULL;
    float* f32 = NULL;
    if (d->rule == CAH_RULE_LENIA) {
        f64 = (double*)malloc((size_t)n * sizeof(double));
        f32 = (float*)malloc((size_t)n * sizeof(float));
    }
    /* the epidemic's field reads the state at the start of the step in every gate */
    int8_t* start = d->rule == CAH_RULE_FIELD_EPIDEMIC ? (int8_t*)malloc((size_t)n) : NULL;
    for (int64_t s = 0; s < nsteps; ++s) {
        int64_t t = t0 + s;
        if (d->rule == CAH_RULE_LENIA) {
            if (d->lenia_precision == 32) field_plain32(&L, (const float*)a, f32);
            else field_plain64(&L, (const float*)a, f64);
        }
        if (start) memcpy(start, a, (size_t)n);
        for (int g = 0; g < ngates(d); ++g) {
            if (d->rule == CAH_RULE_LENIA)
                sweep_plain_lenia(&L, (const float*)a, (float*)b, t, g, f64, f32);
            else
                sweep_plain_i8(&L, (const int8_t*)a, (int8_t*)b, t, g, p, start);
            void* tmp = a;
            a = b;
            b = tmp;
        }
    }
    if (a != u) {
        memcpy(u, a, (size_t)n * es);
        b = a;
    }
    free(b);
    free(f64);
    free(f32);
    free(start);
    return 0;
}

/* update_cell (src/backends.jl:294-307): masked cells keep their state; no gate. */
int ref_update_cell(const cah_desc* d, const void* u, int64_t i1, int64_t i2, int64_t t,
                    const cah_params* p) {
    lat_t L = lat_of(d);
    const int8_t* cur = (const int8_t*)u;
    int64_t lin = i1 + i2 * L.n1;
    int8_t c = cur[lin];
    if (d->mask && !d->mask[lin]) return c;
    int8_t nb[REF_MAX_NB];
    for (int32_t k = 0; k < d->n_offsets; ++k)
        nb[k] = read_i8(&L, cur, i1 + d->offsets[2 * k], i2 + d->offsets[2 * k + 1]);
    ctx_t cx = {d->seed, lin + 1, t, p, 0.0, 0.0f, 1.0};
    if (d->rule == CAH_RULE_FIELD_EPIDEMIC) epi_inputs(&L, cur, i1, i2, t, &cx);
    return rule_i8(d, &cx, c, nb, d->n_offsets);
}

/* ------------------------------------------------------------------ halo engine */

typedef struct {
    lat_t L;
    int32_t h;
    int64_t P1, P2; /* padded dims */
    int64_t* lin_off; /* neighbour offsets as padded linear deltas */
    int64_t* tap_off;
} halo_t;

static int32_t max_abs_offset(const int32_t* o, int32_t m) {
    int32_t h = 0;
    for (int32_t k = 0; k < 2 * m; ++k) h = abs(o[k]) > h ? abs(o[k]) : h;
    return h;
}

/* _fill_halo! (src/backends.jl:425-442): serial, one slab per axis and side,
 * each halo cell written from _boundary_read_padded. */
static void fill_halo(const halo_t* H, void* cur, size_t es) {
    const lat_t* L = &H->L;
    int32_t h = H->h;
    for (int axis = 0; axis < 2; ++axis)
        for (int side = 0; side < 2; ++side) {
            int64_t lo1 = 0, hi1 = H->P1, lo2 = 0, hi2 = H->P2;
            if (axis == 0) {
                lo1 = side == 0 ? 0 : L->n1 + h;
                hi1 = side == 0 ? h : L->n1 + 2 * h;
            } else {
                lo2 = side == 0 ? 0 : L->n2 + h;
                hi2 = side == 0 ? h : L->n2 + 2 * h;
            }
            for (int64_t q = lo2; q < hi2; ++q)
                for (int64_t r = lo1; r < hi1; ++r) {
                    int64_t j = lookup(L, r - h, q - h);
                    int64_t P = r + q * H->P1;
                    if (es == 1) {
                        int8_t v = j >= 0 ? ((int8_t*)cur)[(j % L->n1) + h + (j / L->n1 + h) * H->P1]
                                          : (int8_t)L->d->boundary_value[-j - 1];
                        ((int8_t*)cur)[P] = v;
                    } else {
                        float v = j >= 0 ? ((float*)cur)[(j % L->n1) + h + (j / L->n1 + h) * H->P1]
                                         : (float)L->d->boundary_value[-j - 1];
                        ((float*)cur)[P] = v;
                    }
                }
        }
}

/* _update_cell! over one slab (src/backends.jl:444-479) */
static void sweep_slab_i8(const halo_t* H, const int8_t* cur, int8_t* next, int64_t t,
                          int gate, const cah_params* p, int64_t ylo, int64_t yhi) {
    const lat_t* L = &H->L;
    const cah_desc* d = L->d;
    int8_t nb[REF_MAX_NB];
    int32_t m = d->n_offsets;
    for (int64_t y = ylo; y < yhi; ++y)
        for (int64_t x = 0; x < L->n1; ++x) {
            int64_t lin = x + y * L->n1;
            int64_t P = (x + H->h) + (y + H->h) * H->P1;
            int8_t c = cur[P];
            if ((d->mask && !d->mask[lin]) || !gate_open(d, gate, d->seed, lin + 1, t, x, y)) {
                next[P] = c;
                continue;
            }
            for (int32_t k = 0; k < m; ++k) nb[k] = cur[P + H->lin_off[k]];
            ctx_t cx = {d->seed, lin + 1, t, p, 0.0, 0.0f, 1.0};
            next[P] = rule_i8(d, &cx, c, nb, m);
        }
}

static void sweep_slab_lenia(const halo_t* H, const float* cur, float* next, int64_t t,
                             int gate, int64_t ylo, int64_t yhi) {
    const lat_t* L = &H->L;
    const cah_desc* d = L->d;
    for (int64_t y = ylo; y < yhi; ++y)
        for (int64_t x = 0; x < L->n1; ++x) {
            int64_t lin = x + y * L->n1;
            int64_t P = (x + H->h) + (y + H->h) * H->P1;
            float c = cur[P];
            if ((d->mask && !d->mask[lin]) || !gate_open(d, gate, d->seed, lin + 1, t, x, y)) {
                next[P] = c;
                continue;
            }
            /* _field_at (src/environment.jl:163-169): the field of the step-start
             * state, evaluated here because Lenia reads nothing else from it */
            float out;
            if (d->lenia_precision == 32) {
                float s = 0.0f;
                for (int32_t k = 0; k < d->n_taps; ++k)
                    s += (float)d->tap_weights[k] * cur[P + H->tap_off[k]];
                out = lenia32(d, c, s);
            } else {
                double s = 0.0;
                for (int32_t k = 0; k < d->n_taps; ++k)
                    s += d->tap_weights[k] * (double)cur[P + H->tap_off[k]];
                out = lenia64(d, c, s);
            }
            next[P] = out;
        }
}

int ref_advance_halo(const cah_desc* d, void* u, int64_t t0, int64_t nsteps,
                     const cah_params* p, int32_t nthreads) {
    if (d->n_offsets > REF_MAX_NB) return -1;
    halo_t H;
    H.L = lat_of(d);
    int32_t hn = max_abs_offset(d->offsets, d->n_offsets);
    int32_t hf = d->rule == CAH_RULE_LENIA ? max_abs_offset(d->tap_offsets, d->n_taps) : 0;
    /* _halo_width (src/backends.jl:405) */
    H.h = hn > hf ? hn : hf;
    if (H.h < 1) H.h = 1;
    H.P1 = H.L.n1 + 2 * H.h;
    H.P2 = H.L.n2 + 2 * H.h;
    if (d->rule == CAH_RULE_LENIA && ngates(d) > 1) return -1; /* field reuse across gates */
    if (d->rule == CAH_RULE_FIELD_EPIDEMIC) return -1;  /* restated in the plain engine only */
    size_t es = d->dtype == CAH_DTYPE_F32 ? 4 : 1;
    H.lin_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d->n_offsets + 1));
    H.tap_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d->n_taps + 1));
    for (int32_t k = 0; k < d->n_offsets; ++k)
        H.lin_off[k] = d->offsets[2 * k] + (int64_t)d->offsets[2 * k + 1] * H.P1;
    for (int32_t k = 0; k < d->n_taps; ++k)
        H.tap_off[k] = d->tap_offsets[2 * k] + (int64_t)d->tap_offsets[2 * k + 1] * H.P1;
    size_t pn = (size_t)(H.P1 * H.P2);
    char* a = (char*)calloc(pn, es);
    char* b = (char*)calloc(pn, es);
    for (int64_t y = 0; y < H.L.n2; ++y)
        memcpy(a + ((size_t)H.h + (size_t)(y + H.h) * (size_t)H.P1) * es,
               (const char*)u + (size_t)(y * H.L.n1) * es, (size_t)H.L.n1 * es);
    int32_t nt = nthreads < 1 ? 1 : nthreads;
    int64_t nlast = H.L.n2;
    int64_t nchunks = nlast < 4 * (int64_t)nt ? nlast : 4 * (int64_t)nt;
    for (int64_t s = 0; s < nsteps; ++s) {
        int64_t t = t0 + s;
        for (int g = 0; g < ngates(d); ++g) {
            fill_halo(&H, a, es);
            int64_t c;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
            for (c = 0; c < nchunks; ++c) {
                int64_t lo = (c * nlast) / nchunks, hi = ((c + 1) * nlast) / nchunks;
                if (d->rule == CAH_RULE_LENIA)
                    sweep_slab_lenia(&H, (const float*)a, (float*)b, t, g, lo, hi);
                else
                    sweep_slab_i8(&H, (const int8_t*)a, (int8_t*)b, t, g, p, lo, hi);
            }
            char* tmp = a;
            a = b;
            b = tmp;
        }
    }
    for (int64_t y = 0; y < H.L.n2; ++y)
        memcpy((char*)u + (size_t)(y * H.L.n1) * es,
               a + ((size_t)H.h + (size_t)(y + H.h) * (size_t)H.P1) * es, (size_t)H.L.n1 * es);
    free(a);
    free(b);
    free(H.lin_off);
    free(H.tap_off);
    return 0;
}

/* ------------------------------------------------------------------ synthetic inputs */

static double gen_uniform(uint64_t seed, int64_t cellid) {
    uint64_t id = (uint64_t)cellid;
    uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), 0xFFFFFFFFu, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    ref_philox4x32(ctr, key, r);
    uint64_t bits = ((uint64_t)r[0] << 32) | (uint64_t)r[1];
    return (double)(bits >> 11) * 0x1.0p-53;
}

void ref_fill_bernoulli(uint64_t seed, int64_t n, double p, int8_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = gen_uniform(seed, i + 1) < p ? 1 : 0;
}

void ref_fill_uniform_box(uint64_t seed, int64_t n1, int64_t n2, int64_t lo1, int64_t len1,
                          int64_t lo2, int64_t len2, float* out) {
    for (int64_t y = 0; y < n2; ++y)
        for (int64_t x = 0; x < n1; ++x) {
            int64_t i = x + y * n1;
            int in = x >= lo1 && x < lo1 + len1 && y >= lo2 && y < lo2 + len2;
            out[i] = in ? (float)gen_uniform(seed, i + 1) : 0.0f;
        }
}

/* Layout of cah_desc / cah_params as this C compiler lays them out, for the
 * test that the ctypes mirrors agree: sizeof, then offsetof of every field. */
#include <stddef.h>
int ref_desc_layout(int64_t* out) {
    int k = 0;
    out[k++] = (int64_t)sizeof(cah_desc);
    out[k++] = (int64_t)offsetof(cah_desc, ndims);
    out[k++] = (int64_t)offsetof(cah_desc, dtype);
    out[k++] = (int64_t)offsetof(cah_desc, dims);
    out[k++] = (int64_t)offsetof(cah_desc, boundary);
    out[k++] = (int64_t)offsetof(cah_desc, boundary_value);
    out[k++] = (int64_t)offsetof(cah_desc, scheme);
    out[k++] = (int64_t)offsetof(cah_desc, scheme_p_is_f32);
    out[k++] = (int64_t)offsetof(cah_desc, scheme_p);
    out[k++] = (int64_t)offsetof(cah_desc, seed);
    out[k++] = (int64_t)offsetof(cah_desc, rule);
    out[k++] = (int64_t)offsetof(cah_desc, gen_states);
    out[k++] = (int64_t)offsetof(cah_desc, birth);
    out[k++] = (int64_t)offsetof(cah_desc, survive);
    out[k++] = (int64_t)offsetof(cah_desc, n_offsets);
    out[k++] = (int64_t)offsetof(cah_desc, lenia_precision);
    out[k++] = (int64_t)offsetof(cah_desc, offsets);
    out[k++] = (int64_t)offsetof(cah_desc, n_taps);
    out[k++] = (int64_t)offsetof(cah_desc, field_method);
    out[k++] = (int64_t)offsetof(cah_desc, tap_offsets);
    out[k++] = (int64_t)offsetof(cah_desc, tap_weights);
    out[k++] = (int64_t)offsetof(cah_desc, mu);
    out[k++] = (int64_t)offsetof(cah_desc, sigma);
    out[k++] = (int64_t)offsetof(cah_desc, dt);
    out[k++] = (int64_t)offsetof(cah_desc, mask);
    out[k++] = (int64_t)offsetof(cah_desc, epi_form);
    out[k++] = (int64_t)offsetof(cah_desc, epi_recover);
    out[k++] = (int64_t)offsetof(cah_desc, covariate);
    out[k++] = (int64_t)offsetof(cah_desc, covariate_slices);
    out[k++] = (int64_t)sizeof(cah_params);
    out[k++] = (int64_t)offsetof(cah_params, ff_p);
    out[k++] = (int64_t)offsetof(cah_params, ff_f);
    out[k++] = (int64_t)offsetof(cah_params, ff_p_is_f32);
    out[k++] = (int64_t)offsetof(cah_params, ff_f_is_f32);
    return k;
}
