/*
 * cahip.h -- C ABI of the MI355X backend for the synchronous stencil sweep of
 * CellularAutomataCore.jl.
 *
 * The reference backend interface is a set of Julia generic functions dispatched
 * on the backend type (src/backends.jl:1-14, src/backends.jl:89-148).  Julia
 * reaches this library through `ccall`; every entry point below names the
 * reference function it replaces.  Plain pointers and sizes only: the caller owns
 * host buffers, the library owns device buffers.
 *
 * Conventions
 *   - Lattices are stored column-major (first axis fastest), as Julia arrays are.
 *   - `cellid` is the 1-based linear index of a cell on the unpadded lattice
 *     (src/backends.jl:446, :458); `t` is the time of the state being read.
 *   - Every call returns CAH_OK (0) or a negative CAH_ERR_* code, and leaves a
 *     message for cah_last_error() on failure.
 *   - Calls on distinct handles may run concurrently from different host
 *     threads; each state owns one HIP stream.  Calls on one handle must not
 *     overlap.
 */
#ifndef CAHIP_H
#define CAHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- error codes
 * Mapped by the Julia glue onto the reference's exception types
 * (src/errors.jl:28-132): */
#define CAH_OK                      0
#define CAH_ERR_ARGUMENT           -1  /* ArgumentError                  */
#define CAH_ERR_INCOMPATIBLE_SPACE -2  /* IncompatibleSpaceError         */
#define CAH_ERR_UNSUPPORTED        -3  /* UnsupportedCombinationError    */
#define CAH_ERR_OUT_OF_MEMORY      -4  /* OutOfMemoryError               */
#define CAH_ERR_STATE_SHAPE        -5  /* StateShapeError                */
#define CAH_ERR_INVALID_STATE      -6  /* InvalidStateError              */
#define CAH_ERR_DEVICE             -7  /* a HIP runtime failure          */
#define CAH_ERR_BUSY               -8  /* server: admission refused now  */

/* ---------------------------------------------------------------- enums */
#define CAH_DTYPE_I8   1   /* Int8 states                                 */
#define CAH_DTYPE_F32  2   /* Float32 states (Lenia)                      */

/* Boundary of one lattice axis (src/spaces.jl:16-51). */
#define CAH_BOUNDARY_PERIODIC   0
#define CAH_BOUNDARY_REFLECTING 1
#define CAH_BOUNDARY_CONSTANT   2

/* Gated schemes (src/schemes.jl:22-93). */
#define CAH_SCHEME_SYNCHRONOUS  0
#define CAH_SCHEME_STOCHASTIC   1
#define CAH_SCHEME_CHECKERBOARD 2

/* Rule families the backend runs (check_problem refuses anything else). */
#define CAH_RULE_LIFE        1  /* Life(birth, survive)       src/rules.jl:354-389 */
#define CAH_RULE_GENERATIONS 2  /* Generations(b, s, C)       src/rules.jl:424-456 */
#define CAH_RULE_FOREST_FIRE 3  /* the forest-fire LocalRule of README.md:56-60,
                                   Int8 states empty=0, tree=1, fire=2 */
#define CAH_RULE_LENIA       4  /* Lenia(; R, mu, sigma, dt)  src/lenia.jl:17-47   */
#define CAH_RULE_FIELD_EPIDEMIC 5 /* an SI / SIR epidemic driven by a KernelField
                                   whose source is the infected cells (==(1)),
                                   src/environment.jl:5-69: the kernel-field
                                   LocalRule / ProbabilisticRule cases of
                                   test/backend_conformance.jl:63-68, 82-104.
                                   Int8 states S=0, I=1, R=2.  A susceptible
                                   cell is infected with probability
                                   q = 1 - exp(-beta * covariate * field);
                                   an infected one is removed with probability
                                   gamma when epi_recover is set.            */

/* How the epidemic draws its outcome (src/likelihood.jl:67-98). */
#define CAH_EPI_BERNOULLI 0  /* LocalRule: bernoulli(ctx, q, 1) infects,
                                bernoulli(ctx, gamma, 2) removes            */
#define CAH_EPI_STATED    1  /* ProbabilisticRule outcomes (0 => 1-q, 1 => q),
                                (1 => 1-gamma, 2 => gamma), one inversion
                                draw uniform(ctx, 1) per cell              */

/* How the Lenia potential (its :potential KernelField) is computed. */
#define CAH_FIELD_DIRECT_SEQ  0  /* one sequential sum per cell in tap order, in the
                                    rule's precision (FP64 for Lenia(), FP32 for
                                    precision=Float32), as _field_at
                                    (src/environment.jl:163-169)                   */
#define CAH_FIELD_DIRECT_F32  1  /* FP32 FMAs over a dense weight table             */
#define CAH_FIELD_MFMA_F16    2  /* implicit GEMM on MFMA, FP16 in, FP32 accumulate  */
#define CAH_FIELD_MFMA_BF16   3  /* implicit GEMM on MFMA, BF16 in, FP32 accumulate  */
#define CAH_FIELD_FFT_F32     4  /* FP32 FFT cross-correlation, periodic lattices
                                    (ext/CellularAutomataCoreFFTWExt.jl:26-30)     */
#define CAH_FIELD_MFMA_F16X2  5  /* split-FP16 (hi+lo) MFMA, FP32 accumulate         */

/* ---------------------------------------------------------------- descriptors */

/* Everything a CAProblem fixes for the sweep (src/problem.jl:38-92), resolved on
 * the host: the neighbourhood as column-major offsets (src/neighbourhoods.jl:124-153),
 * the Lenia kernel as read offsets and normalised weights
 * (src/environment.jl:59-69, src/kernels.jl:174-189). */
typedef struct cah_desc {
    int32_t  ndims;              /* 1 or 2                                      */
    int32_t  dtype;              /* CAH_DTYPE_*                                 */
    int64_t  dims[2];            /* size(space); dims[1] = 1 when ndims == 1    */
    int32_t  boundary[2];        /* CAH_BOUNDARY_* per axis                      */
    double   boundary_value[2];  /* Constant(value) per axis                    */
    int32_t  scheme;             /* CAH_SCHEME_*                                */
    int32_t  scheme_p_is_f32;    /* Stochastic{Float32}: the gate uses the
                                    24-bit draw (src/schemes.jl:86-88)          */
    double   scheme_p;           /* Stochastic(p)                               */
    uint64_t seed;               /* prob.seed                                   */
    int32_t  rule;               /* CAH_RULE_*                                  */
    int32_t  gen_states;         /* Generations: C                              */
    uint64_t birth;              /* Life/Generations birth mask (bit n)         */
    uint64_t survive;            /* Life/Generations survive mask               */
    int32_t  n_offsets;          /* length of the resolved neighbourhood        */
    int32_t  lenia_precision;    /* Lenia: 64 (Lenia()) or 32 (precision=Float32) */
    const int32_t* offsets;      /* n_offsets pairs (o1, o2), in order          */
    int32_t  n_taps;             /* Lenia field: number of taps                 */
    int32_t  field_method;       /* CAH_FIELD_*                                 */
    const int32_t* tap_offsets;  /* n_taps pairs: read offsets (-o)             */
    const double*  tap_weights;  /* n_taps FP64 weights, summing to one         */
    double   mu, sigma, dt;      /* Lenia growth parameters                     */
    const uint8_t* mask;         /* NULL, or prod(dims) bytes; 0 = never update
                                    (src/environment.jl:126-148)                */
    /* CAH_RULE_FIELD_EPIDEMIC: the field's taps are tap_offsets / tap_weights
     * (read offsets, normalised weights), its precision lenia_precision (64 or
     * 32, KernelField(...; precision)); beta and gamma are cah_params ff_p and
     * ff_f.  The arithmetic is Float32 when beta, the field and the covariate
     * are (no covariate counts as Float32), else Float64, as Julia promotes. */
    int32_t  epi_form;           /* CAH_EPI_*                                   */
    int32_t  epi_recover;        /* 1: SIR (I -> R with gamma), 0: SI           */
    const double* covariate;     /* NULL, or prod(dims) * covariate_slices values,
                                    column-major, slice-major: the multiplier of
                                    beta at time t is slice t mod slices
                                    (covariate(ctx, name), src/environment.jl:93-113) */
    int64_t  covariate_slices;
} cah_desc;

/* Parameters that may change between steps: the integrator's `p`
 * (src/backends.jl:121-130).  The forest fire reads (p, f), the field epidemic
 * (beta, gamma) from the same two slots. */
typedef struct cah_params {
    double  ff_p;      /* forest fire: growth probability ctx.p.p (draw k = 1);
                          field epidemic: ctx.p.beta                     */
    double  ff_f;      /* forest fire: lightning probability ctx.p.f (draw k = 2);
                          field epidemic: ctx.p.gamma                    */
    int32_t ff_p_is_f32;  /* p::Float32: bernoulli uses the 24-bit draw (src/random.jl:129-130) */
    int32_t ff_f_is_f32;
} cah_params;

typedef struct cah_ctx   cah_ctx;
typedef struct cah_state cah_state;
typedef struct cah_batch cah_batch;

/* ---------------------------------------------------------------- context */

/* Open device `device` (HIP ordinal).  No reference counterpart: it is the
 * device half of GPUBackend(device) (src/backends.jl:83-87). */
int  cah_open(int device, cah_ctx** out);
void cah_close(cah_ctx* ctx);

/* Thread-local message of the last failed call on this thread. */
const char* cah_last_error(void);

/* ---------------------------------------------------------------- backend interface */

/* supports(backend, scheme) (src/backends.jl:95, :153): 1 for the gated
 * schemes, 0 otherwise.  `scheme` is CAH_SCHEME_* or any other integer. */
int cah_supports(int scheme);

/* check_problem(backend, prob, scheme) (src/backends.jl:105;
 * ext/CellularAutomataCoreKernelAbstractionsExt.jl:42-49): CAH_OK, or the error
 * the reference raises for a problem the backend cannot run. */
int cah_check_problem(const cah_desc* d);

/* allocate_state(backend, prob, scheme) (src/backends.jl:119, :408-422;
 * ext/...KernelAbstractionsExt.jl:51-59): device buffers for the state and
 * its double buffer.  The state starts all zero; load u0 with cah_state_set. */
int cah_state_create(cah_ctx* ctx, const cah_desc* d, cah_state** out);

/* set_state!(backend, st, u) (src/backends.jl:148, :525-530): copy the whole
 * state from a host array in the layout of u0 (nbytes = prod(dims)*eltype). */
int cah_state_set(cah_state* st, const void* host, size_t nbytes);

/* current_state(backend, st) (src/backends.jl:139, :521-524): synchronise and
 * copy the whole state into a host array the caller owns. */
int cah_state_get(cah_state* st, void* host, size_t nbytes);

/* The same two copies from / into a dense device array on the state's device
 * (column-major, prod(dims) elements), for callers whose data already sits in
 * HBM. */
int cah_state_set_device(cah_state* st, const void* dptr, size_t nbytes);
int cah_state_get_device(cah_state* st, void* dptr, size_t nbytes);

/* read_cell(backend, st, I) / write_cell!(backend, st, I, v)
 * (src/agents.jl:165, :174; src/backends.jl:531-534): one cell, 0-based
 * coordinates (i1, i2); the value is an Int8 or a Float32 per dtype. */
int cah_state_read_cell(cah_state* st, int64_t i1, int64_t i2, void* value);
int cah_state_write_cell(cah_state* st, int64_t i1, int64_t i2, const void* value);

/* state_counts(current_state(backend, st)) (src/measures.jl:14-20) on the
 * device, for Int8 states: counts[v] = cells holding the byte value v (256
 * entries; index a negative Int8 value by v & 0xFF).  With `other` (an Int8
 * state of the same dims, or NULL), *equal = the cells where the two states
 * agree, the count(predicted .== observed) of cohens_kappa
 * (src/evaluation.jl:43-47).  2 KiB leave the device instead of the state;
 * synchronises. */
int cah_state_counts(cah_state* st, const cah_state* other, int64_t* counts, int64_t* equal);

/* advance!(backend, st, prob, scheme, p, t) (src/backends.jl:130, :494-506;
 * ext/...KernelAbstractionsExt.jl:92-110), `nsteps` times: steps t0, t0+1, ...
 * Work is queued on the state's stream; it returns before the GPU finishes
 * (current_state / cah_state_get synchronise). */
int cah_advance(cah_state* st, int64_t t0, int64_t nsteps, const cah_params* p);

/* Block until the state's queued work is done. */
int cah_state_sync(cah_state* st);

/* The state's HIP stream (a hipStream_t), so callers can time its work with
 * HIP events or order their own work after it. */
int cah_state_stream(cah_state* st, void** stream);

/* Launch shape of the band-walk kernels (no reference counterpart; set by the
 * autotuner): vb = dwords per lane (1, 2 or 4), rows = rows per wave; 0 keeps
 * the default.  Results do not depend on it. */
int cah_state_tune(cah_state* st, int32_t vb, int32_t rows);

/* SparseBackend(background; tilesize) (src/sparse.jl:6-40, :126-211) on the
 * device: from now on each step updates only the tiles (tilesize^2 cells, a
 * tilesize-long segment in 1-D) within reach of a tile holding a cell that is not
 * `background`, plus the tiles one step earlier held and the tiles beside a
 * Constant edge whose value is not the background; every other cell stays the
 * background.  The result equals the dense sweep's.  The rule must be quiescent
 * (a background cell with background neighbours stays background): Life and
 * Generations are checked here, the forest fire at every cah_advance (background
 * 0 needs p = 0, background 1 needs f = 0); Synchronous and Stochastic only; no
 * kernel fields (the field epidemic) and no server states.  tilesize in 1..128. */
int cah_state_sparse(cah_state* st, int8_t background, int32_t tilesize);
/* stored_tiles (src/sparse.jl): the tiles that may hold a non-background cell
 * (exact after a step; a write_cell marks its tile). */
int cah_state_active_tiles(cah_state* st, int64_t* n);

/* Release a state; safe from any host thread (Julia finalizers). */
int cah_state_destroy(cah_state* st);

/* ---------------------------------------------------------------- synthetic inputs
 * No reference counterpart.  The build's documented generator of initial states:
 * cell `cellid` gets the draw of an independent Philox4x32-10 stream,
 * ctr = (cellid lo32, cellid hi32, 0xFFFFFFFF, 0), key = (seed lo32, seed hi32),
 * turned into u = ((r0 << 32 | r1) >> 11) * 2^-53. */

/* Int8 state: cell = (u < p) ? 1 : 0. */
int cah_state_fill_bernoulli(cah_state* st, uint64_t seed, double p);
/* Float32 state: cell = Float32(u) inside the box [lo1, lo1+len1) x [lo2, lo2+len2)
 * (0-based), 0 elsewhere. */
int cah_state_fill_uniform_box(cah_state* st, uint64_t seed, int64_t lo1, int64_t len1,
                               int64_t lo2, int64_t len2);

/* ---------------------------------------------------------------- batched ensembles
 * No single reference counterpart: the device form of
 * solve(EnsembleProblem(prob), alg, EnsembleThreads(); trajectories=n)
 * (src/problem.jl:571-589) for trajectories that share a rule and a scheme and
 * differ in seed (and in lattice size).  Trajectory i advances exactly as a
 * cah_state with descriptor descs[i] would. */

/* n descriptors of one rule family (Int8 rules only). */
int cah_batch_create(cah_ctx* ctx, int32_t n, const cah_desc* descs, cah_batch** out);
int cah_batch_set(cah_batch* b, int32_t i, const void* host, size_t nbytes);
int cah_batch_get(cah_batch* b, int32_t i, void* host, size_t nbytes);
int cah_batch_fill_bernoulli(cah_batch* b, int32_t i, uint64_t seed, double p);
/* One launch per step over every trajectory; with use_graph != 0 the nsteps
 * launches are captured once into a hipGraph and replayed. */
int cah_batch_advance(cah_batch* b, int64_t t0, int64_t nsteps, const cah_params* p,
                      int32_t use_graph);
/* cah_state_counts of every trajectory: counts[i*256 + v], equal[i] (other may
 * be NULL; it must match every trajectory's dims otherwise), as an ensemble's
 * output_func that reduces each trajectory to its state counts or its agreement
 * with an observed map (src/calibration.jl:76) would compute them. */
int cah_batch_counts(cah_batch* b, const cah_state* other, int64_t* counts, int64_t* equal);
/* No counterpart in the reference: cah_batch_counts for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same outputs, and it calls no hipMalloc
 * or hipFree: its counters live in a fixed device array of about 2.3 MiB that
 * every device the library opens holds from cah_open until the process ends.
 * That array holds the counters of CAH_COUNTS_NOALLOC_MAX trajectories; a batch
 * of more is refused with CAH_ERR_ARGUMENT.  Every call on one device shares the
 * array: calls that overlap on one device race, so the caller serializes them. */
#define CAH_COUNTS_NOALLOC_MAX 512
int cah_batch_counts_noalloc(cah_batch* b, const cah_state* other, int64_t* counts, int64_t* equal);
/* No reference counterpart: the autotuner of the batched band walk.  Times
 * `steps` steps from t0 for each candidate tile (4 or 8 cells a lane, *vb = 1 or
 * 2, radius-5 Life 8 only; 4, 8, 16, 32 or 64 rows a wave; one tile for every
 * trajectory) with HIP events, keeps the fastest, and leaves every trajectory's
 * state and time as it was (the trial steps run on the lattices and are undone
 * from a device copy, which needs device memory for one copy of every lattice).
 * *vb and *rows are the choice, or 0 when the batch has no band walk to tune
 * (trajectories that advance one by one, Moore(1) Life, which runs k_life_tb).
 * The layout holds until the next cah_batch_autotune. */
int cah_batch_autotune(cah_batch* b, int64_t t0, int32_t steps, const cah_params* p, int32_t* vb,
                       int32_t* rows);
/* No reference counterpart: sets the band walk's tile (vb = 1 or 2 dwords a lane,
 * rows a wave; 0 for the default of either), as a tile cah_batch_autotune chose
 * for an earlier batch of the same shapes.  No effect on batches without a band
 * walk. */
int cah_batch_tune(cah_batch* b, int32_t vb, int32_t rows);
int cah_batch_sync(cah_batch* b);
int cah_batch_stream(cah_batch* b, void** stream);
int cah_batch_destroy(cah_batch* b);

/* ---------------------------------------------------------------- launch checks
 * No reference counterpart: a debugging aid.  Every kernel launch the library
 * makes can be checked on the host before it is queued: a model of the
 * kernel's addressing (the kernel's own index arithmetic, over every wave, lane,
 * block and row of the grid) gives the lowest and highest byte each buffer
 * argument can be read or written at, which must lie inside that buffer's
 * allocation; the grid must cover the work and the LDS size must fit.  A failed
 * check refuses the launch with CAH_ERR_DEVICE.  With the environment variable
 * CAH_CHECK_LAUNCHES set (to anything but 0) when the library loads, every
 * launch is checked against the device allocations.  The two calls below need
 * no device: they build the state (or batch) with fake device addresses and
 * run the launch sequence of cah_advance (cah_batch_advance, without a graph)
 * through the checks only.  `report` (may be NULL) receives one line per kernel:
 * its launches and checks made. */
int cah_check_launches(const cah_desc* d, int64_t t0, int64_t nsteps, const cah_params* p, char* report,
                       size_t cap);
/* vb, rows: the band walk's tile as cah_batch_tune takes it (0, 0: the default). */
int cah_check_batch_launches(int32_t n, const cah_desc* descs, int64_t t0, int64_t nsteps, const cah_params* p,
                             int32_t vb, int32_t rows, char* report, size_t cap);

/* ---------------------------------------------------------------- serving
 * No single reference counterpart: a simulation server for the traffic that
 * ensembles and abc_rejection generate (src/calibration.jl:66-91), with
 * continuous batching.  Requests (any rule the backend runs, any lattice, any
 * tspan) wait in a FIFO until the device-memory budget admits them; admitted
 * lattices stay resident in HBM.  Each tick advances the resident requests by up
 * to `chunk` steps, in arrival order while the tick's work budget lasts, and the
 * requests of one kernel family advance together, one batched launch per step
 * over all their lattices; so requests that arrive while others run start as
 * soon as memory frees up and their work overlaps on the device.  A request's
 * result is exactly what a cah_state with the same descriptor would give.
 * Calls on one server may come from several host threads and are serialized,
 * except that a fetch copies its result out without holding the server: one
 * thread can fetch results while another submits and ticks.  No call may
 * overlap cah_server_destroy. */
typedef struct cah_server cah_server;

#define CAH_REQ_PENDING 0   /* waiting for memory            */
#define CAH_REQ_ACTIVE  1   /* resident, steps in flight     */
#define CAH_REQ_DONE    2   /* finished; result on the device */
#define CAH_REQ_FETCHED 3   /* result copied out, memory freed */

/* budget_bytes: device memory the resident lattices may hold; chunk: steps
 * queued per request per tick at most (0 = no limit but the tick's work). */
int cah_server_create(cah_ctx* ctx, size_t budget_bytes, int32_t chunk, cah_server** out);
/* Steps per request per tick at most (as in create) and the estimated device
 * time one tick may hold, in microseconds (0, the default: no limit, every
 * resident request advances `chunk` steps a tick; a limit gives the steps to the
 * oldest requests first). */
int cah_server_tune(cah_server* s, int32_t chunk, double tick_us);
/* Allocate the server's device arena (lattices, tick tables; at most the
 * budget) and pinned host arena (initial states waiting for admission, uploads,
 * results) up front until they hold device_bytes and pinned_bytes, so that a
 * server sized for its traffic pays no allocation while it serves (each arena
 * otherwise grows by slabs as requests need them).  No counterpart in the
 * reference: a serving loop's start-up sizing. */
int cah_server_reserve(cah_server* s, size_t device_bytes, size_t pinned_bytes);
/* Queue a request: the problem (copied), its initial state (host, copied; NULL =
 * all zero), steps t0 .. t1-1 and parameters.  Returns its id. */
int cah_server_submit(cah_server* s, const cah_desc* d, const void* u0, size_t nbytes,
                      int64_t t0, int64_t t1, const cah_params* p, uint64_t* id);
/* One scheduling pass: retire finished ticks, admit waiting requests that fit
 * the budget, queue the next tick.  Never blocks on the GPU.  *active / *pending
 * (may be NULL) return the counts after the pass. */
int cah_server_tick(cah_server* s, int32_t* active, int32_t* pending);
/* The ids of requests that finished since the last poll, in the order their
 * ticks finished: up to cap of them into ids, their number into *n (the rest
 * wait for the next poll). */
int cah_server_poll(cah_server* s, uint64_t* ids, int32_t cap, int32_t* n);
/* CAH_REQ_* of a request. */
int cah_server_status(cah_server* s, uint64_t id, int32_t* status);
/* Copy a finished request's result out and free its device memory
 * (CAH_ERR_BUSY while another thread fetches it). */
int cah_server_fetch(cah_server* s, uint64_t id, void* host, size_t nbytes);
/* Host timestamps (seconds, steady clock) of submission, admission and the tick
 * that saw the request finish. */
int cah_server_times(cah_server* s, uint64_t id, double* t_submit, double* t_admit, double* t_done);
/* The steady clock the timestamps use, now. */
double cah_server_clock(void);
int cah_server_destroy(cah_server* s);

#ifdef __cplusplus
}
#endif

#endif /* CAHIP_H */
