// cahip.hip -- the C ABI (include/cahip.h) over HIP for gfx950.
//
// One state = two dense device buffers (current / next) with a 64-byte-aligned
// row pitch, one HIP stream, and the resolved rule, neighbourhood, kernel-field
// taps and mask in device memory.  advance queues one sweep launch per gate and
// swaps the buffers; boundary conditions are applied inside the sweep kernels
// while tiles or rows are loaded, so there is no halo buffer and no halo launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/cahip.h"
#include "device.hpp"
#include "host_copy.hpp"
#include "kernels_counts.hpp"
#include "kernels_epidemic.hpp"
#include "kernels_life_tb.hpp"
#include "kernels_box.hpp"
#include "kernels_generic.hpp"
#include "launch.hpp"
#include "lenia_launch.hpp"
#include "kernels_lenia.hpp"
#include "launch_check.hpp"

// ---------------------------------------------------------------- Life dispatch

// Dwords a lane of the Life band walk holds at least: a neighbour lane must hold
// the R guard cells of each side (radius 5 and more take two).
inline int life_min_vb(int R) { return R >= 5 ? 2 : 1; }

inline void launch_life(int R, int vb, bool bin, int niv, int blocks, const LifeArgs& a,
                        hipStream_t st) {
    switch (R) {
    case 1: launch_life_r<1>(vb, bin, niv, blocks, a, st); break;
    case 2: launch_life_r<2>(vb, bin, niv, blocks, a, st); break;
    case 3: launch_life_r<3>(vb, bin, niv, blocks, a, st); break;
    case 4: launch_life_r<4>(vb, bin, niv, blocks, a, st); break;
    case 5: launch_life_r<5>(vb, bin, niv, blocks, a, st); break;
    default: launch_life_r<6>(vb, bin, niv, blocks, a, st); break;
    }
}

inline void launch_life_gated(int R, int vb, int cells, int niv, bool masked, int blocks, const LifeArgs& a,
                              hipStream_t st) {
    switch (R) {
    case 1: launch_life_gated_r<1>(vb, cells, niv, masked, blocks, a, st); break;
    case 2: launch_life_gated_r<2>(vb, cells, niv, masked, blocks, a, st); break;
    case 3: launch_life_gated_r<3>(vb, cells, niv, masked, blocks, a, st); break;
    case 4: launch_life_gated_r<4>(vb, cells, niv, masked, blocks, a, st); break;
    case 5: launch_life_gated_r<5>(vb, cells, niv, masked, blocks, a, st); break;
    default: launch_life_gated_r<6>(vb, cells, niv, masked, blocks, a, st); break;
    }
}

inline int life_gated_blocks_per_cu(int R, int vb, int cells, int niv, bool masked) {
    switch (R) {
    case 1: return life_gated_blocks_per_cu_r<1>(vb, cells, niv, masked);
    case 2: return life_gated_blocks_per_cu_r<2>(vb, cells, niv, masked);
    case 3: return life_gated_blocks_per_cu_r<3>(vb, cells, niv, masked);
    case 4: return life_gated_blocks_per_cu_r<4>(vb, cells, niv, masked);
    case 5: return life_gated_blocks_per_cu_r<5>(vb, cells, niv, masked);
    default: return life_gated_blocks_per_cu_r<6>(vb, cells, niv, masked);
    }
}

inline void launch_life_batch(int R, int vb, bool bin, int niv, int blocks, const LifeArgs* it,
                              const int32_t* wi, const int32_t* wb, int32_t nw, int32_t parity,
                              hipStream_t st) {
    switch (R) {
    case 1: launch_life_batch_r<1>(vb, bin, niv, blocks, it, wi, wb, nw, parity, st); break;
    case 2: launch_life_batch_r<2>(vb, bin, niv, blocks, it, wi, wb, nw, parity, st); break;
    case 3: launch_life_batch_r<3>(vb, bin, niv, blocks, it, wi, wb, nw, parity, st); break;
    case 4: launch_life_batch_r<4>(vb, bin, niv, blocks, it, wi, wb, nw, parity, st); break;
    case 5: launch_life_batch_r<5>(vb, bin, niv, blocks, it, wi, wb, nw, parity, st); break;
    default: launch_life_batch_r<6>(vb, bin, niv, blocks, it, wi, wb, nw, parity, st); break;
    }
}

// ---------------------------------------------------------------- errors

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? CAH_ERR_OUT_OF_MEMORY : CAH_ERR_DEVICE, \
                        "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

extern "C" const char* cah_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------- launch checks
//
// The launches the checks have seen on this thread since the last reset, by
// kernel: (launches, individual checks), for the dry entry points' report.
static thread_local std::map<std::string, std::pair<int64_t, int64_t>> g_tally;

// Runs the launch checks `body` (a function of an lchk::Chk) when they are on:
// for a dry state or batch, or everywhere with CAH_CHECK_LAUNCHES set.  The
// launch must not run when this returns an error.
template <class F>
static int checked(bool dry, const char* kernel, F&& body) {
    if (!dry && !lchk::enabled()) return CAH_OK;
    lchk::Chk c(kernel);
    body(c);
    auto& t = g_tally[kernel];
    t.first += 1;
    t.second += c.checks;
    if (!c.ok()) return fail(CAH_ERR_DEVICE, "launch check failed: %s", c.err.c_str());
    return CAH_OK;
}


// ---------------------------------------------------------------- context

struct cah_ctx {
    int device;
};

// The counters of cah_batch_counts_noalloc: a fixed device array (each device
// the library opens holds its own copy until the process ends), laid out as
// the partial sums of one k_counts launch, then CAH_COUNTS_NOALLOC_MAX lattices'
// CNT_SLOTS counters, then their counts and agreements in the caller's layout.
// Host code reaches it only through the address hipGetSymbolAddress gives for
// the current device (g_scratch_ptr), never through the symbol's name.
constexpr size_t SCRATCH_PART = (size_t)CNT_BLOCKS * CNT_PART;
constexpr size_t SCRATCH_WORDS = SCRATCH_PART + (size_t)CAH_COUNTS_NOALLOC_MAX * (CNT_SLOTS + 256 + 1);
static __device__ __align__(256) unsigned long long scratch[SCRATCH_WORDS];
constexpr int SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned long long* g_scratch_ptr[SCRATCH_DEVICES] = {};

// The scratch's address on `device` (the current device), looked up once.
static int scratch_setup(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return CAH_OK;  // the no-alloc counts refuse it
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return CAH_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (unsigned long long*)p;
    lchk::reg(p, SCRATCH_WORDS * sizeof(unsigned long long), "no-alloc counts scratch");
    return CAH_OK;
}

extern "C" int cah_open(int device, cah_ctx** out) {
    if (!out) return fail(CAH_ERR_ARGUMENT, "cah_open: out is NULL");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(CAH_ERR_ARGUMENT, "device %d does not exist (%d visible)", device, n);
    HIPCHK(hipSetDevice(device));
    if (int rc = scratch_setup(device)) return rc;
    *out = new cah_ctx{device};
    return CAH_OK;
}

extern "C" void cah_close(cah_ctx* ctx) { delete ctx; }

extern "C" int cah_supports(int scheme) {
    return scheme == CAH_SCHEME_SYNCHRONOUS || scheme == CAH_SCHEME_STOCHASTIC ||
           scheme == CAH_SCHEME_CHECKERBOARD;
}

// ---------------------------------------------------------------- problem checks

static int32_t max_abs(const int32_t* o, int32_t n) {
    int32_t h = 0;
    for (int32_t k = 0; k < 2 * n; ++k) h = std::max(h, std::abs(o[k]));
    return h;
}

constexpr int32_t GEN_MAX_HALO = 32;
constexpr int32_t LENIA_MAX_HALO = 16;

extern "C" int cah_check_problem(const cah_desc* d) {
    if (!d) return fail(CAH_ERR_ARGUMENT, "descriptor is NULL");
    if (d->ndims != 1 && d->ndims != 2)
        return fail(CAH_ERR_INCOMPATIBLE_SPACE,
                    "the HIP backend runs one- and two-dimensional square lattices, got %d dimensions",
                    d->ndims);
    const int64_t n2 = d->ndims == 2 ? d->dims[1] : 1;
    if (d->dims[0] < 1 || n2 < 1)
        return fail(CAH_ERR_ARGUMENT, "lattice dimensions must be positive");
    if (d->dims[0] >= (1LL << 30) || n2 > (1LL << 20))
        return fail(CAH_ERR_ARGUMENT, "lattice sides must be below 2^30 (first axis) and 2^20 (second)");
    for (int a = 0; a < d->ndims; ++a)
        if (d->boundary[a] < CAH_BOUNDARY_PERIODIC || d->boundary[a] > CAH_BOUNDARY_CONSTANT)
            return fail(CAH_ERR_ARGUMENT, "unknown boundary on axis %d", a + 1);
    if (!cah_supports(d->scheme))
        return fail(CAH_ERR_UNSUPPORTED, "the HIP backend runs Synchronous, Stochastic and Checkerboard");
    if (d->scheme == CAH_SCHEME_STOCHASTIC && !(d->scheme_p > 0.0 && d->scheme_p <= 1.0))
        return fail(CAH_ERR_ARGUMENT, "update probability must be in (0, 1], got %g", d->scheme_p);
    switch (d->rule) {
    case CAH_RULE_LIFE:
    case CAH_RULE_GENERATIONS:
    case CAH_RULE_FOREST_FIRE:
        if (d->dtype != CAH_DTYPE_I8) return fail(CAH_ERR_ARGUMENT, "this rule runs on Int8 states");
        if (d->n_offsets < 1 || !d->offsets)
            return fail(CAH_ERR_ARGUMENT, "a neighbourhood needs at least one offset");
        if (max_abs(d->offsets, d->n_offsets) > GEN_MAX_HALO)
            return fail(CAH_ERR_ARGUMENT, "neighbourhoods wider than %d cells are not supported",
                        GEN_MAX_HALO);
        if (d->rule == CAH_RULE_GENERATIONS && d->gen_states < 2)
            return fail(CAH_ERR_ARGUMENT, "a Generations rule needs at least 2 states");
        for (int a = 0; a < d->ndims; ++a)
            if (d->boundary[a] == CAH_BOUNDARY_CONSTANT &&
                (d->boundary_value[a] != std::floor(d->boundary_value[a]) ||
                 d->boundary_value[a] < -128 || d->boundary_value[a] > 127))
                return fail(CAH_ERR_INCOMPATIBLE_SPACE,
                            "the boundary Constant(%g): a state of Int8 cannot hold it",
                            d->boundary_value[a]);
        break;
    case CAH_RULE_LENIA:
        if (d->dtype != CAH_DTYPE_F32)
            return fail(CAH_ERR_INVALID_STATE, "Lenia runs on Float32 states here");
        if (d->n_taps < 1 || !d->tap_offsets || !d->tap_weights)
            return fail(CAH_ERR_ARGUMENT, "Lenia needs its kernel taps");
        if (max_abs(d->tap_offsets, d->n_taps) > LENIA_MAX_HALO)
            return fail(CAH_ERR_ARGUMENT, "kernel fields wider than %d cells are not supported",
                        LENIA_MAX_HALO);
        if (d->lenia_precision != 32 && d->lenia_precision != 64)
            return fail(CAH_ERR_ARGUMENT, "Lenia precision must be 32 or 64");
        if (!(d->sigma > 0)) return fail(CAH_ERR_ARGUMENT, "sigma must be positive");
        if (d->field_method < CAH_FIELD_DIRECT_SEQ || d->field_method > CAH_FIELD_MFMA_F16X2)
            return fail(CAH_ERR_ARGUMENT, "unknown field method %d", d->field_method);
        if (d->field_method == CAH_FIELD_FFT_F32) {
            // _resolve_field (src/environment.jl:64-66)
            for (int a = 0; a < d->ndims; ++a)
                if (d->boundary[a] != CAH_BOUNDARY_PERIODIC)
                    return fail(CAH_ERR_ARGUMENT,
                                "field :potential uses method=:fft, which needs periodic boundaries on every axis");
            if (d->ndims != 2)
                return fail(CAH_ERR_UNSUPPORTED, "the FFT field runs on two-dimensional lattices");
        }
        break;
    case CAH_RULE_FIELD_EPIDEMIC:
        if (d->dtype != CAH_DTYPE_I8) return fail(CAH_ERR_ARGUMENT, "the field epidemic runs on Int8 states");
        if (d->n_taps < 1 || !d->tap_offsets || !d->tap_weights)
            return fail(CAH_ERR_ARGUMENT, "the field epidemic needs its kernel field's taps");
        if (max_abs(d->tap_offsets, d->n_taps) > GEN_MAX_HALO || max_abs(d->offsets, d->n_offsets) > GEN_MAX_HALO)
            return fail(CAH_ERR_ARGUMENT, "kernel fields wider than %d cells are not supported", GEN_MAX_HALO);
        if (d->n_offsets < 0 || (d->n_offsets > 0 && !d->offsets))
            return fail(CAH_ERR_ARGUMENT, "bad neighbourhood");
        if (d->lenia_precision != 32 && d->lenia_precision != 64)
            return fail(CAH_ERR_ARGUMENT, "the field's precision must be 32 or 64");
        if (d->epi_form != CAH_EPI_BERNOULLI && d->epi_form != CAH_EPI_STATED)
            return fail(CAH_ERR_ARGUMENT, "unknown epidemic form %d", d->epi_form);
        if (d->epi_recover != 0 && d->epi_recover != 1)
            return fail(CAH_ERR_ARGUMENT, "epi_recover must be 0 or 1");
        if (d->covariate && d->covariate_slices < 1)
            return fail(CAH_ERR_ARGUMENT, "a covariate needs at least one slice");
        for (int a = 0; a < d->ndims; ++a)
            if (d->boundary[a] == CAH_BOUNDARY_CONSTANT &&
                (d->boundary_value[a] != std::floor(d->boundary_value[a]) ||
                 d->boundary_value[a] < -128 || d->boundary_value[a] > 127))
                return fail(CAH_ERR_INCOMPATIBLE_SPACE,
                            "the boundary Constant(%g): a state of Int8 cannot hold it",
                            d->boundary_value[a]);
        break;
    default:
        return fail(CAH_ERR_ARGUMENT,
                    "the HIP backend runs Life, Generations, the forest fire, Lenia and the field "
                    "epidemic (rule %d)",
                    d->rule);
    }
    return CAH_OK;
}

// ---------------------------------------------------------------- state

// P_LIFE_GATED and P_FOREST_GATED: the band walks of Life / Generations and of the
// forest fire under a Stochastic or Checkerboard gate or a mask (Generations
// under any scheme), one sweep a launch.  P_BOX: Life / Generations with the
// neighbourhoods the band walks do not take, Moore(R <= 32) and VonNeumann(r <= 8)
// (k_box_i8, any scheme and mask)
enum Path { P_GENERIC, P_LIFE_SWAR, P_FOREST, P_LENIA_SEQ, P_LENIA_F32COL, P_LENIA_MFMA,
            P_LENIA_FFT, P_LIFE_GATED, P_FOREST_GATED, P_BOX };

struct cah_state {
    int device = 0;
    hipStream_t stream = nullptr;
    cah_desc d{};
    std::vector<int32_t> offsets, taps;
    std::vector<double> weights;
    int64_t n1 = 0, n2 = 0, pitch = 0;  // pitch in elements
    size_t es = 1;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    uint8_t* mask_d = nullptr;
    int32_t* ofs_d = nullptr;
    int32_t* taps_d = nullptr;
    double* w64_d = nullptr;
    float* w32_d = nullptr;    // per-tap Float32 weights
    float* wtab_d = nullptr;   // dense table for k_lenia_f32col
    void* field_d = nullptr;   // field buffer for the sequential path
    void* frag_d = nullptr;    // MFMA A fragments
    size_t frag_bytes = 0;     //   ... bytes of one plane of them
    unsigned long long* counts_d = nullptr;  // CNT_SLOTS counters of cah_state_counts
    double* cov_d = nullptr;   // field epidemic: covariate slices
    // the sparse sweep (cah_state_sparse): two candidate lists used in alternate
    // steps, three counters for them in rotation, per-tile stamps and edge flags,
    // the edge tiles, two counters of non-background tiles, and while the state
    // is rescanned, per-tile flags and the list of non-background tiles
    bool sparse = false, sp_dirty = false;  // dirty: a cell was written since the scan
    int8_t sp_bg = 0;
    int32_t sp_ts = 0, sp_tsy = 0, sp_ntx = 0, sp_nty = 0, sp_rx = 0, sp_ry = 0, sp_nedge = 0;
    int32_t* sp_cand[2] = {nullptr, nullptr};
    int32_t* sp_cnt = nullptr;     // 3 candidate counters, then 2 active counters, then 1 scan counter
    int32_t* sp_stamp = nullptr;
    uint8_t* sp_edgef = nullptr;
    int32_t* sp_edges = nullptr;
    uint8_t* sp_tmp = nullptr;
    int32_t* sp_act = nullptr;
    int32_t sp_step = 0;           // steps since the scan
    std::vector<double> cov;   //   ... host copy while the state is created
    int32_t epi_h = 0;         //   ... its field's halo
    int lenia_mode = 0;        // LM_F16 / LM_BF16 / LM_F16X2
    LeniaFFT* fft = nullptr;
    int32_t halo = 1;
    int32_t lenia_H = 0;
    int32_t lenia_dy0 = 0, lenia_np = 0;  // MFMA field: first row offset, row-offset pairs
    Path path = P_GENERIC;
    int32_t life_R = 0;
    bool box_vn = false;     // P_BOX: von Neumann (else Moore) of radius halo
    bool bin_clean = false;  // every cell is 0 or 1 (Life only)
    bool tri_clean = false;  // every cell is 0, 1 or 2 (the forest fire's k_forest_vn1 variant)
    bool small_clean = false;  // every cell is 0, 1, 2 or 3 (k_life_gated's small-cell form)
    LifeArgs life{};
    int32_t tune_vb = 0, tune_rows = 0;
    // k_life_gated blocks a CU holds (by cell kind, and with a mask), -1: not asked yet
    int gated_occ[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
    struct Arena* arena = nullptr;        // device memory source (nullptr: hipMalloc)
    struct Arena* staging = nullptr;      // pinned host memory of a server's uploads
    bool own_stream = true;               // false: the stream belongs to a server
    std::vector<void*> keep;              // pinned sources of queued uploads
    bool dry = false;  // fake device addresses, no HIP call: launches are checked, not run
};

static Lat lat_of(const cah_state* s) {
    Lat L;
    L.pitch = s->pitch;
    L.n1 = (int32_t)s->n1;
    L.n2 = (int32_t)s->n2;
    L.b1 = s->d.boundary[0];
    L.b2 = s->d.ndims == 2 ? s->d.boundary[1] : CAH_BOUNDARY_PERIODIC;
    L.two_d = s->d.ndims == 2;
    return L;
}

// Offsets as a set equal to Moore(R) (radius returned) or VonNeumann(1).
static int moore_radius(const std::vector<int32_t>& o) {
    const int m = (int)o.size() / 2;
    int r = 0;
    for (int k = 0; k < 2 * m; ++k) r = std::max(r, std::abs(o[k]));
    if (r < 1 || m != (2 * r + 1) * (2 * r + 1) - 1) return 0;
    std::set<std::pair<int, int>> s;
    for (int k = 0; k < m; ++k) s.insert({o[2 * k], o[2 * k + 1]});
    if ((int)s.size() != m || s.count({0, 0})) return 0;
    return r;
}

// r when the offsets are exactly VonNeumann(r) (every nonzero offset of L1 norm
// at most r, in any order), else 0.
static int vn_radius(const std::vector<int32_t>& o) {
    const int m = (int)o.size() / 2;
    int r = 0;
    for (int k = 0; k < m; ++k) r = std::max(r, std::abs(o[2 * k]) + std::abs(o[2 * k + 1]));
    if (r < 1 || m != 2 * r * (r + 1)) return 0;
    std::set<std::pair<int, int>> s;
    for (int k = 0; k < m; ++k) s.insert({o[2 * k], o[2 * k + 1]});
    if ((int)s.size() != m || s.count({0, 0})) return 0;
    return r;
}

static bool is_vn1(const std::vector<int32_t>& o) {
    if (o.size() != 8) return false;
    std::set<std::pair<int, int>> s;
    for (int k = 0; k < 4; ++k) s.insert({o[2 * k], o[2 * k + 1]});
    return s == std::set<std::pair<int, int>>{{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
}

// ---------------------------------------------------------------- device memory
//
// A state's buffers come from hipMalloc, or, for the states a server owns, from
// the server's arena: slabs taken with hipMalloc and carved first-fit in 256-byte
// units, so that admitting and retiring requests neither allocates nor frees
// device memory (hipFree waits for the whole device to go idle).  A block is
// only handed out again after every launch that used it has completed; the
// server guarantees that by releasing a request's blocks after its result left.
// A pinned arena does the same with page-locked host memory (hipHostMalloc), the
// staging area of the server's host <-> device copies.

struct Arena {
    bool pinned = false;
    struct Blk {
        size_t size;
        int slab;
    };
    size_t slab_min = (size_t)256 << 20;
    std::vector<char*> slabs;
    std::map<char*, Blk> free_;
    std::unordered_map<char*, Blk> used;
    size_t in_use = 0;

    void* alloc(size_t n) {
        n = std::max<size_t>(256, (n + 255) / 256 * 256);
        for (int attempt = 0; attempt < 2; ++attempt) {
            for (auto it = free_.begin(); it != free_.end(); ++it) {
                if (it->second.size < n) continue;
                char* p = it->first;
                const Blk b = it->second;
                free_.erase(it);
                if (b.size > n) free_[p + n] = Blk{b.size - n, b.slab};
                used[p] = Blk{n, b.slab};
                in_use += n;
                if (!pinned) lchk::reg(p, n, "server arena block");
                return p;
            }
            const size_t sz = std::max(slab_min, n);
            char* slab = nullptr;
            const hipError_t e = pinned ? hipHostMalloc((void**)&slab, sz, hipHostMallocDefault)
                                        : hipMalloc((void**)&slab, sz);
            if (e != hipSuccess) return nullptr;
            slabs.push_back(slab);
            free_[slab] = Blk{sz, (int)slabs.size() - 1};
        }
        return nullptr;
    }

    // Slabs allocated up front until the arena holds `bytes` in all (one more
    // slab of the shortfall); false when the allocation fails.
    bool reserve(size_t bytes) {
        size_t have = 0;
        for (const auto& kv : free_) have += kv.second.size;
        for (const auto& kv : used) have += kv.second.size;
        if (have >= bytes) return true;
        const size_t sz = (bytes - have + 255) / 256 * 256;
        char* slab = nullptr;
        const hipError_t e = pinned ? hipHostMalloc((void**)&slab, sz, hipHostMallocDefault)
                                    : hipMalloc((void**)&slab, sz);
        if (e != hipSuccess) return false;
        slabs.push_back(slab);
        free_[slab] = Blk{sz, (int)slabs.size() - 1};
        return true;
    }

    void release(void* vp) {
        char* p = (char*)vp;
        auto u = used.find(p);
        if (u == used.end()) return;
        Blk b = u->second;
        used.erase(u);
        in_use -= b.size;
        if (!pinned) lchk::unreg(p);
        auto nx = free_.find(p + b.size);
        if (nx != free_.end() && nx->second.slab == b.slab) {
            b.size += nx->second.size;
            free_.erase(nx);
        }
        auto it = free_.emplace(p, b).first;
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second.size == p && pv->second.slab == b.slab) {
                pv->second.size += b.size;
                free_.erase(it);
            }
        }
    }

    ~Arena() {
        for (char* s : slabs) {
            if (pinned) hipHostFree(s);
            else hipFree(s);
        }
    }
};

// hipMalloc, and the span the launch checks know the block by.
static hipError_t dmalloc(void** p, size_t n, const char* what) {
    const hipError_t e = hipMalloc(p, std::max<size_t>(n, 1));
    if (e == hipSuccess) lchk::reg(*p, std::max<size_t>(n, 1), what);
    return e;
}

static void dfree(void* p, bool dry = false) {
    if (!p) return;
    lchk::unreg(p);
    if (!dry) (void)hipFree(p);  // teardown: nothing to report to
}

static int dev_alloc(cah_state* s, void** p, size_t n) {
    if (s->dry) {
        *p = lchk::dry_alloc(n, "state buffer (dry)");
        return CAH_OK;
    }
    if (s->arena) {
        *p = s->arena->alloc(n);
        if (!*p) return fail(CAH_ERR_OUT_OF_MEMORY, "the device arena cannot hold %zu more bytes", n);
        return CAH_OK;
    }
    HIPCHK(dmalloc(p, n, "state buffer"));
    return CAH_OK;
}

static void dev_free(cah_state* s, void* p) {
    if (!p) return;
    if (s->arena) s->arena->release(p);
    else dfree(p, s->dry);
}

// The pinned sources of a server state's uploads, once those have completed.
static void release_staging(cah_state* s) {
    for (void* p : s->keep) s->staging->release(p);
    s->keep.clear();
}

static void free_state(cah_state* s) {
    if (!s) return;
    if (!s->dry) hipSetDevice(s->device);
    if (s->staging) release_staging(s);
    if (s->stream && !s->arena) hipStreamSynchronize(s->stream);
    for (void* p : {s->buf[0], s->buf[1], (void*)s->mask_d, (void*)s->ofs_d, (void*)s->taps_d,
                    (void*)s->w64_d, (void*)s->w32_d, (void*)s->wtab_d, s->field_d, s->frag_d,
                    (void*)s->counts_d, (void*)s->cov_d, (void*)s->sp_cand[0], (void*)s->sp_cand[1],
                    (void*)s->sp_cnt, (void*)s->sp_stamp, (void*)s->sp_edgef, (void*)s->sp_edges,
                    (void*)s->sp_tmp, (void*)s->sp_act})
        dev_free(s, p);
    if (s->fft) {
        lchk::unreg(lenia_fft_buffer(s->fft));
        lenia_fft_destroy(s->fft);
    }
    if (s->stream && s->own_stream) hipStreamDestroy(s->stream);
    delete s;
}

template <typename T>
static int upload(T** dst, const T* src, size_t n, bool dry = false) {
    if (dry) {
        *dst = (T*)lchk::dry_alloc(std::max<size_t>(n, 1) * sizeof(T), "table (dry)");
        return CAH_OK;
    }
    HIPCHK(dmalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(T), "table"));
    if (n) HIPCHK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return CAH_OK;
}

// A pinned copy of n host bytes that stays valid until release_staging.
static int stage(cah_state* s, const void* src, size_t n, void** out) {
    *out = s->staging->alloc(n);
    if (!*out) return fail(CAH_ERR_OUT_OF_MEMORY, "cannot stage %zu bytes in pinned memory", n);
    hostcopy::copy(*out, src, n);
    s->keep.push_back(*out);
    return CAH_OK;
}

// A state's device copy of n host values: synchronous for a state of its own,
// queued on the state's stream from pinned memory for a server's.
template <typename T>
static int upload(cah_state* s, T** dst, const T* src, size_t n) {
    if (int rc = dev_alloc(s, (void**)dst, std::max<size_t>(n, 1) * sizeof(T))) return rc;
    if (!n || s->dry) return CAH_OK;
    if (s->staging) {
        void* h;
        if (int rc = stage(s, src, n * sizeof(T), &h)) return rc;
        HIPCHK(hipMemcpyAsync(*dst, h, n * sizeof(T), hipMemcpyHostToDevice, s->stream));
    } else {
        HIPCHK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    }
    return CAH_OK;
}

// arena, staging and stream non-null: a server's state, on the server's stream
// and memory
static int create_state(cah_ctx* ctx, const cah_desc* d, cah_state* s, Arena* arena = nullptr,
                        Arena* staging = nullptr, hipStream_t stream = nullptr) {
    s->device = ctx->device;
    if (!s->dry) HIPCHK(hipSetDevice(s->device));
    s->arena = arena;
    s->staging = staging;
    if (stream) {
        s->stream = stream;
        s->own_stream = false;
    } else if (s->dry) {
        s->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    s->d = *d;
    s->n1 = d->dims[0];
    s->n2 = d->ndims == 2 ? d->dims[1] : 1;
    s->es = d->dtype == CAH_DTYPE_F32 ? 4 : 1;
    s->pitch = d->dtype == CAH_DTYPE_F32 ? (s->n1 + 15) / 16 * 16 : (s->n1 + 63) / 64 * 64;
    s->offsets.assign(d->offsets, d->offsets + 2 * (size_t)d->n_offsets);
    if (d->rule == CAH_RULE_LENIA || d->rule == CAH_RULE_FIELD_EPIDEMIC) {
        s->taps.assign(d->tap_offsets, d->tap_offsets + 2 * (size_t)d->n_taps);
        s->weights.assign(d->tap_weights, d->tap_weights + d->n_taps);
    }
    if (d->rule == CAH_RULE_FIELD_EPIDEMIC && d->covariate)
        s->cov.assign(d->covariate, d->covariate + (size_t)(s->n1 * s->n2) * (size_t)d->covariate_slices);
    s->d.offsets = nullptr;
    s->d.tap_offsets = nullptr;
    s->d.tap_weights = nullptr;
    s->d.mask = nullptr;
    s->d.covariate = nullptr;
    const size_t bytes = (size_t)s->pitch * (size_t)s->n2 * s->es;
    for (int b = 0; b < 2; ++b) {
        if (int rc = dev_alloc(s, &s->buf[b], bytes)) return rc;
        if (!s->dry) HIPCHK(hipMemsetAsync(s->buf[b], 0, bytes, s->stream));
    }
    if (d->mask) {
        int rc = upload(s, &s->mask_d, d->mask, (size_t)(s->n1 * s->n2));
        if (rc) return rc;
    }
    if (int rc = upload(s, &s->ofs_d, s->offsets.data(), s->offsets.size())) return rc;
    s->halo = std::max(1, max_abs(s->offsets.data(), d->n_offsets));

    if (d->rule == CAH_RULE_LENIA) {
        s->halo = std::max(s->halo, max_abs(s->taps.data(), d->n_taps));
        if (int rc = upload(s, &s->taps_d, s->taps.data(), s->taps.size())) return rc;
        if (int rc = upload(s, &s->w64_d, s->weights.data(), s->weights.size())) return rc;
        std::vector<float> w32(s->weights.begin(), s->weights.end());
        if (int rc = upload(s, &s->w32_d, w32.data(), w32.size())) return rc;
        const int ht = max_abs(s->taps.data(), d->n_taps);
        s->lenia_H = ht <= 4 ? 4 : ht <= 8 ? 8 : ht <= 12 ? 12 : 16;
        const int K = 2 * s->lenia_H + 1;
        std::vector<float> tab((size_t)K * K, 0.0f);
        for (int k = 0; k < d->n_taps; ++k)
            tab[(size_t)(s->taps[2 * k + 1] + s->lenia_H) * K + (s->taps[2 * k] + s->lenia_H)] +=
                (float)s->weights[k];
        if (int rc = upload(s, &s->wtab_d, tab.data(), tab.size())) return rc;
        // fused field + growth kernels need the field of the step-start state for
        // every cell they update, so Checkerboard (two half-sweeps reading the
        // same field) runs on the separate field kernels
        const bool cb = d->scheme == CAH_SCHEME_CHECKERBOARD;
        const int fm = d->field_method;
        if (fm == CAH_FIELD_FFT_F32) {
            s->path = P_LENIA_FFT;
            const size_t nspec = (size_t)(s->n1 / 2 + 1) * (size_t)s->n2;
            if (s->dry) {
                s->fft = lenia_fft_make_dry((int)s->n1, (int)s->n2, lchk::dry_alloc(nspec * 8, "FFT spectrum (dry)"),
                                            lchk::dry_alloc(nspec * 8, "FFT kernel spectrum (dry)"),
                                            (float*)lchk::dry_alloc((size_t)(s->n1 * s->n2) * 4, "FFT field (dry)"));
                return CAH_OK;
            }
            if (int rc = lenia_fft_create(&s->fft, (int)s->n1, (int)s->n2, s->pitch, s->taps.data(),
                                          s->weights.data(), d->n_taps, s->stream))
                return fail(rc, "cannot set up the FFT field (hipFFT)");
            lchk::reg(lenia_fft_buffer(s->fft), (size_t)(s->n1 * s->n2) * 4, "FFT field");
            return CAH_OK;
        }
        if ((fm == CAH_FIELD_MFMA_F16 || fm == CAH_FIELD_MFMA_BF16 || fm == CAH_FIELD_MFMA_F16X2) &&
            !cb && ht <= LM_HMAX) {
            s->path = P_LENIA_MFMA;
            s->lenia_mode = fm == CAH_FIELD_MFMA_BF16 ? LM_BF16 : fm == CAH_FIELD_MFMA_F16X2 ? LM_F16X2 : LM_F16;
            const int Kd = 2 * ht + 1;
            std::vector<double> wd((size_t)Kd * Kd, 0.0);
            for (int k = 0; k < d->n_taps; ++k)
                wd[(size_t)(s->taps[2 * k + 1] + ht) * Kd + (s->taps[2 * k] + ht)] += s->weights[k];
            s->lenia_H = ht;
            lenia_mfma_pairs(ht, wd.data(), s->lenia_dy0, s->lenia_np);
            std::vector<char> frag;
            lenia_mfma_fragments(s->lenia_mode, ht, wd.data(), s->lenia_dy0, s->lenia_np, frag);
            s->frag_bytes = frag.size() / 2;  // one plane
            return upload(s, (char**)&s->frag_d, frag.data(), frag.size());
        }
        const bool fused = fm == CAH_FIELD_DIRECT_F32 && !cb;
        s->path = fused ? P_LENIA_F32COL : P_LENIA_SEQ;
        if (!fused)
            if (int rc = dev_alloc(s, &s->field_d, (size_t)(s->n1 * s->n2) * (d->lenia_precision == 32 ? 4 : 8)))
                return rc;
        return CAH_OK;
    }

    s->path = P_GENERIC;
    if (d->rule == CAH_RULE_FIELD_EPIDEMIC) {
        // the field's taps and weights in its precision, its buffer, the covariate
        s->epi_h = std::max(1, max_abs(s->taps.data(), d->n_taps));
        if (int rc = upload(s, &s->taps_d, s->taps.data(), s->taps.size())) return rc;
        if (d->lenia_precision == 32) {
            std::vector<float> w32(s->weights.begin(), s->weights.end());
            if (int rc = upload(s, &s->w32_d, w32.data(), w32.size())) return rc;
        } else if (int rc = upload(s, &s->w64_d, s->weights.data(), s->weights.size())) {
            return rc;
        }
        if (int rc = dev_alloc(s, &s->field_d, (size_t)(s->n1 * s->n2) * (d->lenia_precision == 32 ? 4 : 8)))
            return rc;
        if (!s->cov.empty()) {
            if (int rc = upload(s, &s->cov_d, s->cov.data(), s->cov.size())) return rc;
            std::vector<double>().swap(s->cov);  // copied (or staged in pinned memory)
        }
        s->bin_clean = true;
        return CAH_OK;
    }
    // the band walks: two-dimensional lattices, every gated scheme (cah_supports)
    const bool plain = d->scheme == CAH_SCHEME_SYNCHRONOUS && !d->mask;
    if (d->ndims == 2 && (d->rule == CAH_RULE_LIFE || d->rule == CAH_RULE_GENERATIONS)) {
        const int R = moore_radius(s->offsets);
        LifeArgs& la = s->life;
        int nb = life_intervals(d->birth, la.b_lo, la.b_hi);
        int ns = life_intervals(d->survive, la.s_lo, la.s_hi);
        if (R >= 1 && R <= 6 && nb >= 0 && ns >= 0) {
            la.nb = nb;
            la.ns = ns;
            if (d->rule == CAH_RULE_GENERATIONS) {
                const int C = d->gen_states;
                la.gen_C = C;
                la.gen_k = (uint32_t)std::max(0, 129 - C) * 0x01010101u;
                la.gen_dying = (C > 2 ? 2u : 0u) * 0x01010101u;
            }
            s->life_R = R;
            s->path = plain && d->rule == CAH_RULE_LIFE ? P_LIFE_SWAR : P_LIFE_GATED;
        }
        if (s->path == P_GENERIC) {
            const int rv = vn_radius(s->offsets);
            if (R >= 1 && R <= BOX_MOORE_MAX) {
                s->path = P_BOX;
                s->box_vn = false;
            } else if (rv >= 1 && rv <= BOX_VN_MAX) {
                s->path = P_BOX;
                s->box_vn = true;
            }
        }
    } else if (d->ndims == 2 && d->rule == CAH_RULE_FOREST_FIRE && is_vn1(s->offsets)) {
        s->path = plain ? P_FOREST : P_FOREST_GATED;
    }
    s->bin_clean = s->tri_clean = s->small_clean = true;  // the buffers start all zero
    return CAH_OK;
}

static int new_state(cah_ctx* ctx, const cah_desc* d, cah_state** out, bool dry) {
    if (!ctx || !out) return fail(CAH_ERR_ARGUMENT, "cah_state_create: NULL argument");
    if (int rc = cah_check_problem(d)) return rc;
    cah_state* s = new cah_state();
    s->dry = dry;
    int rc = create_state(ctx, d, s);
    if (rc) {
        std::string keep = g_err;
        free_state(s);
        g_err = keep;
        return rc;
    }
    *out = s;
    return CAH_OK;
}

extern "C" int cah_state_create(cah_ctx* ctx, const cah_desc* d, cah_state** out) {
    return new_state(ctx, d, out, false);
}

extern "C" int cah_state_destroy(cah_state* s) {
    free_state(s);
    return CAH_OK;
}

extern "C" int cah_state_tune(cah_state* s, int32_t vb, int32_t rows) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "state is NULL");
    if (vb != 0 && vb != 1 && vb != 2 && vb != 4)
        return fail(CAH_ERR_ARGUMENT, "vb must be 0, 1, 2 or 4");
    if (rows < 0) return fail(CAH_ERR_ARGUMENT, "rows must be >= 0");
    s->tune_vb = vb;
    s->tune_rows = rows;
    return CAH_OK;
}

static size_t state_bytes(const cah_state* s) { return (size_t)(s->n1 * s->n2) * s->es; }

// the sparse sweep's bookkeeping (defined with it below)
static int sparse_refresh(cah_state* s);

static bool scan_binary(const int8_t* h, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if ((uint8_t)h[i] > 1) return false;
    return true;
}

// Whether every byte is 0, 1, 2 or 3, eight bytes a step.
static bool scan_small(const int8_t* h, size_t n) {
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, h + i, 8);
        if (w & 0xFCFCFCFCFCFCFCFCull) return false;
    }
    for (; i < n; ++i)
        if ((uint8_t)h[i] > 3) return false;
    return true;
}


// Rows of rb bytes from a pitched to a dense layout or back, on a stream.
static int copy_rows(const void* src, int64_t sp, void* dst, int64_t dp, int64_t rb, int64_t rows,
                     hipStream_t st) {
    const uint64_t a = (uint64_t)(uintptr_t)src | (uint64_t)(uintptr_t)dst | (uint64_t)sp |
                       (uint64_t)dp | (uint64_t)rb;
    const int W = a % 16 == 0 ? 16 : a % 4 == 0 ? 4 : 1;
    const int64_t n = rb / W * rows;
    if (n == 0) return CAH_OK;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 16384);
    const uint8_t* s8 = (const uint8_t*)src;
    uint8_t* d8 = (uint8_t*)dst;
    if (int rc = checked(false, "k_copy_rows", [&](lchk::Chk& c) {
            c.need(rb % W == 0 && sp % W == 0 && dp % W == 0, "rows of %lld bytes in %d-byte pieces", (long long)rb, W);
            c.range(src, 0, (rows - 1) * sp + rb, "source rows");
            c.range(dst, 0, (rows - 1) * dp + rb, "destination rows");
        }))
        return rc;
    if (W == 16) k_copy_rows<16><<<blocks, 256, 0, st>>>(s8, sp, d8, dp, rb, rows);
    else if (W == 4) k_copy_rows<4><<<blocks, 256, 0, st>>>(s8, sp, d8, dp, rb, rows);
    else k_copy_rows<1><<<blocks, 256, 0, st>>>(s8, sp, d8, dp, rb, rows);
    HIPCHK(hipGetLastError());
    return CAH_OK;
}

// Host <-> device copies of a state's contents move the dense form (n1 * es
// bytes a row) with linear copies through one pinned staging buffer, in pieces
// of whole rows of at most STAGE_MAX bytes; the pitched form is made on the
// device from the dense one in the state's other buffer (between steps it holds
// nothing, and no kernel reads the pitch padding).  A pitched copy from or to
// pageable memory is staged row by row by the runtime, and the DMA engines split
// one whose rows are not 4-byte multiples into a transfer per row: milliseconds
// for a lattice of a thousand rows.
constexpr size_t STAGE_MAX = (size_t)64 << 20;
static std::mutex g_stage_mu;
static char* g_stage = nullptr;
static size_t g_stage_cap = 0;

static int stage_rows(const cah_state* s, int64_t* rows_per) {
    const size_t row = (size_t)s->n1 * s->es;
    *rows_per = std::max<int64_t>(1, (int64_t)(STAGE_MAX / row));
    const size_t need = row * (size_t)std::min<int64_t>(*rows_per, s->n2);
    if (need > g_stage_cap) {
        if (g_stage) hipHostFree(g_stage);
        g_stage = nullptr;
        g_stage_cap = 0;
        HIPCHK(hipHostMalloc((void**)&g_stage, need, hipHostMallocDefault));
        g_stage_cap = need;
    }
    return CAH_OK;
}

static int copy_in(cah_state* s, const void* host) {
    const size_t row = (size_t)s->n1 * s->es, pitchb = (size_t)s->pitch * s->es;
    char* dense = (char*)(row == pitchb ? s->buf[s->cur] : s->buf[1 - s->cur]);
    std::lock_guard<std::mutex> lock(g_stage_mu);
    int64_t per;
    if (int rc = stage_rows(s, &per)) return rc;
    for (int64_t y = 0; y < s->n2; y += per) {
        const size_t nb = (size_t)std::min<int64_t>(per, s->n2 - y) * row;
        hostcopy::copy(g_stage, (const char*)host + (size_t)y * row, nb);
        HIPCHK(hipMemcpyAsync(dense + (size_t)y * row, g_stage, nb, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    if (row != pitchb)
        if (int rc = copy_rows(dense, (int64_t)row, s->buf[s->cur], (int64_t)pitchb, (int64_t)row, s->n2,
                               s->stream))
            return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return CAH_OK;
}

static int copy_out(cah_state* s, void* host) {
    const size_t row = (size_t)s->n1 * s->es, pitchb = (size_t)s->pitch * s->es;
    const char* dense = (const char*)s->buf[s->cur];
    if (row != pitchb) {
        dense = (const char*)s->buf[1 - s->cur];
        if (int rc = copy_rows(s->buf[s->cur], (int64_t)pitchb, s->buf[1 - s->cur], (int64_t)row,
                               (int64_t)row, s->n2, s->stream))
            return rc;
    }
    std::lock_guard<std::mutex> lock(g_stage_mu);
    int64_t per;
    if (int rc = stage_rows(s, &per)) return rc;
    for (int64_t y = 0; y < s->n2; y += per) {
        const size_t nb = (size_t)std::min<int64_t>(per, s->n2 - y) * row;
        HIPCHK(hipMemcpyAsync(g_stage, dense + (size_t)y * row, nb, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        hostcopy::copy((char*)host + (size_t)y * row, g_stage, nb);
    }
    // the sparse sweep leaves tiles of the other buffer unwritten and needs them
    // to hold the background: after its use as scratch, refill it and forget the
    // previous step's tiles (none of them is left in it)
    if (row != pitchb && s->sparse) {
        HIPCHK(hipMemsetAsync(s->buf[1 - s->cur], (int)(uint8_t)s->sp_bg, (size_t)s->pitch * (size_t)s->n2,
                              s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return CAH_OK;
}

extern "C" int cah_state_set(cah_state* s, const void* host, size_t nbytes) {
    if (!s || !host) return fail(CAH_ERR_ARGUMENT, "cah_state_set: NULL argument");
    if (nbytes != state_bytes(s))
        return fail(CAH_ERR_STATE_SHAPE, "state has %zu bytes, got %zu", state_bytes(s), nbytes);
    HIPCHK(hipSetDevice(s->device));
    if (int rc = copy_in(s, host)) return rc;
    if (s->es == 1) s->bin_clean = scan_binary((const int8_t*)host, nbytes);
    // a state with 2s takes the forest fire's exact byte test for its first step
    // (a scan for them here would cost a full pass over the host copy); the
    // gated walk of Generations with at most 3 states, which keeps such a state
    // on its exact path, has it scanned
    if (s->es == 1) s->tri_clean = s->small_clean = s->bin_clean;
    if (s->es == 1 && !s->bin_clean && s->path == P_LIFE_GATED && s->d.rule == CAH_RULE_GENERATIONS &&
        s->d.gen_states <= 4)
        s->small_clean = scan_small((const int8_t*)host, nbytes);
    if (s->sparse) return sparse_refresh(s);
    return CAH_OK;
}

extern "C" int cah_state_get(cah_state* s, void* host, size_t nbytes) {
    if (!s || !host) return fail(CAH_ERR_ARGUMENT, "cah_state_get: NULL argument");
    if (nbytes != state_bytes(s))
        return fail(CAH_ERR_STATE_SHAPE, "state has %zu bytes, got %zu", state_bytes(s), nbytes);
    HIPCHK(hipSetDevice(s->device));
    return copy_out(s, host);
}

extern "C" int cah_state_set_device(cah_state* s, const void* dptr, size_t nbytes) {
    if (!s || !dptr) return fail(CAH_ERR_ARGUMENT, "cah_state_set_device: NULL argument");
    if (nbytes != state_bytes(s))
        return fail(CAH_ERR_STATE_SHAPE, "state has %zu bytes, got %zu", state_bytes(s), nbytes);
    HIPCHK(hipSetDevice(s->device));
    lchk::reg(dptr, nbytes, "caller's device array");
    const int rc = copy_rows(dptr, (int64_t)(s->n1 * s->es), s->buf[s->cur], (int64_t)(s->pitch * s->es),
                             (int64_t)(s->n1 * s->es), s->n2, s->stream);
    lchk::unreg(dptr);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->es == 1) s->bin_clean = s->tri_clean = s->small_clean = false;  // not scanned
    if (s->sparse) return sparse_refresh(s);
    return CAH_OK;
}

extern "C" int cah_state_get_device(cah_state* s, void* dptr, size_t nbytes) {
    if (!s || !dptr) return fail(CAH_ERR_ARGUMENT, "cah_state_get_device: NULL argument");
    if (nbytes != state_bytes(s))
        return fail(CAH_ERR_STATE_SHAPE, "state has %zu bytes, got %zu", state_bytes(s), nbytes);
    HIPCHK(hipSetDevice(s->device));
    lchk::reg(dptr, nbytes, "caller's device array");
    const int rc = copy_rows(s->buf[s->cur], (int64_t)(s->pitch * s->es), dptr, (int64_t)(s->n1 * s->es),
                             (int64_t)(s->n1 * s->es), s->n2, s->stream);
    lchk::unreg(dptr);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return CAH_OK;
}

// ---------------------------------------------------------------- counts

// States the rule produces (0 .. K-1), which k_counts counts in registers.
static int rule_values(const cah_state* s) {
    switch (s->d.rule) {
    case CAH_RULE_LIFE: return 2;
    case CAH_RULE_FOREST_FIRE:
    case CAH_RULE_FIELD_EPIDEMIC: return 3;
    case CAH_RULE_GENERATIONS: return std::max(2, s->d.gen_states);
    default: return 16;
    }
}

static int counts_check(const cah_state* s, const cah_state* o) {
    if (s->d.dtype != CAH_DTYPE_I8)
        return fail(CAH_ERR_ARGUMENT, "cell counts on the device take Int8 states");
    if (o && (o->d.dtype != CAH_DTYPE_I8 || o->n1 != s->n1 || o->n2 != s->n2 || o->pitch != s->pitch))
        return fail(CAH_ERR_STATE_SHAPE, "the states compared must be Int8 lattices of the same dims");
    return CAH_OK;
}

// Device words of a counts buffer: the CNT_SLOTS results of `n` lattices and
// one set of per-block partial sums that their launches share in turn.
static size_t counts_words(size_t n) { return n * CNT_SLOTS + (size_t)CNT_BLOCKS * CNT_PART; }

template <int K>
static void launch_counts(int blocks, const uint8_t* a, const uint8_t* b, const cah_state* s,
                          unsigned long long* part, unsigned long long* out, hipStream_t st) {
    hipLaunchKernelGGL(k_counts<K>, blocks, 256, 0, st, a, b, s->pitch, s->n1, s->n2, part);
    hipLaunchKernelGGL(k_counts_reduce<K>, 1, CNT_RED, 0, st, part, blocks, out);
}

// Queue the counts of the values 0 .. K-1 of s, and its agreement with o, into
// out_d[0 .. CNT_SLOTS) (every slot written), through the partial sums part_d
// (CNT_BLOCKS * CNT_PART words; launches on one stream may share them).
static int queue_counts(const cah_state* s, const cah_state* o, unsigned long long* out_d,
                        unsigned long long* part_d, hipStream_t st) {
    const int64_t work = (s->n1 + CNT_SEG - 1) / CNT_SEG * s->n2;  // wave-iterations
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((work + 15) / 16, CNT_BLOCKS));
    const uint8_t* a = (const uint8_t*)s->buf[s->cur];
    const uint8_t* b = o ? (const uint8_t*)o->buf[o->cur] : nullptr;
    const int K = rule_values(s);
    if (int rc = checked(false, "k_counts", [&](lchk::Chk& c) {
            // each lane reads 16 bytes from a column inside the row: to the row's
            // n1 rounded up to 16, within the pitch
            const int64_t end = (s->n2 - 1) * s->pitch + (s->n1 + 15) / 16 * 16;
            c.need(s->pitch % 16 == 0, "pitch %lld", (long long)s->pitch);
            c.range(a, 0, end, "lattice");
            if (b) c.range(b, 0, end, "second lattice");
            c.range(part_d, 0, 8LL * CNT_BLOCKS * CNT_PART, "partial sums");
            c.range(out_d, 0, 8LL * CNT_SLOTS, "counters");
            c.need(blocks >= 1 && blocks <= CNT_BLOCKS, "%d blocks", blocks);
        }))
        return rc;
    if (K <= 2) launch_counts<2>(blocks, a, b, s, part_d, out_d, st);
    else if (K <= 3) launch_counts<3>(blocks, a, b, s, part_d, out_d, st);
    else if (K <= 4) launch_counts<4>(blocks, a, b, s, part_d, out_d, st);
    else if (K <= 8) launch_counts<8>(blocks, a, b, s, part_d, out_d, st);
    else launch_counts<16>(blocks, a, b, s, part_d, out_d, st);
    HIPCHK(hipGetLastError());
    return CAH_OK;
}

// h: the CNT_SLOTS counters of s copied back.  When the counted values do not
// cover every cell, the lattice holds values outside 0 .. K-1: count all 256
// byte values instead (h[0 .. 256) replaced; h[256] kept).
static int finish_counts(const cah_state* s, unsigned long long* out_d, unsigned long long* h, hipStream_t st) {
    unsigned long long sum = 0;
    for (int v = 0; v < 256; ++v) sum += h[v];
    const int64_t n = s->n1 * s->n2;
    if (sum == (unsigned long long)n) return CAH_OK;
    HIPCHK(hipMemsetAsync(out_d, 0, 256 * sizeof(unsigned long long), st));
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
    if (int rc = checked(false, "k_hist256", [&](lchk::Chk& c) {
            lchk::lattice(c, s->buf[s->cur], lat_of(s), 1, "lattice");
            c.range(out_d, 0, 8LL * 256, "histogram");
        }))
        return rc;
    hipLaunchKernelGGL(k_hist256, blocks, 256, 0, st, (const uint8_t*)s->buf[s->cur], s->pitch, s->n1, s->n2,
                       out_d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, out_d, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CAH_OK;
}

static void store_counts(const unsigned long long* h, int64_t* counts, int64_t* equal) {
    for (int v = 0; v < 256; ++v) counts[v] = (int64_t)h[v];
    if (equal) *equal = (int64_t)h[256];
}

extern "C" int cah_state_counts(cah_state* s, const cah_state* other, int64_t* counts, int64_t* equal) {
    if (!s || !counts || (other && !equal)) return fail(CAH_ERR_ARGUMENT, "cah_state_counts: NULL argument");
    if (int rc = counts_check(s, other)) return rc;
    HIPCHK(hipSetDevice(s->device));
    if (other && other->stream != s->stream) HIPCHK(hipStreamSynchronize(other->stream));
    if (!s->counts_d)
        if (int rc = dev_alloc(s, (void**)&s->counts_d, counts_words(1) * sizeof(unsigned long long))) return rc;
    if (int rc = queue_counts(s, other, s->counts_d, s->counts_d + CNT_SLOTS, s->stream)) return rc;
    unsigned long long h[CNT_SLOTS];
    HIPCHK(hipMemcpyAsync(h, s->counts_d, sizeof h, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (int rc = finish_counts(s, s->counts_d, h, s->stream)) return rc;
    store_counts(h, counts, equal);
    return CAH_OK;
}

static int cell_ptr(cah_state* s, int64_t i1, int64_t i2, char** p) {
    if (i1 < 0 || i1 >= s->n1 || i2 < 0 || i2 >= s->n2)
        return fail(CAH_ERR_ARGUMENT, "cell (%lld, %lld) is off the lattice", (long long)i1,
                    (long long)i2);
    *p = (char*)s->buf[s->cur] + ((size_t)i2 * (size_t)s->pitch + (size_t)i1) * s->es;
    return CAH_OK;
}

extern "C" int cah_state_read_cell(cah_state* s, int64_t i1, int64_t i2, void* value) {
    if (!s || !value) return fail(CAH_ERR_ARGUMENT, "NULL argument");
    char* p;
    if (int rc = cell_ptr(s, i1, i2, &p)) return rc;
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipMemcpyAsync(value, p, s->es, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return CAH_OK;
}

extern "C" int cah_state_write_cell(cah_state* s, int64_t i1, int64_t i2, const void* value) {
    if (!s || !value) return fail(CAH_ERR_ARGUMENT, "NULL argument");
    char* p;
    if (int rc = cell_ptr(s, i1, i2, &p)) return rc;
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipMemcpyAsync(p, value, s->es, hipMemcpyHostToDevice, s->stream));
    if (s->sparse) s->sp_dirty = true;  // the lists are rebuilt before the next step
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->es == 1 && *(const uint8_t*)value > 1) s->bin_clean = false;
    if (s->es == 1 && *(const uint8_t*)value > 2) s->tri_clean = false;
    if (s->es == 1 && *(const uint8_t*)value > 3) s->small_clean = false;
    return CAH_OK;
}

extern "C" int cah_state_sync(cah_state* s) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "state is NULL");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    return CAH_OK;
}

extern "C" int cah_state_stream(cah_state* s, void** stream) {
    if (!s || !stream) return fail(CAH_ERR_ARGUMENT, "NULL argument");
    *stream = (void*)s->stream;
    return CAH_OK;
}

// ---------------------------------------------------------------- launches

static int waves_to_blocks(int64_t nwaves) { return (int)((nwaves + 3) / 4); }

// Compute units of a device (256 on MI355X), asked once.
static int device_cus(int device) {
    static int cus[64] = {};
    if (device < 0 || device >= 64) return 256;
    if (!cus[device]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0)
            n = 256;
        cus[device] = n;
    }
    return cus[device];
}

// Rows per wave.  The forest fire and Life (R = 1) are bound by the VALU and
// like many short chunks: enough waves to fill the device several times over,
// 32 rows at most.  Larger-than-Life (R >= 2) is bound by HBM and the VALU
// together, and all its waves are resident at once, so the grid is packed to
// a whole number of four-wave blocks per CU: measured on MI355X (Moore(5),
// 8192^2, 17 bands), the step time follows rows x the most blocks on any CU,
// and four blocks per CU is fastest (35 rows: 23.4-24.5 us; 46 rows, three
// blocks: 24.8-25.5 us; 45 rows, a few CUs with four: 28.5 us).  A small
// lattice gets short chunks (4 rows at least): its waves are few, and more of
// them finish sooner despite the 2R-row warm-up each.
// waves_cu: the walk's waves a CU holds at once (16: four blocks of four waves).
static int32_t band_rows(const cah_state* s, int64_t nbands, int R, int waves_cu = 16) {
    if (s->tune_rows > 0) return s->tune_rows;
    int64_t rows;
    if (R == 1) {
        rows = (s->n2 * nbands + 4095) / 4096;
        rows = std::min<int64_t>(std::max<int64_t>(rows, 4), 32);
    } else {
        const int64_t chunks = std::max<int64_t>(1, (int64_t)waves_cu * device_cus(s->device) / nbands);
        rows = (s->n2 + chunks - 1) / chunks;
        rows = std::min<int64_t>(std::max<int64_t>(rows, 4), 256);
    }
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(rows, s->n2));
}

// Dwords a lane of the Life walks.  Two on wide lattices; the gated Moore(1)
// walk, bound by the VALU, keeps one (Brian's Brain 8192^2 on MI355X: 33.2 us a
// step with one, 36.2 with two).
static int life_vb(const cah_state* s) {
    const bool wide = s->n1 >= 2048 && !(s->path == P_LIFE_GATED && s->life_R == 1);
    int vb = s->tune_vb ? std::min(s->tune_vb, 2) : (wide ? 2 : 1);
    if (vb < life_min_vb(s->life_R)) vb = life_min_vb(s->life_R);
    return vb;
}

static int life_niv(const LifeArgs& a) { return std::max(1, std::max(a.nb, a.ns)); }

// The gate and mask of sweep `gate` of a step (the gated band walks).
static BandGate band_gate(const cah_state* s, int gate) {
    BandGate g{};
    g.mask = s->mask_d;
    g.scheme = s->d.scheme;
    g.gate = gate;
    g.sp_f32 = s->d.scheme_p_is_f32;
    g.sp = s->d.scheme_p;
    return g;
}

// One sweep of a Moore(R) Life rule (k_life_swar), or of Life or Generations
// under sweep `gate` of a gated scheme and the mask (k_life_gated), at step t.
static int step_life(cah_state* s, int64_t t, int gate) {
    LifeArgs a = s->life;
    a.cur = (const int8_t*)s->buf[s->cur];
    a.next = (int8_t*)s->buf[1 - s->cur];
    a.L = lat_of(s);
    a.cv1 = (int32_t)s->d.boundary_value[0];
    a.cv2 = (int32_t)s->d.boundary_value[1];
    const bool gated = s->path == P_LIFE_GATED;
    if (gated) {
        a.g = band_gate(s, gate);
        a.seed = s->d.seed;
        a.t = t;
    }
    // the no-conversion kernel needs every value it may read, the Constant
    // edges included, to be 0 or 1
    const bool edges01 = (s->d.boundary[0] != CAH_BOUNDARY_CONSTANT || (a.cv1 == 0 || a.cv1 == 1)) &&
                         (s->d.boundary[1] != CAH_BOUNDARY_CONSTANT || (a.cv2 == 0 || a.cv2 == 1));
    const bool saved = s->bin_clean;
    s->bin_clean = saved && edges01;
    // the gated walk's cell kinds: 0 and 1 only, or 0, 1 and 2 only (edges
    // included) under Life or a Generations rule of at most 3 states
    auto edge_small = [&](int k) {
        const double v = s->d.boundary_value[k];
        return s->d.boundary[k] != CAH_BOUNDARY_CONSTANT || v == 0 || v == 1 || v == 2 || v == 3;
    };
    const bool writes_small = a.gen_C <= 4;  // Life (0) or Generations with C <= 4
    const int cells = s->bin_clean ? GATED_BIN
                      : (gated && writes_small && s->life_R <= 2 && s->small_clean && edge_small(0) && edge_small(1))
                          ? GATED_SMALL
                          : GATED_ANY;
    const bool masked = gated && a.g.mask != nullptr;
    const int vb = life_vb(s);
    const int64_t ow = 62 * 4 * vb;
    a.nbands = (int32_t)((s->n1 + ow - 1) / ow);
    // the gated walk of a wide radius holds fewer waves a CU (registers): its
    // chunks are sized so that every wave is resident at once all the same
    int waves_cu = 16;
    if (gated && !s->dry) {
        int& occ = s->gated_occ[cells][masked ? 1 : 0];
        if (occ < 0) occ = life_gated_blocks_per_cu(s->life_R, vb, cells, life_niv(a), masked);
        if (occ > 0) waves_cu = std::min(16, 4 * occ);
    }
    a.rows = band_rows(s, a.nbands, s->life_R, waves_cu);
    // the gated Moore(1) walk: chunks of at most 16 rows (8192^2 on MI355X, one
    // dword a lane: Brian's Brain 29.6 -> 28.7 us a step, Star Wars 36.0 -> 34.9;
    // 8 and 64 rows are slower)
    if (gated && s->life_R == 1 && s->tune_rows <= 0) a.rows = std::min<int32_t>(a.rows, 16);
    a.nwaves = (int32_t)(a.nbands * ((s->n2 + a.rows - 1) / a.rows));
    const int blocks = waves_to_blocks(a.nwaves);
    const int R = s->life_R;
    int rc = checked(s->dry, gated ? "k_life_gated" : "k_life_swar", [&](lchk::Chk& c) {
        if (vb == 1) lchk::band_walk<1>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, 4LL * blocks, R, 2 * R + 1);
        else lchk::band_walk<2>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, 4LL * blocks, R, 2 * R + 1);
        if (gated && a.g.mask) {
            if (vb == 1) lchk::band_mask<1>(c, a.L, a.g.mask, a.nbands, a.rows, 2 * R + 1);
            else lchk::band_mask<2>(c, a.L, a.g.mask, a.nbands, a.rows, 2 * R + 1);
        }
    });
    if (!rc && !s->dry) {
        if (gated) launch_life_gated(R, vb, cells, life_niv(a), masked, blocks, a, s->stream);
        else launch_life(R, vb, s->bin_clean, life_niv(a), blocks, a, s->stream);
    }
    // Life keeps a lattice of 0 and 1 so (the gate and mask keep cells as they
    // are); Generations writes 2 and more.  Life writes 0 or 1 from any value, so
    // a sweep over every cell leaves cells below 4 only; Generations with at most
    // 4 states keeps a lattice of cells below 4 so, but moves a negative cell up
    // by one, and with more states may write 4 and more.
    s->bin_clean = saved && s->d.rule == CAH_RULE_LIFE;
    if (!writes_small) s->small_clean = false;
    else if (a.gen_C == 0 && (!gated || !band_gated(a.g))) s->small_clean = true;
    return rc;
}

// k_life_tb over several lattices (batches, the server's Moore(1) Life family).
// The tile shapes, TXW words x TY rows; a launch uses one for all its lattices.
struct TbShape {
    int txw, ty;
};
constexpr TbShape TB_SHAPES[3] = {{2, 16}, {4, 32}, {16, 128}};

static int32_t tb_tiles(const cah_state* s, int shape) {
    const TbShape& t = TB_SHAPES[shape];
    return (int32_t)(((s->n1 + 32 * t.txw - 1) / (32 * t.txw)) * ((s->n2 + t.ty - 1) / t.ty));
}

// The shape with the least LDS work ((TXW + 2) x (TY + 2K) words a tile) among
// those that give at least 1024 tiles (4 a CU), else the one with the most tiles.
// A single lattice gets the shape step_life_tb picks for it.
static int tb_pick_shape(const std::vector<const cah_state*>& m) {
    int best = -1, most = 0;
    double best_work = 0;
    int64_t most_tiles = -1;
    for (int sh = 0; sh < 3; ++sh) {
        int64_t tiles = 0;
        for (const cah_state* s : m) tiles += tb_tiles(s, sh);
        const double work = (double)tiles * (TB_SHAPES[sh].txw + 2) * (TB_SHAPES[sh].ty + 2 * TB_K);
        if (tiles > most_tiles) {
            most_tiles = tiles;
            most = sh;
        }
        if (tiles >= 1024 && (best < 0 || work < best_work)) {
            best = sh;
            best_work = work;
        }
    }
    return best >= 0 ? best : most;
}

static LifeTbArgs tb_item(const cah_state* s, int cur, int32_t steps) {
    LifeTbArgs a{};
    a.cur = (const int8_t*)s->buf[cur];
    a.next = (int8_t*)s->buf[1 - cur];
    a.L = lat_of(s);
    a.cv1 = (int8_t)s->d.boundary_value[0];
    a.cv2 = (int8_t)s->d.boundary_value[1];
    a.birth = (uint32_t)(s->d.birth & 0x1FFu);
    a.survive = (uint32_t)(s->d.survive & 0x1FFu);
    a.steps = steps;
    return a;
}

// `steps` (1..TB_K) steps of a Moore(1) Life rule in one launch (k_life_tb): tiles
// sized so that small lattices still spread over the device.
static int step_life_tb(cah_state* s, int steps) {
    const LifeTbArgs a = tb_item(s, s->cur, steps);
    const int64_t cells = s->n1 * s->n2;
    auto launch = [&](auto kernel, int txw, int ty) -> int {
        dim3 grid((unsigned)((s->n1 + 32 * txw - 1) / (32 * txw)), (unsigned)((s->n2 + ty - 1) / ty));
        if (int rc = checked(s->dry, "k_life_tb", [&](lchk::Chk& c) {
                lchk::grid_y(c, grid.y);
                lchk::life_tb(c, a.L, a.cur, a.next, txw, ty, TB_K, a.steps, grid.x, grid.y);
            }))
            return rc;
        if (!s->dry) kernel<<<grid, 256, 0, s->stream>>>(a);
        return CAH_OK;
    };
    int rc;
    if (cells <= ((int64_t)1 << 20)) rc = launch(k_life_tb<2, 16, TB_K>, 2, 16);
    else if (cells <= ((int64_t)1 << 24)) rc = launch(k_life_tb<4, 32, TB_K>, 4, 32);
    else rc = launch(k_life_tb<16, 128, TB_K>, 16, 128);
    if (rc) return rc;
    if (!s->dry) HIPCHK(hipGetLastError());
    s->cur = 1 - s->cur;
    return CAH_OK;
}

static void launch_life_tb_batch(int shape, int32_t nblocks, const LifeTbArgs* items, const int32_t* block_item,
                                 const int32_t* block_base, int32_t parity, int32_t step0, hipStream_t st) {
    switch (shape) {
    case 0: k_life_tb_batch<2, 16, TB_K><<<nblocks, 256, 0, st>>>(items, block_item, block_base, parity, step0); break;
    case 1: k_life_tb_batch<4, 32, TB_K><<<nblocks, 256, 0, st>>>(items, block_item, block_base, parity, step0); break;
    default:
        k_life_tb_batch<16, 128, TB_K><<<nblocks, 256, 0, st>>>(items, block_item, block_base, parity, step0);
    }
}

// One sweep of the forest fire (k_forest_vn1), or under sweep `gate` of a gated
// scheme and the mask (k_forest_gated), at step t.
static int step_forest(cah_state* s, int64_t t, const cah_params* p, int gate) {
    ForestArgs a{};
    a.cur = (const int8_t*)s->buf[s->cur];
    a.next = (int8_t*)s->buf[1 - s->cur];
    a.L = lat_of(s);
    a.cv1 = (int32_t)s->d.boundary_value[0];
    a.cv2 = (int32_t)s->d.boundary_value[1];
    a.seed = s->d.seed;
    a.t = t;
    a.p = p->ff_p;
    a.f = p->ff_f;
    a.p_f32 = p->ff_p_is_f32;
    a.f_f32 = p->ff_f_is_f32;
    const bool gated = s->path == P_FOREST_GATED;
    if (gated) a.g = band_gate(s, gate);
    // 4 bytes a lane: 78 VGPRs, six waves a SIMD to hide the Philox chains' latency
    // (8 bytes a lane takes 125 VGPRs, four waves: 2-3% slower at 16384^2 on MI355X)
    const int vb = s->tune_vb && !gated ? std::min(s->tune_vb, 2) : 1;
    const int64_t ow = 62 * 4 * vb;
    a.nbands = (int32_t)((s->n1 + ow - 1) / ow);
    a.rows = band_rows(s, a.nbands, 1);
    a.nwaves = (int32_t)(a.nbands * ((s->n2 + a.rows - 1) / a.rows));
    const int blocks = waves_to_blocks(a.nwaves);
    const bool small = (uint64_t)(s->n1 * s->n2) < 0xFFFFFFFFull;
    auto edge012 = [&](int k) {
        const double v = s->d.boundary_value[k];
        return s->d.boundary[k] != CAH_BOUNDARY_CONSTANT || v == 0 || v == 1 || v == 2;
    };
    if (int rc = checked(s->dry, gated ? "k_forest_gated" : "k_forest_vn1", [&](lchk::Chk& c) {
            if (vb == 1) lchk::band_walk<1>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, 4LL * blocks, 1, 3);
            else lchk::band_walk<2>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, 4LL * blocks, 1, 3);
            if (gated && a.g.mask) lchk::band_mask<1>(c, a.L, a.g.mask, a.nbands, a.rows, 3);
        }))
        return rc;
    const bool tri = s->tri_clean && edge012(0) && edge012(1);
    if (!s->dry) {
        if (gated) launch_forest_gated(small, tri, blocks, a, s->stream);
        else launch_forest(vb, small, tri, blocks, a, s->stream);
    }
    // the rule writes 0, 1 and 2 only; a gate or mask keeps other values it finds
    if (!gated) s->tri_clean = true;
    return CAH_OK;
}

static GenArgs generic_args(cah_state* s, int64_t t, int gate, const cah_params* p) {
    GenArgs a{};
    a.cur = (const int8_t*)s->buf[s->cur];
    a.next = (int8_t*)s->buf[1 - s->cur];
    a.mask = s->mask_d;
    a.ofs = s->ofs_d;
    a.L = lat_of(s);
    a.h = s->halo;
    a.nofs = s->d.n_offsets;
    a.cv1 = (int32_t)s->d.boundary_value[0];
    a.cv2 = (int32_t)s->d.boundary_value[1];
    a.rule = s->d.rule;
    a.C = s->d.gen_states;
    a.birth = s->d.birth;
    a.survive = s->d.survive;
    a.seed = s->d.seed;
    a.t = t;
    a.scheme = s->d.scheme;
    a.gate = gate;
    a.sp = s->d.scheme_p;
    a.sp_f32 = s->d.scheme_p_is_f32;
    if (p) {
        a.ffp = p->ff_p;
        a.fff = p->ff_f;
        a.ffp_f32 = p->ff_p_is_f32;
        a.fff_f32 = p->ff_f_is_f32;
    }
    if (s->d.rule == CAH_RULE_FIELD_EPIDEMIC) {
        // the rule reads its field, not its neighbours: no count, no halo
        a.nofs = 0;
        a.h = 0;
        a.field = s->field_d;
        a.field32 = s->d.lenia_precision == 32;
        a.epi_form = s->d.epi_form;
        a.epi_recover = s->d.epi_recover;
        a.cov = s->cov_d;
        a.cov_slices = (int32_t)s->d.covariate_slices;
    }
    return a;
}

// The checks of a general sweep's arguments beyond its tile: the neighbourhood
// within the halo, the offsets, mask, field and covariate buffers, the output.
static void check_generic_args(lchk::Chk& c, const cah_state* s, const GenArgs& a) {
    const int64_t cells = s->n1 * s->n2;
    lchk::offsets_within(c, s->offsets.data(), a.nofs, a.h, "neighbourhood");
    if (a.nofs) c.range(a.ofs, 0, 8LL * a.nofs, "offsets");
    lchk::lattice(c, a.next, a.L, 1, "state written");
    if (a.mask) c.range(a.mask, 0, cells, "mask");
    if (a.field) c.range(a.field, 0, cells * (a.field32 ? 4 : 8), "field");
    if (a.cov) c.range(a.cov, 0, cells * 8 * (int64_t)a.cov_slices, "covariate");
}

static int launch_generic(cah_state* s, int64_t t, int gate, const cah_params* p) {
    const GenArgs a = generic_args(s, t, gate, p);
    dim3 grid((unsigned)((s->n1 + GEN_TX - 1) / GEN_TX), (unsigned)((s->n2 + GEN_TY - 1) / GEN_TY));
    const size_t lds = (size_t)gen_w(a.h) * gen_rows(a.h);
    if (int rc = checked(s->dry, "k_generic_i8", [&](lchk::Chk& c) {
            lchk::grid_y(c, grid.y);
            lchk::generic_tile(c, a.L, a.cur, a.h, grid.x, grid.y, lds);
            check_generic_args(c, s, a);
        }))
        return rc;
    if (!s->dry) k_generic_i8<<<grid, 256, lds, s->stream>>>(a);
    return CAH_OK;
}

static int launch_box(cah_state* s, int64_t t, int gate, const cah_params* p) {
    const GenArgs a = generic_args(s, t, gate, p);
    dim3 grid((unsigned)((s->n1 + BOX_T - 1) / BOX_T), (unsigned)((s->n2 + BOX_T - 1) / BOX_T));
    const size_t lds = box_lds(s->box_vn, a.h);
    if (int rc = checked(s->dry, "k_box_i8", [&](lchk::Chk& c) {
            lchk::grid_y(c, grid.y);
            lchk::box_tile(c, a.L, a.cur, a.next, s->box_vn, a.h, grid.x, grid.y, lds);
            check_generic_args(c, s, a);
        }))
        return rc;
    if (!s->dry) {
        const bool gated = a.mask || a.scheme != CAH_SCHEME_SYNCHRONOUS;
        if (s->box_vn) {
            if (gated) k_box_i8<true, true><<<grid, 256, lds, s->stream>>>(a);
            else k_box_i8<true, false><<<grid, 256, lds, s->stream>>>(a);
        } else {
            if (gated) k_box_i8<false, true><<<grid, 256, lds, s->stream>>>(a);
            else k_box_i8<false, false><<<grid, 256, lds, s->stream>>>(a);
        }
    }
    return CAH_OK;
}

// ---------------------------------------------------------------- sparse sweep

static SparseArgs sparse_args(cah_state* s) {
    SparseArgs sa{};
    const int32_t t = s->sp_step;
    sa.stamp = s->sp_stamp;
    sa.edge = s->sp_edgef;
    sa.list = s->sp_cand[t % 2];
    sa.count = s->sp_cnt + t % 3;
    sa.list_next = s->sp_cand[(t + 1) % 2];
    sa.count_next = s->sp_cnt + (t + 1) % 3;
    sa.count_clear = s->sp_cnt + (t + 2) % 3;
    sa.nactive = s->sp_cnt + 3 + (t + 1) % 2;
    sa.nactive_clear = s->sp_cnt + 3 + t % 2;
    sa.step = t;
    sa.ntx = s->sp_ntx;
    sa.nty = s->sp_nty;
    sa.reach1 = s->sp_rx;
    sa.reach2 = s->sp_ry;
    sa.per1 = s->d.boundary[0] == CAH_BOUNDARY_PERIODIC;
    sa.per2 = s->d.ndims == 2 && s->d.boundary[1] == CAH_BOUNDARY_PERIODIC;
    return sa;
}

// After the state is set, a cell written or the sweep switched on: the other
// buffer all background, the non-background tiles found by a scan, fresh
// stamps, and the first candidate list from them and the edge tiles.
static int sparse_refresh(cah_state* s) {
    const int32_t nt = s->sp_ntx * s->sp_nty;
    HIPCHK(hipMemsetAsync(s->buf[1 - s->cur], (int)(uint8_t)s->sp_bg, (size_t)s->pitch * (size_t)s->n2, s->stream));
    HIPCHK(hipMemsetAsync(s->sp_tmp, 0, (size_t)nt, s->stream));
    HIPCHK(hipMemsetAsync(s->sp_cnt, 0, 6 * sizeof(int32_t), s->stream));
    HIPCHK(hipMemsetAsync(s->sp_stamp, 0xFF, (size_t)nt * sizeof(int32_t), s->stream));  // -1
    dim3 grid((unsigned)((s->n1 + 255) / 256), (unsigned)std::min<int64_t>(s->n2, 65535));
    if (int rc = checked(false, "k_sparse_flags", [&](lchk::Chk& c) {
            lchk::lattice(c, s->buf[s->cur], lat_of(s), 1, "state");
            c.range(s->sp_tmp, 0, nt, "tile flags");
            c.range(s->sp_act, 0, 4LL * nt, "tile list");
            c.range(s->sp_edges, 0, 4LL * s->sp_nedge, "edge tiles");
            c.range(s->sp_cand[0], 0, 4LL * nt, "candidate list");
            c.range(s->sp_stamp, 0, 4LL * nt, "stamps");
            c.range(s->sp_cnt, 0, 24, "counters");
        }))
        return rc;
    k_sparse_flags<<<grid, 256, 0, s->stream>>>((const int8_t*)s->buf[s->cur], lat_of(s), s->sp_ts, s->sp_tsy,
                                                s->sp_ntx, s->sp_bg, s->sp_tmp);
    k_sparse_compact<<<(unsigned)((nt + 255) / 256), 256, 0, s->stream>>>(s->sp_tmp, nt, s->sp_act, s->sp_cnt + 5);
    s->sp_step = 0;
    // step 0 reads candidate list 0 with counter 0; the state's non-background
    // tiles are counted in active counter 0
    SparseArgs sa = sparse_args(s);
    sa.list_next = s->sp_cand[0];
    sa.count_next = s->sp_cnt + 0;
    k_sparse_mark<<<(unsigned)(2 * device_cus(s->device)), 256, 0, s->stream>>>(sa, s->sp_act, s->sp_cnt + 5,
                                                                               s->sp_edges, s->sp_nedge);
    HIPCHK(hipMemcpyAsync(s->sp_cnt + 3, s->sp_cnt + 5, sizeof(int32_t), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->sp_dirty = false;
    return CAH_OK;
}

// Whether a background cell whose neighbours are all background keeps the
// background (the SparseBackend's requirement, src/sparse.jl:49-69), decided
// exactly for the rule families the sparse sweep runs; the forest fire depends on
// its parameters.
static bool quiescent(const cah_state* s, int8_t bg, const cah_params* p) {
    const cah_desc& d = s->d;
    const int n1 = bg == 1 ? d.n_offsets : 0;  // neighbours equal to 1
    switch (d.rule) {
    case CAH_RULE_LIFE: {
        const uint64_t m = bg == 1 ? d.survive : d.birth;
        return (n1 < 64 ? (int)((m >> n1) & 1u) : 0) == bg;
    }
    case CAH_RULE_GENERATIONS: {
        int out;
        if (bg == 0) out = n1 < 64 ? (int)((d.birth >> n1) & 1u) : 0;
        else if (bg == 1) out = (n1 < 64 && ((d.survive >> n1) & 1u)) ? 1 : (d.gen_states > 2 ? 2 : 0);
        else out = bg + 1 < d.gen_states ? bg + 1 : 0;
        return out == bg;
    }
    case CAH_RULE_FOREST_FIRE:
        if (!p) return false;
        if (bg == 0) return !(p->ff_p > 0);  // no growth
        if (bg == 1) return !(p->ff_f > 0);  // no lightning
        return false;
    default: return false;
    }
}

extern "C" int cah_state_sparse(cah_state* s, int8_t background, int32_t tilesize) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "state is NULL");
    const cah_desc& d = s->d;
    if (d.dtype != CAH_DTYPE_I8) return fail(CAH_ERR_ARGUMENT, "the sparse sweep runs Int8 states");
    if (d.rule == CAH_RULE_FIELD_EPIDEMIC)
        return fail(CAH_ERR_ARGUMENT, "SparseBackend does not support kernel fields; use ArrayBackend");
    if (d.rule != CAH_RULE_LIFE && d.rule != CAH_RULE_GENERATIONS && d.rule != CAH_RULE_FOREST_FIRE)
        return fail(CAH_ERR_ARGUMENT, "the sparse sweep runs Life, Generations and the forest fire");
    if (d.scheme != CAH_SCHEME_SYNCHRONOUS && d.scheme != CAH_SCHEME_STOCHASTIC)
        return fail(CAH_ERR_UNSUPPORTED, "SparseBackend runs Synchronous and Stochastic");
    if (tilesize < 1 || tilesize > 128) return fail(CAH_ERR_ARGUMENT, "tilesize must be in 1..128, got %d", tilesize);
    if (s->arena) return fail(CAH_ERR_ARGUMENT, "a server's states run dense");
    if (d.rule != CAH_RULE_FOREST_FIRE && !quiescent(s, background, nullptr))
        return fail(CAH_ERR_ARGUMENT,
                    "SparseBackend needs a quiescent rule: a background cell (%d) with background neighbours "
                    "changes", (int)background);
    const bool two = d.ndims == 2;
    const int64_t ntx = (s->n1 + tilesize - 1) / tilesize, nty = two ? (s->n2 + tilesize - 1) / tilesize : s->n2;
    if (ntx * nty >= (int64_t)1 << 31) return fail(CAH_ERR_ARGUMENT, "too many tiles; use a larger tilesize");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (void* q : {(void*)s->sp_cand[0], (void*)s->sp_cand[1], (void*)s->sp_cnt, (void*)s->sp_stamp,
                    (void*)s->sp_edgef, (void*)s->sp_edges, (void*)s->sp_tmp, (void*)s->sp_act})
        dev_free(s, q);
    s->sp_ts = tilesize;
    s->sp_tsy = two ? tilesize : 1;
    s->sp_ntx = (int32_t)ntx;
    s->sp_nty = (int32_t)nty;
    const int h = s->halo;
    // tiles within reach of a cell's neighbourhood; one more on a periodic axis
    // whose last tile is partial (a wrapped halo can reach past it)
    s->sp_rx = (h + s->sp_ts - 1) / s->sp_ts +
               (d.boundary[0] == CAH_BOUNDARY_PERIODIC && s->n1 % s->sp_ts ? 1 : 0);
    s->sp_ry = two ? (h + s->sp_tsy - 1) / s->sp_tsy +
                         (d.boundary[1] == CAH_BOUNDARY_PERIODIC && s->n2 % s->sp_tsy ? 1 : 0)
                   : 0;
    const size_t nt = (size_t)s->sp_ntx * (size_t)s->sp_nty;
    // every tile holding a cell within h of a Constant edge that is not the background
    std::vector<uint8_t> edge(nt, 0);
    for (int ax = 0; ax < (two ? 2 : 1); ++ax) {
        if (d.boundary[ax] != CAH_BOUNDARY_CONSTANT || (int8_t)d.boundary_value[ax] == background) continue;
        const int64_t n = ax == 0 ? s->n1 : s->n2, ts = ax == 0 ? s->sp_ts : s->sp_tsy;
        for (int64_t c = 0; c < n; ++c) {
            if (c >= h && c < n - h) continue;
            const int64_t tc = c / ts;
            if (ax == 0)
                for (int64_t ty = 0; ty < s->sp_nty; ++ty) edge[(size_t)(ty * s->sp_ntx + tc)] = 1;
            else
                for (int64_t tx = 0; tx < s->sp_ntx; ++tx) edge[(size_t)(tc * s->sp_ntx + tx)] = 1;
        }
    }
    std::vector<int32_t> edges;
    for (size_t k = 0; k < nt; ++k)
        if (edge[k]) edges.push_back((int32_t)k);
    s->sp_nedge = (int32_t)edges.size();
    for (int32_t*& l : s->sp_cand)
        if (int rc = dev_alloc(s, (void**)&l, nt * sizeof(int32_t))) return rc;
    if (int rc = dev_alloc(s, (void**)&s->sp_cnt, 6 * sizeof(int32_t))) return rc;
    if (int rc = dev_alloc(s, (void**)&s->sp_stamp, nt * sizeof(int32_t))) return rc;
    if (int rc = upload(s, &s->sp_edgef, edge.data(), nt)) return rc;
    if (int rc = upload(s, &s->sp_edges, edges.data(), edges.size())) return rc;
    if (int rc = dev_alloc(s, (void**)&s->sp_tmp, nt)) return rc;
    if (int rc = dev_alloc(s, (void**)&s->sp_act, nt * sizeof(int32_t))) return rc;
    s->sparse = true;
    s->sp_bg = background;
    return sparse_refresh(s);
}

extern "C" int cah_state_active_tiles(cah_state* s, int64_t* n) {
    if (!s || !n) return fail(CAH_ERR_ARGUMENT, "NULL argument");
    if (!s->sparse) return fail(CAH_ERR_ARGUMENT, "the state does not run the sparse sweep");
    HIPCHK(hipSetDevice(s->device));
    if (s->sp_dirty)
        if (int rc = sparse_refresh(s)) return rc;
    int32_t c = 0;
    HIPCHK(hipMemcpyAsync(&c, s->sp_cnt + 3 + s->sp_step % 2, sizeof c, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *n = c;
    return CAH_OK;
}

// One sparse step: sweep the listed tiles, which lists the next step's.
static int sparse_step(cah_state* s, int64_t t, const cah_params* p) {
    if (s->sp_dirty || s->sp_step >= INT32_MAX - 1)
        if (int rc = sparse_refresh(s)) return rc;
    const SparseArgs sa = sparse_args(s);
    const GenArgs a = generic_args(s, t, 0, p);
    const size_t lds = (size_t)(s->sp_ts + 2 * a.h) * (size_t)(s->sp_tsy + 2 * a.h);
    // persistent blocks: the number of listed tiles is known only on the device
    const int blocks = (int)std::min<int64_t>((int64_t)s->sp_ntx * s->sp_nty, 4 * device_cus(s->device));
    if (int rc = checked(false, "k_sparse_sweep", [&](lchk::Chk& c) {
            const int64_t nt = (int64_t)s->sp_ntx * s->sp_nty;
            c.need((int64_t)s->sp_ntx * s->sp_ts >= s->n1 && (int64_t)s->sp_nty * s->sp_tsy >= s->n2,
                   "tiles do not cover the lattice");
            lchk::tile_kernel(c, a.L, a.cur, 1, s->sp_ts, s->sp_tsy, a.h, s->sp_ntx, s->sp_nty, lds);
            check_generic_args(c, s, a);
            for (const int32_t* l : {sa.list, (const int32_t*)sa.list_next, (const int32_t*)sa.stamp})
                c.range(l, 0, 4 * nt, "tile list");
            c.range(sa.edge, 0, nt, "edge flags");
        }))
        return rc;
    k_sparse_sweep<<<blocks, 256, lds, s->stream>>>(a, sa, s->sp_ts, s->sp_tsy, s->sp_bg);
    HIPCHK(hipGetLastError());
    ++s->sp_step;
    s->cur = 1 - s->cur;
    return CAH_OK;
}

// The field epidemic's kernel field of the step-start state, read by every gate.
static int launch_epi_field(cah_state* s) {
    const Lat L = lat_of(s);
    dim3 grid((unsigned)((s->n1 + EPI_TX - 1) / EPI_TX), (unsigned)((s->n2 + EPI_TY - 1) / EPI_TY));
    const size_t lds = (size_t)(EPI_TX + 2 * s->epi_h) * (EPI_TY + 2 * s->epi_h);
    const int8_t* cur = (const int8_t*)s->buf[s->cur];
    const int8_t cv1 = (int8_t)s->d.boundary_value[0], cv2 = (int8_t)s->d.boundary_value[1];
    const int64_t es = s->d.lenia_precision == 32 ? 4 : 8;
    if (int rc = checked(s->dry, "k_epi_field", [&](lchk::Chk& c) {
            lchk::grid_y(c, grid.y);
            lchk::tile_kernel(c, L, cur, 1, EPI_TX, EPI_TY, s->epi_h, grid.x, grid.y, lds);
            lchk::offsets_within(c, s->taps.data(), s->d.n_taps, s->epi_h, "kernel field taps");
            c.range(s->taps_d, 0, 8LL * s->d.n_taps, "taps");
            c.range(es == 4 ? (const void*)s->w32_d : (const void*)s->w64_d, 0, es * s->d.n_taps, "weights");
            c.range(s->field_d, 0, es * s->n1 * s->n2, "field");
        }))
        return rc;
    if (s->dry) return CAH_OK;
    if (s->d.lenia_precision == 32)
        k_epi_field<float><<<grid, 256, lds, s->stream>>>(cur, L, cv1, cv2, s->taps_d, s->w32_d, s->d.n_taps,
                                                          s->epi_h, (float*)s->field_d);
    else
        k_epi_field<double><<<grid, 256, lds, s->stream>>>(cur, L, cv1, cv2, s->taps_d, s->w64_d, s->d.n_taps,
                                                           s->epi_h, (double*)s->field_d);
    return CAH_OK;
}

static LeniaArgs lenia_args(cah_state* s, int64_t t, int gate) {
    LeniaArgs a{};
    a.cur = (const float*)s->buf[s->cur];
    a.next = (float*)s->buf[1 - s->cur];
    a.mask = s->mask_d;
    a.taps = s->taps_d;
    a.w64 = s->w64_d;
    a.w32 = s->w32_d;
    a.field = s->field_d;
    a.L = lat_of(s);
    a.h = s->halo;
    a.ntaps = s->d.n_taps;
    a.cv1 = (float)s->d.boundary_value[0];
    a.cv2 = (float)s->d.boundary_value[1];
    a.prec32 = s->d.lenia_precision == 32;
    a.mu = s->d.mu;
    a.sigma = s->d.sigma;
    a.dt = s->d.dt;
    a.seed = s->d.seed;
    a.t = t;
    a.scheme = s->d.scheme;
    a.gate = gate;
    a.sp = s->d.scheme_p;
    a.sp_f32 = s->d.scheme_p_is_f32;
    return a;
}

static int launch_lenia_f32col(cah_state* s, int64_t t) {
    LeniaArgs a = lenia_args(s, t, 0);
    a.w32 = s->wtab_d;
    dim3 grid((unsigned)((s->n1 + 63) / 64), (unsigned)((s->n2 + 31) / 32));
    const int H = s->lenia_H;
    if (int rc = checked(s->dry, "k_lenia_f32col", [&](lchk::Chk& c) {
            c.need(H == 4 || H == 8 || H == 12 || H == 16, "H = %d", H);
            lchk::grid_y(c, grid.y);
            lchk::tile_kernel(c, a.L, a.cur, 4, 64, 32, H, grid.x, grid.y, (size_t)(64 + 2 * H) * (32 + 2 * H) * 4);
            lchk::offsets_within(c, s->taps.data(), s->d.n_taps, H, "kernel taps");
            c.range(a.w32, 0, 4LL * (2 * H + 1) * (2 * H + 1), "weight table");
            lchk::lattice(c, a.next, a.L, 4, "state written");
            if (a.mask) c.range(a.mask, 0, s->n1 * s->n2, "mask");
        }))
        return rc;
    if (s->dry) return CAH_OK;
    switch (s->lenia_H) {
    case 4: k_lenia_f32col<4><<<grid, 256, 0, s->stream>>>(a); break;
    case 8: k_lenia_f32col<8><<<grid, 256, 0, s->stream>>>(a); break;
    case 12: k_lenia_f32col<12><<<grid, 256, 0, s->stream>>>(a); break;
    default: k_lenia_f32col<16><<<grid, 256, 0, s->stream>>>(a); break;
    }
    return CAH_OK;
}

static int launch_lenia_field_seq(cah_state* s, int64_t t) {
    LeniaArgs a = lenia_args(s, t, 0);
    dim3 grid((unsigned)((s->n1 + LEN_TX - 1) / LEN_TX), (unsigned)((s->n2 + LEN_TY - 1) / LEN_TY));
    const size_t lds = (size_t)(LEN_TX + 2 * a.h) * (LEN_TY + 2 * a.h) * sizeof(float);
    if (int rc = checked(s->dry, "k_lenia_field_seq", [&](lchk::Chk& c) {
            const int64_t es = a.prec32 ? 4 : 8;
            lchk::grid_y(c, grid.y);
            lchk::tile_kernel(c, a.L, a.cur, 4, LEN_TX, LEN_TY, a.h, grid.x, grid.y, lds);
            lchk::offsets_within(c, s->taps.data(), a.ntaps, a.h, "kernel taps");
            c.range(a.taps, 0, 8LL * a.ntaps, "taps");
            c.range(a.prec32 ? (const void*)a.w32 : (const void*)a.w64, 0, es * a.ntaps, "weights");
            c.range(a.field, 0, es * s->n1 * s->n2, "field");
        }))
        return rc;
    if (s->dry) return CAH_OK;
    if (a.prec32) k_lenia_field_seq<float><<<grid, 256, lds, s->stream>>>(a);
    else k_lenia_field_seq<double><<<grid, 256, lds, s->stream>>>(a);
    return CAH_OK;
}

static int launch_lenia_growth(cah_state* s, int64_t t, int gate) {
    LeniaArgs a = lenia_args(s, t, gate);
    if (s->fft) {
        a.field = lenia_fft_buffer(s->fft);
        a.field32 = 1;
    }
    dim3 grid((unsigned)((s->n1 + 255) / 256), (unsigned)std::min<int64_t>(s->n2, 65535));
    if (int rc = checked(s->dry, "k_lenia_growth", [&](lchk::Chk& c) {
            c.need((int64_t)grid.x * 256 >= s->n1, "grid.x = %u", grid.x);
            lchk::grid_y(c, grid.y);
            lchk::lattice(c, a.cur, a.L, 4, "state read");
            lchk::lattice(c, a.next, a.L, 4, "state written");
            c.range(a.field, 0, (a.prec32 || a.field32 ? 4 : 8) * s->n1 * s->n2, "field");
            if (a.mask) c.range(a.mask, 0, s->n1 * s->n2, "mask");
        }))
        return rc;
    if (!s->dry) k_lenia_growth<<<grid, 256, 0, s->stream>>>(a);
    return CAH_OK;
}

static int launch_lenia_mfma_step(cah_state* s, int64_t t) {
    LeniaMfmaArgs a{};
    a.cur = (const float*)s->buf[s->cur];
    a.next = (float*)s->buf[1 - s->cur];
    a.mask = s->mask_d;
    a.afrag = s->frag_d;
    a.L = lat_of(s);
    a.H = s->lenia_H;
    a.dy0 = s->lenia_dy0;
    a.npairs = s->lenia_np;
    a.cv1 = (float)s->d.boundary_value[0];
    a.cv2 = (float)s->d.boundary_value[1];
    a.prec32 = s->d.lenia_precision == 32;
    a.mu = s->d.mu;
    a.sigma = s->d.sigma;
    a.dt = s->d.dt;
    a.seed = s->d.seed;
    a.t = t;
    a.scheme = s->d.scheme;
    a.sp = s->d.scheme_p;
    a.sp_f32 = s->d.scheme_p_is_f32;
    const int64_t nbx = (s->n1 + LM_BX - 1) / LM_BX, nby = (s->n2 + LM_BY - 1) / LM_BY;
    a.nbx = (int32_t)nbx;
    const int32_t ntiles = (int32_t)(nbx * nby);
    if (int rc = checked(s->dry, "k_lenia_mfma", [&](lchk::Chk& c) {
            lchk::lenia_mfma(c, a, s->lenia_mode, nbx, nby, s->frag_bytes);
        }))
        return rc;
    if (!s->dry) launch_lenia_mfma(s->lenia_mode, ntiles, a, s->stream);
    return CAH_OK;
}

static int ngates(const cah_desc& d) { return d.scheme == CAH_SCHEME_CHECKERBOARD ? 2 : 1; }

extern "C" int cah_advance(cah_state* s, int64_t t0, int64_t nsteps, const cah_params* p) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "state is NULL");
    if (nsteps < 0 || t0 < 0) return fail(CAH_ERR_ARGUMENT, "t0 and nsteps must be >= 0");
    if (s->d.rule == CAH_RULE_FOREST_FIRE && !p)
        return fail(CAH_ERR_ARGUMENT, "the forest fire needs its parameters (p, f)");
    if (s->d.rule == CAH_RULE_FIELD_EPIDEMIC && !p)
        return fail(CAH_ERR_ARGUMENT, "the field epidemic needs its parameters (beta, gamma)");
    if (!s->dry) HIPCHK(hipSetDevice(s->device));
    if (s->sparse) {
        if (!quiescent(s, s->sp_bg, p))
            return fail(CAH_ERR_ARGUMENT,
                        "SparseBackend needs a quiescent rule: with these parameters a background cell (%d) "
                        "with background neighbours changes", (int)s->sp_bg);
        for (int64_t k = 0; k < nsteps; ++k)
            if (int rc = sparse_step(s, t0 + k, p)) return rc;
        if (s->d.rule != CAH_RULE_LIFE && nsteps > 0) s->bin_clean = false;
        return CAH_OK;
    }
    // Moore(1) Life: up to TB_K steps a launch in LDS, one bit a cell (the band walk
    // takes a last single step; a launch shape set with cah_state_tune keeps it)
    if (s->path == P_LIFE_SWAR && s->life_R == 1 && !s->tune_vb && !s->tune_rows) {
        int64_t left = nsteps;
        while (left >= 2) {
            const int k = (int)std::min<int64_t>(left, TB_K);
            if (int rc = step_life_tb(s, k)) return rc;
            left -= k;
        }
        nsteps = left;
        s->bin_clean = true;  // Life writes 0 and 1 only
    }
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t t = t0 + k;
        if (s->path == P_LENIA_SEQ)
            if (int rc = launch_lenia_field_seq(s, t)) return rc;
        if (s->d.rule == CAH_RULE_FIELD_EPIDEMIC)
            if (int rc = launch_epi_field(s)) return rc;
        if (s->path == P_LENIA_FFT) {
            // the R2C transform reads the pitched state (n1 of every pitch floats)
            if (int rc = checked(s->dry, "hipfftExecR2C", [&](lchk::Chk& c) {
                    lchk::lattice(c, s->buf[s->cur], lat_of(s), 4, "state transformed");
                }))
                return rc;
            if (!s->dry && lenia_fft_field(s->fft, (const float*)s->buf[s->cur], s->stream))
                return fail(CAH_ERR_DEVICE, "hipFFT failed in the Lenia field");
        }
        for (int g = 0; g < ngates(s->d); ++g) {
            int rc = CAH_OK;
            switch (s->path) {
            case P_LIFE_SWAR:
            case P_LIFE_GATED: rc = step_life(s, t, g); break;
            case P_FOREST:
            case P_FOREST_GATED: rc = step_forest(s, t, p, g); break;
            case P_LENIA_SEQ:
            case P_LENIA_FFT: rc = launch_lenia_growth(s, t, g); break;
            case P_LENIA_F32COL: rc = launch_lenia_f32col(s, t); break;
            case P_LENIA_MFMA: rc = launch_lenia_mfma_step(s, t); break;
            case P_BOX: rc = launch_box(s, t, g, p); break;
            default: rc = launch_generic(s, t, g, p); break;
            }
            if (rc) return rc;
            s->cur = 1 - s->cur;
        }
    }
    if (!s->dry) HIPCHK(hipGetLastError());
    // Life maps {0,1} to {0,1}; any other rule may leave other values
    if (s->d.rule != CAH_RULE_LIFE && nsteps > 0) s->bin_clean = false;
    return CAH_OK;
}

// ---------------------------------------------------------------- synthetic inputs

extern "C" int cah_state_fill_bernoulli(cah_state* s, uint64_t seed, double p) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "state is NULL");
    if (s->es != 1) return fail(CAH_ERR_ARGUMENT, "fill_bernoulli needs an Int8 state");
    HIPCHK(hipSetDevice(s->device));
    dim3 grid((unsigned)((s->n1 + 255) / 256), (unsigned)std::min<int64_t>(s->n2, 65535));
    if (int rc = checked(false, "k_fill_bernoulli", [&](lchk::Chk& c) {
            lchk::lattice(c, s->buf[s->cur], lat_of(s), 1, "state");
        }))
        return rc;
    k_fill_bernoulli<<<grid, 256, 0, s->stream>>>((int8_t*)s->buf[s->cur], lat_of(s), seed, p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->bin_clean = s->tri_clean = s->small_clean = true;
    if (s->sparse) return sparse_refresh(s);
    return CAH_OK;
}

extern "C" int cah_state_fill_uniform_box(cah_state* s, uint64_t seed, int64_t lo1, int64_t len1,
                                          int64_t lo2, int64_t len2) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "state is NULL");
    if (s->es != 4) return fail(CAH_ERR_ARGUMENT, "fill_uniform_box needs a Float32 state");
    HIPCHK(hipSetDevice(s->device));
    dim3 grid((unsigned)((s->n1 + 255) / 256), (unsigned)std::min<int64_t>(s->n2, 65535));
    if (int rc = checked(false, "k_fill_uniform_box", [&](lchk::Chk& c) {
            lchk::lattice(c, s->buf[s->cur], lat_of(s), 4, "state");
        }))
        return rc;
    k_fill_uniform_box<<<grid, 256, 0, s->stream>>>((float*)s->buf[s->cur], lat_of(s), seed,
                                                    (int32_t)lo1, (int32_t)len1, (int32_t)lo2,
                                                    (int32_t)len2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return CAH_OK;
}

// ---------------------------------------------------------------- batches
//
// A batch owns one state per trajectory (its buffers, pitch and rule) and
// advances all of them with one launch per step on the batch's stream: the
// trajectories' band-walk waves are concatenated into one flat wave list.
// Trajectories that no band-walk kernel runs advance one after another.

enum BatchKind { B_LOOP, B_FOREST, B_LIFE };

struct cah_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<cah_state*> members;
    BatchKind kind = B_LOOP;
    int vb = 1, life_R = 0, niv = 1;
    bool small = true;
    int parity = 0;
    int32_t rows = 0;  // rows per wave of the band walk; 0: batch_rows' choice
    int32_t nwaves = 0;
    void* items_d = nullptr;
    int32_t* wave_item_d = nullptr;
    int32_t* wave_base_d = nullptr;
    // the steps of one cah_batch_advance, captured as a graph, and what it was
    // captured for: step count, buffer parity before and after, the short byte
    // tests; the forest fire's starting time is read from t0_d, so one graph
    // serves every t0
    hipGraphExec_t exec = nullptr;
    int64_t g_nsteps = -1;
    int g_parity0 = -1, g_parity1 = 0;
    bool g_clean = false;
    int64_t* t0_d = nullptr;
    std::vector<ForestArgs> fitems;  // host copy of the forest items (p and f change)
    bool fp_set = false;
    cah_params fp{};
    unsigned long long* counts_d = nullptr;  // CNT_SLOTS counters per member
    // Moore(1) Life: k_life_tb tables (items with steps = TB_K, block -> item map)
    bool tb = false;
    int tb_shape = 0;
    int32_t tb_blocks = 0;
    LifeTbArgs* tb_items_d = nullptr;
    int32_t* tb_item_d = nullptr;
    int32_t* tb_base_d = nullptr;
    // host copies of the tables, for the launch checks (made once per layout and
    // buffer parity)
    std::vector<LifeArgs> litems;
    std::vector<int32_t> wave_item_h, wave_base_h;
    std::vector<LifeTbArgs> tb_items_h;
    std::vector<int32_t> tb_item_h, tb_base_h;
    bool checked_parity[2] = {false, false};
    bool dry = false;  // members are dry states: launches are checked, not run
};

static void free_batch(cah_batch* b) {
    if (!b) return;
    if (!b->dry) hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    for (void* p : {b->items_d, (void*)b->wave_item_d, (void*)b->wave_base_d, (void*)b->counts_d,
                    (void*)b->tb_items_d, (void*)b->tb_item_d, (void*)b->tb_base_d, (void*)b->t0_d})
        dfree(p, b->dry);
    for (auto* m : b->members) free_state(m);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

// Rows per wave of a batched launch, one value for all its lattices so that every
// wave carries about the same work: `work` = the sum over lattices of bands x
// rows; about 4096 waves, 4 to 64 rows each (a lone small lattice gets short
// chunks, a large batch long ones).
static int64_t batch_rows(int64_t work) {
    return std::min<int64_t>(64, std::max<int64_t>(4, (work + 4095) / 4096));
}

static int batch_layout(cah_batch* b) {
    std::vector<int32_t> wave_item, wave_base;
    std::vector<ForestArgs> fitems;
    std::vector<LifeArgs> litems;
    const int64_t ow = 62 * 4 * b->vb;
    int64_t work = 0;
    for (cah_state* s : b->members) work += (s->n1 + ow - 1) / ow * s->n2;
    const int64_t brows = b->rows > 0 ? b->rows : batch_rows(work);
    for (size_t k = 0; k < b->members.size(); ++k) {
        cah_state* s = b->members[k];
        const int32_t nbands = (int32_t)((s->n1 + ow - 1) / ow);
        const int32_t rows = (int32_t)std::min<int64_t>(brows, s->n2);
        const int32_t nw = nbands * (int32_t)((s->n2 + rows - 1) / rows);
        wave_base.push_back((int32_t)wave_item.size());
        for (int32_t w = 0; w < nw; ++w) wave_item.push_back((int32_t)k);
        if (b->kind == B_FOREST) {
            ForestArgs a{};
            a.cur = (const int8_t*)s->buf[0];
            a.next = (int8_t*)s->buf[1];
            a.L = lat_of(s);
            a.cv1 = (int32_t)s->d.boundary_value[0];
            a.cv2 = (int32_t)s->d.boundary_value[1];
            a.nbands = nbands;
            a.rows = rows;
            a.nwaves = nw;
            a.seed = s->d.seed;
            fitems.push_back(a);
        } else {
            LifeArgs a = s->life;
            a.cur = (const int8_t*)s->buf[0];
            a.next = (int8_t*)s->buf[1];
            a.L = lat_of(s);
            a.cv1 = (int32_t)s->d.boundary_value[0];
            a.cv2 = (int32_t)s->d.boundary_value[1];
            a.nbands = nbands;
            a.rows = rows;
            a.nwaves = nw;
            litems.push_back(a);
        }
    }
    b->nwaves = (int32_t)wave_item.size();
    b->checked_parity[0] = b->checked_parity[1] = false;
    if (b->kind == B_FOREST) {
        b->fitems = fitems;
        if (int rc = upload((ForestArgs**)&b->items_d, fitems.data(), fitems.size(), b->dry)) return rc;
    } else {
        b->litems = litems;
        if (int rc = upload((LifeArgs**)&b->items_d, litems.data(), litems.size(), b->dry)) return rc;
    }
    if (int rc = upload(&b->wave_item_d, wave_item.data(), wave_item.size(), b->dry)) return rc;
    if (int rc = upload(&b->wave_base_d, wave_base.data(), wave_base.size(), b->dry)) return rc;
    b->wave_item_h = wave_item;
    b->wave_base_h = wave_base;
    if (b->kind == B_LIFE && b->life_R == 1) {  // several steps a launch (k_life_tb)
        std::vector<const cah_state*> m(b->members.begin(), b->members.end());
        b->tb = true;
        b->tb_shape = tb_pick_shape(m);
        std::vector<LifeTbArgs> items;
        std::vector<int32_t> block_item, block_base;
        for (size_t k = 0; k < m.size(); ++k) {
            items.push_back(tb_item(m[k], 0, TB_K));
            block_base.push_back((int32_t)block_item.size());
            block_item.insert(block_item.end(), (size_t)tb_tiles(m[k], b->tb_shape), (int32_t)k);
        }
        b->tb_blocks = (int32_t)block_item.size();
        if (int rc = upload(&b->tb_items_d, items.data(), items.size(), b->dry)) return rc;
        if (int rc = upload(&b->tb_item_d, block_item.data(), block_item.size(), b->dry)) return rc;
        if (int rc = upload(&b->tb_base_d, block_base.data(), block_base.size(), b->dry)) return rc;
        b->tb_items_h = items;
        b->tb_item_h = block_item;
        b->tb_base_h = block_base;
    }
    return CAH_OK;
}

static int new_state(cah_ctx* ctx, const cah_desc* d, cah_state** out, bool dry);

static int batch_create(cah_ctx* ctx, int32_t n, const cah_desc* descs, cah_batch** out, bool dry) {
    if (!ctx || !descs || !out || n < 1) return fail(CAH_ERR_ARGUMENT, "cah_batch_create: bad arguments");
    cah_batch* b = new cah_batch();
    b->device = ctx->device;
    b->dry = dry;
    auto bail = [&](int rc) {
        std::string keep = g_err;
        free_batch(b);
        g_err = keep;
        return rc;
    };
    if (!dry && (hipSetDevice(b->device) != hipSuccess ||
                 hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess))
        return bail(fail(CAH_ERR_DEVICE, "cannot create the batch stream"));
    for (int32_t i = 0; i < n; ++i) {
        if (descs[i].rule != descs[0].rule || descs[i].dtype != CAH_DTYPE_I8)
            return bail(fail(CAH_ERR_ARGUMENT, "a batch holds trajectories of one Int8 rule"));
        cah_state* s = nullptr;
        if (int rc = new_state(ctx, &descs[i], &s, dry)) return bail(rc);
        b->members.push_back(s);
    }
    bool all_forest = true, all_life = true;
    int R = b->members[0]->life_R;
    uint64_t cells = 0;
    for (auto* s : b->members) {
        all_forest &= s->path == P_FOREST;
        all_life &= s->path == P_LIFE_SWAR && s->life_R == R;
        cells = std::max<uint64_t>(cells, (uint64_t)(s->n1 * s->n2));
    }
    b->small = cells < 0xFFFFFFFFull;
    if (all_forest) {
        b->kind = B_FOREST;
        b->vb = 1;
    } else if (all_life) {
        b->kind = B_LIFE;
        b->life_R = R;
        b->vb = life_min_vb(R);
        for (auto* s : b->members) b->niv = std::max(b->niv, life_niv(s->life));
    }
    if (b->kind != B_LOOP)
        if (int rc = batch_layout(b)) return bail(rc);
    if (!dry)
        for (auto* s : b->members) HIPCHK(hipStreamSynchronize(s->stream));
    *out = b;
    return CAH_OK;
}

extern "C" int cah_batch_create(cah_ctx* ctx, int32_t n, const cah_desc* descs, cah_batch** out) {
    return batch_create(ctx, n, descs, out, false);
}

static int member(cah_batch* b, int32_t i, cah_state** s) {
    if (!b || i < 0 || i >= (int32_t)b->members.size())
        return fail(CAH_ERR_ARGUMENT, "no trajectory %d in the batch", i);
    *s = b->members[i];
    return CAH_OK;
}

extern "C" int cah_batch_set(cah_batch* b, int32_t i, const void* host, size_t nbytes) {
    cah_state* s;
    if (int rc = member(b, i, &s)) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    return cah_state_set(s, host, nbytes);
}

extern "C" int cah_batch_get(cah_batch* b, int32_t i, void* host, size_t nbytes) {
    cah_state* s;
    if (int rc = member(b, i, &s)) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    return cah_state_get(s, host, nbytes);
}

extern "C" int cah_batch_counts(cah_batch* b, const cah_state* other, int64_t* counts, int64_t* equal) {
    if (!b || !counts || (other && !equal)) return fail(CAH_ERR_ARGUMENT, "cah_batch_counts: NULL argument");
    for (cah_state* m : b->members)
        if (int rc = counts_check(m, other)) return rc;
    HIPCHK(hipSetDevice(b->device));
    if (other && other->stream != b->stream) HIPCHK(hipStreamSynchronize(other->stream));
    // a batch without a batched kernel advances its members on their own streams
    for (cah_state* m : b->members)
        if (m->stream != b->stream) HIPCHK(hipStreamSynchronize(m->stream));
    const size_t n = b->members.size();
    if (!b->counts_d) HIPCHK(dmalloc((void**)&b->counts_d, counts_words(n) * sizeof(unsigned long long), "batch counters"));
    // every member's counts on the batch stream, then one copy back
    for (size_t i = 0; i < n; ++i)
        if (int rc = queue_counts(b->members[i], other, b->counts_d + i * CNT_SLOTS, b->counts_d + n * CNT_SLOTS,
                                  b->stream))
            return rc;
    std::vector<unsigned long long> h(n * CNT_SLOTS);
    HIPCHK(hipMemcpyAsync(h.data(), b->counts_d, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < n; ++i) {
        unsigned long long* hi = h.data() + i * CNT_SLOTS;
        if (int rc = finish_counts(b->members[i], b->counts_d + i * CNT_SLOTS, hi, b->stream)) return rc;
        store_counts(hi, counts + i * 256, equal ? equal + i : nullptr);
    }
    return CAH_OK;
}

extern "C" int cah_batch_counts_noalloc(cah_batch* b, const cah_state* other, int64_t* counts, int64_t* equal) {
    if (!b || !counts || (other && !equal))
        return fail(CAH_ERR_ARGUMENT, "cah_batch_counts_noalloc: NULL argument");
    const int64_t n = (int64_t)b->members.size();
    if (n > CAH_COUNTS_NOALLOC_MAX)
        return fail(CAH_ERR_ARGUMENT,
                    "cah_batch_counts_noalloc counts at most %d trajectories (its fixed device array holds "
                    "their counters); this batch has %lld: use cah_batch_counts",
                    CAH_COUNTS_NOALLOC_MAX, (long long)n);
    for (cah_state* m : b->members)
        if (int rc = counts_check(m, other)) return rc;
    if (b->device < 0 || b->device >= SCRATCH_DEVICES || !g_scratch_ptr[b->device])
        return fail(CAH_ERR_DEVICE, "device %d has no counts scratch (it is set up by cah_open)", b->device);
    HIPCHK(hipSetDevice(b->device));
    if (other && other->stream != b->stream) HIPCHK(hipStreamSynchronize(other->stream));
    for (cah_state* m : b->members)
        if (m->stream != b->stream) HIPCHK(hipStreamSynchronize(m->stream));
    unsigned long long* part = g_scratch_ptr[b->device];
    unsigned long long* slots = part + SCRATCH_PART;
    unsigned long long* packed = slots + (size_t)CAH_COUNTS_NOALLOC_MAX * CNT_SLOTS;
    unsigned long long* peq = packed + (size_t)CAH_COUNTS_NOALLOC_MAX * 256;
    // every member's counters (k_counts_reduce writes all CNT_SLOTS of a member's
    // slot), then the outputs' layout, then one copy of each output
    for (int64_t i = 0; i < n; ++i)
        if (int rc = queue_counts(b->members[i], other, slots + i * CNT_SLOTS, part, b->stream)) return rc;
    const int pblocks = (int)std::min<int64_t>((n * 257 + 255) / 256, 1024);
    if (int rc = checked(false, "k_counts_pack", [&](lchk::Chk& c) {
            c.range(slots, 0, 8 * n * CNT_SLOTS, "counters");
            c.range(packed, 0, 8 * n * 256, "packed counts");
            c.range(peq, 0, 8 * n, "packed agreements");
        }))
        return rc;
    k_counts_pack<<<pblocks, 256, 0, b->stream>>>(slots, (int32_t)n, packed, peq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(counts, packed, (size_t)n * 256 * sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
    if (equal) HIPCHK(hipMemcpyAsync(equal, peq, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    // a lattice holding values outside the rule's own: all 256 byte values
    for (int64_t i = 0; i < n; ++i) {
        const cah_state* m = b->members[i];
        int64_t sum = 0;
        for (int v = 0; v < 256; ++v) sum += counts[i * 256 + v];
        if (sum == m->n1 * m->n2) continue;
        unsigned long long* row = packed + i * 256;
        const int hb = (int)std::max<int64_t>(1, std::min<int64_t>((m->n1 * m->n2 + 255) / 256, 4096));
        if (int rc = checked(false, "k_hist256", [&](lchk::Chk& c) {
                lchk::lattice(c, m->buf[m->cur], lat_of(m), 1, "lattice");
                c.range(row, 0, 8LL * 256, "histogram");
            }))
            return rc;
        k_zero_u64<<<1, 256, 0, b->stream>>>(row, 256);
        k_hist256<<<hb, 256, 0, b->stream>>>((const uint8_t*)m->buf[m->cur], m->pitch, m->n1, m->n2, row);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(counts + i * 256, row, 256 * sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return CAH_OK;
}

extern "C" int cah_batch_fill_bernoulli(cah_batch* b, int32_t i, uint64_t seed, double p) {
    cah_state* s;
    if (int rc = member(b, i, &s)) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    return cah_state_fill_bernoulli(s, seed, p);
}

static int forest_prec(const cah_params& p) { return (p.ff_p_is_f32 ? 1 : 0) | (p.ff_f_is_f32 ? 2 : 0); }

// clean: every lattice and Constant edge holds only 0 and 1 (Life) or only 0, 1
// and 2 (the forest fire), which lets the kernels skip the exact byte tests.
// The checks of a batch's launches at buffer parity `par` (every member's
// geometry, both kernels' tables), once per layout and parity.
static int check_batch(cah_batch* b, int par) {
    if ((!b->dry && !lchk::enabled()) || b->checked_parity[par]) return CAH_OK;
    const int64_t n = (int64_t)b->members.size();
    auto swap = [&](auto a) {
        if (par) {
            const auto c = a.cur;
            a.cur = a.next;
            a.next = const_cast<decltype(a.next)>(c);
        }
        return a;
    };
    if (b->tb) {
        const TbShape sh = TB_SHAPES[b->tb_shape];
        if (int rc = checked(b->dry, "k_life_tb_batch", [&](lchk::Chk& c) {
                lchk::item_table(c, b->tb_item_h.data(), b->tb_base_h.data(), b->tb_blocks, n,
                                 [&](int64_t k) { return (int64_t)tb_tiles(b->members[k], b->tb_shape); });
                for (int64_t k = 0; k < n && c.ok(); ++k) {
                    const LifeTbArgs a = swap(b->tb_items_h[k]);
                    const int64_t gx = (a.L.n1 + 32 * sh.txw - 1) / (32 * sh.txw), gy = (a.L.n2 + sh.ty - 1) / sh.ty;
                    lchk::life_tb(c, a.L, a.cur, a.next, sh.txw, sh.ty, TB_K, 1, gx, gy);
                }
            }))
            return rc;
    }
    if (int rc = checked(b->dry, b->kind == B_FOREST ? "k_forest_batch" : "k_life_batch", [&](lchk::Chk& c) {
            lchk::item_table(c, b->wave_item_h.data(), b->wave_base_h.data(), b->nwaves, n, [&](int64_t k) {
                return (int64_t)(b->kind == B_FOREST ? b->fitems[k].nwaves : b->litems[k].nwaves);
            });
            for (int64_t k = 0; k < n && c.ok(); ++k) {
                const int R = b->kind == B_FOREST ? 1 : b->life_R;
                if (b->kind == B_FOREST) {
                    const ForestArgs a = swap(b->fitems[k]);
                    if (k == 0) c.range(b->t0_d, 0, 8, "batch time base");
                    if (b->vb == 1) lchk::band_walk<1>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, 1, 3);
                    else lchk::band_walk<2>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, 1, 3);
                } else {
                    const LifeArgs a = swap(b->litems[k]);
                    if (b->vb == 1)
                        lchk::band_walk<1>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, R, 2 * R + 1);
                    else
                        lchk::band_walk<2>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, R, 2 * R + 1);
                }
            }
        }))
        return rc;
    b->checked_parity[par] = true;
    return CAH_OK;
}

// One 8-byte value into device memory in stream order.
__global__ void k_store_i64(int64_t* p, int64_t v) {
    if (threadIdx.x == 0) *p = v;
}

// Step k of a cah_batch_advance (the forest fire's time: *t0_d + k).
static void launch_batch_step(cah_batch* b, int64_t k, bool clean) {
    const int blocks = waves_to_blocks(b->nwaves);
    if (b->dry) {
        b->parity ^= 1;
        return;
    }
    if (b->kind == B_FOREST)
        launch_forest_batch(b->vb, b->small, clean, forest_prec(b->fp), blocks, (const ForestArgs*)b->items_d,
                            b->wave_item_d, b->wave_base_d, b->nwaves, b->parity, k, b->t0_d, b->stream);
    else
        launch_life_batch(b->life_R, b->vb, clean, b->niv, blocks, (const LifeArgs*)b->items_d,
                          b->wave_item_d, b->wave_base_d, b->nwaves, b->parity, b->stream);
    b->parity ^= 1;
}

extern "C" int cah_batch_advance(cah_batch* b, int64_t t0, int64_t nsteps, const cah_params* p,
                                 int32_t use_graph) {
    if (!b) return fail(CAH_ERR_ARGUMENT, "batch is NULL");
    if (nsteps < 0 || t0 < 0) return fail(CAH_ERR_ARGUMENT, "t0 and nsteps must be >= 0");
    if (!b->dry) HIPCHK(hipSetDevice(b->device));
    if (b->kind == B_LOOP) {
        for (auto* s : b->members)
            if (int rc = cah_advance(s, t0, nsteps, p)) return rc;
        return CAH_OK;
    }
    if (b->kind == B_FOREST && !p)
        return fail(CAH_ERR_ARGUMENT, "the forest fire needs its parameters (p, f)");
    if (b->kind == B_FOREST && !b->t0_d) {
        if (b->dry) b->t0_d = (int64_t*)lchk::dry_alloc(8, "batch time base (dry)");
        else HIPCHK(dmalloc((void**)&b->t0_d, 8, "batch time base"));
    }
    // members may have been set through their own streams
    if (!b->dry)
        for (auto* s : b->members) HIPCHK(hipStreamSynchronize(s->stream));
    if (b->kind == B_FOREST &&
        (!b->fp_set || b->fp.ff_p != p->ff_p || b->fp.ff_f != p->ff_f ||
         b->fp.ff_p_is_f32 != p->ff_p_is_f32 || b->fp.ff_f_is_f32 != p->ff_f_is_f32)) {
        // the items carry p and f; the host copy is rewritten only once no
        // queued copy still reads it
        if (!b->dry) HIPCHK(hipStreamSynchronize(b->stream));
        for (ForestArgs& a : b->fitems) {
            a.p = p->ff_p;
            a.f = p->ff_f;
            a.p_f32 = p->ff_p_is_f32;
            a.f_f32 = p->ff_f_is_f32;
        }
        if (!b->dry)
            HIPCHK(hipMemcpyAsync(b->items_d, b->fitems.data(), b->fitems.size() * sizeof(ForestArgs),
                                  hipMemcpyHostToDevice, b->stream));
        b->fp = *p;
        b->fp_set = true;
    }
    bool bin = true, tri = true;
    for (auto* s : b->members) {
        auto edge = [&](int k, double top) {
            const double v = s->d.boundary_value[k];
            return s->d.boundary[k] != CAH_BOUNDARY_CONSTANT || (v >= 0 && v <= top && v == (double)(int)v);
        };
        bin &= s->bin_clean && edge(0, 1) && edge(1, 1);
        tri &= s->tri_clean && edge(0, 2) && edge(1, 2);
    }
    if (b->kind == B_FOREST) bin = tri;  // the forest kernels' short byte tests
    // Moore(1) Life: up to TB_K steps a launch (the items run TB_K - step0 steps),
    // a last single step through the band walk, as a single lattice advances
    // the launch checks of both buffer parities, before anything is queued
    for (int par = 0; par < 2; ++par)
        if (int rc = check_batch(b, par)) return rc;
    auto queue = [&]() {
        int64_t k = 0;
        if (b->tb)
            for (; nsteps - k >= 2; b->parity ^= 1) {
                const int32_t n = (int32_t)std::min<int64_t>(nsteps - k, TB_K);
                if (!b->dry)
                    launch_life_tb_batch(b->tb_shape, b->tb_blocks, b->tb_items_d, b->tb_item_d, b->tb_base_d,
                                         b->parity, TB_K - n, b->stream);
                k += n;
            }
        for (; k < nsteps; ++k) launch_batch_step(b, k, bin);
    };
    // the starting time, read by the forest kernels on the device (a kernel
    // argument, so the value is taken when the launch is queued)
    if (b->kind == B_FOREST && !b->dry) k_store_i64<<<1, 64, 0, b->stream>>>(b->t0_d, t0);
    if (use_graph && nsteps > 1 && !b->dry) {
        if (b->exec && b->g_nsteps == nsteps && b->g_parity0 == b->parity && b->g_clean == bin) {
            b->parity = b->g_parity1;  // the captured launches, replayed
        } else {
            if (b->exec) {
                HIPCHK(hipStreamSynchronize(b->stream));
                HIPCHK(hipGraphExecDestroy(b->exec));
                b->exec = nullptr;
            }
            b->g_nsteps = nsteps;
            b->g_parity0 = b->parity;
            b->g_clean = bin;
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            queue();
            HIPCHK(hipStreamEndCapture(b->stream, &graph));
            HIPCHK(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(graph));
            b->g_parity1 = b->parity;
        }
        HIPCHK(hipGraphLaunch(b->exec, b->stream));
    } else {
        queue();
    }
    if (!b->dry) HIPCHK(hipGetLastError());
    for (auto* s : b->members) {
        s->cur = b->parity;
        if (b->kind == B_FOREST && nsteps > 0) {
            s->bin_clean = false;
            s->tri_clean = true;  // the forest fire writes 0, 1 and 2 only
        }
        if (b->kind == B_LIFE && nsteps > 0) s->bin_clean = true;  // Life writes 0 and 1 only
    }
    return CAH_OK;
}

// The band walk's tables, freed before batch_layout builds them again.
static int free_layout(cah_batch* b) {
    if (!b->dry) HIPCHK(hipStreamSynchronize(b->stream));
    if (b->exec) HIPCHK(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
    b->g_nsteps = -1;
    for (void** q : {&b->items_d, (void**)&b->wave_item_d, (void**)&b->wave_base_d}) {
        dfree(*q, b->dry);
        *q = nullptr;
    }
    b->fp_set = false;  // the new forest items carry no p and f yet
    return CAH_OK;
}

extern "C" int cah_batch_tune(cah_batch* b, int32_t vb, int32_t rows) {
    if (!b || vb < 0 || vb > 2 || rows < 0) return fail(CAH_ERR_ARGUMENT, "cah_batch_tune: bad arguments");
    if (b->kind == B_LOOP || b->tb) return CAH_OK;  // no band walk
    if (b->kind == B_LIFE && life_min_vb(b->life_R) == 2 && vb == 1)
        return fail(CAH_ERR_ARGUMENT, "radius-5 Life needs 8 cells a lane (vb = 2)");
    if (!b->dry) HIPCHK(hipSetDevice(b->device));
    b->vb = vb ? vb : (b->kind == B_LIFE ? life_min_vb(b->life_R) : 1);
    b->rows = rows;
    if (int rc = free_layout(b)) return rc;
    return batch_layout(b);
}

extern "C" int cah_batch_autotune(cah_batch* b, int64_t t0, int32_t steps, const cah_params* p, int32_t* vb,
                                  int32_t* rows) {
    if (!b || !vb || !rows || steps < 1 || t0 < 0) return fail(CAH_ERR_ARGUMENT, "cah_batch_autotune: bad arguments");
    *vb = *rows = 0;
    if (b->kind == B_LOOP || b->tb) return CAH_OK;  // no band walk to tune
    if (b->kind == B_FOREST && !p) return fail(CAH_ERR_ARGUMENT, "the forest fire needs its parameters (p, f)");
    HIPCHK(hipSetDevice(b->device));
    for (auto* s : b->members) HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    // every trajectory's lattice, to undo the trial steps
    const int par0 = b->parity;
    std::vector<void*> saved(b->members.size(), nullptr);
    std::vector<bool> bin0, tri0;
    auto release = [&]() {
        for (void* q : saved)
            if (q) hipFree(q);
    };
    size_t total = 0;
    for (auto* s : b->members) total += (size_t)(s->pitch * s->n2) * s->es;
    for (size_t i = 0; i < b->members.size(); ++i) {
        cah_state* s = b->members[i];
        const size_t bytes = (size_t)(s->pitch * s->n2) * s->es;
        if (hipMalloc(&saved[i], bytes) != hipSuccess) {
            release();
            return fail(CAH_ERR_OUT_OF_MEMORY, "cah_batch_autotune: no device memory for a copy of the batch "
                                               "(%zu bytes)", total);
        }
        HIPCHK(hipMemcpyAsync(saved[i], s->buf[s->cur], bytes, hipMemcpyDeviceToDevice, b->stream));
        bin0.push_back(s->bin_clean);
        tri0.push_back(s->tri_clean);
    }
    auto restore = [&]() -> int {
        HIPCHK(hipStreamSynchronize(b->stream));
        b->parity = par0;
        for (size_t i = 0; i < b->members.size(); ++i) {
            cah_state* s = b->members[i];
            s->cur = par0;
            s->bin_clean = bin0[i];
            s->tri_clean = tri0[i];
            HIPCHK(hipMemcpyAsync(s->buf[par0], saved[i], (size_t)(s->pitch * s->n2) * s->es,
                                  hipMemcpyDeviceToDevice, b->stream));
        }
        HIPCHK(hipStreamSynchronize(b->stream));
        return CAH_OK;
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int32_t keep = b->rows, keep_vb = b->vb;
    int32_t best = 0, best_vb = 0;
    float best_ms = 0;
    int rc = CAH_OK;
    // cells a lane: 4 or 8 (radius 5 needs 8); rows a wave: 4 .. 64
    std::vector<std::pair<int, int32_t>> cands;
    for (int v : {1, 2})
        if (!(b->kind == B_LIFE && life_min_vb(b->life_R) == 2 && v == 1))
            for (int32_t r : {4, 8, 16, 32, 64}) cands.push_back({v, r});
    for (auto [cand_vb, cand] : cands) {
        b->vb = cand_vb;
        b->rows = cand;
        if ((rc = free_layout(b)) || (rc = batch_layout(b))) break;
        if ((rc = cah_batch_advance(b, t0, 1, p, 0))) break;  // warm-up: code and caches
        if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = fail(CAH_ERR_DEVICE, "hipEventRecord failed"); break; }
        if ((rc = cah_batch_advance(b, t0 + 1, steps, p, 0))) break;
        float ms = 0;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(CAH_ERR_DEVICE, "timing a trial layout failed");
            break;
        }
        if (!best || ms < best_ms) {
            best = cand;
            best_vb = cand_vb;
            best_ms = ms;
        }
        if ((rc = restore())) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!rc) {
        b->vb = best_vb;
        b->rows = best;
        if (!(rc = free_layout(b))) rc = batch_layout(b);
    } else {
        b->vb = keep_vb;
        b->rows = keep;
        restore();
        if (!free_layout(b)) batch_layout(b);
    }
    if (!rc) rc = restore();
    release();
    if (!rc) {
        *vb = best_vb;
        *rows = best;
    }
    return rc;
}

extern "C" int cah_batch_sync(cah_batch* b) {
    if (!b) return fail(CAH_ERR_ARGUMENT, "batch is NULL");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto* s : b->members)
        if (int rc = cah_state_sync(s)) return rc;
    return CAH_OK;
}

extern "C" int cah_batch_stream(cah_batch* b, void** stream) {
    if (!b || !stream) return fail(CAH_ERR_ARGUMENT, "NULL argument");
    *stream = (void*)b->stream;
    return CAH_OK;
}

extern "C" int cah_batch_destroy(cah_batch* b) {
    free_batch(b);
    return CAH_OK;
}

// ---------------------------------------------------------------- launch checks without a device
//
// A dry state or batch has fake device addresses and makes no HIP call: its
// advance runs the launch sequence through the launch checks only.

static cah_ctx g_dry_ctx{0};

static void tally_report(char* report, size_t cap) {
    if (!report || !cap) return;
    std::string r;
    for (const auto& kv : g_tally) {
        char line[256];
        snprintf(line, sizeof line, "%s: %lld launches, %lld checks\n", kv.first.c_str(), (long long)kv.second.first,
                 (long long)kv.second.second);
        r += line;
    }
    snprintf(report, cap, "%s", r.c_str());
}

extern "C" int cah_check_launches(const cah_desc* d, int64_t t0, int64_t nsteps, const cah_params* p, char* report,
                                  size_t cap) {
    g_tally.clear();
    cah_state* s = nullptr;
    int rc = new_state(&g_dry_ctx, d, &s, true);
    if (!rc) rc = cah_advance(s, t0, nsteps, p);
    const std::string keep = g_err;
    free_state(s);
    g_err = keep;
    tally_report(report, cap);
    return rc;
}

extern "C" int cah_check_batch_launches(int32_t n, const cah_desc* descs, int64_t t0, int64_t nsteps,
                                        const cah_params* p, int32_t vb, int32_t rows, char* report, size_t cap) {
    g_tally.clear();
    cah_batch* b = nullptr;
    int rc = batch_create(&g_dry_ctx, n, descs, &b, true);
    if (!rc && (vb || rows)) rc = cah_batch_tune(b, vb, rows);
    if (!rc) rc = cah_batch_advance(b, t0, nsteps, p, 0);
    const std::string keep = g_err;
    free_batch(b);
    g_err = keep;
    tally_report(report, cap);
    return rc;
}

// ---------------------------------------------------------------- serving
//
// Continuous batching.  Requests wait in a FIFO until the device-memory budget
// admits them; an admitted request is a state whose buffers come from the
// server's arena.  Every tick queues the next chunk of steps of every resident
// request: requests of one kernel family (Larger-than-Life of one radius and
// interval count, the forest fire of one precision, MFMA Lenia of one mode)
// advance together, one batched launch per step over all their lattices, on the
// family's own stream, so the families overlap on the device; the others run
// their own launches on a fourth stream.  A request that arrives mid-run joins
// the next tick.  At most two ticks are in flight, so the device never waits for
// the host between them and a new request waits at most about one tick.  The
// per-tick item and wave tables are written to pinned host memory and copied in
// one transfer on the upload stream, which also carries the admitted states'
// initial contents; every family stream waits for that copy.  (A pipeline of
// ticks per family, so that a family with little work per tick need not wait
// for one with much, measured slower on ensemble traffic on MI355X.)  Results
// stay on the device until fetched, through a separate copy stream.

static double steady_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" double cah_server_clock(void) { return steady_now(); }

enum SrvClass { SC_LIFE = 0, SC_FOREST = 1, SC_LENIA = 2, SC_LOOP = 3 };
constexpr int SRV_STREAMS = 4;
constexpr int SRV_SLOTS = 2;

struct SrvReq {
    uint64_t id = 0;
    cah_desc d{};
    std::vector<int32_t> offsets, taps;
    std::vector<double> weights;
    std::vector<uint8_t> mask;
    std::vector<double> cov;             // field epidemic: covariate slices
    std::vector<char> u0;                // initial state, when it is not in pinned memory
    char* u0_pinned = nullptr;           // initial state in the server's pinned arena
    size_t u0_bytes = 0;
    bool u0_binary = false;              //   ... every byte of it 0 or 1 (found while copying it)
    int64_t t0 = 0, t1 = 0, queued = 0;  // queued: the time its queued steps reach
    cah_params p{};
    cah_state* st = nullptr;
    int cls = SC_LOOP, key = 0;          // family and kernel variant
    bool tail_queued = false;            // its last step is in a queued tick
    bool in_tick = false;                // some queued tick follows its uploads
    int32_t status = CAH_REQ_PENDING;
    bool fetching = false;               // a fetch is copying its result out
    size_t bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int32_t steps = 0;                   // steps in the tick being built
    bool tb = false;                     // its tick runs k_life_tb launches (TB_K steps each)
};

struct SrvSlot {
    hipEvent_t up = nullptr;
    hipEvent_t done[SRV_STREAMS] = {};
    char* host = nullptr;  // pinned
    size_t host_cap = 0;
    char* dev = nullptr;
    size_t dev_cap = 0;
    std::vector<SrvReq*> finishing;
    std::vector<SrvReq*> staged;  // requests whose initial uploads precede this tick
    bool busy = false;
};

// Pinned memory that the initial states of waiting requests may hold; beyond it
// they wait in pageable memory and are staged when admitted.
constexpr size_t SRV_PENDING_PINNED = (size_t)4 << 30;
// Initial-state bytes one tick call admits at most (beyond its first request).
constexpr size_t SRV_ADMIT_BYTES = (size_t)256 << 20;

struct cah_server {
    // every entry point holds it, but a fetch drops it while its result is copied
    // out, so one thread can fetch while another submits and ticks
    std::mutex mu;
    cah_ctx* ctx = nullptr;
    size_t budget = 0, used = 0;
    size_t pending_pinned = 0;  // pinned bytes of waiting requests' initial states
    int32_t chunk = 0;
    Arena arena;   // device memory of the resident lattices and tick tables
    Arena pinned;  // staging of initial states and results
    hipStream_t up = nullptr, copy = nullptr;
    hipStream_t st[SRV_STREAMS] = {};
    SrvSlot slot[SRV_SLOTS];
    int oldest = 0, nbusy = 0;
    std::deque<SrvReq*> pending;
    std::vector<SrvReq*> active;  // admitted and not yet done
    std::unordered_map<uint64_t, SrvReq*> all;
    std::deque<uint64_t> finished;  // done since the last poll
    uint64_t next = 1;
    double tick_ps = 0;             // estimated device work per tick (0: no limit)
};

// Device bytes a request holds while resident: both buffers, the mask, and the
// Lenia field or spectra of the paths that keep one.
static size_t request_bytes(const cah_desc& d) {
    const int64_t n1 = d.dims[0], n2 = d.ndims == 2 ? d.dims[1] : 1;
    const bool f = d.dtype == CAH_DTYPE_F32;
    const int64_t pitch = f ? (n1 + 15) / 16 * 16 : (n1 + 63) / 64 * 64;
    size_t b = 2 * (size_t)pitch * (size_t)n2 * (f ? 4 : 1);
    if (d.mask) b += (size_t)(n1 * n2);
    if (d.rule == CAH_RULE_LENIA) {
        const int fm = d.field_method;
        const bool mfma = fm == CAH_FIELD_MFMA_F16 || fm == CAH_FIELD_MFMA_BF16 || fm == CAH_FIELD_MFMA_F16X2;
        b += (size_t)(n1 * n2) * (fm == CAH_FIELD_FFT_F32 ? 16 : mfma ? 0 : 8);
        b += (size_t)1 << 20;  // taps, weights, tables, fragments
    }
    if (d.rule == CAH_RULE_FIELD_EPIDEMIC) {
        b += (size_t)(n1 * n2) * (d.lenia_precision == 32 ? 4 : 8);
        if (d.covariate) b += (size_t)(n1 * n2) * (size_t)d.covariate_slices * 8;
        b += (size_t)d.n_taps * 16;
    }
    return b;
}

static void free_slot(SrvSlot& sl) {
    if (sl.up) hipEventDestroy(sl.up);
    for (hipEvent_t e : sl.done)
        if (e) hipEventDestroy(e);
    if (sl.host) hipHostFree(sl.host);
}

extern "C" int cah_server_destroy(cah_server* s) {
    if (!s) return CAH_OK;
    hipSetDevice(s->ctx->device);
    for (hipStream_t q : s->st)
        if (q) hipStreamSynchronize(q);
    if (s->up) hipStreamSynchronize(s->up);
    if (s->copy) hipStreamSynchronize(s->copy);
    for (auto& kv : s->all) {
        if (kv.second->st) free_state(kv.second->st);
        delete kv.second;
    }
    for (SrvSlot& sl : s->slot) free_slot(sl);
    for (hipStream_t q : s->st)
        if (q) hipStreamDestroy(q);
    if (s->up) hipStreamDestroy(s->up);
    if (s->copy) hipStreamDestroy(s->copy);
    delete s;  // the arena's slabs go with it
    return CAH_OK;
}

extern "C" int cah_server_create(cah_ctx* ctx, size_t budget_bytes, int32_t chunk, cah_server** out) {
    if (!ctx || !out) return fail(CAH_ERR_ARGUMENT, "cah_server_create: NULL argument");
    if (chunk < 0) return fail(CAH_ERR_ARGUMENT, "chunk must be >= 0");
    HIPCHK(hipSetDevice(ctx->device));
    cah_server* s = new cah_server();
    s->ctx = ctx;
    s->budget = budget_bytes;
    s->chunk = chunk;
    s->arena.slab_min = std::min<size_t>(s->arena.slab_min, std::max<size_t>(budget_bytes, 1 << 20));
    s->pinned.pinned = true;
    s->pinned.slab_min = (size_t)64 << 20;
    auto make = [&]() -> int {
        for (hipStream_t& q : s->st) HIPCHK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking));
        for (SrvSlot& sl : s->slot) {
            HIPCHK(hipEventCreateWithFlags(&sl.up, hipEventDisableTiming));
            for (hipEvent_t& e : sl.done) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        return CAH_OK;
    };
    if (int rc = make()) {
        std::string keep = g_err;
        cah_server_destroy(s);
        g_err = keep;
        return rc;
    }
    *out = s;
    return CAH_OK;
}

extern "C" int cah_server_submit(cah_server* s, const cah_desc* d, const void* u0, size_t nbytes,
                                 int64_t t0, int64_t t1, const cah_params* p, uint64_t* id) {
    if (!s || !d || !id) return fail(CAH_ERR_ARGUMENT, "cah_server_submit: NULL argument");
    if (int rc = cah_check_problem(d)) return rc;
    if (t0 < 0 || t1 < t0) return fail(CAH_ERR_ARGUMENT, "need 0 <= t0 <= t1");
    if (d->rule == CAH_RULE_FOREST_FIRE && !p)
        return fail(CAH_ERR_ARGUMENT, "the forest fire needs its parameters (p, f)");
    const int64_t cells = d->dims[0] * (d->ndims == 2 ? d->dims[1] : 1);
    const size_t want = (size_t)cells * (d->dtype == CAH_DTYPE_F32 ? 4 : 1);
    if (u0 && nbytes != want) return fail(CAH_ERR_STATE_SHAPE, "state has %zu bytes, got %zu", want, nbytes);
    const size_t bytes = request_bytes(*d);
    std::unique_lock<std::mutex> lk(s->mu);
    if (bytes > s->budget)
        return fail(CAH_ERR_BUSY, "the request needs %zu bytes, more than the whole budget", bytes);
    SrvReq* r = new SrvReq();
    r->d = *d;
    r->offsets.assign(d->offsets, d->offsets + 2 * (size_t)d->n_offsets);
    r->d.offsets = r->offsets.data();
    if (d->rule == CAH_RULE_LENIA || d->rule == CAH_RULE_FIELD_EPIDEMIC) {
        r->taps.assign(d->tap_offsets, d->tap_offsets + 2 * (size_t)d->n_taps);
        r->weights.assign(d->tap_weights, d->tap_weights + d->n_taps);
        r->d.tap_offsets = r->taps.data();
        r->d.tap_weights = r->weights.data();
    }
    if (d->rule == CAH_RULE_FIELD_EPIDEMIC && d->covariate) {
        r->cov.assign(d->covariate, d->covariate + (size_t)cells * (size_t)d->covariate_slices);
        r->d.covariate = r->cov.data();
    }
    if (d->mask) {
        r->mask.assign(d->mask, d->mask + cells);
        r->d.mask = r->mask.data();
    }
    if (u0 && s->pending_pinned + nbytes <= SRV_PENDING_PINNED) {
        hipSetDevice(s->ctx->device);
        r->u0_pinned = (char*)s->pinned.alloc(nbytes);
        if (r->u0_pinned) {
            r->u0_bytes = nbytes;
            s->pending_pinned += nbytes;
        }
    }
    // the copy runs without the lock; the request is not visible until it ends
    lk.unlock();
    if (r->u0_pinned) {
        bool above = true;
        hostcopy::copy(r->u0_pinned, u0, nbytes, d->dtype == CAH_DTYPE_I8 ? &above : nullptr);
        r->u0_binary = !above;
    } else if (u0)
        r->u0.assign((const char*)u0, (const char*)u0 + nbytes);
    lk.lock();
    r->t0 = r->queued = t0;
    r->t1 = t1;
    if (p) r->p = *p;
    r->bytes = bytes;
    r->id = s->next++;
    r->t_submit = steady_now();
    s->pending.push_back(r);
    s->all[r->id] = r;
    *id = r->id;
    return CAH_OK;
}

static bool edges012(const cah_state* st) {
    const cah_desc& d = st->d;
    auto ok = [&](int k) {
        const double v = d.boundary_value[k];
        return d.boundary[k] != CAH_BOUNDARY_CONSTANT || v == 0 || v == 1 || v == 2;
    };
    return ok(0) && ok(1);
}

static bool edges01(const cah_state* st) {
    const cah_desc& d = st->d;
    return (d.boundary[0] != CAH_BOUNDARY_CONSTANT || d.boundary_value[0] == 0 || d.boundary_value[0] == 1) &&
           (d.boundary[1] != CAH_BOUNDARY_CONSTANT || d.boundary_value[1] == 0 || d.boundary_value[1] == 1);
}

// Create the request's state on the upload stream, queue its initial contents
// and pick its family.
static int admit(cah_server* s, SrvReq* r) {
    cah_state* st = new cah_state();
    if (int rc = create_state(s->ctx, &r->d, st, &s->arena, &s->pinned, s->up)) {
        std::string keep = g_err;
        free_state(st);
        g_err = keep;
        return rc;
    }
    r->st = st;
    if (r->u0_pinned || !r->u0.empty()) {
        void* h;
        size_t n;
        const bool pinned = r->u0_pinned != nullptr;
        if (pinned) {
            // released with the state's other staging once the upload has completed
            h = r->u0_pinned;
            n = r->u0_bytes;
            st->keep.push_back(h);
            s->pending_pinned -= n;
            r->u0_pinned = nullptr;
        } else {
            n = r->u0.size();
            if (int rc = stage(st, r->u0.data(), n, &h)) return rc;
        }
        const size_t row = (size_t)st->n1 * st->es, pitchb = (size_t)st->pitch * st->es;
        void* dense = row == pitchb ? st->buf[st->cur] : st->buf[1 - st->cur];
        HIPCHK(hipMemcpyAsync(dense, h, n, hipMemcpyHostToDevice, s->up));
        if (row != pitchb)
            if (int rc = copy_rows(dense, (int64_t)row, st->buf[st->cur], (int64_t)pitchb, (int64_t)row,
                                   st->n2, s->up))
                return rc;
        if (st->es == 1) st->bin_clean = pinned ? r->u0_binary : scan_binary((const int8_t*)h, n);
        if (st->es == 1) st->tri_clean = st->small_clean = st->bin_clean;
        std::vector<char>().swap(r->u0);
    }
    switch (st->path) {
    case P_LIFE_SWAR:
        r->cls = SC_LIFE;
        // Moore(1) runs k_life_tb, which takes any B/S masks and contents: one group
        r->key = st->life_R == 1 ? 4
                                 : st->life_R * 4 + (life_niv(st->life) - 1) * 2 + (st->bin_clean && edges01(st) ? 1 : 0);
        break;
    case P_FOREST:
        r->cls = SC_FOREST;
        r->key = forest_prec(r->p) * 2 + ((uint64_t)(st->n1 * st->n2) < 0xFFFFFFFFull ? 1 : 0);
        break;
    case P_LENIA_MFMA:
        r->cls = SC_LENIA;
        r->key = st->lenia_mode;
        break;
    default:
        r->cls = SC_LOOP;
        r->key = 0;
    }
    // launches of its own (the fourth family) go on the loop stream
    st->stream = s->st[SC_LOOP];
    r->status = CAH_REQ_ACTIVE;
    r->t_admit = steady_now();
    return CAH_OK;
}

// Retire the oldest in-flight ticks that have completed.
static int retire(cah_server* s) {
    while (s->nbusy > 0) {
        SrvSlot& sl = s->slot[s->oldest];
        for (hipEvent_t e : sl.done) {
            const hipError_t q = hipEventQuery(e);
            if (q == hipErrorNotReady) return CAH_OK;
            if (q != hipSuccess) return fail(CAH_ERR_DEVICE, "server tick: %s", hipGetErrorString(q));
        }
        const double now = steady_now();
        // the uploads before the tick's table copy have completed
        for (SrvReq* r : sl.staged) release_staging(r->st);
        sl.staged.clear();
        for (SrvReq* r : sl.finishing) {
            r->status = CAH_REQ_DONE;
            r->t_done = now;
            s->finished.push_back(r->id);
            s->active.erase(std::find(s->active.begin(), s->active.end(), r));
        }
        sl.finishing.clear();
        sl.busy = false;
        s->oldest = (s->oldest + 1) % SRV_SLOTS;
        --s->nbusy;
    }
    return CAH_OK;
}

// One family's requests of one kernel variant, ordered by steps in this tick
// (most first): the items the launch of step k advances are a prefix.
struct SrvGroup {
    int cls = 0, key = 0;
    std::vector<SrvReq*> reqs;
    size_t off_items = 0, off_map = 0, off_base = 0;  // in the tick's table buffer
    int64_t rows = 64;                                 // rows per wave (band walks)
    std::vector<int32_t> base;                         // first wave/block of each item, + total
    bool tb = false;                                   // Moore(1) Life through k_life_tb
    int shape = 0;                                     // its tile shape
    bool tri_edges = false, tri_cells = false;         // forest fire: edges / lattices hold only 0, 1, 2
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Estimated device time of one cell update in picoseconds, per family, from the
// batched kernels' rates on MI355X (Life and Larger-than-Life about 2400 GCUPS,
// the forest fire about 600, MFMA Lenia with a 27-row kernel about 95); only the
// size of a tick depends on it.
static double cell_ps(const SrvReq* r) {
    switch (r->cls) {
    case SC_LIFE: return 0.5;
    case SC_FOREST: return 1.7;
    case SC_LENIA: return 11.0 * (2 * r->st->lenia_H + 1) / 27.0;
    default: return 10.0;
    }
}

static int build_tick(cah_server* s) {
    // Steps go to the resident requests in arrival order: each gets up to chunk
    // steps while the tick's work budget (if one is set) lasts, the oldest at
    // least one, and a younger request only the steps that still fit.  Without a
    // budget every resident request advances chunk steps a tick; measured on
    // MI355X that serves both large-lattice and many-small-lattice traffic best,
    // since the batched launches fill the device with the small ones at little
    // cost to the large.
    std::vector<SrvReq*> run;
    double budget = s->tick_ps > 0 ? s->tick_ps : 1e30;
    bool oldest = true;
    for (SrvReq* r : s->active) {
        if (r->tail_queued) continue;
        const int64_t left = r->t1 - r->queued;
        int64_t steps = s->chunk > 0 ? std::min<int64_t>(s->chunk, left) : left;
        if (left > 0) {
            const double per = (double)(r->st->n1 * r->st->n2) * cell_ps(r);
            const int64_t fit = (int64_t)std::min(budget / per, 1e18);
            if (fit < steps) steps = oldest ? std::max<int64_t>(1, fit) : fit;
            if (steps == 0) continue;
            budget -= (double)steps * per;
            oldest = false;
        }
        r->steps = (int32_t)steps;
        run.push_back(r);
    }
    if (run.empty()) return CAH_OK;
    SrvSlot& sl = s->slot[(s->oldest + s->nbusy) % SRV_SLOTS];

    // group by family and variant
    std::map<std::pair<int, int>, SrvGroup> groups;
    for (SrvReq* r : run) {
        r->tb = false;
        if (r->cls == SC_LOOP || r->steps == 0) continue;
        SrvGroup& g = groups[{r->cls, r->key}];
        g.cls = r->cls;
        g.key = r->key;
        g.tb = r->tb = r->cls == SC_LIFE && r->key == 4;
        if (g.reqs.empty()) g.tri_edges = g.tri_cells = true;
        g.tri_edges &= edges012(r->st);
        g.tri_cells &= r->st->tri_clean;
        g.reqs.push_back(r);
    }
    // table layout: per group items, wave (block) -> item map, item bases
    size_t off = 0;
    std::vector<std::vector<int32_t>> maps;
    for (auto& kv : groups) {
        SrvGroup& g = kv.second;
        std::stable_sort(g.reqs.begin(), g.reqs.end(),
                         [](const SrvReq* a, const SrvReq* b) { return a->steps > b->steps; });
        std::vector<int32_t> map;
        int64_t work = 0;
        for (const SrvReq* r : g.reqs) {
            const int64_t ow = 62 * 4 * (g.cls == SC_LIFE ? life_min_vb(r->st->life_R) : 1);
            work += (r->st->n1 + ow - 1) / ow * r->st->n2;
        }
        g.rows = batch_rows(work);
        if (g.tb) {
            std::vector<const cah_state*> m;
            for (const SrvReq* r : g.reqs) m.push_back(r->st);
            g.shape = tb_pick_shape(m);
        }
        for (size_t k = 0; k < g.reqs.size(); ++k) {
            const cah_state* st = g.reqs[k]->st;
            int32_t n;
            if (g.tb) {
                n = tb_tiles(st, g.shape);
            } else if (g.cls == SC_LENIA) {
                n = (int32_t)(((st->n1 + LM_BX - 1) / LM_BX) * ((st->n2 + LM_BY - 1) / LM_BY));
            } else {
                const int vb = g.cls == SC_LIFE ? life_min_vb(st->life_R) : 1;
                const int64_t ow = 62 * 4 * vb;
                const int64_t rows = std::min<int64_t>(g.rows, st->n2);
                n = (int32_t)(((st->n1 + ow - 1) / ow) * ((st->n2 + rows - 1) / rows));
            }
            g.base.push_back((int32_t)map.size());
            map.insert(map.end(), (size_t)n, (int32_t)k);
        }
        g.base.push_back((int32_t)map.size());
        const size_t isz = g.tb ? sizeof(LifeTbArgs)
                           : g.cls == SC_LIFE ? sizeof(LifeArgs)
                           : g.cls == SC_FOREST ? sizeof(ForestArgs)
                                                : sizeof(LeniaMfmaArgs);
        g.off_items = off;
        off = align256(off + isz * g.reqs.size());
        g.off_map = off;
        off = align256(off + 4 * map.size());
        g.off_base = off;
        off = align256(off + 4 * g.base.size());
        maps.push_back(std::move(map));
    }
    if (off > sl.host_cap) {
        if (sl.host) HIPCHK(hipHostFree(sl.host));
        sl.host = nullptr;
        const size_t cap = std::max<size_t>(off * 2, 1 << 16);
        HIPCHK(hipHostMalloc((void**)&sl.host, cap, hipHostMallocDefault));
        sl.host_cap = cap;
    }
    if (off > sl.dev_cap) {
        if (sl.dev) s->arena.release(sl.dev);
        const size_t cap = std::max<size_t>(off * 2, 1 << 16);
        sl.dev = (char*)s->arena.alloc(cap);
        if (!sl.dev) return fail(CAH_ERR_OUT_OF_MEMORY, "no device memory for the tick tables");
        sl.dev_cap = cap;
    }
    size_t gi = 0;
    for (auto& kv : groups) {
        SrvGroup& g = kv.second;
        for (size_t k = 0; k < g.reqs.size(); ++k) {
            const SrvReq* r = g.reqs[k];
            const cah_state* st = r->st;
            const int32_t n = g.base[k + 1] - g.base[k];
            if (g.tb) {
                const LifeTbArgs a = tb_item(st, st->cur, r->steps);
                std::memcpy(sl.host + g.off_items + k * sizeof a, &a, sizeof a);
            } else if (g.cls == SC_LIFE) {
                LifeArgs a = st->life;
                const int vb = life_min_vb(st->life_R);
                a.cur = (const int8_t*)st->buf[st->cur];
                a.next = (int8_t*)st->buf[1 - st->cur];
                a.L = lat_of(st);
                a.cv1 = (int32_t)st->d.boundary_value[0];
                a.cv2 = (int32_t)st->d.boundary_value[1];
                a.nbands = (int32_t)((st->n1 + 62 * 4 * vb - 1) / (62 * 4 * vb));
                a.rows = (int32_t)std::min<int64_t>(g.rows, st->n2);
                a.nwaves = n;
                std::memcpy(sl.host + g.off_items + k * sizeof a, &a, sizeof a);
            } else if (g.cls == SC_FOREST) {
                ForestArgs a{};
                a.cur = (const int8_t*)st->buf[st->cur];
                a.next = (int8_t*)st->buf[1 - st->cur];
                a.L = lat_of(st);
                a.cv1 = (int32_t)st->d.boundary_value[0];
                a.cv2 = (int32_t)st->d.boundary_value[1];
                a.nbands = (int32_t)((st->n1 + 247) / 248);
                a.rows = (int32_t)std::min<int64_t>(g.rows, st->n2);
                a.nwaves = n;
                a.seed = st->d.seed;
                a.t = r->queued;
                a.p = r->p.ff_p;
                a.f = r->p.ff_f;
                a.p_f32 = r->p.ff_p_is_f32;
                a.f_f32 = r->p.ff_f_is_f32;
                std::memcpy(sl.host + g.off_items + k * sizeof a, &a, sizeof a);
            } else {
                LeniaMfmaArgs a{};
                a.cur = (const float*)st->buf[st->cur];
                a.next = (float*)st->buf[1 - st->cur];
                a.mask = st->mask_d;
                a.afrag = st->frag_d;
                a.L = lat_of(st);
                a.H = st->lenia_H;
                a.dy0 = st->lenia_dy0;
                a.npairs = st->lenia_np;
                a.cv1 = (float)st->d.boundary_value[0];
                a.cv2 = (float)st->d.boundary_value[1];
                a.prec32 = st->d.lenia_precision == 32;
                a.mu = st->d.mu;
                a.sigma = st->d.sigma;
                a.dt = st->d.dt;
                a.seed = st->d.seed;
                a.t = r->queued;
                a.scheme = st->d.scheme;
                a.sp = st->d.scheme_p;
                a.sp_f32 = st->d.scheme_p_is_f32;
                a.nbx = (int32_t)((st->n1 + LM_BX - 1) / LM_BX);
                std::memcpy(sl.host + g.off_items + k * sizeof a, &a, sizeof a);
            }
        }
        std::memcpy(sl.host + g.off_map, maps[gi].data(), 4 * maps[gi].size());
        std::memcpy(sl.host + g.off_base, g.base.data(), 4 * g.base.size());
        ++gi;
    }
    if (lchk::enabled()) {
        // every group's tables and every item's geometry, at both buffer parities
        // (the launch of step k swaps an item's buffers when k is odd)
        size_t gj = 0;
        for (auto& kv : groups) {
            SrvGroup& g = kv.second;
            const int64_t n = (int64_t)g.reqs.size();
            const int32_t* map = maps[gj++].data();
            const char* name = g.tb ? "k_life_tb_batch (server)" : g.cls == SC_LIFE ? "k_life_batch (server)"
                               : g.cls == SC_FOREST ? "k_forest_batch (server)" : "k_lenia_mfma_batch (server)";
            int rc = checked(false, name, [&](lchk::Chk& c) {
                lchk::item_table(c, map, g.base.data(), g.base.back(), n,
                                 [&](int64_t k) { return (int64_t)(g.base[k + 1] - g.base[k]); });
                for (int64_t k = 0; k < n && c.ok(); ++k)
                    for (int par = 0; par < 2; ++par) {
                        const char* item = sl.host + g.off_items;
                        auto swap = [&](auto a) {
                            if (par) {
                                const auto q = a.cur;
                                a.cur = a.next;
                                a.next = const_cast<decltype(a.next)>(q);
                            }
                            return a;
                        };
                        if (g.tb) {
                            LifeTbArgs a;
                            std::memcpy(&a, item + k * sizeof a, sizeof a);
                            a = swap(a);
                            const TbShape sh = TB_SHAPES[g.shape];
                            lchk::life_tb(c, a.L, a.cur, a.next, sh.txw, sh.ty, TB_K, std::min(a.steps, TB_K),
                                          (a.L.n1 + 32 * sh.txw - 1) / (32 * sh.txw), (a.L.n2 + sh.ty - 1) / sh.ty);
                        } else if (g.cls == SC_LIFE) {
                            LifeArgs a;
                            std::memcpy(&a, item + k * sizeof a, sizeof a);
                            a = swap(a);
                            const int R = g.key / 4;
                            if (life_min_vb(R) == 2)
                                lchk::band_walk<2>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, R, 2 * R + 1);
                            else
                                lchk::band_walk<1>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, R, 2 * R + 1);
                        } else if (g.cls == SC_FOREST) {
                            ForestArgs a;
                            std::memcpy(&a, item + k * sizeof a, sizeof a);
                            a = swap(a);
                            lchk::band_walk<1>(c, a.L, a.cur, a.next, a.nbands, a.rows, a.nwaves, a.nwaves, 1, 3);
                        } else {
                            LeniaMfmaArgs a;
                            std::memcpy(&a, item + k * sizeof a, sizeof a);
                            a = swap(a);
                            lchk::lenia_mfma(c, a, g.key, a.nbx, (a.L.n2 + LM_BY - 1) / LM_BY,
                                             g.reqs[k]->st->frag_bytes);
                        }
                    }
            });
            if (rc) return rc;
        }
    }
    if (off) HIPCHK(hipMemcpyAsync(sl.dev, sl.host, off, hipMemcpyHostToDevice, s->up));
    HIPCHK(hipEventRecord(sl.up, s->up));
    for (hipStream_t q : s->st) HIPCHK(hipStreamWaitEvent(q, sl.up, 0));

    // the batched families: step k advances the items with more than k steps
    // (k_life_tb: the launch at step k = j * TB_K advances them by up to TB_K)
    int32_t kmax = 0;
    for (auto& kv : groups) kmax = std::max(kmax, kv.second.reqs.front()->steps);
    for (int32_t k = 0; k < kmax; ++k) {
        for (auto& kv : groups) {
            SrvGroup& g = kv.second;
            if (g.tb && k % TB_K) continue;
            size_t na = 0;
            while (na < g.reqs.size() && g.reqs[na]->steps > k) ++na;
            const int32_t nw = g.base[na];
            if (nw == 0) continue;
            const int32_t* map = (const int32_t*)(sl.dev + g.off_map);
            const int32_t* base = (const int32_t*)(sl.dev + g.off_base);
            const void* items = sl.dev + g.off_items;
            if (g.tb) {
                launch_life_tb_batch(g.shape, nw, (const LifeTbArgs*)items, map, base, (k / TB_K) & 1, k,
                                     s->st[SC_LIFE]);
            } else if (g.cls == SC_LIFE) {
                const int R = g.key / 4, niv = (g.key / 2) % 2 + 1;
                launch_life_batch(R, life_min_vb(R), g.key % 2, niv, waves_to_blocks(nw),
                                  (const LifeArgs*)items, map, base, nw, k & 1, s->st[SC_LIFE]);
            } else if (g.cls == SC_FOREST) {
                // the short byte tests once every lattice holds only 0, 1 and 2: from
                // the tick's second step on, every item has taken a step of the rule
                const bool tri = g.tri_edges && (g.tri_cells || k > 0);
                launch_forest_batch(1, g.key % 2, tri, g.key / 2, waves_to_blocks(nw), (const ForestArgs*)items,
                                    map, base, nw, k & 1, k, nullptr, s->st[SC_FOREST]);
            } else {
                launch_lenia_mfma_batch(g.key, nw, (const LeniaMfmaArgs*)items, map, base, k & 1, k,
                                        s->st[SC_LENIA]);
            }
        }
    }
    HIPCHK(hipGetLastError());
    // the rest: launches of their own
    for (SrvReq* r : run)
        if (r->cls == SC_LOOP && r->steps > 0)
            if (int rc = cah_advance(r->st, r->queued, r->steps, &r->p)) return rc;
    for (SrvReq* r : run) {
        if (!r->in_tick) sl.staged.push_back(r);
        r->in_tick = true;
        if (r->tb) {
            r->st->cur ^= ((r->steps + TB_K - 1) / TB_K) & 1;  // one buffer swap a launch
            r->st->bin_clean = true;                             // Life writes 0 and 1 only
        } else if (r->cls != SC_LOOP) {
            r->st->cur ^= r->steps & 1;
            if (r->cls == SC_FOREST && r->steps > 0) {
                r->st->bin_clean = false;
                r->st->tri_clean = true;  // the forest fire writes 0, 1 and 2 only
            }
        }
        r->queued += r->steps;
        if (r->queued == r->t1) {
            r->tail_queued = true;
            sl.finishing.push_back(r);
        }
    }
    for (int c = 0; c < SRV_STREAMS; ++c) HIPCHK(hipEventRecord(sl.done[c], s->st[c]));
    sl.busy = true;
    ++s->nbusy;
    return CAH_OK;
}

extern "C" int cah_server_tick(cah_server* s, int32_t* active, int32_t* pending) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    HIPCHK(hipSetDevice(s->ctx->device));
    if (int rc = retire(s)) return rc;
    // admit in arrival order while the budget allows, at most SRV_ADMIT_BYTES of
    // initial states a call (at least one request), so that the device starts on
    // the first requests while later ones are still being set up
    size_t admitted = 0;
    while (!s->pending.empty() && s->used + s->pending.front()->bytes <= s->budget &&
           (admitted == 0 || admitted < SRV_ADMIT_BYTES)) {
        SrvReq* r = s->pending.front();
        if (int rc = admit(s, r)) return rc;
        s->pending.pop_front();
        s->used += r->bytes;
        admitted += std::max<size_t>(r->bytes / 2, 1);
        s->active.push_back(r);
    }
    if (s->nbusy < SRV_SLOTS)
        if (int rc = build_tick(s)) return rc;
    if (active) *active = (int32_t)s->active.size();
    if (pending) *pending = (int32_t)s->pending.size();
    return CAH_OK;
}

extern "C" int cah_server_reserve(cah_server* s, size_t device_bytes, size_t pinned_bytes) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    HIPCHK(hipSetDevice(s->ctx->device));
    if (!s->arena.reserve(std::min(device_bytes, s->budget)))
        return fail(CAH_ERR_OUT_OF_MEMORY, "cannot reserve %zu bytes of device memory", device_bytes);
    if (!s->pinned.reserve(pinned_bytes))
        return fail(CAH_ERR_OUT_OF_MEMORY, "cannot reserve %zu bytes of pinned host memory", pinned_bytes);
    return CAH_OK;
}

extern "C" int cah_server_tune(cah_server* s, int32_t chunk, double tick_us) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "server is NULL");
    if (chunk < 0) return fail(CAH_ERR_ARGUMENT, "chunk must be >= 0");
    if (!(tick_us >= 0)) return fail(CAH_ERR_ARGUMENT, "tick_us must be >= 0");
    std::lock_guard<std::mutex> lk(s->mu);
    s->chunk = chunk;
    s->tick_ps = tick_us * 1e6;
    return CAH_OK;
}

extern "C" int cah_server_poll(cah_server* s, uint64_t* ids, int32_t cap, int32_t* n) {
    if (!s || !n || (cap > 0 && !ids)) return fail(CAH_ERR_ARGUMENT, "cah_server_poll: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t k = 0;
    while (k < cap && !s->finished.empty()) {
        ids[k++] = s->finished.front();
        s->finished.pop_front();
    }
    *n = k;
    return CAH_OK;
}

static int find_req(cah_server* s, uint64_t id, SrvReq** r) {
    auto it = s ? s->all.find(id) : decltype(s->all.end()){};
    if (!s || it == s->all.end()) return fail(CAH_ERR_ARGUMENT, "no request %llu", (unsigned long long)id);
    *r = it->second;
    return CAH_OK;
}

extern "C" int cah_server_status(cah_server* s, uint64_t id, int32_t* status) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    SrvReq* r;
    if (int rc = find_req(s, id, &r)) return rc;
    if (!status) return fail(CAH_ERR_ARGUMENT, "status is NULL");
    *status = r->status;
    return CAH_OK;
}

extern "C" int cah_server_fetch(cah_server* s, uint64_t id, void* host, size_t nbytes) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    SrvReq* r;
    if (int rc = find_req(s, id, &r)) return rc;
    if (r->status != CAH_REQ_DONE || r->fetching)
        return fail(CAH_ERR_BUSY, "request %llu is not finished or is being fetched", (unsigned long long)id);
    if (!host) return fail(CAH_ERR_ARGUMENT, "cah_server_fetch: NULL argument");
    cah_state* st = r->st;
    if (nbytes != state_bytes(st))
        return fail(CAH_ERR_STATE_SHAPE, "state has %zu bytes, got %zu", state_bytes(st), nbytes);
    HIPCHK(hipSetDevice(s->ctx->device));
    const size_t row = (size_t)st->n1 * st->es, pitchb = (size_t)st->pitch * st->es;
    const void* dense = st->buf[st->cur];
    if (row != pitchb) {
        dense = st->buf[1 - st->cur];
        if (int rc = copy_rows(st->buf[st->cur], (int64_t)pitchb, st->buf[1 - st->cur], (int64_t)row,
                               (int64_t)row, st->n2, s->copy))
            return rc;
    }
    char* h = (char*)s->pinned.alloc(nbytes);
    if (!h) return fail(CAH_ERR_OUT_OF_MEMORY, "cannot stage %zu bytes in pinned memory", nbytes);
    hipEvent_t ev = nullptr;
    hipError_t e = hipMemcpyAsync(h, dense, nbytes, hipMemcpyDeviceToHost, s->copy);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev, s->copy);
    // wait and copy without the lock: the request stays DONE and is not freed
    // meanwhile, and a second fetch of it is refused
    r->fetching = true;
    lk.unlock();
    if (e == hipSuccess) e = hipEventSynchronize(ev);
    if (e == hipSuccess) hostcopy::copy(host, h, nbytes);
    if (ev) hipEventDestroy(ev);
    lk.lock();
    r->fetching = false;
    s->pinned.release(h);
    HIPCHK(e);
    // every launch that used its blocks has completed: back to the arena
    free_state(st);
    r->st = nullptr;
    s->used -= r->bytes;
    r->status = CAH_REQ_FETCHED;
    return CAH_OK;
}

extern "C" int cah_server_times(cah_server* s, uint64_t id, double* t_submit, double* t_admit,
                                double* t_done) {
    if (!s) return fail(CAH_ERR_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    SrvReq* r;
    if (int rc = find_req(s, id, &r)) return rc;
    if (t_submit) *t_submit = r->t_submit;
    if (t_admit) *t_admit = r->t_admit;
    if (t_done) *t_done = r->t_done;
    return CAH_OK;
}
