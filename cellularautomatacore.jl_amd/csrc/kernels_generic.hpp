// kernels_generic.hpp -- the general Int8 sweep and the synthetic-input generators.
//
//   k_generic_i8  any neighbourhood (halo <= 32), any Int8 rule, every gated
//                 scheme, masks.  An LDS tile with the boundary conditions applied
//                 while it is loaded (no halo buffer, no separate halo launch);
//                 four cells a thread, counted a dword at a time.
#pragma once
#include <type_traits>

#include "device.hpp"

// ---------------------------------------------------------------- generic

struct GenArgs {
    const int8_t* cur;
    int8_t* next;
    const uint8_t* mask;
    const int32_t* ofs;  // n_offsets pairs (o1, o2)
    Lat L;
    int32_t h, nofs;
    int32_t cv1, cv2;
    int32_t rule, C;
    uint64_t birth, survive;
    uint64_t seed;
    int64_t t;
    int32_t scheme, gate, sp_f32;
    double sp;
    double ffp, fff;
    int32_t ffp_f32, fff_f32;
    // the field epidemic: its field (dense, FP64 or FP32), covariate slices
    const void* field;
    int32_t field32, epi_form, epi_recover, cov_slices;
    const double* cov;
};

// Outcome inversion over (a => 1 - q, b => q) with the draw u (_pick,
// src/likelihood.jl:74-83).
template <typename T>
CAH_DEV int pick2(int a, int b, T q, T u) {
    const T pa = T(1) - q;
    if (pa > T(0) && u < pa) return a;
    return q > T(0) ? b : a;
}

// The field epidemic on S=0, I=1, R=2 (include/cahip.h, CAH_RULE_FIELD_EPIDEMIC).
CAH_DEV int epidemic_rule(const GenArgs& a, int c, int64_t lin, uint64_t cellid) {
    if (c == 1) {
        if (!a.epi_recover) return 1;
        if (a.epi_form == CAH_EPI_STATED) {
            if (a.fff_f32) return pick2<float>(1, 2, (float)a.fff, uniform32(a.seed, cellid, a.t, 1));
            return pick2<double>(1, 2, a.fff, uniform64(a.seed, cellid, a.t, 1));
        }
        return bernoulli(a.seed, cellid, a.t, 2, a.fff, a.fff_f32) ? 2 : 1;
    }
    if (c != 0) return c;
    if (a.ffp_f32 && a.field32 && !a.cov) {
        // beta, field and q in Float32; the 24-bit draw
        const float lam = (float)a.ffp * static_cast<const float*>(a.field)[lin];
        const float q = 1.0f - expf(-lam);
        const float u = uniform32(a.seed, cellid, a.t, 1);
        return a.epi_form == CAH_EPI_STATED ? pick2<float>(0, 1, q, u) : (u < q ? 1 : 0);
    }
    const double beta = a.ffp_f32 ? (double)(float)a.ffp : a.ffp;
    const double f = a.field32 ? (double)static_cast<const float*>(a.field)[lin]
                               : static_cast<const double*>(a.field)[lin];
    double lam;
    if (a.cov) {
        const int64_t n = (int64_t)a.L.n1 * a.L.n2;
        lam = (beta * a.cov[(a.t % a.cov_slices) * n + lin]) * f;
    } else {
        lam = beta * f;
    }
    const double q = 1.0 - exp(-lam);
    const double u = uniform64(a.seed, cellid, a.t, 1);
    return a.epi_form == CAH_EPI_STATED ? pick2<double>(0, 1, q, u) : (u < q ? 1 : 0);
}

constexpr int GEN_TX = 64, GEN_TY = 16;

// The rule of one cell of the general sweep given its neighbour counts n1 (cells
// equal to 1) and n2 (equal to 2): the gate and mask first (_update_cell!,
// src/backends.jl:444-455); (x, y) are its 0-based coordinates.
CAH_DEV int generic_rule(const GenArgs& a, int c, int n1, int n2, int x, int y) {
    const int64_t lin = (int64_t)y * a.L.n1 + x;  // unpadded, 0-based
    const uint64_t cellid = (uint64_t)lin + 1;
    const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                      gate_open(a.scheme, a.gate, a.sp, a.sp_f32, a.seed, cellid, a.t, x, y, a.L.two_d);
    if (!open) return c;
    if (a.rule == CAH_RULE_LIFE) {
        // local_rule(::Life) (src/rules.jl:385-389)
        const uint64_t m = c == 1 ? a.survive : a.birth;
        return n1 < 64 ? (int)((m >> n1) & 1u) : 0;
    }
    if (a.rule == CAH_RULE_GENERATIONS) {
        // local_rule(::Generations) (src/rules.jl:446-456)
        if (c == 0) return n1 < 64 ? (int)((a.birth >> n1) & 1u) : 0;
        if (c == 1) return (n1 < 64 && ((a.survive >> n1) & 1u)) ? 1 : (a.C > 2 ? 2 : 0);
        return c + 1 < a.C ? c + 1 : 0;
    }
    if (a.rule == CAH_RULE_FIELD_EPIDEMIC) return epidemic_rule(a, c, lin, cellid);
    // the forest-fire LocalRule (README.md:56-60), Int8 encoding
    if (c == 2) return 0;
    if (c == 0) return bernoulli(a.seed, cellid, a.t, 1, a.ffp, a.ffp_f32) ? 1 : 0;
    if (n2 > 0) return 2;
    return bernoulli(a.seed, cellid, a.t, 2, a.fff, a.fff_f32) ? 2 : 1;
}

// One cell of the general sweep: cp points at the cell in an LDS tile of width W
// that holds its neighbourhood (the sparse sweep's form, a cell at a time).
CAH_DEV int generic_cell(const GenArgs& a, const int8_t* cp, int W, int x, int y) {
    int n1 = 0, n2 = 0;
    for (int k = 0; k < a.nofs; ++k) {
        const int v = cp[a.ofs[2 * k] + a.ofs[2 * k + 1] * W];
        n1 += (v == 1);
        n2 += (v == 2);
    }
    return generic_rule(a, *cp, n1, n2, x, y);
}

// The general sweep's LDS tile for halo h: columns x0 - hp .. (hp = h rounded up
// to 4, so that each thread's four cells start a dword), W bytes a row (the 64
// cells, hp on each side and 8 more, so that the dword after the last one a
// count reads is in the row), 16 + 2h rows.
// The rows x W bytes (W a multiple of 4) of the state at (xs, ys) into tile32,
// sixteen loads of a thread in flight at a time: a dword at a time where the
// whole tile lies inside the lattice, else cell by cell through the boundary map.
constexpr int TILE_FILL_Q = 16;
CAH_DEV void fill_tile_i8(const GenArgs& a, uint32_t* tile32, int xs, int ys, int W, int rows) {
    constexpr int Q = TILE_FILL_Q;
    const int nt = blockDim.x;
    if (xs >= 0 && xs + W <= a.L.n1 && ys >= 0 && ys + rows <= a.L.n2) {
        const int8_t* src = a.cur + (int64_t)ys * a.L.pitch + xs;
        const int WD = W / 4, n = rows * WD;
        for (int base = 0; base < n; base += Q * nt) {
            uint32_t v[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int i = base + (int)threadIdx.x + nt * k, ly = i / WD, d = i - ly * WD;
                v[k] = i < n ? *reinterpret_cast<const uint32_t*>(src + (int64_t)ly * a.L.pitch + 4 * d) : 0u;
            }
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int i = base + (int)threadIdx.x + nt * k;
                if (i < n) tile32[i] = v[k];
            }
        }
    } else {
        int8_t* tile = reinterpret_cast<int8_t*>(tile32);
        const int n = W * rows;
        for (int base = 0; base < n; base += Q * nt) {
            int8_t v[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int i = base + (int)threadIdx.x + nt * k, lx = i % W, ly = i / W;
                v[k] = i < n ? read_mapped<int8_t>(a.cur, a.L, (int8_t)a.cv1, (int8_t)a.cv2, xs + lx, ys + ly)
                             : (int8_t)0;
            }
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int i = base + (int)threadIdx.x + nt * k;
                if (i < n) tile[i] = v[k];
            }
        }
    }
}

CAH_HD int gen_hp(int h) { return (h + 3) & ~3; }
CAH_HD int gen_w(int h) { return GEN_TX + 2 * gen_hp(h) + 8; }
CAH_HD int gen_rows(int h) { return GEN_TY + 2 * h; }

// The general sweep, four cells a thread.  Block (bx, by) computes the 64 x 16
// cells at (64 bx, 16 by): thread tid the dword of cells 4 (tid % 16) .. + 3 of
// row tid / 16.  The tile is filled a dword at a time where it lies inside the
// lattice, else cell by cell through the boundary map.  A neighbour offset
// (o1, o2) is the same for all threads, so each thread reads the four
// neighbours of its four cells as one dword (two aligned dwords and a byte
// shift by o1 mod 4), tests all four for == 1 and == 2 at once (bytes_eq) and
// adds the results into byte counters (flushed into 32-bit counts every 255
// offsets); the rule then runs cell by cell.
__global__ __launch_bounds__(256) void k_generic_i8(GenArgs a) {
    extern __shared__ uint32_t tile32[];
    const int h = a.h, hp = gen_hp(h), W = gen_w(h), rows = gen_rows(h);
    const int x0 = blockIdx.x * GEN_TX, y0 = blockIdx.y * GEN_TY;
    const int xs = x0 - hp, ys = y0 - h;
    fill_tile_i8(a, tile32, xs, ys, W, rows);
    __syncthreads();
    const int tx = 4 * (threadIdx.x % 16), ty = threadIdx.x / 16;
    const int x = x0 + tx, y = y0 + ty;
    if (y >= a.L.n2 || x >= a.L.n1) return;
    const int at = (ty + h) * W + hp + tx;  // byte of the first cell, a multiple of 4
    uint32_t n1[4] = {0, 0, 0, 0}, n2[4] = {0, 0, 0, 0};
    for (int k0 = 0; k0 < a.nofs; k0 += 255) {
        const int k1 = min(a.nofs, k0 + 255);
        uint32_t b1 = 0, b2 = 0;  // byte counters, at most 255 each
        for (int k = k0; k < k1; ++k) {
            const int o1 = a.ofs[2 * k], o2 = a.ofs[2 * k + 1];
            const int b = at + o2 * W + o1, sh = b & 3, dw = b >> 2;
            const uint32_t lo = tile32[dw];
            const uint32_t v = sh ? __builtin_amdgcn_alignbyte(tile32[dw + 1], lo, sh) : lo;
            b1 += bytes_eq(v, 0x01010101u);
            b2 += bytes_eq(v, 0x02020202u);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            n1[e] += (b1 >> (8 * e)) & 0xFFu;
            n2[e] += (b2 >> (8 * e)) & 0xFFu;
        }
    }
    const uint32_t cw = tile32[at >> 2];
    uint32_t out = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = (int)(int8_t)(cw >> (8 * e));
        const int o = x + e < a.L.n1 ? generic_rule(a, c, (int)n1[e], (int)n2[e], x + e, y) : 0;
        out |= (uint32_t)(uint8_t)o << (8 * e);
    }
    int8_t* dst = a.next + (int64_t)y * a.L.pitch + x;
    if (x + 4 <= a.L.n1) {
        *reinterpret_cast<uint32_t*>(dst) = out;
    } else {
        for (int e = 0; x + e < a.L.n1; ++e) dst[e] = (int8_t)(out >> (8 * e));
    }
}

// ---------------------------------------------------------------- the sparse sweep
//
// The SparseBackend's active-tile step (src/sparse.jl:126-211) on a dense lattice,
// with work proportional to the active area and one launch per step.  Tiles are
// ts x tsy cells (tsy = ts in 2-D, 1 in 1-D).  A step sweeps the candidate tiles of
// a list: the tiles within reach of a tile holding a non-background cell, the
// tiles that held one at the previous step (their cells in the buffer being
// written still hold that step's values) and the tiles within reach of a
// Constant edge that is not the background.  While it sweeps, it builds the next
// step's list: a tile that ends with a non-background cell claims its reach
// window, and a tile that began with one, or is an edge tile, claims itself.  A
// claim is an atomicMax of the step number on the tile's stamp, so each tile is
// listed once per step with nothing to clear.  Every other tile is background in
// both buffers, so the result is the dense sweep's.  k_sparse_mark makes the
// first list after the state is set (from the list of non-background tiles).

struct SparseArgs {
    int32_t* stamp;            // per tile: the last step that listed it
    const uint8_t* edge;       // per tile: always swept
    const int32_t* list;       // this step's candidates
    const int32_t* count;      //   ... their number
    int32_t* list_next;        // the next step's candidates
    int32_t* count_next;       //   ... their number (cleared before this step)
    int32_t* count_clear;      // the counter the step after next claims into: cleared here
    int32_t* nactive;          // tiles of the written state holding a non-background cell
    int32_t* nactive_clear;    // the other such counter: cleared here
    int32_t step;              // this step's number (the next list is stamped step + 1)
    int32_t ntx, nty, reach1, reach2;  // tiles a neighbourhood reaches, per axis
    int32_t per1, per2;                // periodic tile axes
};

CAH_DEV void sparse_claim(int32_t* stamp, int32_t step, int32_t* list, int32_t* count, int32_t k) {
    if (atomicMax(stamp + k, step) < step) list[atomicAdd(count, 1)] = k;
}

// Tile d (0 <= d < window) of the reach window around tile k, or -1 off the lattice.
CAH_DEV int32_t sparse_window(const SparseArgs& s, int32_t k, int32_t d) {
    const int32_t wx = 2 * s.reach1 + 1;
    int32_t x = k % s.ntx + d % wx - s.reach1, y = k / s.ntx + d / wx - s.reach2;
    if (s.per1) x = (x % s.ntx + s.ntx) % s.ntx;
    if (s.per2) y = (y % s.nty + s.nty) % s.nty;
    if (x < 0 || x >= s.ntx || y < 0 || y >= s.nty) return -1;
    return y * s.ntx + x;
}

// The first list (stamp `step`) from the non-background tiles `act` and the edge
// tiles `edges`.
__global__ __launch_bounds__(256) void k_sparse_mark(SparseArgs s, const int32_t* act, const int32_t* nact,
                                                     const int32_t* edges, int32_t nedge) {
    const int32_t win = (2 * s.reach1 + 1) * (2 * s.reach2 + 1);
    const int64_t a = (int64_t)*nact * win, total = a + nedge;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = i < a ? sparse_window(s, act[i / win], (int32_t)(i % win)) : edges[i - a];
        if (k >= 0) sparse_claim(s.stamp, s.step, s.list_next, s.count_next, k);
    }
}

// Blocks of 256 threads walk the candidate list; a tile's cells and their halo
// go through an LDS tile of (ts + 2h) x (tsy + 2h) bytes.
__global__ __launch_bounds__(256) void k_sparse_sweep(GenArgs a, SparseArgs s, int32_t ts, int32_t tsy,
                                                      int8_t bg) {
    extern __shared__ int8_t stile[];
    __shared__ int in_any, out_any;
    const int W = ts + 2 * a.h, H = tsy + 2 * a.h;
    const int32_t n = *s.count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // nobody uses these two during this step
        *s.count_clear = 0;
        *s.nactive_clear = 0;
    }
    const int32_t win = (2 * s.reach1 + 1) * (2 * s.reach2 + 1);
    for (int32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t k = s.list[i];
        const int x0 = (k % s.ntx) * ts, y0 = (k / s.ntx) * tsy;
        if (threadIdx.x == 0) in_any = out_any = 0;
        for (int j = threadIdx.x; j < W * H; j += blockDim.x) {
            const int lx = j % W, ly = j / W;
            stile[j] = read_mapped<int8_t>(a.cur, a.L, (int8_t)a.cv1, (int8_t)a.cv2, x0 - a.h + lx, y0 - a.h + ly);
        }
        __syncthreads();
        int inm = 0, outm = 0;
        for (int j = threadIdx.x; j < ts * tsy; j += blockDim.x) {
            const int lx = j % ts, ly = j / ts;
            const int x = x0 + lx, y = y0 + ly;
            if (x >= a.L.n1 || y >= a.L.n2) continue;
            const int8_t* cp = stile + (lx + a.h) + (ly + a.h) * W;
            const int out = generic_cell(a, cp, W, x, y);
            a.next[(int64_t)y * a.L.pitch + x] = (int8_t)out;
            inm |= *cp != bg;
            outm |= out != bg;
        }
        if (inm) in_any = 1;
        if (outm) out_any = 1;
        __syncthreads();
        if (out_any) {  // the written tile holds a non-background cell: list its reach window
            if (threadIdx.x == 0) atomicAdd(s.nactive, 1);
            for (int d = threadIdx.x; d < win; d += blockDim.x) {
                const int32_t kk = sparse_window(s, k, d);
                if (kk >= 0) sparse_claim(s.stamp, s.step + 1, s.list_next, s.count_next, kk);
            }
        } else if (threadIdx.x == 0 && (in_any || s.edge[k])) {
            sparse_claim(s.stamp, s.step + 1, s.list_next, s.count_next, k);
        }
        __syncthreads();  // the tile and the flags are reused by the next candidate
    }
}

// flags[k] = 1 when tile k holds a cell that is not the background.
__global__ __launch_bounds__(256) void k_sparse_flags(const int8_t* u, Lat L, int32_t ts, int32_t tsy,
                                                      int32_t ntx, int8_t bg, uint8_t* flags) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= L.n1) return;
    for (int32_t y = blockIdx.y; y < L.n2; y += gridDim.y)
        if (u[(int64_t)y * L.pitch + x] != bg) flags[(int64_t)(y / tsy) * ntx + x / ts] = 1;
}

// The flagged tiles, appended to a list (after a state is set).
__global__ __launch_bounds__(256) void k_sparse_compact(const uint8_t* flags, int32_t nt, int32_t* list,
                                                        int32_t* n) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nt && flags[k]) list[atomicAdd(n, 1)] = k;
}

// ---------------------------------------------------------------- synthetic inputs

// The documented generator: ctr = (cellid lo32, cellid hi32, 0xFFFFFFFF, 0).
CAH_DEV double gen_uniform(uint64_t seed, uint64_t cellid) {
    uint32_t c0 = (uint32_t)cellid, c1 = (uint32_t)(cellid >> 32), c2 = 0xFFFFFFFFu, c3 = 0u;
    philox10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t bits = ((uint64_t)c0 << 32) | (uint64_t)c1;
    return (double)(bits >> 11) * 0x1.0p-53;
}

__global__ void k_fill_bernoulli(int8_t* u, Lat L, uint64_t seed, double p) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= L.n1) return;
    for (int32_t y = blockIdx.y; y < L.n2; y += gridDim.y) {
        const uint64_t id = (uint64_t)y * (uint64_t)L.n1 + (uint64_t)x + 1;
        u[(int64_t)y * L.pitch + x] = gen_uniform(seed, id) < p ? 1 : 0;
    }
}

__global__ void k_fill_uniform_box(float* u, Lat L, uint64_t seed, int32_t lo1, int32_t len1,
                                   int32_t lo2, int32_t len2) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= L.n1) return;
    for (int32_t y = blockIdx.y; y < L.n2; y += gridDim.y) {
        const uint64_t id = (uint64_t)y * (uint64_t)L.n1 + (uint64_t)x + 1;
        const bool in = x >= lo1 && x < lo1 + len1 && y >= lo2 && y < lo2 + len2;
        u[(int64_t)y * L.pitch + x] = in ? (float)gen_uniform(seed, id) : 0.0f;
    }
}

// ---------------------------------------------------------------- row layout copies
//
// Rows of rb bytes between a pitched and a dense layout (pitches sp and dp in
// bytes), W bytes per thread; W divides rb, sp, dp and both base addresses.
// Host <-> device transfers of a state move the dense form with one linear
// copy: the DMA engines split a pitched copy whose rows are not 4-byte multiples
// into one transfer per row.
template <int W>
__global__ __launch_bounds__(256) void k_copy_rows(const uint8_t* __restrict__ src, int64_t sp,
                                                   uint8_t* __restrict__ dst, int64_t dp, int64_t rb,
                                                   int64_t rows) {
    using V = std::conditional_t<W == 16, uint4, std::conditional_t<W == 4, uint32_t, uint8_t>>;
    const int64_t per = rb / W, n = per * rows;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = i / per, x = i - y * per;
        *reinterpret_cast<V*>(dst + y * dp + x * W) = *reinterpret_cast<const V*>(src + y * sp + x * W);
    }
}
