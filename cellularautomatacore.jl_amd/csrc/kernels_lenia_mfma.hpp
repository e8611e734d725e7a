// kernels_lenia_mfma.hpp -- the Lenia potential as an implicit GEMM on MFMA.
//
// U(x, y) = sum over (dx, dy) of W(dx, dy) S(x + dx, y + dy).  Output columns come
// in 16-wide strips (M), output rows in 16-row tiles (N).  K = 32 of
// v_mfma_f32_16x16x32_{f16,bf16} is a pair of row offsets (dy, dy + 1) times a
// 16-column chunk of input: k = 8 g + i with g = 0..3 takes row dy + (g >> 1) and
// chunk column 8 (g & 1) + i, so each lane's 8 K values are 16 contiguous bytes
// of one LDS row.  With chunk c of the tile starting 16 (c - HC) columns left of
// strip c - r's first column, the A operand of (pair p, chunk offset r) is
//   A[m][k] = W(16 (r - HC) + 8 (g & 1) + i - m, dy0 + 2 p + (g >> 1)),
// zero off the band; it does not depend on the strip or the tile, so the A
// fragments are made once per problem in operand order and read from global
// memory (they stay in L1 / L2).  With H <= 16 (HC = 1) a strip needs chunk
// offsets r = 0, 1, 2 (16 + 2H input columns), and a wave computes all four
// strips of a 64-column block for one 16-row tile: an input chunk loaded from LDS
// serves every strip whose offset range holds it, so a 64 x 16 output tile reads
// 6 chunks a pair where one strip at a time would read 12.  Row
// offsets whose weights are all zero are not visited (the rows dy0 .. dy0 +
// 2 npairs - 1 cover the nonzero ones; R = 13: 25 of 27 rows).  Split FP16 keeps
// the residual of both operands (S = hi + lo, W = whi + wlo) and adds whi*lo and
// wlo*hi.
//
// A block is one 64 x 64 output tile: its input tile (rows y0 - H .. y0 + 64 + H,
// columns x0 - 16 .. x0 + 79) goes to LDS in the MFMA input type,
// the 64 x 64 centre loaded first so that each thread keeps the FP32 cells its
// growth updates in registers; the potentials go through LDS so that the growth
// writes whole 256-byte rows.
#pragma once
#include <type_traits>

#include "device.hpp"
#include "lenia_growth.hpp"

typedef _Float16 lm_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 lm_b8 __attribute__((ext_vector_type(8)));
typedef float lm_f4 __attribute__((ext_vector_type(4)));

enum { LM_F16 = 0, LM_BF16 = 1, LM_F16X2 = 2 };

constexpr int LM_BX = 64;    // output block: 64 x 64 cells, 4 waves of 16 rows
constexpr int LM_BY = 64;
constexpr int LM_HMAX = 16;  // kernel radius the MFMA forms take
constexpr int LM_NT = 256;   // threads a block

// The tile geometry: 16-column chunks of halo a side (HC), of a row (NC) and
// columns a row, the most rows (65 + 2H), float4 chunks a row, and the float4
// loads a thread makes for the largest tile.
constexpr int LM_HC = 1;
constexpr int LM_NC = 4 + 2 * LM_HC;
constexpr int LM_NR = 2 * LM_HC + 1;  // chunk offsets a strip needs
constexpr int LM_COLS = 16 * LM_NC;
constexpr int LM_ROWS = LM_BY + 2 * LM_HMAX + 1;
constexpr int LM_NCH = LM_COLS / 4;
constexpr int LM_FILL_Q = (LM_ROWS * LM_NCH + LM_NT - 1) / LM_NT;

// Bytes of one plane of the A fragments: npairs x LM_NR fragments of 64 lanes x
// 8 two-byte values.
inline size_t lenia_mfma_fragment_bytes(int32_t npairs) { return (size_t)npairs * LM_NR * 64 * 8 * 2; }

struct LeniaMfmaArgs {
    const float* cur;
    float* next;
    const uint8_t* mask;
    const void* afrag;  // [npairs][LM_NR][64][8] (and the low parts after it for split FP16)
    Lat L;
    int32_t H;
    int32_t dy0, npairs;  // first row offset visited, row-offset pairs
    float cv1, cv2;
    int32_t prec32;
    double mu, sigma, dt;
    uint64_t seed;
    int64_t t;
    int32_t scheme, sp_f32;
    double sp;
    int32_t nbx;  // output blocks (tiles) along the first axis
};

template <int MODE>
struct LmT {
    using E = std::conditional_t<MODE == LM_BF16, __bf16, _Float16>;
    using V8 = std::conditional_t<MODE == LM_BF16, lm_b8, lm_h8>;
    static CAH_DEV lm_f4 mfma(V8 a, V8 b, lm_f4 c) {
        if constexpr (MODE == LM_BF16) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
        else return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// The LDS row stride in elements, a whole number of 16-byte chunks.  Measured
// at 2048^2, R = 13 (SQ_LDS_BANK_CONFLICT): 224-byte rows take the B reads of the
// MFMA loop with 0.11 M conflict cycles a launch, 208-, 240- and 272-byte rows
// with 1.39 M (the time is the same: the loop waits on the MFMAs).  Split FP16
// keeps 208-byte rows, with which four blocks of its two planes fit a CU's
// 160 KB of LDS (with 224 three do: 26.0 against 24.5 us).
template <int MODE>
constexpr int lm_stride() { return MODE == LM_F16X2 ? LM_COLS + 8 : LM_COLS + 16; }

template <int MODE>
using LmTile = typename LmT<MODE>::E[MODE == LM_F16X2 ? 2 : 1][LM_ROWS * lm_stride<MODE>()];

// ---------------------------------------------------------------- the input tile
//
// Tile row ly, column lx hold the state at (x0 - 16 HC + lx, y0 - H + ly) in the
// MFMA input type (and, for split FP16, the residual in the second plane), rows
// 0 .. 64 + 2H.  The tile is read as float4 chunks (rows are 16-float aligned
// and x0 - 16 is a multiple of 16), each loaded whole where it lies inside the
// lattice.  A chunk beyond an edge maps, as _map_axis does cell by cell:
//  - when the first side is a multiple of 4 and at least the tile's width
//    (lenia_cols_whole), onto an aligned chunk inside, wrapped for Periodic,
//    mirrored and read back to front for Reflecting, or onto the Constant value;
//  - otherwise cell by cell (four loads; a chunk may also straddle the edge).
// Rows map by one formula for every edge kind when the second side is at least
// the tile's height (lenia_rows_near: no row lies more than one period beyond
// an edge), else through map_axis0.

CAH_HD bool lenia_cols_whole(const Lat& L) { return L.n1 % 4 == 0 && L.n1 >= LM_COLS; }

CAH_HD bool lenia_rows_near(const Lat& L, int H) { return L.n2 >= LM_BY + 2 * H + 1; }

CAH_HD bool lenia_tile_wide(const Lat& L, int H, int x0) {
    const int xa = x0 - 16 * LM_HC;
    return lenia_rows_near(L, H) && (lenia_cols_whole(L) || (xa >= 0 && xa + LM_COLS <= L.n1));
}

// Load q of a thread: q < 4 are the 64 x 64 centre (row H + i / 16, chunk
// 4 + i % 16 for i = tid + 256 q: the float4 the growth of that thread
// updates), then the halo: the 2H + 1 rows above and below the centre, whole,
// then the centre rows' left and right chunks.  Returns false past the tile.
CAH_HD bool lenia_fill_at(int H, int tid, int q, int& ly, int& j) {
    if (q < 4) {
        const int i = tid + LM_NT * q;
        ly = H + i / 16;
        j = 4 * LM_HC + i % 16;
        return true;
    }
    int h = tid + LM_NT * (q - 4);
    const int band = (2 * H + 1) * LM_NCH;
    if (h < band) {
        const int r = h / LM_NCH;
        ly = r < H ? r : r + LM_BY;
        j = h - r * LM_NCH;
        return true;
    }
    h -= band;
    constexpr int side = LM_NCH - 16;
    if (h >= LM_BY * side) return false;
    const int r = h / side, cj = h - r * side;
    ly = H + r;
    j = cj < 4 * LM_HC ? cj : cj + 16;
    return true;
}

// A boundary map for coordinates at most one period beyond an edge, as the same
// arithmetic for every edge kind: m = alo + slo g below 0, ahi + shi g at n or
// above (Periodic g + n, g - n; Reflecting -w - g, 2n - w - g for a span of w
// cells read back to front; Constant 0, the value being set apart).
struct LmMap {
    int32_t alo, slo, ahi, shi;
};

CAH_HD LmMap lm_axis(int32_t kind, int32_t n, int32_t w) {
    if (kind == CAH_BOUNDARY_PERIODIC) return {n, 1, -n, 1};
    if (kind == CAH_BOUNDARY_REFLECTING) return {-w, -1, 2 * n - w, -1};
    return {0, 0, 0, 0};
}

// The row of tile row ly (lattice row y0 - H + ly) as an element offset, and
// whether it lies beyond a Constant edge.
CAH_DEV int64_t lenia_fill_row(const LeniaMfmaArgs& a, int y, bool near, const LmMap& my_, bool& oy) {
    int my;
    if (near) {
        const bool ylo = y < 0, yhi = y >= a.L.n2;
        my = ylo ? my_.alo + my_.slo * y : (yhi ? my_.ahi + my_.shi * y : y);
        oy = a.L.b2 == CAH_BOUNDARY_CONSTANT && (ylo || yhi);
    } else {
        my = map_axis0(a.L.b2, y, a.L.n2, oy);
    }
    return oy ? 0 : (int64_t)my * a.L.pitch;
}

// The chunks this thread stages, loaded into registers with one float4 load
// each, all issued before any is used, from addresses mapped without branches.
// What a chunk still needs goes in fix, two bits a load: 1 read back to front
// (a whole chunk mirrored), 2 the first axis's Constant value, 3 the second's
// (axis 2 wins, as in _boundary_read); and in by_cell, one bit a load, a chunk
// off the lattice that is not whole (lenia_fill_cells maps it cell by cell).
// WIDE: rows near (lenia_rows_near) and every column chunk of the tile either
// inside the lattice or whole (lenia_cols_whole): the tiles of every lattice of
// 96 x (65 + 2H) cells or more whose first side is a multiple of 4, and the
// tiles clear of the first axis's edges on any lattice (the test is made once
// per block, lenia_tile_wide).
template <bool WIDE>
CAH_DEV void lenia_fill_load(const LeniaMfmaArgs& a, int x0, int y0, float4 (&v)[LM_FILL_Q], uint32_t& fix,
                             uint32_t& by_cell) {
    const int xa = x0 - 16 * LM_HC, n1 = a.L.n1;
    const bool whole = WIDE || lenia_cols_whole(a.L), near = WIDE || lenia_rows_near(a.L, a.H);
    const LmMap my_ = lm_axis(a.L.b2, a.L.n2, 1), mx_ = lm_axis(a.L.b1, n1, 4);
    const bool c1 = a.L.b1 == CAH_BOUNDARY_CONSTANT;
    const uint32_t fx = c1 ? 2 : (a.L.b1 == CAH_BOUNDARY_REFLECTING ? 1 : 0);  // a whole chunk off the lattice
    fix = by_cell = 0;
#pragma unroll
    for (int q = 0; q < LM_FILL_Q; ++q) {
        int ly, j;
        const bool in = lenia_fill_at(a.H, (int)threadIdx.x, q, ly, j);
        bool oy;
        const int64_t row = lenia_fill_row(a, y0 - a.H + ly, near, my_, oy);
        const int x = xa + 4 * j;
        const bool inr = x >= 0 && x + 4 <= n1;
        int xm = inr ? x : 0;
        uint32_t f = oy ? 3u : 0u;
        if (whole) {
            if (!inr && !c1) xm = x < 0 ? mx_.alo + mx_.slo * x : mx_.ahi + mx_.shi * x;
            if (!inr && !oy) f = fx;
        } else if (in && !inr && !oy) {
            by_cell |= 1u << q;
        }
        v[q] = *reinterpret_cast<const float4*>(a.cur + (in && !oy ? row + xm : 0));
        fix |= f << (2 * q);
    }
}

// The chunks of by_cell (lenia_fill_load) through the boundary map cell by
// cell: a lattice whose first side is not a multiple of 4 or is narrower than
// the tile.
CAH_DEV void lenia_fill_cells(const LeniaMfmaArgs& a, int x0, int y0, float4 (&v)[LM_FILL_Q], uint32_t by_cell) {
    if (!by_cell) return;
    const bool near = lenia_rows_near(a.L, a.H);
    const LmMap my_ = lm_axis(a.L.b2, a.L.n2, 1);
#pragma unroll
    for (int q = 0; q < LM_FILL_Q; ++q) {
        if (!((by_cell >> q) & 1)) continue;
        int ly, j;
        lenia_fill_at(a.H, (int)threadIdx.x, q, ly, j);
        bool oy;
        const float* row = a.cur + lenia_fill_row(a, y0 - a.H + ly, near, my_, oy);
        const int x = x0 - 16 * LM_HC + 4 * j;
        float e[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bool o1;
            const int m = map_axis0(a.L.b1, x + k, a.L.n1, o1);
            e[k] = o1 ? a.cv1 : row[m];
        }
        v[q] = make_float4(e[0], e[1], e[2], e[3]);
    }
}

// What the loads still need (fix, from lenia_fill_load), once they have landed.
CAH_DEV void lenia_fill_fix(const LeniaMfmaArgs& a, float4 (&v)[LM_FILL_Q], uint32_t fix) {
#pragma unroll
    for (int q = 0; q < LM_FILL_Q; ++q) {
        const uint32_t f = (fix >> (2 * q)) & 3;
        const float4 w = v[q];
        const float c = f == 2 ? a.cv1 : a.cv2;
        v[q] = f == 0 ? w : (f == 1 ? make_float4(w.w, w.z, w.y, w.x) : make_float4(c, c, c, c));
    }
}

// The loaded chunks into the tile, four values a store.
template <int MODE>
CAH_DEV void lenia_fill_store(const LeniaMfmaArgs& a, const float4 (&v)[LM_FILL_Q], LmTile<MODE>& tile) {
    using E = typename LmT<MODE>::E;
    typedef E e4 __attribute__((ext_vector_type(4)));
    constexpr int S = lm_stride<MODE>();
#pragma unroll
    for (int q = 0; q < LM_FILL_Q; ++q) {
        int ly, j;
        if (!lenia_fill_at(a.H, (int)threadIdx.x, q, ly, j)) continue;
        const float vv[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
        E h[4], l[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            h[e] = (E)vv[e];
            l[e] = (E)(vv[e] - (float)h[e]);
        }
        *reinterpret_cast<e4*>(&tile[0][ly * S + 4 * j]) = e4{h[0], h[1], h[2], h[3]};
        if constexpr (MODE == LM_F16X2)
            *reinterpret_cast<e4*>(&tile[1][ly * S + 4 * j]) = e4{l[0], l[1], l[2], l[3]};
    }
}

// ---------------------------------------------------------------- the growth

// The growth of the 64 x 64 block at (x0, y0) from its potentials U (row stride
// LM_UST floats, in LDS), in FP32 (lenia_fast32): mu, -1 / (2 sigma^2), dt.
// The cells come from the registers the fill loaded them into (the centre is
// its first four loads), so the state is read once a step.  A block wholly
// inside the lattice writes whole 256-byte rows.
constexpr int LM_UST = LM_BX + 4;
constexpr int LM_CELLS_Q = LM_BX * LM_BY / 4 / LM_NT;  // float4s of state a thread

CAH_HD bool lenia_full_block(const Lat& L, int x0, int y0) {
    return x0 + LM_BX <= L.n1 && y0 + LM_BY <= L.n2;
}

// The cell (x, y) of the block after the growth from c and potential u, or c
// where the gate or the mask keeps it.
CAH_DEV float lenia_grow(const LeniaMfmaArgs& a, int x, int y, float c, float u, float mu, float nk, float dt) {
    const int64_t lin = (int64_t)y * a.L.n1 + x;
    const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                      gate_open(a.scheme, 0, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1, a.t, x, y, a.L.two_d);
    return open ? lenia_fast32(c, u, mu, nk, dt) : c;
}

// A block wholly inside the lattice: float4 k of the thread is row q / 16,
// columns 4 (q % 16) .. + 3 for q = tid + 256 k, its cells c[k].
CAH_DEV void lenia_growth_full(const LeniaMfmaArgs& a, int x0, int y0, const float* U, const float4 (&c)[LM_CELLS_Q]) {
    const float mu = (float)a.mu, nk = (float)(-1.0 / (2.0 * a.sigma * a.sigma)), dt = (float)a.dt;
    if (a.mask == nullptr && a.scheme == CAH_SCHEME_SYNCHRONOUS) {
        // every cell updates: no per-cell gate or mask
#pragma unroll
        for (int k = 0; k < LM_CELLS_Q; ++k) {
            const int q = threadIdx.x + LM_NT * k;
            const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
            const float4 u4 = *reinterpret_cast<const float4*>(&U[ly * LM_UST + lx]);
            const float4 c4 = c[k];
            *reinterpret_cast<float4*>(a.next + (int64_t)(y0 + ly) * a.L.pitch + x0 + lx) =
                make_float4(lenia_fast32(c4.x, u4.x, mu, nk, dt), lenia_fast32(c4.y, u4.y, mu, nk, dt),
                            lenia_fast32(c4.z, u4.z, mu, nk, dt), lenia_fast32(c4.w, u4.w, mu, nk, dt));
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < LM_CELLS_Q; ++k) {
        const int q = threadIdx.x + LM_NT * k;
        const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
        const int y = y0 + ly, x = x0 + lx;
        const float4 u4 = *reinterpret_cast<const float4*>(&U[ly * LM_UST + lx]);
        const float4 c4 = c[k];
        *reinterpret_cast<float4*>(a.next + (int64_t)y * a.L.pitch + x) =
            make_float4(lenia_grow(a, x, y, c4.x, u4.x, mu, nk, dt), lenia_grow(a, x + 1, y, c4.y, u4.y, mu, nk, dt),
                        lenia_grow(a, x + 2, y, c4.z, u4.z, mu, nk, dt), lenia_grow(a, x + 3, y, c4.w, u4.w, mu, nk, dt));
    }
}

// c: the block's cells from the fill (float4 k of the thread: row q / 16,
// columns 4 (q % 16) .. + 3 for q = tid + 256 k), right wherever they lie on
// the lattice; a block past an edge writes only those.
CAH_DEV void lenia_growth_block(const LeniaMfmaArgs& a, int x0, int y0, const float* U,
                                const float4 (&c)[LM_CELLS_Q]) {
    if (lenia_full_block(a.L, x0, y0)) {
        lenia_growth_full(a, x0, y0, U, c);
        return;
    }
    const float mu = (float)a.mu, nk = (float)(-1.0 / (2.0 * a.sigma * a.sigma)), dt = (float)a.dt;
    const bool every = a.mask == nullptr && a.scheme == CAH_SCHEME_SYNCHRONOUS;
#pragma unroll
    for (int k = 0; k < LM_CELLS_Q; ++k) {
        const int q = threadIdx.x + LM_NT * k;
        const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
        const int y = y0 + ly, x = x0 + lx;
        if (y >= a.L.n2) continue;
        const float4 u4 = *reinterpret_cast<const float4*>(&U[ly * LM_UST + lx]);
        const float uu[4] = {u4.x, u4.y, u4.z, u4.w}, cc[4] = {c[k].x, c[k].y, c[k].z, c[k].w};
        const int64_t at = (int64_t)y * a.L.pitch + x;
        float r[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            r[e] = x + e >= a.L.n1 ? cc[e]  // off the lattice: not written (and no mask byte read)
                   : every ? lenia_fast32(cc[e], uu[e], mu, nk, dt) : lenia_grow(a, x + e, y, cc[e], uu[e], mu, nk, dt);
        if (x + 4 <= a.L.n1) {
            *reinterpret_cast<float4*>(a.next + at) = make_float4(r[0], r[1], r[2], r[3]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (x + e < a.L.n1) a.next[at + e] = r[e];
        }
    }
}

// ---------------------------------------------------------------- the MFMA loop

// The potentials of the wave's 64 x 16 tile (rows 16 w .. 16 w + 15 of the block)
// into acc: strip s, D[m][n] in lane n + 16 (m / 4), register m % 4.
template <int MODE>
CAH_DEV void lenia_mfma_loop(const LeniaMfmaArgs& a, const LmTile<MODE>& tile, lm_f4 (&acc)[4]) {
    using T = LmT<MODE>;
    using V8 = typename T::V8;
    using E = typename T::E;
    constexpr int NR = LM_NR;
    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int g = lane / 16, n = lane % 16;
    const V8* A = static_cast<const V8*>(a.afrag) + lane;
    const int nfrag = a.npairs * NR * 64;
    // this lane's row of the first pair, and its half of a chunk
    const int base = (16 * w + n + a.H + a.dy0 + (g >> 1)) * lm_stride<MODE>() + 8 * (g & 1);
    const E* b0 = &tile[0][base];
    const E* b1 = &tile[MODE == LM_F16X2 ? 1 : 0][base];
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[s] = lm_f4{0.f, 0.f, 0.f, 0.f};
    for (int p = 0; p < a.npairs; ++p) {
        V8 ahi[NR], alo[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            ahi[r] = A[(p * NR + r) * 64];
            if constexpr (MODE == LM_F16X2) alo[r] = A[nfrag + (p * NR + r) * 64];
        }
        const int po = 2 * p * lm_stride<MODE>();
#pragma unroll
        for (int c = 0; c < LM_NC; ++c) {
            const V8 b = *reinterpret_cast<const V8*>(b0 + po + 16 * c);
            V8 blo;
            if constexpr (MODE == LM_F16X2) blo = *reinterpret_cast<const V8*>(b1 + po + 16 * c);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int r = c - s;
                if (r < 0 || r >= NR) continue;
                acc[s] = T::mfma(ahi[r], b, acc[s]);
                if constexpr (MODE == LM_F16X2) {
                    acc[s] = T::mfma(ahi[r], blo, acc[s]);
                    acc[s] = T::mfma(alo[r], b, acc[s]);
                }
            }
        }
    }
}

// The potentials into LDS (the tile buffer they were computed from is free once
// every wave is past the loop), row-major with stride LM_UST.
CAH_DEV void lenia_store_potentials(const lm_f4 (&acc)[4], float* U) {
    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int g = lane / 16, n = lane % 16;
#pragma unroll
    for (int s = 0; s < 4; ++s)
        *reinterpret_cast<float4*>(&U[(16 * w + n) * LM_UST + 16 * s + 4 * g]) =
            make_float4(acc[s][0], acc[s][1], acc[s][2], acc[s][3]);
}

// ---------------------------------------------------------------- kernels

// Tile k of one lattice (at (k % nbx, k / nbx)): the fill, the MFMA loop, the
// potentials through LDS, the growth.
template <int MODE>
CAH_DEV void lenia_mfma_tile(const LeniaMfmaArgs& a, int k) {
    static_assert(LM_BY * LM_UST * 4 <= (int)sizeof(LmTile<MODE>), "the potentials must fit the tile buffer");
    __shared__ __attribute__((aligned(16))) LmTile<MODE> tile;
    const int x0 = (k % a.nbx) * LM_BX, y0 = (k / a.nbx) * LM_BY;
    float4 cells[LM_CELLS_Q];
    {
        float4 v[LM_FILL_Q];
        uint32_t fix, by_cell;
        if (lenia_tile_wide(a.L, a.H, x0)) {
            lenia_fill_load<true>(a, x0, y0, v, fix, by_cell);
            lenia_fill_fix(a, v, fix);
        } else {
            lenia_fill_load<false>(a, x0, y0, v, fix, by_cell);
            lenia_fill_fix(a, v, fix);
            lenia_fill_cells(a, x0, y0, v, by_cell);
        }
#pragma unroll
        for (int q = 0; q < LM_CELLS_Q; ++q) cells[q] = v[q];
        lenia_fill_store<MODE>(a, v, tile);
    }
    __syncthreads();
    lm_f4 acc[4];
    lenia_mfma_loop<MODE>(a, tile, acc);
    __syncthreads();  // every wave is past the loop: the tile buffer takes the potentials
    float* U = reinterpret_cast<float*>(&tile[0][0]);
    lenia_store_potentials(acc, U);
    __syncthreads();
    lenia_growth_block(a, x0, y0, U, cells);
}

// One lattice: block b computes tile block_id_xcd(b), so that each XCD's L2
// holds a run of tiles and the edges they share (at 2048^2, R = 13: 20.0 ->
// 17.1 us a step against tile b on block b).
template <int MODE>
__global__ __launch_bounds__(LM_NT) void k_lenia_mfma(LeniaMfmaArgs a) {
    lenia_mfma_tile<MODE>(a, block_id_xcd());
}

// Batched: block b (in the XCD-aware order) advances tile (b - block_base[i])
// of item i = block_item[b], with the item's buffers swapped when parity is set
// and its time base + dt.
template <int MODE>
__global__ __launch_bounds__(LM_NT) void k_lenia_mfma_batch(const LeniaMfmaArgs* __restrict__ items,
                                                           const int32_t* __restrict__ block_item,
                                                           const int32_t* __restrict__ block_base,
                                                           int32_t parity, int64_t dt) {
    const int b = block_id_xcd();
    const int i = block_item[b];
    LeniaMfmaArgs a = items[i];
    if (parity) {
        const float* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<float*>(c);
    }
    a.t += dt;
    lenia_mfma_tile<MODE>(a, b - block_base[i]);
}
