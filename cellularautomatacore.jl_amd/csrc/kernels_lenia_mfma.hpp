// kernels_lenia_mfma.hpp -- the Lenia potential as an implicit GEMM on MFMA.
//
// U(x, y) = sum over (dx, dy) of W(dx, dy) S(x + dx, y + dy).  For each row offset
// dy the sum over dx is a banded (Toeplitz) matrix times the input: with the
// output x of a 16-wide strip as M, the input x positions as K and 16 output rows
// as N,
//   D[m][n] += A_dy[m][k] * B_dy[k][n],  A_dy[m][k] = W(k - H - m, dy),
//   B_dy[k][n] = S(x0 - H + k, y0 + n + dy),
// accumulated over the 2H+1 row offsets into FP32 (v_mfma_f32_16x16x32_{f16,bf16}).
// A wave covers one 16-column strip and four row tiles (64 rows), so each A
// fragment serves four MFMAs.  A_dy does not depend on the tile, so it is
// precomputed once per problem in fragment order and read from global memory (it
// stays in L1/L2); the input tile is staged in LDS in the MFMA input type.
// Split-FP16 keeps the residual of both operands (S = hi + lo, W = whi + wlo) and
// adds whi*lo and wlo*hi.
//
// Measured on MI355X (R = 13, 2048^2, per-block timestamps): a block's fill
// (HBM, 5.3 us), MFMA loop (LDS reads of B, 12 us) and growth (4 us) ran one
// after the other, and with every block of the lattice resident at once the
// whole device went through the three phases in lock step (27 us a step).  The
// loop is bound by the LDS reads of B with the four waves a SIMD that a 2048^2
// lattice offers (half as many waves, two tiles a block through double-buffered
// LDS, took twice as long).  v_mfma_f32_32x32x16 forms, which halve the LDS bytes
// per multiply-add, measured slower: their A fragment loads, twice the bytes a
// wave, then bound the loop through L1.  Loading the next row offset's A
// fragments ahead of the MFMAs gained nothing.
#pragma once
#include <type_traits>

#include "device.hpp"
#include "lenia_growth.hpp"

typedef _Float16 lm_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 lm_b8 __attribute__((ext_vector_type(8)));
typedef float lm_f4 __attribute__((ext_vector_type(4)));

enum { LM_F16 = 0, LM_BF16 = 1, LM_F16X2 = 2 };

constexpr int LM_BX = 64;           // output block: 64 x 64 cells, 4 waves of 16 columns
constexpr int LM_BY = 64;
constexpr int LM_HMAX = 24;         // 16 + 2H input columns must fit two K = 32 chunks
constexpr int LM_COLS = LM_BX + 48; // input columns staged (past 64 + 2H they only meet zeros)
constexpr int LM_ROWS = LM_BY + 2 * LM_HMAX;
// LDS row stride of the tile in elements, by mode.  224-byte rows take the B
// reads of the MFMA loop (16 lanes on 16 rows, 16 bytes each) without bank
// conflicts, where 240-byte rows cost 2.4 M conflict cycles a launch and 13% of
// the loop's time on MI355X; measured over the whole kernel at R = 13, 2048^2,
// the split-FP16 form gains from them (48 -> 43 us) and the one-plane forms do not
// (28 -> 31 us), so those keep 240-byte rows.
template <int MODE>
constexpr int lm_stride() { return MODE == LM_F16X2 ? LM_COLS : 120; }
constexpr int LM_STRIDE_MAX = 120;
constexpr int LM_KCHUNKS = 2;       // K = 32 chunks a 16-column strip's 16 + 2H inputs need
constexpr int LM_NT = 256;          // threads a block
constexpr int LM_NCH = LM_COLS / 4 + 1;  // float4 chunks covering LM_COLS + 3 columns
// float4 loads a thread makes for one tile (rows x chunks over the block)
constexpr int LM_FILL_Q = (LM_ROWS * LM_NCH + LM_NT - 1) / LM_NT;

struct LeniaMfmaArgs {
    const float* cur;
    float* next;
    const uint8_t* mask;
    const void* afrag;  // [2H+1][2][64][8] (and the low parts after it for split-FP16)
    Lat L;
    int32_t H;
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

// ---------------------------------------------------------------- the input tile
//
// Tile column lx, row ly hold the state at (x0 - H + lx, y0 - H + ly) in the MFMA
// input type (and, for split FP16, the residual in the second plane), rows
// 0 .. 64 + 2H - 1, columns 0 .. LM_COLS - 1.  The tile is read as aligned float4
// chunks starting at xa = (x0 - H) rounded down to 4 (rows are 16-float aligned
// and x0 is a multiple of 64).  With a first side that is a multiple of 4 and a
// lattice larger than the tile, every chunk lies wholly inside the lattice or
// wholly beyond one edge, where it maps onto an aligned chunk inside (wrapped for
// Periodic, mirrored and read back to front for Reflecting) or onto the Constant
// value, as _map_axis does cell by cell; rows map the same way.  Columns at or
// past 64 + 2H only meet zero weights: they are staged as zeros without a load
// (they must still be finite).  Smaller lattices go through the boundary map
// cell by cell (lenia_fill_direct).

CAH_DEV bool lenia_chunked(const LeniaMfmaArgs& a) {
    return a.L.n1 % 4 == 0 && a.L.n1 >= 4 * LM_NCH && a.L.n2 >= LM_BY + 2 * a.H;
}

// The chunks of the tile at (x0, y0) this thread stages, loaded into registers
// (all loads issued before any is used).
CAH_DEV void lenia_fill_load(const LeniaMfmaArgs& a, int x0, int y0, float4 (&v)[LM_FILL_Q]) {
    const int H = a.H, rows = LM_BY + 2 * H;
    const int xs = x0 - H, sh = xs & 3, xa = xs - sh;
    const int nload = (LM_BX + 2 * H + sh + 3) / 4;
#pragma unroll
    for (int k = 0; k < LM_FILL_Q; ++k) {
        const int i = threadIdx.x + LM_NT * k;
        v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i >= rows * LM_NCH) continue;
        const int ly = i / LM_NCH, j = i - ly * LM_NCH;
        if (j >= nload) continue;
        bool o2;
        const int my = map_axis0(a.L.b2, y0 - H + ly, a.L.n2, o2);
        const int x = xa + 4 * j;
        const float* row = a.cur + (int64_t)my * a.L.pitch;
        if (o2) {
            v[k] = make_float4(a.cv2, a.cv2, a.cv2, a.cv2);
        } else if (x >= 0 && x + 4 <= a.L.n1) {
            v[k] = *reinterpret_cast<const float4*>(row + x);
        } else if (a.L.b1 == CAH_BOUNDARY_PERIODIC) {
            v[k] = *reinterpret_cast<const float4*>(row + (x < 0 ? x + a.L.n1 : x - a.L.n1));
        } else if (a.L.b1 == CAH_BOUNDARY_REFLECTING) {
            const float4 f = *reinterpret_cast<const float4*>(row + (x < 0 ? -x - 4 : 2 * a.L.n1 - x - 4));
            v[k] = make_float4(f.w, f.z, f.y, f.x);
        } else {
            v[k] = make_float4(a.cv1, a.cv1, a.cv1, a.cv1);
        }
    }
}

// The loaded chunks into a tile buffer; four values go to LDS at once when the
// chunk lands on a 4-column boundary.
template <int MODE, typename E>
CAH_DEV void lenia_fill_store(const LeniaMfmaArgs& a, int x0, const float4 (&v)[LM_FILL_Q],
                              E (*tile)[LM_ROWS * lm_stride<MODE>()]) {
    constexpr int NP = MODE == LM_F16X2 ? 2 : 1;
    const int H = a.H, rows = LM_BY + 2 * H;
    const int sh = (x0 - H) & 3;
#pragma unroll
    for (int k = 0; k < LM_FILL_Q; ++k) {
        const int i = threadIdx.x + LM_NT * k;
        if (i >= rows * LM_NCH) continue;
        const int ly = i / LM_NCH, j = i - ly * LM_NCH;
        const float vv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
        const int lx0 = 4 * j - sh;
        if (sh == 0 && lx0 + 4 <= LM_COLS) {
            E h[4], l[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                h[e] = (E)vv[e];
                l[e] = (E)(vv[e] - (float)h[e]);
            }
            typedef E e4 __attribute__((ext_vector_type(4)));
            *reinterpret_cast<e4*>(&tile[0][ly * lm_stride<MODE>() + lx0]) = e4{h[0], h[1], h[2], h[3]};
            if constexpr (NP == 2)
                *reinterpret_cast<e4*>(&tile[1][ly * lm_stride<MODE>() + lx0]) = e4{l[0], l[1], l[2], l[3]};
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int lx = lx0 + e;
                if (lx >= 0 && lx < LM_COLS) {
                    const E hi = (E)vv[e];
                    tile[0][ly * lm_stride<MODE>() + lx] = hi;
                    if constexpr (NP == 2) tile[1][ly * lm_stride<MODE>() + lx] = (E)(vv[e] - (float)hi);
                }
            }
        }
    }
}

// The tile cell by cell through the boundary map (lattices too small for chunks).
template <int MODE, typename E>
CAH_DEV void lenia_fill_direct(const LeniaMfmaArgs& a, int x0, int y0, E (*tile)[LM_ROWS * lm_stride<MODE>()]) {
    constexpr int NP = MODE == LM_F16X2 ? 2 : 1;
    const int H = a.H, rows = LM_BY + 2 * H;
    for (int i = threadIdx.x; i < rows * LM_COLS; i += LM_NT) {
        const int lx = i % LM_COLS, ly = i / LM_COLS;
        const float v = read_mapped<float>(a.cur, a.L, a.cv1, a.cv2, x0 - H + lx, y0 - H + ly);
        const E hi = (E)v;
        tile[0][ly * lm_stride<MODE>() + lx] = hi;
        if constexpr (NP == 2) tile[1][ly * lm_stride<MODE>() + lx] = (E)(v - (float)hi);
    }
}

// ---------------------------------------------------------------- the growth

// The growth of the 64 x 64 block at (x0, y0) from its potentials U (row stride
// LM_UST floats, in LDS), in FP32 (lenia_fast32): mu, -1 / (2 sigma^2), dt.  The
// block reads the state and writes the result in whole 256-byte rows; in a block
// wholly inside the lattice each thread issues all its state loads before it
// uses the first.  (Issuing them before the MFMA loop instead, to hide their
// latency there, made the kernel slower: 28.6 against 24.3 us at R = 13, 2048^2,
// FP16, as the loop's first fragment loads then waited behind them.)
constexpr int LM_UST = LM_BX + 4;
constexpr int LM_CELLS_Q = LM_BX * LM_BY / 4 / LM_NT;  // float4s of state a thread

CAH_DEV bool lenia_full_block(const LeniaMfmaArgs& a, int x0, int y0) {
    return x0 + LM_BX <= a.L.n1 && y0 + LM_BY <= a.L.n2;
}

CAH_DEV void lenia_state_load(const LeniaMfmaArgs& a, int x0, int y0, float4 (&c)[LM_CELLS_Q]) {
#pragma unroll
    for (int k = 0; k < LM_CELLS_Q; ++k) {
        const int q = threadIdx.x + LM_NT * k;
        const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
        c[k] = *reinterpret_cast<const float4*>(a.cur + (int64_t)(y0 + ly) * a.L.pitch + x0 + lx);
    }
}

CAH_DEV void lenia_growth_block(const LeniaMfmaArgs& a, int x0, int y0, const float* U) {
    const float mu = (float)a.mu, nk = (float)(-1.0 / (2.0 * a.sigma * a.sigma)), dt = (float)a.dt;
    auto cell = [&](int x, int y, float cv, float u) {
        const int64_t lin = (int64_t)y * a.L.n1 + x;
        const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                          gate_open(a.scheme, 0, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1, a.t, x, y,
                                    a.L.two_d);
        return open ? lenia_fast32(cv, u, mu, nk, dt) : cv;
    };
    if (lenia_full_block(a, x0, y0)) {
        float4 c[LM_CELLS_Q];
        lenia_state_load(a, x0, y0, c);
#pragma unroll
        for (int k = 0; k < LM_CELLS_Q; ++k) {
            const int q = threadIdx.x + LM_NT * k;
            const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
            const int y = y0 + ly, x = x0 + lx;
            const float4 u4 = *reinterpret_cast<const float4*>(&U[ly * LM_UST + lx]);
            const float4 c4 = c[k];
            *reinterpret_cast<float4*>(a.next + (int64_t)y * a.L.pitch + x) =
                make_float4(cell(x, y, c4.x, u4.x), cell(x + 1, y, c4.y, u4.y), cell(x + 2, y, c4.z, u4.z),
                            cell(x + 3, y, c4.w, u4.w));
        }
        return;
    }
    for (int q = threadIdx.x; q < LM_BY * LM_BX / 4; q += LM_NT) {
        const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
        const int y = y0 + ly, x = x0 + lx;
        if (y >= a.L.n2) continue;
        const float4 u4 = *reinterpret_cast<const float4*>(&U[ly * LM_UST + lx]);
        const float uu[4] = {u4.x, u4.y, u4.z, u4.w};
        const int64_t at = (int64_t)y * a.L.pitch + x;
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (x + e < a.L.n1) a.next[at + e] = cell(x + e, y, a.cur[at + e], uu[e]);
    }
}

// ---------------------------------------------------------------- the MFMA loop

// The potentials of the tile in `tile` into acc (tile t of the wave's strip:
// D[m][n] in lane n + 16 (m / 4), register m % 4).
template <int MODE, typename E>
CAH_DEV void lenia_mfma_loop(const LeniaMfmaArgs& a, const E (*tile)[LM_ROWS * lm_stride<MODE>()], lm_f4 (&acc)[4]) {
    using T = LmT<MODE>;
    using V8 = typename T::V8;
    constexpr int NP = MODE == LM_F16X2 ? 2 : 1;
    const int H = a.H;
    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int g = lane / 16, n = lane % 16;
    const V8* A = static_cast<const V8*>(a.afrag);
    const int nfrag = (2 * H + 1) * LM_KCHUNKS * 64;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = lm_f4{0.f, 0.f, 0.f, 0.f};
    for (int dyi = 0; dyi <= 2 * H; ++dyi) {
#pragma unroll
        for (int c = 0; c < LM_KCHUNKS; ++c) {
            const int f = (dyi * LM_KCHUNKS + c) * 64 + lane;
            const V8 ahi = A[f];
            V8 alo;
            if constexpr (NP == 2) alo = A[nfrag + f];
            const int col = 16 * w + 32 * c + 8 * g;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = 16 * t + n + dyi;
                const V8 b = *reinterpret_cast<const V8*>(&tile[0][row * lm_stride<MODE>() + col]);
                acc[t] = T::mfma(ahi, b, acc[t]);
                if constexpr (NP == 2) {
                    const V8 blo = *reinterpret_cast<const V8*>(&tile[1][row * lm_stride<MODE>() + col]);
                    acc[t] = T::mfma(ahi, blo, acc[t]);
                    acc[t] = T::mfma(alo, b, acc[t]);
                }
            }
        }
    }
}

// The potentials through LDS (the tile buffer they were computed from is free
// once every wave is past the loop), so that the growth reads the state and
// writes the result in whole 256-byte rows: from the MFMA layout each wave would
// touch 16 rows 64 bytes at a time.
CAH_DEV void lenia_store_potentials(const lm_f4 (&acc)[4], float* U) {
    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int g = lane / 16, n = lane % 16;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        *reinterpret_cast<float4*>(&U[(16 * t + n) * LM_UST + 16 * w + 4 * g]) =
            make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
}

// ---------------------------------------------------------------- kernels

// Tile k of one lattice (at (k % nbx, k / nbx)): the fill, the MFMA loop, the
// potentials through LDS, the growth.
template <int MODE>
CAH_DEV void lenia_mfma_tile(const LeniaMfmaArgs& a, int k) {
    using E = typename LmT<MODE>::E;
    constexpr int NP = MODE == LM_F16X2 ? 2 : 1;
    static_assert(LM_BY * LM_UST * 4 <= (int)sizeof(E) * NP * LM_ROWS * lm_stride<MODE>(),
                  "the potentials must fit the tile buffer");
    __shared__ __attribute__((aligned(16))) E tile[NP][LM_ROWS * lm_stride<MODE>()];
    const int x0 = (k % a.nbx) * LM_BX, y0 = (k / a.nbx) * LM_BY;
    if (lenia_chunked(a)) {
        float4 v[LM_FILL_Q];
        lenia_fill_load(a, x0, y0, v);
        lenia_fill_store<MODE>(a, x0, v, tile);
    } else {
        lenia_fill_direct<MODE>(a, x0, y0, tile);
    }
    __syncthreads();
    lm_f4 acc[4];
    lenia_mfma_loop<MODE>(a, tile, acc);
    __syncthreads();  // every wave is past the loop: the tile buffer takes the potentials
    float* U = reinterpret_cast<float*>(&tile[0][0]);
    lenia_store_potentials(acc, U);
    __syncthreads();
    lenia_growth_block(a, x0, y0, U);
}

// One lattice: block b computes tile b.
template <int MODE>
__global__ __launch_bounds__(LM_NT) void k_lenia_mfma(LeniaMfmaArgs a) {
    lenia_mfma_tile<MODE>(a, (int)blockIdx.x);
}

// Batched: block b advances tile (b - block_base[i]) of item i = block_item[b],
// with the item's buffers swapped when parity is set and its time base + dt.
template <int MODE>
__global__ __launch_bounds__(LM_NT) void k_lenia_mfma_batch(const LeniaMfmaArgs* __restrict__ items,
                                                           const int32_t* __restrict__ block_item,
                                                           const int32_t* __restrict__ block_base,
                                                           int32_t parity, int64_t dt) {
    const int i = block_item[blockIdx.x];
    LeniaMfmaArgs a = items[i];
    if (parity) {
        const float* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<float*>(c);
    }
    a.t += dt;
    lenia_mfma_tile<MODE>(a, (int)blockIdx.x - block_base[i]);
}
