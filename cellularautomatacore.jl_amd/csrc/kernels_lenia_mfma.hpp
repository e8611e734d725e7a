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
// adds whi*lo and wlo*hi.  Measured on MI355X (R = 13, 2048^2): per row offset the
// LDS reads of B, the A loads and the MFMAs are within 2x of one another; giving
// each wave two strips halves the LDS reads but doubles the A loads and was
// slower, as was staging A in LDS.
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
constexpr int LM_STRIDE = 120;      // LDS row stride in elements (16-byte aligned rows)

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
    int32_t nbx;  // output blocks along the first axis (batched launches)
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

// One 64 x 64 output block (bx, by) of one lattice; 256 threads.
template <int MODE>
CAH_DEV void lenia_mfma_block(const LeniaMfmaArgs& a, int bx, int by) {
    using T = LmT<MODE>;
    using E = typename T::E;
    using V8 = typename T::V8;
    constexpr int NP = MODE == LM_F16X2 ? 2 : 1;  // planes: value, and residual
    __shared__ __attribute__((aligned(16))) E tile[NP][LM_ROWS * LM_STRIDE];

    const int H = a.H;
    const int x0 = bx * LM_BX, y0 = by * LM_BY;
    const int rows = LM_BY + 2 * H;
    auto put = [&](int lx, int ly, float v) {
        const E hi = (E)v;
        tile[0][ly * LM_STRIDE + lx] = hi;
        if constexpr (NP == 2) tile[1][ly * LM_STRIDE + lx] = (E)(v - (float)hi);
    };
    // Tile column lx is lattice column x0 - H + lx.  The tile is read as aligned
    // float4 chunks starting at xa = (x0 - H) rounded down to 4 (rows are 16-float
    // aligned and x0 is a multiple of 64).  With a first side that is a multiple
    // of 4 and a lattice larger than the tile, every chunk lies wholly inside the
    // lattice or wholly beyond one edge, where it maps onto an aligned chunk inside
    // (wrapped for Periodic, mirrored and read back to front for Reflecting) or
    // onto the Constant value, as _map_axis does cell by cell; rows map the same
    // way.  Smaller lattices go through the boundary map cell by cell.
    constexpr int NCH = LM_COLS / 4 + 1;  // chunks covering LM_COLS + 3 columns
    const int xs = x0 - H, sh = xs & 3, xa = xs - sh;
    const bool chunked = a.L.n1 % 4 == 0 && a.L.n1 >= 4 * NCH && a.L.n2 >= rows;
    if (chunked) {
        // every load of the thread is issued before the first is used, then four
        // values go to LDS at once when the chunk lands on a 4-column boundary
        // Columns at or past 64 + 2H only meet zero weights: they are staged as
        // zeros without a load (they must still be finite).  The loads go in
        // passes of PASS per thread, all issued before the first is used (one load
        // in flight at a time leaves the fill waiting on HBM latency); four values
        // then go to LDS at once when the chunk lands on a 4-column boundary.
        constexpr int PASS = 8;
        const int nload = (LM_BX + 2 * H + sh + 3) / 4;
        for (int base = 0; base < rows * NCH; base += 256 * PASS) {
            float4 v4[PASS];
#pragma unroll
            for (int k = 0; k < PASS; ++k) {
                const int i = base + threadIdx.x + 256 * k;
                v4[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (i >= rows * NCH) continue;
                const int ly = i / NCH, j = i - ly * NCH;
                if (j >= nload) continue;
                bool o2;
                const int my = map_axis0(a.L.b2, y0 - H + ly, a.L.n2, o2);
                const int x = xa + 4 * j;
                const float* row = a.cur + (int64_t)my * a.L.pitch;
                if (o2) {
                    v4[k] = make_float4(a.cv2, a.cv2, a.cv2, a.cv2);
                } else if (x >= 0 && x + 4 <= a.L.n1) {
                    v4[k] = *reinterpret_cast<const float4*>(row + x);
                } else if (a.L.b1 == CAH_BOUNDARY_PERIODIC) {
                    v4[k] = *reinterpret_cast<const float4*>(row + (x < 0 ? x + a.L.n1 : x - a.L.n1));
                } else if (a.L.b1 == CAH_BOUNDARY_REFLECTING) {
                    const float4 f = *reinterpret_cast<const float4*>(row + (x < 0 ? -x - 4 : 2 * a.L.n1 - x - 4));
                    v4[k] = make_float4(f.w, f.z, f.y, f.x);
                } else {
                    v4[k] = make_float4(a.cv1, a.cv1, a.cv1, a.cv1);
                }
            }
#pragma unroll
            for (int k = 0; k < PASS; ++k) {
                const int i = base + threadIdx.x + 256 * k;
                if (i >= rows * NCH) continue;
                const int ly = i / NCH, j = i - ly * NCH;
                const float vv[4] = {v4[k].x, v4[k].y, v4[k].z, v4[k].w};
                const int lx0 = 4 * j - sh;
                if (sh == 0 && lx0 + 4 <= LM_COLS) {
                    E h[4], l[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        h[e] = (E)vv[e];
                        l[e] = (E)(vv[e] - (float)h[e]);
                    }
                    typedef E e4 __attribute__((ext_vector_type(4)));
                    *reinterpret_cast<e4*>(&tile[0][ly * LM_STRIDE + lx0]) = e4{h[0], h[1], h[2], h[3]};
                    if constexpr (NP == 2)
                        *reinterpret_cast<e4*>(&tile[1][ly * LM_STRIDE + lx0]) = e4{l[0], l[1], l[2], l[3]};
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int lx = lx0 + e;
                        if (lx >= 0 && lx < LM_COLS) put(lx, ly, vv[e]);
                    }
                }
            }
        }
    } else {
        for (int i = threadIdx.x; i < rows * LM_COLS; i += blockDim.x) {
            const int lx = i % LM_COLS, ly = i / LM_COLS;
            put(lx, ly, read_mapped<float>(a.cur, a.L, a.cv1, a.cv2, xs + lx, y0 - H + ly));
        }
    }
    __syncthreads();

    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int g = lane / 16, n = lane % 16;
    const V8* A = static_cast<const V8*>(a.afrag);
    const int nfrag = (2 * H + 1) * 2 * 64;
    lm_f4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = lm_f4{0.f, 0.f, 0.f, 0.f};
    for (int dyi = 0; dyi <= 2 * H; ++dyi) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int f = (dyi * 2 + c) * 64 + lane;
            const V8 ahi = A[f];
            V8 alo;
            if constexpr (NP == 2) alo = A[nfrag + f];
            const int col = 16 * w + 32 * c + 8 * g;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = 16 * t + n + dyi;
                const V8 b = *reinterpret_cast<const V8*>(&tile[0][row * LM_STRIDE + col]);
                acc[t] = T::mfma(ahi, b, acc[t]);
                if constexpr (NP == 2) {
                    const V8 blo = *reinterpret_cast<const V8*>(&tile[1][row * LM_STRIDE + col]);
                    acc[t] = T::mfma(ahi, blo, acc[t]);
                    acc[t] = T::mfma(alo, b, acc[t]);
                }
            }
        }
    }

    // The potentials go through LDS (the input tile is no longer needed), so the
    // block then reads the state and writes the result in whole 256-byte rows:
    // from the MFMA layout each wave would touch 16 rows 64 bytes at a time.
    __syncthreads();
    constexpr int UST = LM_BX + 4;  // row stride in floats
    static_assert(LM_BY * UST * 4 <= (int)sizeof(tile), "the potentials must fit the tile buffer");
    float* U = reinterpret_cast<float*>(&tile[0][0]);
    // D[m][n] of tile t sits in lane (n + 16 * (m / 4)), register m % 4
#pragma unroll
    for (int t = 0; t < 4; ++t)
        *reinterpret_cast<float4*>(&U[(16 * t + n) * UST + 16 * w + 4 * g]) =
            make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
    __syncthreads();

    // the growth in FP32 (lenia_fast32): mu, -1 / (2 sigma^2), dt
    const float mu = (float)a.mu, nk = (float)(-1.0 / (2.0 * a.sigma * a.sigma)), dt = (float)a.dt;
    auto cell = [&](int x, int y, float c, float u) {
        const int64_t lin = (int64_t)y * a.L.n1 + x;
        const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                          gate_open(a.scheme, 0, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1, a.t, x, y,
                                    a.L.two_d);
        return open ? lenia_fast32(c, u, mu, nk, dt) : c;
    };
    const bool full = x0 + LM_BX <= a.L.n1 && y0 + LM_BY <= a.L.n2;
    for (int q = threadIdx.x; q < LM_BY * LM_BX / 4; q += blockDim.x) {
        const int ly = q / (LM_BX / 4), lx = 4 * (q % (LM_BX / 4));
        const int y = y0 + ly, x = x0 + lx;
        const float4 u4 = *reinterpret_cast<const float4*>(&U[ly * UST + lx]);
        const int64_t at = (int64_t)y * a.L.pitch + x;
        if (full) {
            const float4 c4 = *reinterpret_cast<const float4*>(a.cur + at);
            *reinterpret_cast<float4*>(a.next + at) =
                make_float4(cell(x, y, c4.x, u4.x), cell(x + 1, y, c4.y, u4.y), cell(x + 2, y, c4.z, u4.z),
                            cell(x + 3, y, c4.w, u4.w));
        } else if (y < a.L.n2) {
            const float uu[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (x + e < a.L.n1) a.next[at + e] = cell(x + e, y, a.cur[at + e], uu[e]);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_lenia_mfma(LeniaMfmaArgs a) {
    lenia_mfma_block<MODE>(a, blockIdx.x, blockIdx.y);
}

// Batched: block b advances block (b - block_base[i]) of item i = block_item[b],
// with the item's buffers swapped when parity is set and its time base + dt.
template <int MODE>
__global__ __launch_bounds__(256) void k_lenia_mfma_batch(const LeniaMfmaArgs* __restrict__ items,
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
    const int local = (int)blockIdx.x - block_base[i];
    lenia_mfma_block<MODE>(a, local % a.nbx, local / a.nbx);
}
