// kernels_lenia_mfma.hpp -- the Lenia potential as an implicit GEMM on MFMA.
//
// U(x, y) = sum over (dx, dy) of W(dx, dy) S(x + dx, y + dy).  For each row offset
// dy the sum over dx is a banded (Toeplitz) matrix times the input: with the
// output x of a 16-wide strip as M, the input x positions as K and 16 output rows
// as N,
//   D[m][n] += A_dy[m][k] * B_dy[k][n],  A_dy[m][k] = W(k - H - m, dy),
//   B_dy[k][n] = S(x0 - H + k, y0 + n + dy),
// accumulated over the 2H+1 row offsets into FP32 (v_mfma_f32_16x16x32_{f16,bf16}).
// A_dy does not depend on the tile, so it is precomputed once per problem in
// fragment order and read from global memory (it stays in L1/L2); the input
// tile is staged in LDS in the MFMA input type.  Split-FP16 keeps the residual of
// both operands (S = hi + lo, W = whi + wlo) and adds whi*lo and wlo*hi.
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
constexpr int LM_COLS = LM_BX + 48; // input columns staged (past 16 + 2H they only meet zeros)
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
    for (int i = threadIdx.x; i < rows * LM_COLS; i += blockDim.x) {
        const int lx = i % LM_COLS, ly = i / LM_COLS;
        const float v = read_mapped<float>(a.cur, a.L, a.cv1, a.cv2, x0 - H + lx, y0 - H + ly);
        const E hi = (E)v;
        tile[0][ly * LM_STRIDE + lx] = hi;
        if constexpr (NP == 2) tile[1][ly * LM_STRIDE + lx] = (E)(v - (float)hi);
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

    // D[m][n] of tile t sits in lane (n + 16 * (m / 4)), register m % 4
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int y = y0 + 16 * t + n;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = x0 + 16 * w + 4 * g + j;
            if (x >= a.L.n1 || y >= a.L.n2) continue;
            const int64_t lin = (int64_t)y * a.L.n1 + x;
            const float c = a.cur[(int64_t)y * a.L.pitch + x];
            float out = c;
            const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                              gate_open(a.scheme, 0, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1, a.t,
                                        x, y, a.L.two_d);
            if (open) {
                if (a.prec32)
                    out = lenia_new32(c, acc[t][j], (float)a.mu, (float)a.sigma, (float)a.dt);
                else
                    out = lenia_new64(c, (double)acc[t][j], a.mu, a.sigma, a.dt);
            }
            a.next[(int64_t)y * a.L.pitch + x] = out;
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
