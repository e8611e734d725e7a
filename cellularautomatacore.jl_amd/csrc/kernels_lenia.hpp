// kernels_lenia.hpp -- the Lenia potential (a KernelField) and growth step.
//
//   k_lenia_field_seq  the potential in tap order, one sequential sum per cell
//                      (FP64 for Lenia(), FP32 for precision=Float32), exactly the
//                      operation order of _field_at (src/environment.jl:163-169).
//   k_lenia_growth     local_rule(::Lenia) (src/lenia.jl:42-47) with the gate and
//                      the mask, reading a field buffer.
//   k_lenia_f32col     field and growth fused: FP32 FMAs over a dense
//                      (2H+1)^2 weight table, each thread holding one column of
//                      CB+2H tile values in registers for CB outputs.
#pragma once
#include "device.hpp"
#include "lenia_growth.hpp"

struct LeniaArgs {
    const float* cur;
    float* next;
    const uint8_t* mask;
    const int32_t* taps;  // n_taps read offsets (o1, o2)
    const double* w64;
    const float* w32;     // Float32(w) per tap, or the dense table for k_lenia_f32col
    void* field;          // dense prod(dims) FP64 or FP32 field buffer
    Lat L;
    int32_t h, ntaps;
    float cv1, cv2;
    int32_t prec32;
    double mu, sigma, dt;
    uint64_t seed;
    int64_t t;
    int32_t scheme, gate, sp_f32;
    double sp;
    int32_t field32;      // the field buffer holds FP32 values (FFT path) for a Lenia() rule
};

constexpr int LEN_TX = 64, LEN_TY = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_lenia_field_seq(LeniaArgs a) {
    extern __shared__ float ftile[];
    const int W = LEN_TX + 2 * a.h, H = LEN_TY + 2 * a.h;
    const int x0 = blockIdx.x * LEN_TX, y0 = blockIdx.y * LEN_TY;
    for (int i = threadIdx.x; i < W * H; i += blockDim.x) {
        const int lx = i % W, ly = i / W;
        ftile[i] = read_mapped<float>(a.cur, a.L, a.cv1, a.cv2, x0 - a.h + lx, y0 - a.h + ly);
    }
    __syncthreads();
    const int tx = threadIdx.x % LEN_TX;
    for (int ty = threadIdx.x / LEN_TX; ty < LEN_TY; ty += blockDim.x / LEN_TX) {
        const int x = x0 + tx, y = y0 + ty;
        if (x >= a.L.n1 || y >= a.L.n2) continue;
        const float* cp = ftile + (tx + a.h) + (ty + a.h) * W;
        T s = T(0);
        for (int k = 0; k < a.ntaps; ++k) {
            const T w = sizeof(T) == 8 ? (T)a.w64[k] : (T)a.w32[k];
            s = s + w * (T)cp[a.taps[2 * k] + a.taps[2 * k + 1] * W];  // no FMA: built with
                                                                       // -ffp-contract=off
        }
        static_cast<T*>(a.field)[(int64_t)y * a.L.n1 + x] = s;
    }
}

__global__ __launch_bounds__(256) void k_lenia_growth(LeniaArgs a) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.L.n1) return;
    for (int32_t y = blockIdx.y; y < a.L.n2; y += gridDim.y) {
    const int64_t lin = (int64_t)y * a.L.n1 + x;
    const float c = a.cur[(int64_t)y * a.L.pitch + x];
    float out = c;
    const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                      gate_open(a.scheme, a.gate, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1, a.t,
                                x, y, a.L.two_d);
    if (open) {
        if (a.prec32)
            out = lenia_new32(c, static_cast<const float*>(a.field)[lin], (float)a.mu,
                              (float)a.sigma, (float)a.dt);
        else if (a.field32)
            out = lenia_new64(c, (double)static_cast<const float*>(a.field)[lin], a.mu, a.sigma,
                              a.dt);
        else
            out = lenia_new64(c, static_cast<const double*>(a.field)[lin], a.mu, a.sigma, a.dt);
    }
    a.next[(int64_t)y * a.L.pitch + x] = out;
    }
}

// Fused FP32 field + growth.  Tile: 64 columns x 32 rows of outputs; thread
// (tx, g) computes rows 8g .. 8g+7 of column tx.  `a.w32` is the dense table
// wd[(dy + H) * (2H + 1) + (dx + H)] (zero where the kernel has no tap).
template <int H>
__global__ __launch_bounds__(256) void k_lenia_f32col(LeniaArgs a) {
    constexpr int TX = 64, CB = 8, TY = 4 * CB, W = TX + 2 * H, HT = TY + 2 * H, K = 2 * H + 1;
    __shared__ float tile[W * HT];
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    for (int i = threadIdx.x; i < W * HT; i += blockDim.x) {
        const int lx = i % W, ly = i / W;
        tile[i] = read_mapped<float>(a.cur, a.L, a.cv1, a.cv2, x0 - H + lx, y0 - H + ly);
    }
    __syncthreads();
    const int tx = threadIdx.x % TX, g = threadIdx.x / TX;
    float acc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) acc[j] = 0.0f;
    for (int dx = 0; dx < K; ++dx) {
        float col[CB + 2 * H];
#pragma unroll
        for (int m = 0; m < CB + 2 * H; ++m) col[m] = tile[(tx + dx) + (g * CB + m) * W];
#pragma unroll
        for (int dy = 0; dy < K; ++dy) {
            const float w = a.w32[dy * K + dx];
            if (w != 0.0f) {
#pragma unroll
                for (int j = 0; j < CB; ++j) acc[j] = __builtin_fmaf(w, col[j + dy], acc[j]);
            }
        }
    }
    const int x = x0 + tx;
    if (x >= a.L.n1) return;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int y = y0 + g * CB + j;
        if (y >= a.L.n2) break;
        const int64_t lin = (int64_t)y * a.L.n1 + x;
        const float c = tile[(tx + H) + (g * CB + j + H) * W];
        float out = c;
        const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                          gate_open(a.scheme, a.gate, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1,
                                    a.t, x, y, a.L.two_d);
        if (open) {
            if (a.prec32)
                out = lenia_new32(c, acc[j], (float)a.mu, (float)a.sigma, (float)a.dt);
            else
                out = lenia_new64(c, (double)acc[j], a.mu, a.sigma, a.dt);
        }
        a.next[(int64_t)y * a.L.pitch + x] = out;
    }
}
