// kernels_epidemic.hpp -- the kernel field of the field epidemic.
//
//   k_epi_field<T>  field(I) = sum over taps k, in tap order, of w_k * (u(I + o_k) == 1),
//                   in T (double for KernelField precision Float64, float for
//                   Float32): _compute_field_reference! / _field_at
//                   (src/environment.jl:151-169) with the source ==(1).  w * false
//                   is +0, which leaves the running sum unchanged, so the sum adds
//                   the weights of the taps that see an infected cell, in order.
//                   The state is read into an LDS tile through the boundary map
//                   (Constant edges count as infected when their value is 1).
#pragma once
#include "device.hpp"

constexpr int EPI_TX = 64, EPI_TY = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_epi_field(const int8_t* __restrict__ cur, Lat L, int8_t cv1, int8_t cv2,
                                                   const int32_t* __restrict__ taps, const T* __restrict__ w,
                                                   int32_t ntaps, int32_t h, T* __restrict__ field) {
    extern __shared__ int8_t etile[];
    const int W = EPI_TX + 2 * h, H = EPI_TY + 2 * h;
    const int x0 = blockIdx.x * EPI_TX, y0 = blockIdx.y * EPI_TY;
    for (int i = threadIdx.x; i < W * H; i += blockDim.x) {
        const int lx = i % W, ly = i / W;
        etile[i] = read_mapped<int8_t>(cur, L, cv1, cv2, x0 - h + lx, y0 - h + ly);
    }
    __syncthreads();
    // a thread's four cells (rows ty, ty + 4, ty + 8, ty + 12 of the tile) keep four
    // independent sums, so the tap loop issues four LDS reads and four adds a tap;
    // each cell still adds its taps in order
    const int tx = threadIdx.x % EPI_TX, ty = threadIdx.x / EPI_TX;
    constexpr int R = EPI_TY / 4;
    static_assert(EPI_TX * 4 == 256, "four rows of threads a block");
    const int8_t* cp = etile + (tx + h) + (ty + h) * W;
    T s[R];
#pragma unroll
    for (int r = 0; r < R; ++r) s[r] = T(0);
#pragma unroll 4
    for (int k = 0; k < ntaps; ++k) {
        const int off = taps[2 * k] + taps[2 * k + 1] * W;
        const T wk = w[k];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (cp[off + 4 * r * W] == 1) s[r] = s[r] + wk;
    }
    const int x = x0 + tx;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y0 + ty + 4 * r;
        if (x < L.n1 && y < L.n2) field[(int64_t)y * L.n1 + x] = s[r];
    }
}
