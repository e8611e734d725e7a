// kernels_box.hpp -- Life and Generations with a Moore(R) neighbourhood the
// band walks do not take (radius 7..32, or more count intervals than they test)
// or a VonNeumann(r) neighbourhood of radius 1..8: neighbour counts as window
// sums (src/rules.jl:377-389, :446-456).
//
//   k_box_i8<VN>  a 64 x 64 block of cells a launch block, from an LDS tile of
//                 the state with the boundary conditions applied while it is
//                 loaded (as k_generic_i8 loads its tile), in three passes:
//   1. per tile row, the prefix counts of cells equal to 1 (bytes: a row holds
//      at most 128 cells), four cells a dword;
//   2. per tile row, window sums of those prefixes over the row's 64 output
//      columns: half-width R for Moore, each half-width w = 0..r for von Neumann
//      (one plane of sums per w);
//   3. per output cell, the column sums of step 2 over the rows dy = -R..R
//      (Moore: a sliding window down four rows, in 16-bit lanes) or, for von
//      Neumann, the sum of plane r - |dy| at row dy; the count less the cell
//      itself, then the rule (box_rule) and, for GATED, the gate and the mask.
// The count work per cell is independent of the radius in passes 1 and 2 and
// linear in it in pass 3, where the general kernel's is quadratic.
#pragma once
#include "kernels_generic.hpp"

constexpr int BOX_T = 64;                          // output block: BOX_T x BOX_T cells
constexpr int BOX_MOORE_MAX = 32;  // Moore(R <= 6) with at most two count intervals takes the band walks
constexpr int BOX_VN_MAX = 8;

// The LDS layout of a block, all row strides multiples of 4 bytes:
//   tile: box_rows(r) rows of box_w(r) cells (the state at columns x0 - hp ..,
//         rows y0 - r ..);
//   pre:  box_rows(r) rows of box_pw(r) bytes, a zero dword and then the prefix
//         counts of the tile row (byte 4 + i: cells 0..i equal to 1);
//   sums: planes x box_rows(r) rows of BOX_T bytes.
CAH_HD int box_hp(int r) { return (r + 3) & ~3; }
CAH_HD int box_w(int r) { return BOX_T + 2 * box_hp(r); }
CAH_HD int box_pw(int r) { return box_w(r) + 4; }
CAH_HD int box_rows(int r) { return BOX_T + 2 * r; }
CAH_HD int box_planes(bool vn, int r) { return vn ? r + 1 : 1; }
CAH_HD size_t box_lds(bool vn, int r) {
    return (size_t)box_rows(r) * (size_t)(box_w(r) + box_pw(r) + BOX_T * box_planes(vn, r));
}

// Bytes off .. off + 3 of an LDS byte array read as one dword (two aligned
// dwords and a byte shift; the second only contributes when off is unaligned).
CAH_DEV uint32_t box_dword(const uint32_t* p32, int off) {
    const int dw = off >> 2;
    return __builtin_amdgcn_alignbyte(p32[dw + 1], p32[dw], off & 3);
}

// Inclusive prefix sums of x over each half-wave (lanes 0-31 and 32-63) by DPP:
// row shifts within the 16-lane rows, then lane 15 of rows 0 and 2 added to rows 1
// and 3.  Every lane of a half-wave that is used must be active.
CAH_DEV uint32_t box_scan32(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    return x;
}

// local_rule of Life (src/rules.jl:385-389) or Generations (:446-456) for a
// cell of state c with n neighbours equal to 1.
CAH_DEV int box_rule(const GenArgs& a, int c, int n) {
    if (a.rule == CAH_RULE_GENERATIONS) {
        if (c == 0) return n < 64 ? (int)((a.birth >> n) & 1u) : 0;
        if (c == 1) return (n < 64 && ((a.survive >> n) & 1u)) ? 1 : (a.C > 2 ? 2 : 0);
        return c + 1 < a.C ? c + 1 : 0;
    }
    const uint64_t m = c == 1 ? a.survive : a.birth;
    return n < 64 ? (int)((m >> n) & 1u) : 0;
}

// a.h is the radius: R for Moore, r for von Neumann.  GATED: a mask or a
// Stochastic / Checkerboard gate keeps cells (_update_cell!, src/backends.jl:444-455).
template <bool VN, bool GATED>
__global__ __launch_bounds__(256) void k_box_i8(GenArgs a) {
    extern __shared__ uint32_t box32[];
    const int r = a.h, hp = box_hp(r), W = box_w(r), PW = box_pw(r), rows = box_rows(r);
    uint32_t* tile32 = box32;
    uint32_t* pre32 = tile32 + rows * W / 4;
    uint32_t* sum32 = pre32 + rows * PW / 4;
    const int x0 = blockIdx.x * BOX_T, y0 = blockIdx.y * BOX_T;
    const int xs = x0 - hp, ys = y0 - r;

    fill_tile_i8(a, tile32, xs, ys, W, rows);
    __syncthreads();

    // 1. prefix counts, half a wave per tile row (W / 4 <= 32 dwords), lane l
    // taking dword l: within the dword by one multiply (byte k of e * 0x01010101
    // is e_0 + .. + e_k), then the counts of the row's earlier dwords (a scan
    // over the half-wave) added to every byte (at most 128: no byte overflows)
    {
        const int l = threadIdx.x & 31, ND = W / 4;
        for (int ly = threadIdx.x >> 5; ly < rows; ly += blockDim.x >> 5) {
            const uint32_t q = bytes_eq(l < ND ? tile32[ly * ND + l] : 0u, 0x01010101u) * 0x01010101u;
            const uint32_t own = q >> 24;
            const uint32_t run = box_scan32(own);
            uint32_t* p = pre32 + ly * (PW / 4);
            if (l == 0) p[0] = 0;
            if (l < ND) p[1 + l] = q + (run - own) * 0x01010101u;
        }
    }
    __syncthreads();

    // 2. window sums of the output columns: cells hp + x - w .. hp + x + w of a
    // row are prefix bytes 4 + hp + x + w minus 3 + hp + x - w (bytewise: every
    // byte's difference is a count of 0..2w + 1)
    const int planes = VN ? r + 1 : 1;
    for (int i = threadIdx.x; i < planes * rows * (BOX_T / 4); i += blockDim.x) {
        const int k = i % (BOX_T / 4), rest = i / (BOX_T / 4);
        const int ly = rest % rows, w = VN ? rest / rows : r;
        const uint32_t* p = pre32 + ly * (PW / 4);
        const int x = 4 * k;
        sum32[i] = box_dword(p, 4 + hp + x + w) - box_dword(p, 3 + hp + x - w);
    }
    __syncthreads();

    // 3. column sums and the rule: thread tid takes the dword of cells
    // 4 (tid % 16) .. + 3 of output rows 4 (tid / 16) .. + 3
    const int c4 = threadIdx.x % 16, yb = 4 * (threadIdx.x / 16);
    const int x = x0 + 4 * c4;
    if (x >= a.L.n1) return;
    const uint32_t* col = sum32 + c4;  // row ly of plane w: col[(w * rows + ly) * 16]
    uint32_t lo = 0, hi = 0;  // Moore: the window's sums in 16-bit lanes (cells 0, 2 and 1, 3)
    if constexpr (!VN) {
        for (int j = 0; j < 2 * r; ++j) {
            const uint32_t s = col[(yb + j) * 16];
            lo += s & 0x00FF00FFu;
            hi += (s >> 8) & 0x00FF00FFu;
        }
    }
    for (int dyo = 0; dyo < 4; ++dyo) {
        const int ly = yb + dyo, y = y0 + ly;
        if (y >= a.L.n2) break;
        uint32_t n[4];
        if constexpr (VN) {
            // plane r - |j - r| of tile row ly + j, j = 0..2r; at most 2r(r + 1) + 1
            // a byte (145 for r = 8): bytewise sums
            uint32_t v = 0;
            for (int j = 0; j <= 2 * r; ++j) {
                const int w = r - (j < r ? r - j : j - r);
                v += col[(w * rows + ly + j) * 16];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) n[e] = (v >> (8 * e)) & 0xFFu;
        } else {
            const uint32_t s = col[(ly + 2 * r) * 16];  // the row entering the window
            lo += s & 0x00FF00FFu;
            hi += (s >> 8) & 0x00FF00FFu;
            n[0] = lo & 0xFFFFu;
            n[1] = hi & 0xFFFFu;
            n[2] = lo >> 16;
            n[3] = hi >> 16;
            const uint32_t o = col[ly * 16];  // the row leaving it
            lo -= o & 0x00FF00FFu;
            hi -= (o >> 8) & 0x00FF00FFu;
        }
        const uint32_t cw = tile32[((ly + r) * W + hp + 4 * c4) >> 2];
        uint32_t out = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = (int)(int8_t)(cw >> (8 * e));
            const int cnt = (int)n[e] - (c == 1 ? 1 : 0);  // the window holds the cell itself
            int v = 0;
            if (x + e < a.L.n1) {
                bool open = true;
                if constexpr (GATED) {
                    const int64_t lin = (int64_t)y * a.L.n1 + x + e;
                    open = (a.mask == nullptr || a.mask[lin] != 0) &&
                           gate_open(a.scheme, a.gate, a.sp, a.sp_f32, a.seed, (uint64_t)lin + 1, a.t, x + e, y,
                                     a.L.two_d);
                }
                v = open ? box_rule(a, c, cnt) : c;
            }
            out |= (uint32_t)(uint8_t)v << (8 * e);
        }
        int8_t* dst = a.next + (int64_t)y * a.L.pitch + x;
        if (x + 4 <= a.L.n1) {
            *reinterpret_cast<uint32_t*>(dst) = out;
        } else {
            for (int e = 0; x + e < a.L.n1; ++e) dst[e] = (int8_t)(out >> (8 * e));
        }
    }
}
