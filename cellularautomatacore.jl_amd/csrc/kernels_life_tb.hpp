// kernels_life_tb.hpp -- several steps of a Moore(1) Life rule per launch,
// one bit per cell in LDS.
//
// A block owns an output tile of TXW words (32 cells each) x TY rows.  It loads
// the tile with one word of halo on each side and K rows above and below
// through the boundary map, as bits: bit i of a word is the cell 32c + i of its
// row and is set when the cell's byte equals 1 (the only value Life counts as
// alive, src/rules.jl:377-389, so any Int8 contents pack exactly).  It then runs
// up to K synchronous steps in LDS and writes the tile's interior back as
// bytes (Life writes only 0 and 1).
//
// One step of a word: the three rows' words and their neighbours give the eight
// neighbour bit-vectors through funnel shifts; a bit-sliced adder forms the
// count's four bit planes; the birth / survive masks select the new bits.  Cells
// outside the tile's valid region go wrong from the tile's edge inwards, one
// cell a step, and never reach the interior: the halo is K rows and 32
// columns (K <= 32).  Outside the lattice the halo holds the boundary map's image:
// periodic and reflected copies evolve exactly as the cells they copy (Moore(1)
// is symmetric under both), and cells beyond a Constant edge are reset to the
// constant after every step (the row axis's value wins in corners, as
// _boundary_lookup's highest-axis rule has it).
#pragma once
#include "device.hpp"

constexpr int TB_K = 16;  // steps a launch at most: the halo's rows (each memory pass then serves 16 steps)

struct LifeTbArgs {
    const int8_t* cur;
    int8_t* next;
    Lat L;
    int8_t cv1, cv2;          // Constant values per axis
    uint32_t birth, survive;  // Life masks over counts 0..8
    int32_t steps;            // 1..K
};

// Bit i of the result is set when byte i of the 32 bytes at p equals 1.
CAH_DEV uint32_t tb_pack16(uint4 v) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t x = w[j] ^ 0x01010101u;  // zero bytes where the cell is 1
        uint32_t e = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu) >> 7;  // 0x01 per such byte
        e |= e >> 7;
        e |= e >> 14;
        r |= (e & 0xFu) << (4 * j);
    }
    return r;
}

// Four bits (the low nibble of n) as four bytes 0 / 1.
CAH_DEV uint32_t tb_nibble_bytes(uint32_t n) { return ((n & 0xFu) * 0x00204081u) & 0x01010101u; }

// The Life rule on 32 cells: x the cells, n0..n7 their eight neighbours.
CAH_DEV uint32_t tb_rule(uint32_t x, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e, uint32_t f,
                         uint32_t g, uint32_t h, uint32_t birth, uint32_t survive) {
    // bit-sliced count of the eight inputs: full adders (sum = xor3, carry = majority)
    const uint32_t s1 = a ^ b ^ c, c1 = (a & b) | (c & (a ^ b));
    const uint32_t s2 = d ^ e ^ f, c2 = (d & e) | (f & (d ^ e));
    const uint32_t s3 = g ^ h, c3 = g & h;
    const uint32_t q0 = s1 ^ s2 ^ s3, c4 = (s1 & s2) | (s3 & (s1 ^ s2));  // weight 1, carry weight 2
    const uint32_t t = c1 ^ c2 ^ c3, u = (c1 & c2) | (c3 & (c1 ^ c2));    // weight 2, carry weight 4
    const uint32_t q1 = t ^ c4, v = t & c4;                               // weight 2, carry weight 4
    const uint32_t q2 = u ^ v, q3 = u & v;                                // weights 4, 8
    // count == k as (high two bits == k >> 2) & (low two bits == k & 3), and the
    // masks as all-ones / all-zero words: no branch
    const uint32_t lo[4] = {~q1 & ~q0, ~q1 & q0, q1 & ~q0, q1 & q0};
    const uint32_t hi[3] = {~q3 & ~q2, ~q3 & q2, q3 & ~q2};
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k <= 8; ++k) {
        const uint32_t bk = 0u - ((birth >> k) & 1u), sk = 0u - ((survive >> k) & 1u);
        out |= hi[k >> 2] & lo[k & 3] & ((sk & x) | (bk & ~x));
    }
    return out;
}

// Tile (bx, by) of a lattice, by one block of 256 threads.  LDS rows hold a zero
// word before and after the GW words of the grid, so every word has a left and a
// right neighbour to read.
template <int TXW, int TY, int K>
CAH_DEV void life_tb_tile(const LifeTbArgs& a, int bx, int by) {
    static_assert(K >= 1 && K <= 32, "the 32-column halo covers at most 32 steps");
    constexpr int GW = TXW + 2, GH = TY + 2 * K, NW = GW * GH, RS = GW + 2;
    __shared__ uint32_t grid[2][GH * RS];
    __shared__ uint32_t outside[GW];  // per grid column: the bits beyond a Constant first-axis edge
    const int X0 = bx * TXW * 32 - 32;  // column of bit 0 of grid word 0
    const int Y0 = by * TY - K;          // row of grid row 0
    const Lat& L = a.L;
    const bool const1 = L.b1 == CAH_BOUNDARY_CONSTANT, const2 = L.b2 == CAH_BOUNDARY_CONSTANT;
    if (threadIdx.x < GW) {
        const int x = X0 + 32 * (int)threadIdx.x;
        uint32_t m = 0;
        if (const1)
            for (int i = 0; i < 32; ++i) m |= (uint32_t)(x + i < 0 || x + i >= L.n1) << i;
        outside[threadIdx.x] = m;
    }
    for (int r = threadIdx.x; r < GH; r += blockDim.x) {  // the padding words
        grid[0][r * RS] = grid[0][r * RS + RS - 1] = 0u;
        grid[1][r * RS] = grid[1][r * RS + RS - 1] = 0u;
    }
    // load and pack: the row through the boundary map once (a row beyond a
    // Constant edge is that constant, whatever the column: axis 2 wins) and whole
    // in-lattice words with two 16-byte loads
    for (int w = threadIdx.x; w < NW; w += blockDim.x) {
        const int r = w / GW, c = w % GW;
        const int y = Y0 + r, x = X0 + 32 * c;
        bool o2;
        const int32_t ym = map_axis0(L.b2, y, L.n2, o2);
        uint32_t bits = 0;  // a word straddling a column edge is filled below
        if (o2) {
            bits = a.cv2 == 1 ? ~0u : 0u;
        } else if (x >= 0 && x + 32 <= L.n1) {
            const uint4* p = reinterpret_cast<const uint4*>(a.cur + (int64_t)ym * L.pitch + x);
            bits = tb_pack16(p[0]) | (tb_pack16(p[1]) << 16);
        }
        grid[0][r * RS + c + 1] = bits;
    }
    __syncthreads();  // the zeros above come before the words below
    // the columns that straddle or lie beyond a first-axis edge: a wave's 64
    // lanes read 64 bytes (two rows' words) through the column map and form the
    // two words with one ballot
    for (int c = 0; c < GW; ++c) {
        const int x = X0 + 32 * c;
        if (x >= 0 && x + 32 <= L.n1) continue;  // uniform
        for (int j0 = threadIdx.x & ~63; j0 < GH * 32; j0 += blockDim.x) {
            const int j = j0 + (threadIdx.x & 63), r = j >> 5, i = j & 31;
            bool alive = false, o2 = true;
            int32_t ym = 0;
            if (r < GH) ym = map_axis0(L.b2, Y0 + r, L.n2, o2);
            if (!o2) {
                bool o1;
                const int32_t xm = map_axis0(L.b1, x + i, L.n1, o1);
                alive = (o1 ? a.cv1 : a.cur[(int64_t)ym * L.pitch + xm]) == 1;
            }
            const unsigned long long m = __ballot(alive);
            if ((threadIdx.x & 63) == 0 && r < GH && !o2) grid[0][r * RS + c + 1] = (uint32_t)m;
            if ((threadIdx.x & 63) == 32 && r < GH && !o2) grid[0][r * RS + c + 1] = (uint32_t)(m >> 32);
        }
    }
    __syncthreads();
    const uint32_t cbits1 = a.cv1 == 1 ? ~0u : 0u, cbits2 = a.cv2 == 1 ? ~0u : 0u;
    for (int s = 0; s < a.steps; ++s) {
        const uint32_t* src = grid[s & 1];
        uint32_t* dst = grid[(s + 1) & 1];
        for (int w = threadIdx.x + GW; w < NW - GW; w += blockDim.x) {  // rows 1 .. GH-2
            const int r = w / GW, c = w % GW, i = r * RS + c + 1;
            const uint32_t ul = src[i - RS - 1], um = src[i - RS], ur = src[i - RS + 1];
            const uint32_t ml = src[i - 1], mm = src[i], mr = src[i + 1];
            const uint32_t dl = src[i + RS - 1], dm = src[i + RS], dr = src[i + RS + 1];
            // left neighbour of bit i is bit i-1: (w << 1) | (wl >> 31); right: (w >> 1) | (wr << 31)
            uint32_t out = tb_rule(mm, __builtin_amdgcn_alignbit(um, ul, 31), um, __builtin_amdgcn_alignbit(ur, um, 1),
                                   __builtin_amdgcn_alignbit(mm, ml, 31), __builtin_amdgcn_alignbit(mr, mm, 1),
                                   __builtin_amdgcn_alignbit(dm, dl, 31), dm, __builtin_amdgcn_alignbit(dr, dm, 1),
                                   a.birth, a.survive);
            if (const1) {  // cells beyond a Constant first-axis edge keep it
                const uint32_t m = outside[c];
                out = (out & ~m) | (cbits1 & m);
            }
            if (const2) {  // rows beyond a Constant second-axis edge (its value wins in corners)
                const int y = Y0 + r;
                if (y < 0 || y >= L.n2) out = cbits2;
            }
            dst[i] = out;
        }
        __syncthreads();
    }
    // unpack the interior: rows K .. K+TY-1, words 1 .. TXW
    const uint32_t* fin = grid[a.steps & 1];
    for (int w = threadIdx.x; w < TXW * TY; w += blockDim.x) {
        const int r = K + w / TXW, c = 1 + w % TXW;
        const int y = Y0 + r, x = X0 + 32 * c;
        if (y >= L.n2 || x >= L.n1) continue;
        const uint32_t bits = fin[r * RS + c + 1];
        int8_t* q = a.next + (int64_t)y * L.pitch + x;
        if (x + 32 <= L.n1) {
            uint4 lo, hi;
            lo.x = tb_nibble_bytes(bits);
            lo.y = tb_nibble_bytes(bits >> 4);
            lo.z = tb_nibble_bytes(bits >> 8);
            lo.w = tb_nibble_bytes(bits >> 12);
            hi.x = tb_nibble_bytes(bits >> 16);
            hi.y = tb_nibble_bytes(bits >> 20);
            hi.z = tb_nibble_bytes(bits >> 24);
            hi.w = tb_nibble_bytes(bits >> 28);
            reinterpret_cast<uint4*>(q)[0] = lo;
            reinterpret_cast<uint4*>(q)[1] = hi;
        } else {
            for (int i = 0; x + i < L.n1; ++i) q[i] = (int8_t)((bits >> i) & 1u);
        }
    }
}

template <int TXW, int TY, int K>
__global__ __launch_bounds__(256) void k_life_tb(LifeTbArgs a) {
    life_tb_tile<TXW, TY, K>(a, (int)blockIdx.x, (int)blockIdx.y);
}

// ---------------------------------------------------------------- batched lattices
//
// One launch advances several lattices: block b belongs to item block_item[b] and
// is its tile (b - block_base[item]), tiles numbered along the first axis first.
// Each item holds its buffers for parity 0 and its own step count; the launch
// swaps the buffers when parity is set and runs min(K, steps - step0) steps of
// each item (the caller launches only over items with steps > step0, listed
// first).
template <int TXW, int TY, int K>
__global__ __launch_bounds__(256) void k_life_tb_batch(const LifeTbArgs* __restrict__ items,
                                                       const int32_t* __restrict__ block_item,
                                                       const int32_t* __restrict__ block_base, int32_t parity,
                                                       int32_t step0) {
    const int k = __builtin_amdgcn_readfirstlane(block_item[blockIdx.x]);
    LifeTbArgs a = items[k];
    if (parity) {
        const int8_t* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<int8_t*>(c);
    }
    a.steps = min(K, a.steps - step0);
    const int t = (int)blockIdx.x - block_base[k];
    const int ntx = (a.L.n1 + 32 * TXW - 1) / (32 * TXW);
    life_tb_tile<TXW, TY, K>(a, t % ntx, t / ntx);
}
