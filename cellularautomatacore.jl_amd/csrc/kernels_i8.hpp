// kernels_i8.hpp -- sweep kernels for Int8 rules (Life, Generations, forest fire).
//
// Three kernels:
//   k_generic_i8  any neighbourhood (halo <= 32), any Int8 rule, every gated
//                 scheme, masks.  An LDS tile with the boundary conditions applied
//                 while it is loaded (no halo buffer, no separate halo launch).
//   k_life_swar   Life with Moore(R <= 5), Synchronous, no mask.  Each lane owns
//                 4*VB consecutive cells of a row and walks down a column band,
//                 keeping the vertical (2R+1)-row running count in registers and
//                 forming the horizontal (2R+1)-wide sum with byte-parallel (SWAR)
//                 adds and funnel shifts; the rule is a union of count intervals
//                 tested four cells at a time.
//   k_forest_vn1  the forest fire with VonNeumann(1), Synchronous, no mask: the
//                 same band walk with one Philox4x32-10 draw per cell.
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
};

constexpr int GEN_TX = 64, GEN_TY = 16;

__global__ __launch_bounds__(256) void k_generic_i8(GenArgs a) {
    extern __shared__ int8_t tile[];
    const int W = GEN_TX + 2 * a.h, H = GEN_TY + 2 * a.h;
    const int x0 = blockIdx.x * GEN_TX, y0 = blockIdx.y * GEN_TY;
    for (int i = threadIdx.x; i < W * H; i += blockDim.x) {
        const int lx = i % W, ly = i / W;
        tile[i] = read_mapped<int8_t>(a.cur, a.L, (int8_t)a.cv1, (int8_t)a.cv2, x0 - a.h + lx,
                                      y0 - a.h + ly);
    }
    __syncthreads();
    const int tx = threadIdx.x % GEN_TX;
    for (int ty = threadIdx.x / GEN_TX; ty < GEN_TY; ty += blockDim.x / GEN_TX) {
        const int x = x0 + tx, y = y0 + ty;
        if (x >= a.L.n1 || y >= a.L.n2) continue;
        const int64_t lin = (int64_t)y * a.L.n1 + x;  // unpadded, 0-based
        const uint64_t cellid = (uint64_t)lin + 1;
        const int8_t* cp = tile + (tx + a.h) + (ty + a.h) * W;
        const int c = *cp;
        int out = c;
        const bool open = (a.mask == nullptr || a.mask[lin] != 0) &&
                          gate_open(a.scheme, a.gate, a.sp, a.sp_f32, a.seed, cellid, a.t, x, y,
                                    a.L.two_d);
        if (open) {
            int n1 = 0, n2 = 0;  // neighbours equal to 1, and to 2
            for (int k = 0; k < a.nofs; ++k) {
                const int v = cp[a.ofs[2 * k] + a.ofs[2 * k + 1] * W];
                n1 += (v == 1);
                n2 += (v == 2);
            }
            if (a.rule == CAH_RULE_LIFE) {
                // local_rule(::Life) (src/rules.jl:385-389)
                const uint64_t m = c == 1 ? a.survive : a.birth;
                out = n1 < 64 ? (int)((m >> n1) & 1u) : 0;
            } else if (a.rule == CAH_RULE_GENERATIONS) {
                // local_rule(::Generations) (src/rules.jl:446-456)
                if (c == 0) out = n1 < 64 ? (int)((a.birth >> n1) & 1u) : 0;
                else if (c == 1) out = (n1 < 64 && ((a.survive >> n1) & 1u)) ? 1 : (a.C > 2 ? 2 : 0);
                else out = c + 1 < a.C ? c + 1 : 0;
            } else {
                // the forest-fire LocalRule (README.md:56-60), Int8 encoding
                if (c == 2) out = 0;
                else if (c == 0) out = bernoulli(a.seed, cellid, a.t, 1, a.ffp, a.ffp_f32) ? 1 : 0;
                else if (n2 > 0) out = 2;
                else out = bernoulli(a.seed, cellid, a.t, 2, a.fff, a.fff_f32) ? 2 : 1;
            }
        }
        a.next[(int64_t)y * a.L.pitch + x] = (int8_t)out;
    }
}

// ---------------------------------------------------------------- band walk helpers

template <int I, int N>
struct Unroll {
    template <class F>
    static CAH_DEV void run(F&& f) {
        f(std::integral_constant<int, I>{});
        Unroll<I + 1, N>::run(f);
    }
};
template <int N>
struct Unroll<N, N> {
    template <class F>
    static CAH_DEV void run(F&&) {}
};

constexpr int fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// One row of a band: VB dwords per lane starting at byte x.  Off-lattice bytes
// take the boundary value; a row beyond a Constant axis-2 edge is its value.
template <int VB>
CAH_DEV void load_row(const int8_t* __restrict__ u, const Lat& L, int8_t cv1, int8_t cv2,
                      int32_t x, int32_t y, uint32_t (&X)[VB]) {
    bool o2;
    const int32_t my = map_axis0(L.b2, y, L.n2, o2);
    if (o2) {
        const uint32_t v = 0x01010101u * (uint8_t)cv2;
#pragma unroll
        for (int i = 0; i < VB; ++i) X[i] = v;
        return;
    }
    const int8_t* row = u + (int64_t)my * L.pitch;
    if (x >= 0 && x + 4 * VB <= L.n1) {
        if constexpr (VB == 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(row + x);
            X[0] = v.x; X[1] = v.y; X[2] = v.z; X[3] = v.w;
        } else if constexpr (VB == 2) {
            const uint2 v = *reinterpret_cast<const uint2*>(row + x);
            X[0] = v.x; X[1] = v.y;
        } else {
#pragma unroll
            for (int i = 0; i < VB; ++i) X[i] = reinterpret_cast<const uint32_t*>(row + x)[i];
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t w = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            bool o1;
            const int32_t mx = map_axis0(L.b1, x + 4 * i + b, L.n1, o1);
            const uint8_t v = o1 ? (uint8_t)cv1 : (uint8_t)row[mx];
            w |= (uint32_t)v << (8 * b);
        }
        X[i] = w;
    }
}

template <int VB>
CAH_DEV void store_row(int8_t* __restrict__ u, const Lat& L, int32_t x, int32_t y,
                       const uint32_t (&X)[VB]) {
    int8_t* row = u + (int64_t)y * L.pitch;
    if constexpr (VB == 4) {
        *reinterpret_cast<uint4*>(row + x) = make_uint4(X[0], X[1], X[2], X[3]);
    } else if constexpr (VB == 2) {
        *reinterpret_cast<uint2*>(row + x) = make_uint2(X[0], X[1]);
    } else {
#pragma unroll
        for (int i = 0; i < VB; ++i) reinterpret_cast<uint32_t*>(row + x)[i] = X[i];
    }
}

// ---------------------------------------------------------------- Life, Moore(R <= 5)

constexpr int LIFE_MAX_IV = 8;

struct LifeArgs {
    const int8_t* cur;
    int8_t* next;
    Lat L;
    int32_t cv1, cv2;
    int32_t nbands, rows, nwaves;
    int32_t nb, ns;  // birth / survive interval counts
    // interval [lo, hi] of counts: bytes n with n + (128 - lo) >= 128 and
    // n + (127 - hi) < 128, replicated over the four bytes
    uint32_t b_lo[LIFE_MAX_IV], b_hi[LIFE_MAX_IV], s_lo[LIFE_MAX_IV], s_hi[LIFE_MAX_IV];
};

// Horizontal (2R+1)-wide byte sums of the dwords E (G guard dwords on each side
// of the lane's VB dwords): W[i] = sum over d in [-R, R] of the bytes at offset d.
template <int R, int VB, int G>
CAH_DEV void hsum(const uint32_t (&E)[VB + 2 * G], uint32_t (&W)[VB]) {
    // prefix sums over E (bytes stay below 256: at most (VB+2G)(2R+1) <= 88)
    uint32_t P[VB + 2 * G];
    P[0] = E[0];
#pragma unroll
    for (int j = 1; j < VB + 2 * G; ++j) P[j] = P[j - 1] + E[j];
    auto range = [&](int a, int b) -> uint32_t { return a > 0 ? P[b] - P[a - 1] : P[b]; };
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t w = 0;
#pragma unroll
        for (int rho = 0; rho < 4; ++rho) {
            const int qlo = -fdiv(R + rho, 4);   // ceil((-R - rho) / 4)
            const int qhi = fdiv(R - rho, 4);
            if (qlo > qhi) continue;
            const uint32_t a0 = range(G + i + qlo, G + i + qhi);
            if (rho == 0) {
                w += a0;
            } else {
                const uint32_t a1 = range(G + i + 1 + qlo, G + i + 1 + qhi);
                w += funnel(a1, a0, 8 * rho);
            }
        }
        W[i] = w;
    }
}

// One wave's work: band `wid % nbands`, rows chunk `wid / nbands`.
template <int R, int VB, bool BIN>
CAH_DEV void life_band(const LifeArgs& a, int wid) {
    constexpr int P = 2 * R + 1;
    constexpr int G = (R + 3) / 4;  // guard dwords needed on each side
    static_assert(G <= VB, "a lane must hold the guard dwords of its neighbours");
    constexpr int OW = 62 * 4 * VB;  // output bytes per band
    const int lane = threadIdx.x & 63;
    const int band = wid % a.nbands, chunk = wid / a.nbands;
    const int32_t ys = chunk * a.rows;
    const int32_t ye = min(ys + a.rows, a.L.n2);
    const int32_t x = band * OW - 4 * VB + lane * 4 * VB;
    const bool writer = lane >= 1 && lane <= 62 && x < a.L.n1;
    const int8_t cv1 = (int8_t)a.cv1, cv2 = (int8_t)a.cv2;

    auto ones = [&](uint32_t (&X)[VB]) {
        if constexpr (!BIN) {
#pragma unroll
            for (int i = 0; i < VB; ++i) X[i] = bytes_eq(X[i], 0x01010101u);
        }
    };

    uint32_t ring[P][VB];
    uint32_t V[VB];
#pragma unroll
    for (int i = 0; i < VB; ++i) V[i] = 0;
    Unroll<0, 2 * R>::run([&](auto K) {
        constexpr int k = decltype(K)::value;
        load_row<VB>(a.cur, a.L, cv1, cv2, x, ys - R + k, ring[k]);
        ones(ring[k]);
#pragma unroll
        for (int i = 0; i < VB; ++i) V[i] += ring[k][i];
    });
    uint32_t pre[VB];
    load_row<VB>(a.cur, a.L, cv1, cv2, x, ys + R, pre);

    int32_t y = ys;
    while (y < ye) {
        Unroll<0, P>::run([&](auto PH) {
            constexpr int ph = decltype(PH)::value;
            constexpr int s_new = (ph + 2 * R) % P, s_c = (ph + R) % P, s_old = ph;
            if (y < ye) {
#pragma unroll
                for (int i = 0; i < VB; ++i) ring[s_new][i] = pre[i];
                ones(ring[s_new]);
                load_row<VB>(a.cur, a.L, cv1, cv2, x, y + R + 1, pre);
#pragma unroll
                for (int i = 0; i < VB; ++i) V[i] += ring[s_new][i];
                uint32_t E[VB + 2 * G];
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    E[j] = __shfl_up(V[VB - G + j], 1, 64);
                    E[G + VB + j] = __shfl_down(V[j], 1, 64);
                }
#pragma unroll
                for (int i = 0; i < VB; ++i) E[G + i] = V[i];
                uint32_t Wd[VB];
                hsum<R, VB, G>(E, Wd);
                uint32_t out[VB];
#pragma unroll
                for (int i = 0; i < VB; ++i) {
                    const uint32_t c = ring[s_c][i];
                    const uint32_t n = Wd[i] - c;  // Moore leaves the centre out
                    uint32_t inb = 0, ins = 0;
                    for (int k = 0; k < a.nb; ++k) inb |= (n + a.b_lo[k]) & ~(n + a.b_hi[k]);
                    for (int k = 0; k < a.ns; ++k) ins |= (n + a.s_lo[k]) & ~(n + a.s_hi[k]);
                    // (mask >> n) & 1 with mask = survive for a live cell, birth otherwise
                    const uint32_t sel = c << 7;  // 0x80 in the bytes of live cells
                    out[i] = (((sel & ins) | (~sel & inb)) & 0x80808080u) >> 7;
                }
                if (writer) store_row<VB>(a.next, a.L, x, y, out);
#pragma unroll
                for (int i = 0; i < VB; ++i) V[i] -= ring[s_old][i];
                ++y;
            }
        });
    }
}

template <int R, int VB, bool BIN>
__global__ __launch_bounds__(256) void k_life_swar(LifeArgs a) {
    const int wid = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (wid >= a.nwaves) return;  // whole waves only
    life_band<R, VB, BIN>(a, wid);
}

// ---------------------------------------------------------------- forest fire, VN(1)

struct ForestArgs {
    const int8_t* cur;
    int8_t* next;
    Lat L;
    int32_t cv1, cv2;
    int32_t nbands, rows, nwaves;
    uint64_t seed;
    int64_t t;
    double p, f;
    int32_t p_f32, f_f32;
};

// One forest-fire cell (README.md:56-60): c = state, nf = a VonNeumann neighbour is
// on fire.  Exactly one Philox draw is formed per cell, k = 1 (growth) for an
// empty cell and k = 2 (lightning) otherwise; it decides the cell only where the
// reference rule draws it.
template <bool SMALL>
CAH_DEV uint32_t forest_cell(uint32_t c, uint32_t nf, uint64_t cellid, const ForestArgs& a) {
    const bool empty = c == 0;
    uint32_t c0 = (uint32_t)cellid, c1 = SMALL ? 0u : (uint32_t)(cellid >> 32),
             c2 = (uint32_t)(uint64_t)a.t, c3 = empty ? 1u : 2u;
    philox10(c0, c1, c2, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
    const double thr = empty ? a.p : a.f;
    const int f32 = empty ? a.p_f32 : a.f_f32;
    bool hit;
    if (f32) {
        hit = (float)(c0 >> 8) * (1.0f / 16777216.0f) < (float)thr;
    } else {
        const uint64_t bits = ((uint64_t)c0 << 32) | (uint64_t)c1;
        hit = (double)(bits >> 11) * 0x1.0p-53 < thr;
    }
    if (c == 2) return 0;
    if (empty) return hit ? 1u : 0u;
    if (nf) return 2;
    return hit ? 2u : 1u;
}

template <int VB, bool SMALL>
CAH_DEV void forest_band(const ForestArgs& a, int wid) {
    constexpr int OW = 62 * 4 * VB;
    const int lane = threadIdx.x & 63;
    const int band = wid % a.nbands, chunk = wid / a.nbands;
    const int32_t ys = chunk * a.rows;
    const int32_t ye = min(ys + a.rows, a.L.n2);
    const int32_t x = band * OW - 4 * VB + lane * 4 * VB;
    const bool writer = lane >= 1 && lane <= 62 && x < a.L.n1;
    const int8_t cv1 = (int8_t)a.cv1, cv2 = (int8_t)a.cv2;

    // ring of 3 rows: raw states and their fire flags
    uint32_t raw[3][VB], fire[3][VB];
    load_row<VB>(a.cur, a.L, cv1, cv2, x, ys - 1, raw[0]);
    load_row<VB>(a.cur, a.L, cv1, cv2, x, ys, raw[1]);
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        fire[0][i] = bytes_eq(raw[0][i], 0x02020202u);
        fire[1][i] = bytes_eq(raw[1][i], 0x02020202u);
    }
    uint32_t pre[VB];
    load_row<VB>(a.cur, a.L, cv1, cv2, x, ys + 1, pre);
    int32_t y = ys;
    while (y < ye) {
        Unroll<0, 3>::run([&](auto PH) {
            constexpr int ph = decltype(PH)::value;
            constexpr int s_up = ph, s_c = (ph + 1) % 3, s_dn = (ph + 2) % 3;
            if (y < ye) {
#pragma unroll
                for (int i = 0; i < VB; ++i) {
                    raw[s_dn][i] = pre[i];
                    fire[s_dn][i] = bytes_eq(pre[i], 0x02020202u);
                }
                load_row<VB>(a.cur, a.L, cv1, cv2, x, y + 2, pre);
                uint32_t E[VB + 2];
                E[0] = __shfl_up(fire[s_c][VB - 1], 1, 64);
                E[VB + 1] = __shfl_down(fire[s_c][0], 1, 64);
#pragma unroll
                for (int i = 0; i < VB; ++i) E[1 + i] = fire[s_c][i];
                uint32_t out[VB];
                const uint64_t rowid = (uint64_t)y * (uint64_t)a.L.n1 + (uint64_t)x + 1;
#pragma unroll
                for (int i = 0; i < VB; ++i) {
                    // VonNeumann(1): left, right, up, down
                    const uint32_t nf = funnel(E[1 + i], E[i], 24) | funnel(E[2 + i], E[1 + i], 8) |
                                        fire[s_up][i] | fire[s_dn][i];
                    const uint32_t c4 = raw[s_c][i];
                    uint32_t o = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t c = (c4 >> (8 * b)) & 0xFFu;
                        const uint32_t f = (nf >> (8 * b)) & 0xFFu;
                        o |= forest_cell<SMALL>(c, f, rowid + (uint64_t)(4 * i + b), a) << (8 * b);
                    }
                    out[i] = o;
                }
                if (writer) store_row<VB>(a.next, a.L, x, y, out);
                ++y;
            }
        });
    }
}

template <int VB, bool SMALL>
__global__ __launch_bounds__(256) void k_forest_vn1(ForestArgs a) {
    const int wid = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (wid >= a.nwaves) return;
    forest_band<VB, SMALL>(a, wid);
}

// ---------------------------------------------------------------- batched ensembles
//
// One launch advances every trajectory of a batch by one step.  The waves of all
// trajectories form one flat list: wave w belongs to item wave_item[w] and is
// that item's wave (w - wave_base[item]).  Each item is the trajectory's own
// argument block (buffers for parity 0, lattice, seed, band layout); the step's
// time, parameters and buffer parity come from the launch.

CAH_DEV int wave_id() {
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
}

template <int VB, bool SMALL>
__global__ __launch_bounds__(256) void k_forest_batch(const ForestArgs* __restrict__ items,
                                                      const int32_t* __restrict__ wave_item,
                                                      const int32_t* __restrict__ wave_base,
                                                      int32_t nwaves, int32_t parity, int64_t t,
                                                      double p, double f, int32_t p_f32,
                                                      int32_t f_f32) {
    const int wid = wave_id();
    if (wid >= nwaves) return;
    const int k = __builtin_amdgcn_readfirstlane(wave_item[wid]);
    ForestArgs a = items[k];
    if (parity) {
        const int8_t* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<int8_t*>(c);
    }
    a.t = t;
    a.p = p;
    a.f = f;
    a.p_f32 = p_f32;
    a.f_f32 = f_f32;
    forest_band<VB, SMALL>(a, wid - wave_base[k]);
}

template <int R, int VB, bool BIN>
__global__ __launch_bounds__(256) void k_life_batch(const LifeArgs* __restrict__ items,
                                                    const int32_t* __restrict__ wave_item,
                                                    const int32_t* __restrict__ wave_base,
                                                    int32_t nwaves, int32_t parity) {
    const int wid = wave_id();
    if (wid >= nwaves) return;
    const int k = __builtin_amdgcn_readfirstlane(wave_item[wid]);
    LifeArgs a = items[k];
    if (parity) {
        const int8_t* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<int8_t*>(c);
    }
    life_band<R, VB, BIN>(a, wid - wave_base[k]);
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
