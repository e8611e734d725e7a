// launch_check.hpp -- host-side bounds checks of the kernel launches.
//
// Every device allocation the library makes is recorded as a span (its base and
// size).  Before a launch, a model of the kernel's addressing -- the same index
// arithmetic the kernel runs (map_axis0, row_map, lane_map, band_part and the
// tile geometry), evaluated on the host over every wave, lane, block and row the
// grid covers -- computes the lowest and highest byte each buffer argument can
// be read or written at, and checks that range against the span the pointer
// belongs to, the grid against the work it must cover, and the LDS size against
// the hardware limit.  A failed check refuses the launch with CAH_ERR_DEVICE
// instead of running it.
//
// The checks run (a) on every launch when the environment variable
// CAH_CHECK_LAUNCHES is set to a value other than 0, and (b) for dry states:
// cah_check_launches / cah_check_batch_launches build a state or a batch from
// descriptors with fake device addresses and no HIP call, and run the launch
// sequence of cah_advance / cah_batch_advance through the checks only, so that
// any launch can be validated on a machine without a GPU.
#pragma once
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace lchk {

// ---------------------------------------------------------------- spans

struct Span {
    size_t size;
    const char* what;
};

inline std::mutex& span_mu() {
    static std::mutex m;
    return m;
}
inline std::map<uintptr_t, Span>& spans() {
    static std::map<uintptr_t, Span> s;
    return s;
}

inline void reg(const void* p, size_t n, const char* what) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(span_mu());
    spans()[(uintptr_t)p] = Span{n, what};
}

inline void unreg(const void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(span_mu());
    spans().erase((uintptr_t)p);
}

// Whether the launch checks run on every launch (CAH_CHECK_LAUNCHES).
inline bool enabled() {
    static const bool on = [] {
        const char* e = std::getenv("CAH_CHECK_LAUNCHES");
        return e && *e && std::strcmp(e, "0") != 0;
    }();
    return on;
}

// Fake device addresses for dry states: a bump allocator far above any host
// mapping, 4 KiB apart, never dereferenced.
inline void* dry_alloc(size_t n, const char* what) {
    static std::mutex m;
    static uintptr_t next = (uintptr_t)1 << 46;
    std::lock_guard<std::mutex> lk(m);
    const uintptr_t p = next;
    next += ((n ? n : 1) + 8191) / 4096 * 4096;
    reg((void*)p, n ? n : 1, what);
    return (void*)p;
}

// ---------------------------------------------------------------- one launch

struct Chk {
    std::string kernel;
    std::string err;  // the first failed check
    int64_t checks = 0;

    explicit Chk(std::string k) : kernel(std::move(k)) {}
    bool ok() const { return err.empty(); }

    void need(bool cond, const char* fmt, ...) {
        ++checks;
        if (cond || !err.empty()) return;
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = kernel + ": " + buf;
    }

    // The bytes [lo, hi) past `base` lie inside the one allocation `base` points
    // into (hi > lo; an empty range is not checked).
    void range(const void* base, int64_t lo, int64_t hi, const char* what) {
        if (hi <= lo) return;
        ++checks;
        if (!err.empty()) return;
        if (!base) {
            need(false, "%s: NULL buffer", what);
            return;
        }
        const uintptr_t b = (uintptr_t)base;
        std::lock_guard<std::mutex> lk(span_mu());
        auto& s = spans();
        auto it = s.upper_bound(b);
        if (it == s.begin()) {
            char buf[256];
            snprintf(buf, sizeof buf, "%s: pointer %p is in no allocation", what, base);
            err = kernel + ": " + buf;
            return;
        }
        --it;
        const uintptr_t s0 = it->first, s1 = s0 + it->second.size;
        if (b >= s1 || (int64_t)(b - s0) + lo < 0 || b + (uint64_t)hi > s1) {
            char buf[384];
            snprintf(buf, sizeof buf,
                     "%s: bytes [%lld, %lld) past %p leave the allocation %s [%p, +%zu)", what, (long long)lo,
                     (long long)hi, base, it->second.what, (void*)s0, it->second.size);
            err = kernel + ": " + buf;
        }
    }
};

// The mapped indices of g in [g0, g1] on an axis of n cells (kind: CAH_BOUNDARY_*):
// their least and greatest value, and whether any of them is off a Constant
// edge; false if some index maps outside [0, n).
inline bool axis_span(int32_t kind, int32_t g0, int32_t g1, int32_t n, int32_t& lo, int32_t& hi) {
    lo = INT32_MAX;
    hi = INT32_MIN;
    for (int32_t g = g0; g <= g1; ++g) {
        bool out;
        const int32_t m = map_axis0(kind, g, n, out);
        if (out) continue;
        if (m < 0 || m >= n) return false;
        lo = m < lo ? m : lo;
        hi = m > hi ? m : hi;
    }
    return true;
}

// Reads through read_mapped (device.hpp) of the window [x0, x1] x [y0, y1] of
// an element-size-es lattice at `u`.
inline void mapped_window(Chk& c, const void* u, const Lat& L, int64_t es, int32_t x0, int32_t x1, int32_t y0,
                          int32_t y1, const char* what) {
    int32_t a, b, p, q;
    c.need(axis_span(L.b1, x0, x1, L.n1, a, b), "%s: a column of [%d, %d] maps off the lattice", what, x0, x1);
    c.need(axis_span(L.b2, y0, y1, L.n2, p, q), "%s: a row of [%d, %d] maps off the lattice", what, y0, y1);
    if (c.ok() && a <= b && p <= q) c.range(u, ((int64_t)p * L.pitch + a) * es, ((int64_t)q * L.pitch + b + 1) * es, what);
}

// The dense interior of a pitched lattice (rows [0, n2), columns [0, n1)).
inline void lattice(Chk& c, const void* u, const Lat& L, int64_t es, const char* what) {
    c.need(L.n1 >= 1 && L.n2 >= 1 && L.pitch >= L.n1, "%s: bad geometry %d x %d, pitch %lld", what, L.n1, L.n2,
           (long long)L.pitch);
    c.range(u, 0, ((int64_t)(L.n2 - 1) * L.pitch + L.n1) * es, what);
}

inline void lds(Chk& c, size_t bytes) { c.need(bytes <= 65536, "%zu bytes of LDS (64 KiB at most)", bytes); }

inline void grid_y(Chk& c, unsigned gy) { c.need(gy >= 1 && gy <= 65535, "grid.y = %u", gy); }

// ---------------------------------------------------------------- band walks

// k_life_swar / k_forest_vn1 (and one item of their batched forms): `launched`
// waves run; R = 1 for the forest fire.  The walk of a wave reads rows
// [ys - R, ye + R - 1] (mirrored within that range when it walks up) and writes
// rows [ys, ye) of its band.
template <int VB>
inline void band_walk(Chk& c, const Lat& L, const int8_t* cur, const int8_t* next, int32_t nbands, int32_t rows,
                      int32_t nwaves, int64_t launched, int R, int halo_rows) {
    constexpr int W = 4 * VB, OW = 62 * W;
    c.need(nbands >= 1 && rows >= 1, "nbands %d, rows %d", nbands, rows);
    if (!c.ok()) return;
    const int64_t chunks = (L.n2 + rows - 1) / rows;
    c.need((int64_t)nbands * OW >= L.n1, "%d bands of %d bytes leave columns of %d unwritten", nbands, OW, L.n1);
    c.need((int64_t)nwaves == (int64_t)nbands * chunks, "%d waves for %d bands x %lld row chunks", nwaves, nbands,
           (long long)chunks);
    c.need(launched >= nwaves, "%lld waves launched for %d", (long long)launched, nwaves);
    c.need(L.pitch % W == 0, "pitch %lld is not a multiple of the lane's %d bytes", (long long)L.pitch, W);
    if (!c.ok()) return;
    const bool fast = mapped_ok<VB>(L, halo_rows);
    // columns each lane reads, per band (the same for every chunk)
    int64_t col_lo = INT64_MAX, col_hi = INT64_MIN, wcol_hi = INT64_MIN;
    for (int32_t band = 0; band < nbands && c.ok(); ++band) {
        for (int lane = 0; lane < 64; ++lane) {
            const BandPart b = band_part<VB>(nbands, rows, L, band, lane);
            int64_t lo, hi;
            if (fast) {
                const LaneMap m = lane_map<VB>(L, b.x);
                c.need(m.src >= 0 && m.src + W <= L.n1, "lane map of column %d gives %d", b.x, m.src);
                lo = m.src;
                hi = m.src + W;
            } else if (b.x >= 0 && b.x + W <= L.n1) {
                lo = b.x;
                hi = b.x + W;
            } else {
                int32_t a, z;
                c.need(axis_span(L.b1, b.x, b.x + W - 1, L.n1, a, z), "column %d maps off the lattice", b.x);
                if (a > z) continue;  // all beyond a Constant edge
                lo = a;
                hi = z + 1;
            }
            col_lo = lo < col_lo ? lo : col_lo;
            col_hi = hi > col_hi ? hi : col_hi;
            if (b.writer) {
                c.need(b.x >= 0 && b.x % W == 0, "writer lane at column %d", b.x);
                wcol_hi = b.x + W > wcol_hi ? b.x + W : wcol_hi;
            }
        }
    }
    c.need(wcol_hi <= L.pitch, "a writer lane reaches byte %lld of a %lld-byte row", (long long)wcol_hi,
           (long long)L.pitch);
    // rows, per chunk
    int64_t row_lo = INT64_MAX, row_hi = INT64_MIN;
    for (int64_t ch = 0; ch < chunks && c.ok(); ++ch) {
        const BandPart b = band_part<VB>(nbands, rows, L, (int)(ch * nbands), 0);
        c.need(b.ys < b.ye, "chunk %lld holds no row", (long long)ch);
        const int32_t y0 = b.ys - R, y1 = b.ye + R - 1;
        if (fast) {
            c.need(y0 > -L.n2 && y1 < 2 * L.n2, "rows [%d, %d] beyond row_map's range", y0, y1);
            for (int32_t y = y0; y <= y1 && c.ok(); ++y) {
                bool fill;
                const int32_t my = row_map(L, y, fill);
                if (fill) continue;
                c.need(my >= 0 && my < L.n2, "row %d maps to %d", y, my);
                row_lo = my < row_lo ? my : row_lo;
                row_hi = my > row_hi ? my : row_hi;
            }
        } else {
            int32_t p, q;
            c.need(axis_span(L.b2, y0, y1, L.n2, p, q), "a row of [%d, %d] maps off the lattice", y0, y1);
            if (p <= q) {
                row_lo = p < row_lo ? p : row_lo;
                row_hi = q > row_hi ? q : row_hi;
            }
        }
    }
    if (!c.ok()) return;
    if (fast)  // buffer offsets are 32-bit: mapped_ok asks pitch * n2 < 2^31
        c.need(L.pitch * (int64_t)L.n2 < ((int64_t)1 << 31), "32-bit buffer offsets overflow");
    if (row_lo <= row_hi && col_lo <= col_hi)
        c.range(cur, row_lo * L.pitch + col_lo, row_hi * L.pitch + col_hi, "band walk reads");
    if (wcol_hi > INT64_MIN) c.range(next, 0, (int64_t)(L.n2 - 1) * L.pitch + wcol_hi, "band walk writes");
}

// The mask reads of a gated band walk (mask_row, gate.hpp): the mask is n1 x n2
// bytes without pitch, and each wave reads the rows of its chunk, [ys, ye), at
// each lane's 4*VB bytes: the lane map's source group in a mapped walk (32-bit
// buffer offsets), else the bytes of [x, x + 4*VB) that lie in [0, n1).
template <int VB>
inline void band_mask(Chk& c, const Lat& L, const uint8_t* mask, int32_t nbands, int32_t rows, int halo_rows) {
    constexpr int W = 4 * VB;
    const bool fast = mapped_ok<VB>(L, halo_rows);
    int64_t col_lo = INT64_MAX, col_hi = INT64_MIN;
    for (int32_t band = 0; band < nbands && c.ok(); ++band) {
        for (int lane = 0; lane < 64; ++lane) {
            const BandPart b = band_part<VB>(nbands, rows, L, band, lane);
            int64_t lo, hi;
            if (fast) {
                const LaneMap m = lane_map<VB>(L, b.x);
                c.need(m.src >= 0 && m.src + W <= L.n1 && m.src % W == 0, "mask: lane map of column %d gives %d",
                       b.x, m.src);
                if (b.writer) c.need(m.src == b.x, "mask: writer lane at column %d reads group %d", b.x, m.src);
                lo = m.src;
                hi = m.src + W;
            } else {
                lo = b.x < 0 ? 0 : b.x;
                hi = b.x + W > L.n1 ? L.n1 : b.x + W;
                if (lo >= hi) continue;
            }
            col_lo = lo < col_lo ? lo : col_lo;
            col_hi = hi > col_hi ? hi : col_hi;
        }
    }
    const int64_t chunks = (L.n2 + rows - 1) / rows;
    int64_t row_lo = INT64_MAX, row_hi = INT64_MIN;
    for (int64_t ch = 0; ch < chunks && c.ok(); ++ch) {
        const BandPart b = band_part<VB>(nbands, rows, L, (int)(ch * nbands), 0);
        c.need(b.ys >= 0 && b.ys < b.ye && b.ye <= L.n2, "mask: chunk %lld rows [%d, %d)", (long long)ch, b.ys, b.ye);
        row_lo = b.ys < row_lo ? b.ys : row_lo;
        row_hi = b.ye - 1 > row_hi ? b.ye - 1 : row_hi;
    }
    if (!c.ok() || col_lo > col_hi || row_lo > row_hi) return;
    if (fast) c.need((int64_t)L.n1 * L.n2 < ((int64_t)1 << 31), "mask: 32-bit buffer offsets overflow");
    c.range(mask, row_lo * L.n1 + col_lo, row_hi * L.n1 + col_hi, "band walk mask reads");
}

// The wave (or block) tables of a batched launch: entry w of `map` names item
// map[w], of which it is the (w - base[item])-th; items are listed in order and
// each gets exactly `count(item)` consecutive entries.
template <class Count>
inline void item_table(Chk& c, const int32_t* map, const int32_t* base, int64_t nentries, int64_t nitems,
                       Count count) {
    int64_t w = 0;
    for (int64_t k = 0; k < nitems && c.ok(); ++k) {
        c.need(base[k] == w, "item %lld starts at %d, not %lld", (long long)k, base[k], (long long)w);
        const int64_t n = count(k);
        for (int64_t j = 0; j < n && w < nentries; ++j, ++w)
            if (map[w] != k) c.need(false, "entry %lld names item %d, not %lld", (long long)w, map[w], (long long)k);
    }
    c.need(w == nentries, "%lld table entries for %lld launched", (long long)w, (long long)nentries);
}

// ---------------------------------------------------------------- k_life_tb

// One lattice of k_life_tb<TXW, TY, K> over the tiles of `grid` (gx x gy).
inline void life_tb(Chk& c, const Lat& L, const int8_t* cur, const int8_t* next, int txw, int ty, int K, int steps,
                    int64_t gx, int64_t gy) {
    const int64_t ntx = (L.n1 + 32 * txw - 1) / (32 * txw), nty = (L.n2 + ty - 1) / ty;
    c.need(gx == ntx && gy == nty, "grid %lld x %lld for %lld x %lld tiles", (long long)gx, (long long)gy,
           (long long)ntx, (long long)nty);
    c.need(steps >= 1 && steps <= K, "%d steps in a launch of at most %d", steps, K);
    c.need(L.pitch % 64 == 0, "pitch %lld", (long long)L.pitch);
    const int GW = txw + 2, GH = ty + 2 * K;
    lds(c, (size_t)(2 * GH * (GW + 2) + GW) * 4);
    if (!c.ok()) return;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    auto note = [&](int64_t a, int64_t b) {
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    };
    // rows of every tile row, columns of every tile column (the two are independent)
    int32_t rlo = INT32_MAX, rhi = INT32_MIN;
    for (int64_t by = 0; by < gy && c.ok(); ++by) {
        int32_t p, q;
        const int32_t Y0 = (int32_t)(by * ty - K);
        c.need(axis_span(L.b2, Y0, Y0 + GH - 1, L.n2, p, q), "tile rows from %d map off the lattice", Y0);
        if (p <= q) {
            rlo = p < rlo ? p : rlo;
            rhi = q > rhi ? q : rhi;
        }
    }
    int64_t clo = INT64_MAX, chi = INT64_MIN;
    for (int64_t bx = 0; bx < gx && c.ok(); ++bx) {
        const int32_t X0 = (int32_t)(bx * txw * 32 - 32);
        for (int w = 0; w < GW; ++w) {
            const int32_t x = X0 + 32 * w;
            if (x >= 0 && x + 32 <= L.n1) {
                c.need(x % 16 == 0, "a whole word at column %d is not 16-byte aligned", x);
                clo = x < clo ? x : clo;
                chi = x + 32 > chi ? x + 32 : chi;
            } else {
                int32_t a, z;
                c.need(axis_span(L.b1, x, x + 31, L.n1, a, z), "word at column %d maps off the lattice", x);
                if (a <= z) {
                    clo = a < clo ? a : clo;
                    chi = z + 1 > chi ? z + 1 : chi;
                }
            }
        }
    }
    if (c.ok() && rlo <= rhi && clo <= chi) note(rlo * L.pitch + clo, rhi * L.pitch + chi);
    if (lo <= hi) c.range(cur, lo, hi, "k_life_tb reads");
    // writes: interior words 1..txw of rows K..K+ty-1, cells inside the lattice
    // only (a whole 32-byte word only where x + 32 <= n1)
    c.range(next, 0, (int64_t)(L.n2 - 1) * L.pitch + L.n1, "k_life_tb writes");
}

// ---------------------------------------------------------------- LDS-tile kernels

// A kernel that loads a (TX + 2h) x (TY + 2h) tile through read_mapped for every
// block of a gx x gy grid of TX x TY output tiles, and writes its outputs where
// x < n1 and y < n2.
inline void tile_kernel(Chk& c, const Lat& L, const void* cur, int64_t es, int TX, int TY, int h, int64_t gx,
                        int64_t gy, size_t lds_bytes) {
    c.need(gx * TX >= L.n1 && gy * TY >= L.n2, "grid %lld x %lld of %d x %d tiles does not cover %d x %d",
           (long long)gx, (long long)gy, TX, TY, L.n1, L.n2);
    lds(c, lds_bytes);
    c.need(h >= 0 && h <= 64, "halo %d", h);
    if (!c.ok()) return;
    // the union of all tiles' windows: each tile's window is checked by its corner
    // columns and rows (every index between maps as well, checked by axis_span)
    int32_t clo = INT32_MAX, chi = INT32_MIN, rlo = INT32_MAX, rhi = INT32_MIN;
    for (int64_t bx = 0; bx < gx && c.ok(); ++bx) {
        int32_t a, z;
        const int32_t x0 = (int32_t)(bx * TX - h);
        c.need(axis_span(L.b1, x0, x0 + TX + 2 * h - 1, L.n1, a, z), "tile columns from %d map off", x0);
        if (a <= z) clo = a < clo ? a : clo, chi = z > chi ? z : chi;
    }
    for (int64_t by = 0; by < gy && c.ok(); ++by) {
        int32_t a, z;
        const int32_t y0 = (int32_t)(by * TY - h);
        c.need(axis_span(L.b2, y0, y0 + TY + 2 * h - 1, L.n2, a, z), "tile rows from %d map off", y0);
        if (a <= z) rlo = a < rlo ? a : rlo, rhi = z > rhi ? z : rhi;
    }
    if (c.ok() && clo <= chi && rlo <= rhi)
        c.range(cur, ((int64_t)rlo * L.pitch + clo) * es, ((int64_t)rhi * L.pitch + chi + 1) * es, "tile reads");
}

// k_generic_i8 over a gx x gy grid of 64 x 16 tiles, halo h: each block's tile
// (columns from 64 bx - gen_hp(h), gen_w(h) of them; rows from 16 by - h,
// gen_rows(h) of them) read a dword at a time where it lies inside the lattice,
// else cell by cell through the boundary map; the count's LDS reads (the dword
// after a thread's last neighbour dword included) within the tile.
inline void generic_tile(Chk& c, const Lat& L, const void* cur, int h, int64_t gx, int64_t gy, size_t lds_bytes) {
    const int hp = gen_hp(h), W = gen_w(h), rows = gen_rows(h);
    c.need(h >= 0 && h <= 32, "halo %d", h);
    c.need(gx * GEN_TX >= L.n1 && gy * GEN_TY >= L.n2, "grid %lld x %lld does not cover %d x %d", (long long)gx,
           (long long)gy, L.n1, L.n2);
    c.need(lds_bytes >= (size_t)W * rows && W % 4 == 0, "LDS of %zu bytes for a %d x %d tile", lds_bytes, W, rows);
    lds(c, lds_bytes);
    // the last dword a count reads: cell column hp + 60 + h + 3 within the dword
    // after it, all in the row
    c.need(((hp + 60 + h) / 4 + 2) * 4 <= W, "count reads past the tile row (h = %d)", h);
    c.need(L.pitch % 4 == 0, "pitch %lld", (long long)L.pitch);
    if (!c.ok()) return;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t by = 0; by < gy && c.ok(); ++by)
        for (int64_t bx = 0; bx < gx && c.ok(); ++bx) {
            const int32_t xs = (int32_t)(bx * GEN_TX) - hp, ys = (int32_t)(by * GEN_TY) - h;
            int32_t c0, c1, r0, r1;
            if (xs >= 0 && xs + W <= L.n1 && ys >= 0 && ys + rows <= L.n2) {
                c0 = xs, c1 = xs + W - 1, r0 = ys, r1 = ys + rows - 1;
            } else {
                c.need(axis_span(L.b1, xs, xs + W - 1, L.n1, c0, c1), "tile columns from %d map off", xs);
                c.need(axis_span(L.b2, ys, ys + rows - 1, L.n2, r0, r1), "tile rows from %d map off", ys);
                if (c0 > c1 || r0 > r1) continue;
            }
            const int64_t a = (int64_t)r0 * L.pitch + c0, z = (int64_t)r1 * L.pitch + c1 + 1;
            lo = a < lo ? a : lo;
            hi = z > hi ? z : hi;
        }
    if (c.ok() && lo <= hi) c.range(cur, lo, hi, "tile reads");
}

// One lattice of k_box_i8 (kernels_box.hpp) over a gx x gy grid of 64 x 64
// blocks: the radius within the kernel's limits, the grid covering the lattice,
// the LDS holding the tile, the prefix rows and the sums; the window reads of
// the prefix rows (aligned dwords up to the one after the last byte used) inside
// the LDS with every byte used inside its row; the sum and tile rows the last
// pass reads; the tile's global reads (a dword at a time inside the lattice,
// else through the boundary map) and the writes.
inline void box_tile(Chk& c, const Lat& L, const void* cur, const void* next, bool vn, int r, int64_t gx,
                     int64_t gy, size_t lds_bytes) {
    c.need(r >= 1 && r <= (vn ? BOX_VN_MAX : BOX_MOORE_MAX), "radius %d", r);
    if (!c.ok()) return;
    const int hp = box_hp(r), W = box_w(r), PW = box_pw(r), rows = box_rows(r), planes = box_planes(vn, r);
    c.need(gx * BOX_T >= L.n1 && gy * BOX_T >= L.n2, "grid %lld x %lld does not cover %d x %d", (long long)gx,
           (long long)gy, L.n1, L.n2);
    c.need(lds_bytes >= box_lds(vn, r) && W % 4 == 0 && PW == W + 4, "LDS of %zu bytes for radius %d", lds_bytes, r);
    lds(c, lds_bytes);
    c.need(L.pitch % 4 == 0, "pitch %lld", (long long)L.pitch);
    const int64_t pre0 = (int64_t)rows * W, sum0 = pre0 + (int64_t)rows * PW;
    for (int w = vn ? 0 : r; w <= r && c.ok(); ++w) {
        // outputs x = 0..63 read prefix bytes 3 + hp + x - w and 4 + hp + x + w, four at a time (x = 4k)
        c.need(3 + hp - w >= 0, "prefix read before its row (r = %d, w = %d)", r, w);
        c.need(4 + hp + 60 + w + 3 <= PW - 1, "window past its prefix row (r = %d, w = %d)", r, w);
        const int64_t end = pre0 + (int64_t)(rows - 1) * PW + ((4 + hp + 60 + w) / 4 + 2) * 4;
        c.need(end <= sum0 + (int64_t)planes * rows * BOX_T && end <= (int64_t)lds_bytes,
               "prefix read past the LDS (r = %d, w = %d)", r, w);
    }
    c.need(63 + 2 * r <= rows - 1 && 63 + r <= rows - 1 && hp + 63 <= W - 1, "rows of the last pass (r = %d)", r);
    if (!c.ok()) return;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t by = 0; by < gy && c.ok(); ++by)
        for (int64_t bx = 0; bx < gx && c.ok(); ++bx) {
            const int32_t xs = (int32_t)(bx * BOX_T) - hp, ys = (int32_t)(by * BOX_T) - r;
            int32_t c0, c1, r0, r1;
            if (xs >= 0 && xs + W <= L.n1 && ys >= 0 && ys + rows <= L.n2) {
                c0 = xs, c1 = xs + W - 1, r0 = ys, r1 = ys + rows - 1;
            } else {
                c.need(axis_span(L.b1, xs, xs + W - 1, L.n1, c0, c1), "tile columns from %d map off", xs);
                c.need(axis_span(L.b2, ys, ys + rows - 1, L.n2, r0, r1), "tile rows from %d map off", ys);
                if (c0 > c1 || r0 > r1) continue;
            }
            const int64_t a = (int64_t)r0 * L.pitch + c0, z = (int64_t)r1 * L.pitch + c1 + 1;
            lo = a < lo ? a : lo;
            hi = z > hi ? z : hi;
        }
    if (c.ok() && lo <= hi) c.range(cur, lo, hi, "tile reads");
    lattice(c, next, L, 1, "state written");
}

// ---------------------------------------------------------------- k_lenia_mfma

// One lattice of k_lenia_mfma<mode> over a gx x gy grid of 64 x 64 blocks:
// the row-offset pairs and the tile rows and columns the MFMAs read, the fill's
// cover of the tile (every element the MFMAs read is written first; the growth's
// cells are the fill's first four loads), its float4 chunks (whole, whole and
// mapped, or mapped cell by cell at an edge) and rows, the A fragments and the
// outputs.
// gx x gy: the lattice's 64 x 64 tiles; plane: the bytes of one plane of the A
// fragments the state holds.
inline void lenia_mfma(Chk& c, const LeniaMfmaArgs& a, int mode, int64_t gx, int64_t gy, size_t plane) {
    const Lat& L = a.L;
    const int H = a.H;
    c.need(H >= 1 && H <= LM_HMAX, "H = %d", H);
    c.need(gx * LM_BX >= L.n1 && gy * LM_BY >= L.n2, "grid %lld x %lld does not cover %d x %d", (long long)gx,
           (long long)gy, L.n1, L.n2);
    c.need(L.pitch % 4 == 0, "pitch %lld is not a whole number of float4s", (long long)L.pitch);
    c.need(a.nbx == gx, "nbx %d for %lld tiles a row", a.nbx, (long long)gx);
    if (!c.ok()) return;
    const int rows = LM_BY + 2 * H + 1;
    c.need(rows <= LM_ROWS, "%d tile rows of %d", rows, LM_ROWS);
    // lane rows 16 w + n + H + dy0 + 2 p + (g >> 1) <= 63 + H + dy0 + 2 npairs - 1
    c.need(a.npairs >= 1 && a.dy0 >= -H && a.dy0 + 2 * a.npairs - 1 <= H + 1,
           "row-offset pairs from %d, %d of them, for H = %d", a.dy0, a.npairs, H);
    c.need(63 + H + a.dy0 + 2 * a.npairs - 1 < rows, "MFMA rows beyond the tile");
    c.need(16 * (LM_NC - 1) + 15 < (mode == LM_F16X2 ? lm_stride<LM_F16X2>() : lm_stride<LM_F16>()),
           "MFMA columns beyond the LDS row");
    const int np = mode == LM_F16X2 ? 2 : 1;
    lds(c, mode == LM_F16X2 ? sizeof(LmTile<LM_F16X2>) : sizeof(LmTile<LM_F16>));
    c.need(LM_BY * LM_UST * 4 <= (int)(sizeof(LmTile<LM_F16>)), "potentials past the tile buffer");
    c.need(plane == lenia_mfma_fragment_bytes(a.npairs), "fragment plane of %zu bytes for %d pairs", plane,
           a.npairs);
    c.range(a.afrag, 0, (int64_t)np * (int64_t)plane, "A fragments");
    if (!c.ok()) return;
    // the fill's loads cover the tile once, and loads 0..3 are the growth's cells
    std::vector<uint8_t> hit((size_t)rows * LM_NCH, 0);
    for (int tid = 0; tid < LM_NT && c.ok(); ++tid)
        for (int q = 0; q < LM_FILL_Q; ++q) {
            int ly, j;
            if (!lenia_fill_at(H, tid, q, ly, j)) continue;
            c.need(ly >= 0 && ly < rows && j >= 0 && j < LM_NCH, "fill load (%d, %d) off the tile", ly, j);
            if (!c.ok()) break;
            c.need(hit[(size_t)ly * LM_NCH + j]++ == 0, "tile chunk (%d, %d) loaded twice", ly, j);
            if (q < LM_CELLS_Q) {
                const int i = tid + LM_NT * q;
                c.need(ly == H + i / (LM_BX / 4) && 4 * j == 16 * LM_HC + 4 * (i % (LM_BX / 4)),
                       "fill load %d of thread %d is not its growth's cells", q, tid);
            }
        }
    for (size_t k = 0; k < hit.size() && c.ok(); ++k)
        c.need(hit[k] == 1, "tile chunk (%zu, %zu) never loaded", k / LM_NCH, k % LM_NCH);
    // columns: whole chunks inside, whole chunks mapped (lenia_cols_whole), or cells
    const bool whole = lenia_cols_whole(L), near = lenia_rows_near(L, H);
    const LmMap mx = lm_axis(L.b1, L.n1, 4), my = lm_axis(L.b2, L.n2, 1);
    int64_t clo = INT64_MAX, chi = INT64_MIN;
    auto col_span = [&](int64_t lo, int64_t hi) {  // [lo, hi)
        clo = lo < clo ? lo : clo;
        chi = hi > chi ? hi : chi;
    };
    for (int64_t bx = 0; bx < gx && c.ok(); ++bx) {
        const int xa = (int)bx * LM_BX - 16 * LM_HC;
        for (int j = 0; j < LM_NCH && c.ok(); ++j) {
            const int x = xa + 4 * j;
            if (x >= 0 && x + 4 <= L.n1) {
                c.need(x % 4 == 0, "chunk at column %d not float4-aligned", x);
                col_span(x, x + 4);
            } else if (whole) {
                if (L.b1 == CAH_BOUNDARY_CONSTANT) continue;
                const int xx = x < 0 ? mx.alo + mx.slo * x : mx.ahi + mx.shi * x;
                c.need(xx >= 0 && xx + 4 <= L.n1 && xx % 4 == 0, "chunk at column %d read at %d", x, xx);
                col_span(xx, xx + 4);
            } else {
                col_span(0, 4);  // the chunk's float4 load at the row's start, replaced cell by cell
                for (int k = 0; k < 4; ++k) {
                    bool o1;
                    const int m = map_axis0(L.b1, x + k, L.n1, o1);
                    if (o1) continue;
                    c.need(m >= 0 && m < L.n1, "cell at column %d read at %d", x + k, m);
                    col_span(m, m + 1);
                }
            }
        }
    }
    int32_t rlo = INT32_MAX, rhi = INT32_MIN;
    for (int64_t by = 0; by < gy && c.ok(); ++by)
        for (int ly = 0; ly < rows; ++ly) {
            const int y = (int)by * LM_BY - H + ly;
            int r;
            bool oy;
            if (near) {
                const bool ylo = y < 0, yhi = y >= L.n2;
                r = ylo ? my.alo + my.slo * y : (yhi ? my.ahi + my.shi * y : y);
                oy = L.b2 == CAH_BOUNDARY_CONSTANT && (ylo || yhi);
            } else {
                r = map_axis0(L.b2, y, L.n2, oy);
            }
            if (oy) continue;
            c.need(r >= 0 && r < L.n2, "tile row %d read at %d", y, r);
            rlo = r < rlo ? r : rlo;
            rhi = r > rhi ? r : rhi;
        }
    if (c.ok() && clo <= chi && rlo <= rhi)
        c.range(a.cur, ((int64_t)rlo * L.pitch + clo) * 4, ((int64_t)rhi * L.pitch + chi) * 4, "MFMA tile reads");
    c.need(L.pitch >= 4, "pitch %lld below a float4", (long long)L.pitch);
    c.range(a.cur, 0, 16, "the fill's loads past the tile or off a Constant edge (the lattice's first float4)");
    // the growth writes whole float4s of full blocks, cells of partial ones
    lattice(c, a.next, L, 4, "state written");
    if (a.mask) c.range(a.mask, 0, (int64_t)L.n1 * L.n2, "mask");
}

// Offsets (pairs) whose reach fits a halo of h.
inline void offsets_within(Chk& c, const int32_t* o, int32_t n, int32_t h, const char* what) {
    for (int32_t k = 0; k < 2 * n; ++k)
        if (o[k] < -h || o[k] > h) {
            c.need(false, "%s: offset %d beyond the tile's halo %d", what, o[k], h);
            return;
        }
    c.need(true, "");
}

}  // namespace lchk
