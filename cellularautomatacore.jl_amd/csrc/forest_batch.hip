// forest_batch.hip -- instantiations and launcher of the batched forest-fire
// kernels (one launch per step over many lattices).
#include "launch.hpp"

template <bool SMALL, int PREC, bool TRI>
static void forest_batch_go(int vb, int blocks, const ForestArgs* it, const int32_t* wi, const int32_t* wb,
                            int32_t nw, int32_t parity, int64_t dt, const int64_t* tbase, hipStream_t st) {
    if (vb == 2) k_forest_batch<2, SMALL, PREC, TRI><<<blocks, 256, 0, st>>>(it, wi, wb, nw, parity, dt, tbase);
    else k_forest_batch<1, SMALL, PREC, TRI><<<blocks, 256, 0, st>>>(it, wi, wb, nw, parity, dt, tbase);
}

void launch_forest_batch(int vb, bool small, bool tri, int prec, int blocks, const ForestArgs* it,
                         const int32_t* wi, const int32_t* wb, int32_t nw, int32_t parity, int64_t dt,
                         const int64_t* tbase, hipStream_t st) {
    auto go = [&](auto S, auto T) {
        constexpr bool SM = decltype(S)::value, TR = decltype(T)::value;
        switch (prec) {
        case 0: forest_batch_go<SM, 0, TR>(vb, blocks, it, wi, wb, nw, parity, dt, tbase, st); break;
        case 1: forest_batch_go<SM, 1, TR>(vb, blocks, it, wi, wb, nw, parity, dt, tbase, st); break;
        case 2: forest_batch_go<SM, 2, TR>(vb, blocks, it, wi, wb, nw, parity, dt, tbase, st); break;
        default: forest_batch_go<SM, 3, TR>(vb, blocks, it, wi, wb, nw, parity, dt, tbase, st); break;
        }
    };
    if (small) {
        if (tri) go(std::true_type{}, std::true_type{});
        else go(std::true_type{}, std::false_type{});
    } else {
        if (tri) go(std::false_type{}, std::true_type{});
        else go(std::false_type{}, std::false_type{});
    }
}
