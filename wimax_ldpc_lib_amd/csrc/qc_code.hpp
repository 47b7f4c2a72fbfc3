// qc_code.hpp -- compile-time descriptors of the 802.16e QC-LDPC codes.
//
// A code is (rate, z).  Base row l of the rate's base model matrix is decoding
// layer l; its nonzero blocks, in ascending base column, are the row's edge
// slots j = 0..deg(l)-1.  Row r of layer l (expanded row l*z + r) touches
// variable blk(l,j)*z + (r + sh(l,j)) mod z through slot j, in ascending
// variable order -- the layout of the reference's expanded tables
// (lib/matrices/*.cc, written by python_ldpc/ldpc_utils.py:116-153 and
// :244-311).  Lifting follows python_ldpc/ldpc_utils.py:99-110.
#pragma once

#include <stdint.h>
#include <type_traits>
#include <utility>

#include "base_matrices.h"

namespace wl {

constexpr int kNumRates = 6;
constexpr int kNumZ = 19;          // z = 24, 28, ..., 96
constexpr int kMaxLayers = 12;     // rate 1/2
constexpr int kMaxDeg = 20;        // rate 5/6

constexpr int z_of_index(int zi) { return 24 + 4 * zi; }
constexpr int index_of_z(int z) { return (z - 24) / 4; }
constexpr bool valid_z(int z) { return z >= 24 && z <= 96 && (z % 4) == 0; }

constexpr int lift(int rate, int p, int z)
{
    return p <= 0 ? p : (rate == 1 ? p % z : (p * z) / WL_Z0);
}

struct LayerTable {
    int nl = 0;                       // layers (base rows)
    int c = 0;                        // max degree
    int deg[kMaxLayers] = {};
    int blk[kMaxLayers][kMaxDeg] = {};
    int p96[kMaxLayers][kMaxDeg] = {};  // shift at z0 = 96
};

constexpr LayerTable make_layer_table(int rate)
{
    LayerTable t{};
    const wl_base_info& bi = wl_base_table[rate];
    t.nl = bi.mb;
    for (int l = 0; l < bi.mb; l++) {
        int j = 0;
        for (int b = 0; b < WL_NB; b++) {
            int p = bi.p[l][b];
            if (p < 0) continue;
            t.blk[l][j] = b;
            t.p96[l][j] = p;
            j++;
        }
        t.deg[l] = j;
        if (j > t.c) t.c = j;
    }
    return t;
}

template <int RATE>
struct RateTab {
    static constexpr LayerTable t = make_layer_table(RATE);
    static constexpr int NL = t.nl;
    static constexpr int C = t.c;
    static constexpr int edges_per_row_set()   // sum of layer degrees
    {
        int s = 0;
        for (int l = 0; l < t.nl; l++) s += t.deg[l];
        return s;
    }
    static constexpr int EPR = edges_per_row_set();
};

template <int RATE, int Z>
struct Code {
    using R = RateTab<RATE>;
    static constexpr int NL = R::NL;
    static constexpr int C = R::C;
    static constexpr int N = 24 * Z;
    static constexpr int M = NL * Z;
    static constexpr int K = N - M;
    static constexpr int deg(int l) { return R::t.deg[l]; }
    static constexpr int blk(int l, int j) { return R::t.blk[l][j]; }
    static constexpr int sh(int l, int j) { return lift(RATE, R::t.p96[l][j], Z); }
};

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__host__ __device__ __forceinline__ constexpr void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(static_cast<F&&>(f));
    }
}

}  // namespace wl
