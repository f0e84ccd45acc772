// Host-side planning of the NUFFT library in plain C++ (no HIP calls), shared by the
// library and by the address/undefined-behaviour-sanitized host test
// (tests/host_asan.cpp): default tile shapes, the radix plan, digit-reversal map and
// twiddles of the pruned FFT passes, and the spread / gather work items and tile
// occupancy masks built from the bin offsets that set_points computes on the device.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace gn {
namespace host {

constexpr int kGatherCellBytesHost = 16; // a staged gather cell: one coil pair, complex64

// {a, b, c, d} with the layout of HIP's int4 (copied to the device as is)
struct Item {
    int x, y, z, w;
};
static_assert(sizeof(Item) == 16, "work items are int4");

// Default spread (output) and gather (home) tile shapes.  Wide windows shrink the
// gather tile, largest axis first, until tile + halo of a coil pair fits 160 KiB.
inline void default_tiles(int dim, int ns, const int64_t *nf, int *tile, int *gtile)
{
    if (dim == 3) {
        tile[0] = 4;
        tile[1] = 4;
        tile[2] = 8;
        gtile[0] = 10;
        gtile[1] = 10;
        gtile[2] = 16;
    } else {
        tile[0] = 8;
        tile[1] = 16;
        tile[2] = 1;
        gtile[0] = 24;
        gtile[1] = 24;
        gtile[2] = 1;
    }
    auto staged = [&]() {
        int64_t r = kGatherCellBytesHost;
        for (int a = 0; a < dim; ++a) r *= std::min<int64_t>(gtile[a] + 2 * ns - 1, nf[a]);
        return r;
    };
    while (staged() > 160 * 1024) {
        int a = 0;
        for (int b = 1; b < dim; ++b)
            if (gtile[b] > gtile[a]) a = b;
        if (gtile[a] == 1) break;
        --gtile[a];
    }
}

// Radix sequence (8, 4, 2, then 3, 5) of a length-L FFT, applied in reverse order of
// extraction; empty when L has another prime factor.
inline std::vector<int> fft_radices(int L)
{
    std::vector<int> r;
    if (L < 1) return r;
    int n = L;
    while (n % 8 == 0 && n > 1) {
        r.push_back(8);
        n /= 8;
    }
    while (n % 4 == 0 && n > 1) {
        r.push_back(4);
        n /= 4;
    }
    while (n % 2 == 0 && n > 1) {
        r.push_back(2);
        n /= 2;
    }
    for (int p : {3, 5})
        while (n % p == 0 && n > 1) {
            r.push_back(p);
            n /= p;
        }
    if (n != 1) return {};
    std::reverse(r.begin(), r.end());
    return r;
}

// Digit-reversal map (pinv[position] = LDS slot) and twiddles e^{-2 pi i k/L}
// (tw[2k], tw[2k+1]) of a length-L transform with radices r.
inline void fft_tables(int L, const std::vector<int> &r, std::vector<int> &pinv, std::vector<float> &tw)
{
    pinv.assign((size_t)L, 0);
    for (int i = 0; i < L; ++i) {
        int rem = i, rev = 0, scale = L;
        for (int rr : r) {
            const int dgt = rem % rr;
            rem /= rr;
            scale /= rr;
            rev += dgt * scale;
        }
        pinv.at((size_t)rev) = i;
    }
    tw.assign(2 * (size_t)L, 0.f);
    const double two_pi = 6.283185307179586476925286766559;
    for (int k = 0; k < L; ++k) {
        const double ang = -two_pi * (double)k / (double)L;
        tw[2 * (size_t)k] = (float)std::cos(ang);
        tw[2 * (size_t)k + 1] = (float)std::sin(ang);
    }
}

struct WorkItems {
    std::vector<Item> spread;  // {tile, pbeg, pend, slot or -1}: tiles with pairs, then empty tiles
    int full = 0;              // spread items of tiles with pairs (the first part)
    std::vector<Item> reduce;  // {tile, slot0, nslots, 0}
    std::vector<Item> gather;  // {gather tile, sbeg, send, 0}
    std::vector<unsigned long long> occ; // per column of spread tiles: leading-axis tile groups with pairs
    int slots = 0;
    int mask_k = 1;
};

// Gather work items {tile, sbeg, send, 0}: each gather tile's samples (bin offsets
// h_home[Kg + 1]) in slices of at most chunk, in tile order.
inline void build_gather_items(int64_t Kg, const int *h_home, int chunk, std::vector<Item> &items)
{
    items.clear();
    for (int64_t k = 0; k < Kg; ++k) {
        const int hb = h_home[k], he = h_home[k + 1];
        for (int s0 = hb; s0 < he; s0 += chunk) items.push_back({(int)k, s0, std::min(he, s0 + chunk), 0});
    }
}

// Work items from the bin offsets: h_pair[K + 1] (pairs of each spread tile) and
// h_home[Kg + 1] (samples of each gather tile).  Every spread tile gets at least one
// item; tiles with more than spread_chunk pairs are split into slices written to
// partial slots and combined by a reduce item.  The items of tiles with pairs come
// first (w.full of them), in tile order, then one item per empty tile, in tile order.
// ntiles: spread tiles per axis (1 for absent axes).
inline void build_work_items(int64_t K, int64_t Kg, const int *ntiles, const int *h_pair, const int *h_home,
                             int spread_chunk, int gather_chunk, WorkItems &w)
{
    w = WorkItems();
    std::vector<Item> empty;
    for (int64_t k = 0; k < K; ++k) {
        const int b = h_pair[k], e = h_pair[k + 1];
        const int n = e - b;
        if (n == 0) {
            empty.push_back({(int)k, b, e, -1});
        } else if (n <= spread_chunk) {
            w.spread.push_back({(int)k, b, e, -1});
        } else {
            const int nch = (n + spread_chunk - 1) / spread_chunk;
            w.reduce.push_back({(int)k, w.slots, nch, 0});
            for (int c = 0; c < nch; ++c)
                w.spread.push_back({(int)k, b + c * spread_chunk, std::min(e, b + (c + 1) * spread_chunk), w.slots + c});
            w.slots += nch;
        }
    }
    w.full = (int)w.spread.size();
    w.spread.insert(w.spread.end(), empty.begin(), empty.end());
    build_gather_items(Kg, h_home, gather_chunk, w.gather);
    // one mask word per column of spread tiles; bit b: leading-axis tiles [b k, (b + 1) k)
    const int64_t ncolt = (int64_t)ntiles[1] * ntiles[2];
    w.mask_k = (ntiles[0] + 63) / 64;
    w.occ.assign((size_t)ncolt, 0ull);
    for (const Item &it : w.spread)
        if (it.z > it.y) {
            const int64_t t0 = it.x / ncolt;
            w.occ.at((size_t)(it.x - t0 * ncolt)) |= 1ull << (t0 / w.mask_k);
        }
}

} // namespace host
} // namespace gn
