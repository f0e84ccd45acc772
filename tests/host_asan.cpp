// Host planner code of the NUFFT library (mripy_amd/csrc/gn_host_plan.h) under
// AddressSanitizer and UndefinedBehaviorSanitizer: default tiles, FFT radix plans,
// digit-reversal maps and twiddles for every length up to 2048, and the spread /
// gather work items and occupancy masks for bin offsets drawn at many geometries,
// including 3-D (16, 20, 18) with 4000 samples on nf (32, 40, 36).  Every structure
// is also checked for consistency (a permutation, items covering each bin exactly,
// masks matching the occupied tiles).  Built and run by tests/test_host_asan.py;
// prints "OK" and exits 0 on success.
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <set>

#include "gn_host_plan.h"

using namespace gn::host;

#define CHECK(c)                                                                                                       \
    do {                                                                                                               \
        if (!(c)) {                                                                                                    \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c);                                         \
            exit(1);                                                                                                   \
        }                                                                                                              \
    } while (0)

static void check_fft(int L)
{
    std::vector<int> r = fft_radices(L);
    int prod = 1;
    for (int x : r) prod *= x;
    bool smooth = true;
    int n = L;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    smooth = n == 1;
    if (!smooth) {
        CHECK(r.empty());
        return;
    }
    CHECK(prod == L);
    std::vector<int> pinv;
    std::vector<float> tw;
    fft_tables(L, r, pinv, tw);
    std::vector<int> seen((size_t)L, 0);
    for (int i = 0; i < L; ++i) {
        CHECK(pinv[(size_t)i] >= 0 && pinv[(size_t)i] < L);
        seen[(size_t)pinv[(size_t)i]]++;
    }
    for (int i = 0; i < L; ++i) CHECK(seen[(size_t)i] == 1);
    CHECK(tw.size() == 2 * (size_t)L && tw[0] == 1.0f && tw[1] == 0.0f);
}

// bin offsets of n items spread over K bins with random (clustered) counts
static std::vector<int> offsets(int64_t K, int64_t n, std::mt19937_64 &rng, bool clustered)
{
    std::vector<int> cnt((size_t)K, 0);
    std::uniform_int_distribution<int64_t> u(0, K - 1);
    int64_t centre = u(rng);
    for (int64_t i = 0; i < n; ++i) {
        int64_t k = u(rng);
        if (clustered && (i % 3)) k = (centre + (k % 7)) % K; // dense bins: split tiles
        cnt[(size_t)k]++;
    }
    std::vector<int> off((size_t)K + 1, 0);
    for (int64_t k = 0; k < K; ++k) off[(size_t)k + 1] = off[(size_t)k] + cnt[(size_t)k];
    return off;
}

static void check_items(int dim, const int64_t *modes, int ns, int64_t M, int rat, std::mt19937_64 &rng,
                        bool clustered)
{
    int64_t nf[3] = {1, 1, 1};
    for (int a = 0; a < dim; ++a) nf[a] = std::max<int64_t>(rat * modes[a], 2 * ns);
    int tile[3] = {1, 1, 1}, gtile[3] = {1, 1, 1};
    default_tiles(dim, ns, nf, tile, gtile);
    int64_t staged = kGatherCellBytesHost;
    for (int a = 0; a < dim; ++a) staged *= std::min<int64_t>(gtile[a] + 2 * ns - 1, nf[a]);
    CHECK(staged <= 160 * 1024);
    int ntiles[3] = {1, 1, 1}, ngt[3] = {1, 1, 1};
    int64_t K = 1, Kg = 1;
    for (int a = 0; a < dim; ++a) {
        const int t = (int)std::min<int64_t>(tile[a], nf[a]);
        const int gt = (int)std::min<int64_t>(gtile[a], nf[a]);
        ntiles[a] = (int)((nf[a] + t - 1) / t);
        ngt[a] = (int)((nf[a] + gt - 1) / gt);
        K *= ntiles[a];
        Kg *= ngt[a];
    }
    const int64_t P = M * 8; // pairs
    std::vector<int> hp = offsets(K, P, rng, clustered), hh = offsets(Kg, M, rng, clustered);
    WorkItems w;
    build_work_items(K, Kg, ntiles, hp.data(), hh.data(), 512, 512, w);
    // spread items: every tile covered, in order, slices of at most 512 pairs
    std::vector<int64_t> covered((size_t)K, 0);
    std::set<int> slots;
    for (const Item &it : w.spread) {
        CHECK(it.x >= 0 && it.x < K && it.y <= it.z && it.z - it.y <= 512);
        CHECK(it.y >= hp[(size_t)it.x] && it.z <= hp[(size_t)it.x + 1]);
        covered[(size_t)it.x] += it.z - it.y;
        if (it.w >= 0) {
            CHECK(it.w < w.slots);
            CHECK(slots.insert(it.w).second);
        }
    }
    for (int64_t k = 0; k < K; ++k) CHECK(covered[(size_t)k] == hp[(size_t)k + 1] - hp[(size_t)k]);
    // tiles with pairs first, in tile order; then exactly one item per empty tile, in tile order
    CHECK(w.full >= 0 && w.full <= (int)w.spread.size());
    int64_t nempty = 0;
    for (int64_t k = 0; k < K; ++k) nempty += hp[(size_t)k + 1] == hp[(size_t)k];
    CHECK((int64_t)w.spread.size() - w.full == nempty);
    for (size_t i = 0; i < w.spread.size(); ++i) {
        const Item &it = w.spread[i];
        CHECK(((int)i < w.full) == (hp[(size_t)it.x + 1] > hp[(size_t)it.x]));
        if (i > 0 && (int)i != w.full) CHECK(w.spread[i - 1].x <= it.x);
    }
    CHECK((int)slots.size() == w.slots);
    for (const Item &it : w.reduce) CHECK(it.y >= 0 && it.y + it.z <= w.slots && it.z >= 2);
    std::vector<int64_t> gcov((size_t)Kg, 0);
    for (const Item &it : w.gather) {
        CHECK(it.x >= 0 && it.x < Kg && it.z - it.y <= 512 && it.z > it.y);
        gcov[(size_t)it.x] += it.z - it.y;
    }
    for (int64_t k = 0; k < Kg; ++k) CHECK(gcov[(size_t)k] == hh[(size_t)k + 1] - hh[(size_t)k]);
    // occupancy masks: bit set iff a tile of that group has pairs
    const int64_t ncolt = (int64_t)ntiles[1] * ntiles[2];
    CHECK((int64_t)w.occ.size() == ncolt && w.mask_k >= 1 && (ntiles[0] + w.mask_k - 1) / w.mask_k <= 64);
    for (int64_t k = 0; k < K; ++k) {
        const int64_t t0 = k / ncolt, col = k - t0 * ncolt;
        if (hp[(size_t)k + 1] > hp[(size_t)k]) CHECK((w.occ[(size_t)col] >> (t0 / w.mask_k)) & 1ull);
    }
}

int main()
{
    for (int L = 1; L <= 2048; ++L) check_fft(L);
    std::mt19937_64 rng(12345);
    const int64_t g1[3] = {16, 20, 18};
    check_items(3, g1, 3, 4000, 2, rng, false);
    check_items(3, g1, 3, 4000, 2, rng, true);
    const int64_t cases3[][3] = {{128, 128, 128}, {24, 20, 30}, {17, 19, 21}, {8, 8, 8}, {32, 32, 16}, {600, 8, 8}};
    for (auto &m : cases3)
        for (int ns : {2, 3, 5, 6, 8})
            for (int rat : {2, 3}) check_items(3, m, ns, 3000, rat, rng, ns % 2 == 0);
    const int64_t cases2[][3] = {{128, 128, 1}, {7, 9, 1}, {600, 120, 1}, {16, 16, 1}, {320, 320, 1}};
    for (auto &m : cases2)
        for (int ns : {2, 3, 6, 14, 16}) check_items(2, m, ns, 5000, 3, rng, ns % 2 == 1);
    printf("OK\n");
    return 0;
}
