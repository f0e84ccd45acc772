// HIP kernels of the Gaussian-gridding NUFFT for CDNA4 (gfx950).
//
// Data layout in HBM (per plan):
//   grid    : nc oversampled grids, coil-major, C order (nf0, nf1[, nf2]), complex64
//   records : samples sorted by home tile (the tile holding the window start cell)
//             cell = {s0,s1,s2,orig}, frac = {t0,t1,t2,dcf}; bin_start[K+1]
//   pairs   : (output tile, sample) pairs grouped by tile, samples in sorted order
//   items   : spread/gather work items (a tile and a bounded slice of its list)
//   image / samples at the boundary: coil-last (reference) or coil-major.
//
// Window convention (fft/nufft_func.py:293-295, 536-546): xi = x mod 2pi,
// m = 1 + xi // h, taps g = m - ns + j (j = 0..2ns-1) wrapped mod nf, weight
// exp(-(xi - h g)^2/(4 tau)) = exp2(-coef (t + ns - j)^2) with t = xi/h - m.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include <map>
#include <mutex>

#include "gn_internal.h"
#include "gn_host_plan.h"

namespace gn {

namespace {

constexpr double kTwoPi = 6.283185307179586; // 2*np.pi in float64
constexpr int kXCDs = 8;                      // MI355X: 8 XCDs, each with its own L2

// Workgroups are dispatched to the XCDs round-robin by linear block id.  Give each
// XCD runs of kXCDRun consecutive work items (neighbouring tiles: shared halo cells
// and sample values stay in one L2) while the runs still rotate over the XCDs, so a
// dense region of the trajectory is shared out evenly.  Blocks past the last whole
// round of runs keep their order; the map is a bijection on [0, n).
#ifndef GN_XCD_RUN
#define GN_XCD_RUN 32
#endif
constexpr int64_t kXCDRun = GN_XCD_RUN;
// floor(n / d) for 0 <= n < 2^16 (or n < 2^22 with d <= 2^12) given r = 1/d: the
// float estimate is within one of the quotient and one correction step each way
// makes it exact (checked exhaustively over those ranges).
__device__ inline int fdiv_small(int n, int d, float r)
{
    int q = (int)((float)n * r);
    q -= (q * d > n) ? 1 : 0;
    q += ((q + 1) * d <= n) ? 1 : 0;
    return q;
}

constexpr int kStageUnroll = 8; // global loads in flight per thread while staging

// Stage the (R0, R1, R2) region at org (wrapped) of one coil's grid g into LDS:
// kStageUnroll independent loads per thread are issued before any LDS store.
template <int D>
__device__ inline void stage_region(float2 *lds, const float2 *__restrict__ g, const int64_t *nf, const int *org,
                                    int R1, int R2, int RC)
{
    const int R12 = R1 * R2;
    const float rR12 = 1.0f / (float)R12, rR2 = 1.0f / (float)R2;
    for (int base = threadIdx.x; base < RC; base += kStageUnroll * blockDim.x) {
        float2 v[kStageUnroll];
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int rr = base + u * blockDim.x;
            if (rr < RC) {
                const int r0 = fdiv_small(rr, R12, rR12);
                const int rem = rr - r0 * R12;
                const int r1 = D == 3 ? fdiv_small(rem, R2, rR2) : rem;
                const int r2 = D == 3 ? rem - r1 * R2 : 0;
                int q0 = org[0] + r0;
                if (q0 >= nf[0]) q0 -= (int)nf[0];
                int q1 = org[1] + r1;
                if (q1 >= nf[1]) q1 -= (int)nf[1];
                int64_t gi = (int64_t)q0 * nf[1] + q1;
                if (D == 3) {
                    int q2 = org[2] + r2;
                    if (q2 >= nf[2]) q2 -= (int)nf[2];
                    gi = gi * nf[2] + q2;
                }
                v[u] = g[gi];
            }
        }
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int rr = base + u * blockDim.x;
            if (rr < RC) lds[rr] = v[u];
        }
    }
}

__device__ inline uint32_t xcd_swizzle(uint32_t lin, uint32_t n)
{
    // 32-bit: every launch of these kernels has fewer than 2^31 blocks (checked on the host)
    constexpr uint32_t run = (uint32_t)kXCDRun, round = (uint32_t)kXCDs * run;
    if (lin >= (n / round) * round) return lin;
    const uint32_t x = lin % kXCDs, k = lin / kXCDs;
    return ((k / run) * kXCDs + x) * run + k % run;
}

// ---------------------------------------------------------------------------
// set_points: window start cell, offset and home tile of every sample (FP64)
// ---------------------------------------------------------------------------

// CPython float divmod (fmod based), the semantics of numba's x % y and x // y.
__device__ inline double py_divmod(double vx, double wx, double *modp)
{
    double mod = fmod(vx, wx);
    double div = (vx - mod) / wx;
    if (mod != 0.0) {
        if ((wx < 0) != (mod < 0)) {
            mod += wx;
            div -= 1.0;
        }
    } else {
        mod = copysign(0.0, wx);
    }
    double fl;
    if (div != 0.0) {
        fl = floor(div);
        if (div - fl > 0.5) fl += 1.0;
    } else {
        fl = copysign(0.0, vx / wx);
    }
    *modp = mod;
    return fl;
}

struct PrepArgs {
    int64_t M;
    const void *x[3];
    int coord_f64;
    double df;
    int dim;
    int64_t nf[3];
    int ns;
    int tile[3];
    int ntiles[3];
    int4 *cell;
    float4 *frac;
    int *err;
};

__global__ __launch_bounds__(256) void k_prep(PrepArgs a)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.M) return;
    int s[3] = {0, 0, 0};
    float t[3] = {0.f, 0.f, 0.f};
    int key = 0;
    for (int ax = 0; ax < a.dim; ++ax) {
        double x = a.coord_f64 ? ((const double *)a.x[ax])[i] : (double)((const float *)a.x[ax])[i];
        x = x * a.df;
        if (!isfinite(x)) {
            atomicOr(a.err, 1);
            x = 0.0;
        }
        double xi, dummy;
        py_divmod(x, kTwoPi, &xi);                         // xi = x % (2 pi)
        const double h = kTwoPi / (double)a.nf[ax];
        int64_t m = 1 + (int64_t)py_divmod(xi, h, &dummy); // m = 1 + int(xi // h)
        double delta = xi - h * (double)m;
        t[ax] = (float)(delta / h);
        int64_t st = (m - a.ns) % a.nf[ax];
        if (st < 0) st += a.nf[ax];
        s[ax] = (int)st;
        key = key * a.ntiles[ax] + s[ax] / a.tile[ax];
    }
    a.cell[i] = make_int4(s[0], s[1], s[2], key);
    a.frac[i] = make_float4(t[0], t[1], t[2], 0.f);
}

// ---------------------------------------------------------------------------
// Device exclusive scan of ints (local scans, one block over the block sums, add)
// ---------------------------------------------------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// exclusive scan of v over the 256 threads of the block; returns the block total
__device__ inline int block_exclusive_scan(int v, int *lds, int *total)
{
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int add = tid >= off ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    int incl = lds[tid];
    *total = lds[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(int *__restrict__ data, int64_t n, int *__restrict__ bsum)
{
    __shared__ int lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? data[base + k] : 0;
        sum += v[k];
    }
    int total;
    int run = block_exclusive_scan(sum, lds, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) data[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single block: exclusive scan of the block sums; bsum[nb] receives the grand total
__global__ __launch_bounds__(kScanThreads) void k_scan_bsums(int *__restrict__ bsum, int64_t nb)
{
    __shared__ int lds[kScanThreads];
    int carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kScanTile) {
        const int64_t base = b0 + (int64_t)threadIdx.x * kScanItems;
        int v[kScanItems];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (base + k < nb) ? bsum[base + k] : 0;
            sum += v[k];
        }
        int total;
        int run = block_exclusive_scan(sum, lds, &total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < nb) bsum[base + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int *__restrict__ data, int64_t n,
                                                            const int *__restrict__ bsum)
{
    const int add = bsum[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
        if (base + k < n) data[base + k] += add;
}

// ---------------------------------------------------------------------------
// Digit passes of the stable counting sort.  A key in [0, K) is sorted by
// ceil(log2 K / kRadixBits) stable counting sorts on successive kRadixBits-wide
// digits, least significant first.  Each wave owns one chunk of consecutive items:
//   k_radix_count   : per-chunk digit histogram (LDS), written digit-major
//                     cnt[digit * C + chunk]
//   exclusive scan  : cnt -> global start of every (digit, chunk) run; digit-major
//                     order keeps chunks, hence items, in index order
//   k_radix_scatter : the wave walks its chunk 64 items at a time; the lanes holding
//                     the same digit find each other with one ballot per digit bit
//                     (match mask), rank themselves by lane, and the lowest lane of
//                     each group advances that digit's running offset in LDS.
// Every pass is stable, so the result is the stable order by key.
// ---------------------------------------------------------------------------

constexpr int kRadixBits = 10;
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kRadixWaves = 4;

__global__ __launch_bounds__(64 * kRadixWaves) void k_radix_count(int64_t n, const int *__restrict__ keys, int stride,
                                                                  int shift, int bins, int64_t chunk_len, int C,
                                                                  int *__restrict__ cnt)
{
    __shared__ int hist[kRadixWaves][kRadixBins];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int chunk = blockIdx.x * kRadixWaves + wave;
    for (int d = lane; d < bins; d += 64) hist[wave][d] = 0;
    __syncthreads();
    if (chunk < C) {
        const int64_t beg = (int64_t)chunk * chunk_len;
        const int64_t end = min(n, beg + chunk_len);
        for (int64_t i = beg + lane; i < end; i += 64)
            atomicAdd(&hist[wave][(keys[i * stride] >> shift) & (bins - 1)], 1);
    }
    __syncthreads();
    if (chunk < C)
        for (int d = lane; d < bins; d += 64) cnt[(int64_t)d * C + chunk] = hist[wave][d];
}

// vals_in == nullptr: the values are the item indices (first pass)
__global__ __launch_bounds__(64 * kRadixWaves) void k_radix_scatter(int64_t n, const int *__restrict__ keys_in,
                                                                    int stride, const int *__restrict__ vals_in,
                                                                    int shift, int bits, int64_t chunk_len, int C,
                                                                    const int *__restrict__ offs,
                                                                    int *__restrict__ keys_out,
                                                                    int *__restrict__ vals_out)
{
    __shared__ int base[kRadixWaves][kRadixBins];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int chunk = blockIdx.x * kRadixWaves + wave;
    const int bins = 1 << bits;
    if (chunk < C)
        for (int d = lane; d < bins; d += 64) base[wave][d] = offs[(int64_t)d * C + chunk];
    __syncthreads();
    if (chunk >= C) return;
    const int64_t beg = (int64_t)chunk * chunk_len;
    const int64_t end = min(n, beg + chunk_len);
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int64_t g = beg; g < end; g += 64) {
        const int64_t i = g + lane;
        const bool valid = i < end;
        const int key = valid ? keys_in[i * stride] : 0;
        const int val = valid ? (vals_in ? vals_in[i] : (int)i) : 0;
        const int d = (key >> shift) & (bins - 1);
        unsigned long long m = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool on = (d >> b) & 1;
            const unsigned long long bb = __ballot(on);
            m &= on ? bb : ~bb;
        }
        const int rank = __popcll(m & lt);
        const int b0 = base[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) base[wave][d] = b0 + __popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            keys_out[b0 + rank] = key;
            vals_out[b0 + rank] = val;
        }
    }
}

// bin_start[k] = first position of key k in the sorted keys (n where none), k = 0..K
__global__ __launch_bounds__(256) void k_bin_starts_sorted(int64_t n, const int *__restrict__ skeys, int64_t K,
                                                           int *__restrict__ bin_start)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t kprev = i == 0 ? -1 : skeys[i - 1];
    const int64_t knext = i == n ? K : skeys[i];
    for (int64_t k = kprev + 1; k <= knext; ++k) bin_start[k] = (int)i;
}

__global__ __launch_bounds__(256) void k_make_records(int64_t M, const int *__restrict__ sorted,
                                                      const int4 *__restrict__ cell, const float4 *__restrict__ frac,
                                                      const float *__restrict__ dcf, int4 *__restrict__ rec_cell,
                                                      float4 *__restrict__ rec_frac)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M) return;
    const int i = sorted[p];
    const int4 c = cell[i];
    const float4 f = frac[i];
    rec_cell[p] = make_int4(c.x, c.y, c.z, i);
    rec_frac[p] = make_float4(f.x, f.y, f.z, dcf ? dcf[i] : 1.0f);
}

// ---------------------------------------------------------------------------
// (tile, sample) pair records for the output-stationary spread
// ---------------------------------------------------------------------------

struct TileGeom {
    int dim;
    int64_t nf[3];
    int tile[3];
    int ntiles[3];
    int W; // window taps per axis (2 nspread)
};

// Tiles touched by the window [s, s+W) mod nf along one axis, in tap order: the
// consecutive tiles q0, q0 + 1, ... (mod the tile count), n of them.  Tiles are
// chosen so that a window never re-enters a tile it left (T + W - 1 <= nf or T = nf).
__device__ inline void axis_tile_span(int s, int W, int nf, int T, int &q0, int &n)
{
    if (T == nf) {
        q0 = 0;
        n = 1;
        return;
    }
    q0 = s / T;
    const int e = s + W - 1; // last cell, unwrapped
    n = e < nf ? e / T - q0 + 1 : ((nf + T - 1) / T - q0) + (e - nf) / T + 1;
}

// window taps [jlo, jlo+cnt) of a window starting at s that fall in tile q, and the
// tile-local cell of tap jlo
__device__ inline void tap_box(int s, int q, int T, int nf, int W, int &jlo, int &cnt, int &lstart)
{
    const int org = q * T;
    const int teff = min(T, nf - org);
    if (T == nf) {
        jlo = 0;
        cnt = W;
        lstart = s;
        return;
    }
    int v = org - s;
    if (v < 0) v += nf;
    if (v < W) {
        jlo = v;
        cnt = min(W - v, teff);
        lstart = 0;
    } else {
        jlo = 0;
        cnt = min(W, teff - (s - org));
        lstart = s - org;
    }
}

__global__ __launch_bounds__(256) void k_pair_count(int64_t M, const int4 *__restrict__ rec_cell, TileGeom g,
                                                    int *__restrict__ cnt)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M) return;
    const int4 rc = rec_cell[p];
    const int s[3] = {rc.x, rc.y, rc.z};
    int n = 1;
    for (int ax = 0; ax < g.dim; ++ax) {
        int q0, na;
        axis_tile_span(s[ax], g.W, (int)g.nf[ax], g.tile[ax], q0, na);
        n *= na;
    }
    cnt[p] = n;
}

// The pairs of sample record p are offs[p] .. offs[p] + n - 1 with n = the product of
// its per-axis tile spans, enumerated leading axis outermost: local index
// (i * n1 + k) * n2 + m is tile (q0[0] + i, q0[1] + k, q0[2] + m), each wrapped.
// k_pair_keys writes each pair's spread tile (the sort key) and its record position;
// after the stable sort by tile, k_pair_records rebuilds every record in sorted order
// from its sample record, so only 8 bytes per pair are written in sample order.
__global__ __launch_bounds__(256) void k_pair_keys(int64_t M, const int4 *__restrict__ rec_cell, TileGeom g,
                                                   const int *__restrict__ offs, int *__restrict__ pkey,
                                                   int *__restrict__ psamp)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M) return;
    const int4 rc = rec_cell[p];
    int q0[3] = {0, 0, 0}, n[3] = {1, 1, 1};
    axis_tile_span(rc.x, g.W, (int)g.nf[0], g.tile[0], q0[0], n[0]);
    axis_tile_span(rc.y, g.W, (int)g.nf[1], g.tile[1], q0[1], n[1]);
    if (g.dim == 3) axis_tile_span(rc.z, g.W, (int)g.nf[2], g.tile[2], q0[2], n[2]);
    int o = offs[p];
    int t0 = q0[0];
    for (int i = 0; i < n[0]; ++i) {
        int t1 = q0[1];
        for (int k = 0; k < n[1]; ++k) {
            int t2 = q0[2];
            for (int m = 0; m < n[2]; ++m) {
                int key = t0 * g.ntiles[1] + t1;
                if (g.dim == 3) key = key * g.ntiles[2] + t2;
                pkey[o] = key;
                psamp[o] = (int)p;
                ++o;
                if (++t2 == g.ntiles[2]) t2 = 0;
            }
            if (++t1 == g.ntiles[1]) t1 = 0;
        }
        if (++t0 == g.ntiles[0]) t0 = 0;
    }
}

// Pair q of the tile-sorted order: original pair index o = sorted[q], its sample record
// p = psamp[o] and local index o - offs[p] give the tile on each axis; the record is
//   a = {orig, j0 | j1<<8 | j2<<16, c0 | c1<<8 | c2<<16, l0 | l1<<10 | l2<<20}
//   b = {t0, t1, t2, dcf} (float bits)
// (j = first window tap inside the tile, c = taps inside, l = tile-local cell of tap j).
__global__ __launch_bounds__(256) void k_pair_records(int64_t P, const int *__restrict__ sorted,
                                                      const int *__restrict__ psamp, const int *__restrict__ offs,
                                                      const int4 *__restrict__ rec_cell,
                                                      const float4 *__restrict__ rec_frac, TileGeom g,
                                                      int4 *__restrict__ pa, int4 *__restrict__ pb)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const int o = sorted[q];
    const int p = psamp[o];
    const int4 rc = rec_cell[p];
    const float4 rf = rec_frac[p];
    int q0[3] = {0, 0, 0}, n[3] = {1, 1, 1};
    axis_tile_span(rc.x, g.W, (int)g.nf[0], g.tile[0], q0[0], n[0]);
    axis_tile_span(rc.y, g.W, (int)g.nf[1], g.tile[1], q0[1], n[1]);
    if (g.dim == 3) axis_tile_span(rc.z, g.W, (int)g.nf[2], g.tile[2], q0[2], n[2]);
    const int local = o - offs[p];
    const int m = local % n[2], r = local / n[2];
    const int k = r % n[1], i = r / n[1];
    int t0 = q0[0] + i, t1 = q0[1] + k, t2 = q0[2] + m;
    if (t0 >= g.ntiles[0]) t0 -= g.ntiles[0];
    if (t1 >= g.ntiles[1]) t1 -= g.ntiles[1];
    if (t2 >= g.ntiles[2]) t2 -= g.ntiles[2];
    int j0, c0, l0, j1, c1, l1, j2 = 0, c2 = 1, l2 = 0;
    tap_box(rc.x, t0, g.tile[0], (int)g.nf[0], g.W, j0, c0, l0);
    tap_box(rc.y, t1, g.tile[1], (int)g.nf[1], g.W, j1, c1, l1);
    if (g.dim == 3) tap_box(rc.z, t2, g.tile[2], (int)g.nf[2], g.W, j2, c2, l2);
    pa[q] = make_int4(rc.w, j0 | (j1 << 8) | (j2 << 16), c0 | (c1 << 8) | (c2 << 16), l0 | (l1 << 10) | (l2 << 20));
    pb[q] = make_int4(__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z), __float_as_int(rf.w));
}

// Spread work items, built on the device from the tile offsets pstart[K + 1] (the
// rule of the host planner's build_work_items): tile k with n pairs gets one item
// {k, pbeg, pend, -1}, or, when n > chunk, ceil(n / chunk) slices {k, ..., slot} that
// write consecutive partial slots and one reduce item {k, slot0, slices, 0}.  The
// items of tiles with pairs come first, in tile order, then one item per empty tile,
// in tile order (a launch that need not write empty tiles takes only the first
// part).  Items, empty-tile items, slots and reduce items are numbered by exclusive
// scans of the per-tile counts that k_item_counts writes.
__global__ __launch_bounds__(256) void k_item_counts(int64_t K, const int *__restrict__ pstart, int chunk,
                                                     int *__restrict__ nitems, int *__restrict__ nempty,
                                                     int *__restrict__ nslots, int *__restrict__ nsplit)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int n = pstart[k + 1] - pstart[k];
    const int nch = (n + chunk - 1) / chunk;
    const bool split = n > chunk;
    nitems[k] = split ? nch : n > 0 ? 1 : 0;
    nempty[k] = n == 0 ? 1 : 0;
    nslots[k] = split ? nch : 0;
    nsplit[k] = split ? 1 : 0;
}

// nfull: the items of tiles with pairs (grand total of the ioff scan)
__global__ __launch_bounds__(256) void k_item_write(int64_t K, const int *__restrict__ pstart, int chunk,
                                                    const int *__restrict__ ioff, const int *__restrict__ eoff,
                                                    const int *__restrict__ nfull, const int *__restrict__ soff,
                                                    const int *__restrict__ roff, int4 *__restrict__ items,
                                                    int4 *__restrict__ reduce)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int b = pstart[k], e = pstart[k + 1];
    if (e == b) {
        items[*nfull + eoff[k]] = make_int4((int)k, b, e, -1);
        return;
    }
    const int o = ioff[k];
    if (e - b <= chunk) {
        items[o] = make_int4((int)k, b, e, -1);
        return;
    }
    const int nch = (e - b + chunk - 1) / chunk, s0 = soff[k];
    reduce[roff[k]] = make_int4((int)k, s0, nch, 0);
    for (int c = 0; c < nch; ++c) items[o + c] = make_int4((int)k, b + c * chunk, min(e, b + (c + 1) * chunk), s0 + c);
}

// Occupancy row masks: one word per column of spread tiles (ncolt = ntiles[1] *
// ntiles[2] columns), bit b set when one of the leading-axis tiles [b mask_k,
// (b + 1) mask_k) of the column holds pairs; one thread per column, no atomics.
__global__ __launch_bounds__(256) void k_occ_masks(int64_t ncolt, int nt0, int mask_k, const int *__restrict__ pstart,
                                                   unsigned long long *__restrict__ occ)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncolt) return;
    unsigned long long m = 0ull;
    for (int t0 = 0; t0 < nt0; ++t0) {
        const int64_t k = t0 * ncolt + c;
        if (pstart[k + 1] > pstart[k]) m |= 1ull << (t0 / mask_k);
    }
    occ[c] = m;
}

// ---------------------------------------------------------------------------
// Adjoint spread: output-stationary tiles accumulated in registers.
//
// A work item is a tile and a slice of at most kSpreadChunk of the (tile,
// sample) pair records whose windows touch it.  The workgroup splits the tile
// into blocks of 64 cells (4x4x4 in 3-D, 8x8 in 2-D), one block per wave and
// one cell per lane; every lane keeps its cell's value for up to 8 coils in
// registers.  Pair records are staged in LDS 64 at a time: the sample's coil
// values (times dcf) and, per axis, a table of the separable kernel weight at
// every tile-local cell, zero outside the window.  A wave then visits the
// staged pairs whose windows reach its block; each lane forms its weight from
// three table reads and accumulates all coils with FMAs.  Nothing is written
// to shared memory by more than one wave, there are no atomics, and every
// lane adds its contributions in the fixed pair order: the result is
// deterministic.  After each batch the plain FP32 batch sums (at most 64
// terms) are folded into Kahan-compensated (sum, compensation) registers.
// The tile is then written to HBM once, coalesced through LDS; slices of a
// split tile go to partial buffers combined by k_spread_reduce.
// ---------------------------------------------------------------------------



#ifndef GN_SPREAD_BATCH
#define GN_SPREAD_BATCH 64
#endif
constexpr int kSpreadBatch = GN_SPREAD_BATCH; // pairs staged in LDS per step (at most one per lane)
constexpr int kSpreadWavesMax = 8;

// Spread tile t has no pairs and no later pass reads it: its cells need not be written.
__device__ inline bool tile_skippable(const SpreadArgs &a, const int4 item)
{
    if (item.y < item.z || !a.occ) return false;
    const uint32_t nt12 = (uint32_t)a.ntiles[1] * (uint32_t)a.ntiles[2];
    const uint32_t t0 = (uint32_t)item.x / nt12, rest = (uint32_t)item.x - t0 * nt12;
    return ((a.occ[rest] >> (t0 / (uint32_t)a.occ_k)) & 1ull) == 0;
}

template <int D, int NCG>
__device__ __forceinline__ void spread_body(const SpreadArgs &a, const int4 item, const int cgi, const int bz)
{
    extern __shared__ float lds_sp[];
    // LDS: weight tables [batch][TT] | coil values [batch][NCG] float2 | masks [batch][3]
    float *s_tab = lds_sp;
    float2 *s_val = (float2 *)(lds_sp + kSpreadBatch * a.TT);
    int *s_msk = (int *)(s_val + kSpreadBatch * NCG);
    int4 *s_pa = (int4 *)(s_msk + kSpreadBatch * 4);
    float4 *s_pb = (float4 *)(s_pa + kSpreadBatch);

    if (tile_skippable(a, item)) return; // uniform over the workgroup
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int nwaves = blockDim.x >> 6;
    const int cg0 = cgi * NCG;
    const int ncg = (int)min((int64_t)NCG, a.nc - cg0);
    const int T0 = a.tile[0], T1 = a.tile[1], T2 = (D == 3) ? a.tile[2] : 1;

    // this wave's block (block index = blockIdx.z * waves + wave) and this lane's cell
    const int nb = a.nblk[0] * a.nblk[1] * (D == 3 ? a.nblk[2] : 1);
    const int bidx = bz * nwaves + wave;
    const bool has_block = bidx < nb;
    int bq0, bq1, bq2 = 0;
    if (D == 3) {
        bq2 = bidx % a.nblk[2];
        bq1 = (bidx / a.nblk[2]) % a.nblk[1];
        bq0 = bidx / (a.nblk[2] * a.nblk[1]);
    } else {
        bq1 = bidx % a.nblk[1];
        bq0 = bidx / a.nblk[1];
    }
    int L0, L1, L2 = 0;
    if (D == 3) {
        L2 = bq2 * a.blk[2] + (lane % a.blk[2]);
        L1 = bq1 * a.blk[1] + (lane / a.blk[2]) % a.blk[1];
        L0 = bq0 * a.blk[0] + lane / (a.blk[2] * a.blk[1]);
    } else {
        L1 = bq1 * a.blk[1] + (lane % a.blk[1]);
        L0 = bq0 * a.blk[0] + lane / a.blk[1];
    }
    const bool in_tile = has_block && L0 < T0 && L1 < T1 && L2 < T2;
    // lanes of a partial block read a valid table entry; their result is never stored
    L0 = min(L0, T0 - 1);
    L1 = min(L1, T1 - 1);
    L2 = min(L2, T2 - 1);
    // bit masks of the block's cells per axis (tile extents <= 32)
    uint32_t bm0 = 0, bm1 = 0, bm2 = 1;
    if (has_block) {
        for (int i = 0; i < a.blk[0]; ++i) bm0 |= (bq0 * a.blk[0] + i < T0) ? (1u << (bq0 * a.blk[0] + i)) : 0u;
        for (int i = 0; i < a.blk[1]; ++i) bm1 |= (bq1 * a.blk[1] + i < T1) ? (1u << (bq1 * a.blk[1] + i)) : 0u;
        if (D == 3) {
            bm2 = 0;
            for (int i = 0; i < a.blk[2]; ++i) bm2 |= (bq2 * a.blk[2] + i < T2) ? (1u << (bq2 * a.blk[2] + i)) : 0u;
        }
    }
    const int to1 = T0, to2 = T0 + T1; // table offsets of axes 1 and 2

    float sr[NCG], si[NCG], kr[NCG], ki[NCG];
#pragma unroll
    for (int c = 0; c < NCG; ++c) sr[c] = si[c] = kr[c] = ki[c] = 0.f;

    for (int base = item.y; base < item.z; base += kSpreadBatch) {
        const int n = min(kSpreadBatch, item.z - base);
        __syncthreads(); // previous batch fully consumed
        // phase 1: one thread per pair loads its record and coil values, builds the masks
        if (tid < n) {
            const int q = base + tid;
            const int4 pa = a.pair_a[q];
            const int4 pbi = a.pair_b[q];
            const float4 pb = make_float4(__int_as_float(pbi.x), __int_as_float(pbi.y), __int_as_float(pbi.z),
                                          __int_as_float(pbi.w));
            s_pa[tid] = pa;
            s_pb[tid] = pb;
            const float scale = a.use_dcf ? pb.w : 1.0f;
#pragma unroll
            for (int c = 0; c < NCG; ++c) {
                float2 v = make_float2(0.f, 0.f);
                if (c < ncg) {
                    const int64_t coil = cg0 + c;
                    v = (a.layout == GN_COIL_LAST) ? a.in[(int64_t)pa.x * a.nc + coil] : a.in[coil * a.M + pa.x];
                }
                s_val[tid * NCG + c] = make_float2(v.x * scale, v.y * scale);
            }
            // valid tile-local cells of the window per axis: (L - l) mod T < c
            const int Ts[3] = {T0, T1, T2};
            for (int ax = 0; ax < D; ++ax) {
                const int l = (ax == 0 ? pa.w : ax == 1 ? (pa.w >> 10) : (pa.w >> 20)) & 0x3ff;
                const int c = (pa.z >> (8 * ax)) & 0xff;
                uint32_t m = 0;
                for (int r = 0; r < c; ++r) {
                    int L = l + r;
                    if (L >= Ts[ax]) L -= Ts[ax];
                    m |= 1u << L;
                }
                s_msk[tid * 4 + ax] = (int)m;
            }
        }
        __syncthreads();
        // phase 2: weight tables, one entry per (pair, axis, tile-local cell)
        for (int e = tid; e < n * a.TT; e += blockDim.x) {
            const int k = e / a.TT;
            const int i = e - k * a.TT;
            const int ax = i < to1 ? 0 : (i < to2 ? 1 : 2);
            const int L = i - (ax == 0 ? 0 : (ax == 1 ? to1 : to2));
            const int4 pa = s_pa[k];
            const float4 pb = s_pb[k];
            const int Tax = ax == 0 ? T0 : (ax == 1 ? T1 : T2);
            const int l = (ax == 0 ? pa.w : ax == 1 ? (pa.w >> 10) : (pa.w >> 20)) & 0x3ff;
            const int c = (pa.z >> (8 * ax)) & 0xff;
            const int j = (pa.y >> (8 * ax)) & 0xff;
            int rel = L - l;
            if (rel < 0) rel += Tax;
            float w = 0.f;
            if (rel < c) {
                const float t = ax == 0 ? pb.x : (ax == 1 ? pb.y : pb.z);
                const float u = t + (float)(a.ns - j - rel);
                w = exp2f(-a.coef[ax] * u * u);
            }
            s_tab[e] = w;
        }
        __syncthreads();
        // phase 3: each wave visits the pairs that reach its block
        bool hit = false;
        if (lane < n && has_block) {
            const int m0 = s_msk[lane * 4], m1 = s_msk[lane * 4 + 1];
            const int m2 = D == 3 ? s_msk[lane * 4 + 2] : 1;
            hit = ((uint32_t)m0 & bm0) && ((uint32_t)m1 & bm1) && ((uint32_t)m2 & bm2);
        }
        unsigned long long todo = __ballot(hit);
        float ar[NCG], ai[NCG];
#pragma unroll
        for (int c = 0; c < NCG; ++c) ar[c] = ai[c] = 0.f;
        // visits in groups of up to 4 so their LDS reads are in flight together; a
        // missing visit reuses pair 0 with weight 0 (adds exact zeros)
        while (todo) {
            int kk[4];
            float wz[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool some = todo != 0;
                kk[u] = some ? __ffsll((long long)todo) - 1 : 0;
                wz[u] = some ? 1.f : 0.f;
                todo &= todo - 1;
            }
            float w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float *tab = s_tab + kk[u] * a.TT;
                float ww = tab[L0] * tab[to1 + L1];
                if (D == 3) ww *= tab[to2 + L2];
                w[u] = ww * wz[u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float2 *v = s_val + kk[u] * NCG;
#pragma unroll
                for (int c = 0; c < NCG; ++c) {
                    const float2 vc = v[c];
                    ar[c] = fmaf(w[u], vc.x, ar[c]);
                    ai[c] = fmaf(w[u], vc.y, ai[c]);
                }
            }
        }
        // Kahan fold of this batch's partial sums
#pragma unroll
        for (int c = 0; c < NCG; ++c) {
            const float yr = ar[c] - kr[c], yi = ai[c] - ki[c];
            const float tr = sr[c] + yr, ti = si[c] + yi;
            kr[c] = (tr - sr[c]) - yr;
            ki[c] = (ti - si[c]) - yi;
            sr[c] = tr;
            si[c] = ti;
        }
    }

    // write the tile once, straight from registers
    if (!in_tile) return;
    const int TC = a.TC;
    const int cidx = (L0 * T1 + L1) * T2 + L2;
    if (item.w >= 0) {
        float2 *dst = a.partial + ((int64_t)item.w * a.nc + cg0) * TC + cidx;
#pragma unroll
        for (int c = 0; c < NCG; ++c)
            if (c < ncg) dst[(int64_t)c * TC] = make_float2(sr[c] - kr[c], si[c] - ki[c]);
        return;
    }
    int org[3] = {0, 0, 0}, teff[3] = {1, 1, 1};
    {
        uint32_t r = (uint32_t)item.x; // tile index < 2^31
        for (int ax = D - 1; ax >= 0; --ax) {
            const int qq = (int)(r % (uint32_t)a.ntiles[ax]);
            r /= (uint32_t)a.ntiles[ax];
            org[ax] = qq * a.tile[ax];
            teff[ax] = (int)min((int64_t)a.tile[ax], a.nf[ax] - org[ax]);
        }
    }
    if (L0 >= teff[0] || L1 >= teff[1] || (D == 3 && L2 >= teff[2])) return;
    int64_t gi = (int64_t)(org[0] + L0) * a.nf[1] + (org[1] + L1);
    if (D == 3) gi = gi * a.nf[2] + (org[2] + L2);
#pragma unroll
    for (int c = 0; c < NCG; ++c)
        if (c < ncg) a.grid[(int64_t)(cg0 + c) * a.G + gi] = make_float2(sr[c] - kr[c], si[c] - ki[c]);
}

template <int D, int NCG>
__global__ __launch_bounds__(64 * kSpreadWavesMax) void k_spread(SpreadArgs a)
{
    const uint32_t n = gridDim.x * gridDim.y * gridDim.z;
    const uint32_t lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const uint32_t w = xcd_swizzle(lin, n);
    const int bx = (int)(w % gridDim.x);
    const uint32_t yz = w / gridDim.x;
    spread_body<D, NCG>(a, a.items[bx], (int)(yz % gridDim.y), (int)(yz / gridDim.y));
}

// Batched launch over many plans: bitems[b] = {instance, work item, coil group, 0},
// taken in the XCD-aware order of the plain launch.
template <int D, int NCG>
__global__ __launch_bounds__(64 * kSpreadWavesMax) void k_spread_batch(const SpreadArgs *__restrict__ args,
                                                                        const int4 *__restrict__ bitems)
{
    const uint32_t n = gridDim.x * gridDim.y;
    const uint32_t w = xcd_swizzle(blockIdx.x + gridDim.x * blockIdx.y, n);
    const int4 bi = bitems[w % gridDim.x];
    const SpreadArgs &a = args[bi.x]; // read in place: uniform scalar loads
    spread_body<D, NCG>(a, a.items[bi.y], bi.z, (int)(w / gridDim.x));
}

// ---------------------------------------------------------------------------
// Adjoint spread on the matrix cores (v_mfma_f32_16x16x4_f32).
//
// Same work items, pair staging and weight tables as spread_body; the per-wave
// visit loop becomes small dense products.  A wave's 64-cell block (4x4x4 in 3-D,
// 8x8 in 2-D) is split into four slices of 16 cells (in 3-D the 4x4 x-y cells of
// each of the block's 4 z planes; in 2-D pairs of x rows of 8 cells) and
// accumulates
//     D[cell i][real j] += sum over 4 listed pairs k of  w_k(cell i) * v_k[j]
// with one MFMA per slice and 4 pairs: rows are the slice's cells, columns the
// 2*NCG reals (re, im per coil; zero columns beyond), K the pairs.  Per batch the
// wave lists the staged pairs that reach its block (padded to a multiple of 4 with
// an all-zero pair); each step a lane forms the weight factor its row shares across
// the slices once (3-D: the x-y weight; 2-D: the y weight) and multiplies it by the
// per-slice factor.  The FMAs run on the matrix pipe.  The products are summed in a fixed order,
// so the result is deterministic; batch sums are folded into Kahan-compensated
// registers as in spread_body.  Lane layout of the 16x16x4 f32 operation:
// A[i][k] and B[k][j] in lane i + 16k and j + 16k, D[4g + v][j] in register v of
// lane j + 16g.  Each lane writes its cells of the finished tile straight from its
// registers (16-byte stores of two cells along the last axis).  Needs block-aligned
// tiles (multiples of 4 per axis in 3-D, of 8 in 2-D) of at most kSpreadWavesMax
// blocks: one workgroup per tile and coil group.
// ---------------------------------------------------------------------------

typedef float f32x4 __attribute__((ext_vector_type(4)));

inline size_t spread_mfma_lds_bytes(const SpreadArgs &a, int waves)
{
    const int ttp = a.TT <= 16 ? 16 : a.TT <= 32 ? 32 : a.TT <= 64 ? 64 : 128; // table row stride
    return (size_t)(kSpreadBatch + 1) * ttp * 4 * 2 + (size_t)kSpreadBatch * 16 + (size_t)kSpreadBatch * 64 +
           (size_t)waves * kSpreadBatch * 4;
}

template <int D, int NCG>
__device__ __forceinline__ void spread_mfma_body(const SpreadArgs &a, const int4 item, const int cgi, const int bz)
{
    constexpr int B = kSpreadBatch;
    extern __shared__ float lds_sp[];
    if (tile_skippable(a, item)) return; // uniform over the workgroup
    // LDS: weight tables [B + 1][ttp] | reals [B + 1][ttp] (16 used) | masks [4][B] |
    //      per-axis window parameters [4][B] {l, c, ns - j, t} | per-wave pair lists
    //      [waves][B]; row B of the tables and reals is the all-zero pair.  Table and
    //      value rows have the same stride, so a list entry is the byte offset of a
    //      pair's rows in both.  Masks and parameters are axis-major, so that a wave's
    //      lanes touch consecutive entries.
    const int TT = a.TT;
    // table rows padded to a power of two: a pair's row starts at kk << lgt
    const int ttp = TT <= 16 ? 16 : TT <= 32 ? 32 : TT <= 64 ? 64 : 128;
    const int lgt = __builtin_ctz(ttp);
    float *s_tab = lds_sp;
    float *s_val = s_tab + (B + 1) * ttp;
    int *s_msk = (int *)(s_val + (B + 1) * ttp);
    int4 *s_par = (int4 *)(s_msk + B * 4);
    int *s_list = (int *)(s_par + B * 4);

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6); // uniform: block set-up on the scalar unit
    const int lane = tid & 63;
    const int nwaves = blockDim.x >> 6;
    const bool split = nwaves >= 2; // phase 1 spread over two waves
    static_assert(kSpreadBatch <= 64, "a batch's pairs are one wave's lanes");
    const int cg0 = cgi * NCG;
    const int ncg = (int)min((int64_t)NCG, a.nc - cg0);
    const int T0 = a.tile[0], T1 = a.tile[1], T2 = (D == 3) ? a.tile[2] : 1;
    const int to1 = T0, to2 = T0 + T1; // table offsets of axes 1 and 2

    const int nb = a.nblk[0] * a.nblk[1] * (D == 3 ? a.nblk[2] : 1);
    const int bidx = bz * nwaves + wave;
    const bool has_block = bidx < nb;
    // block coordinates (wave-uniform integer arithmetic)
    int bq0, bq1, bq2 = 0;
    if (D == 3) {
        const int q2 = bidx / a.nblk[2];
        bq2 = bidx - q2 * a.nblk[2];
        bq0 = q2 / a.nblk[1];
        bq1 = q2 - bq0 * a.nblk[1];
    } else {
        bq0 = bidx / a.nblk[1];
        bq1 = bidx - bq0 * a.nblk[1];
    }
    // this lane's MFMA row ri (a cell of a slice) and its pair slot k = lane / 16
    const int ri = lane & 15, kg = lane >> 4;
    // the wave's block: 4x4x4 cells in 3-D, 8x8 in 2-D, as bit masks per axis
    const uint32_t bm0 = D == 3 ? (0xfu << (bq0 * 4)) : (0xffu << (bq0 * 8));
    const uint32_t bm1 = D == 3 ? (0xfu << (bq1 * 4)) : (0xffu << (bq1 * 8));
    const uint32_t bm2 = D == 3 ? (0xfu << (bq2 * 4)) : 1u;
    // this lane's table columns and value in the rows of pair 0; a list entry (byte
    // offset of a pair's rows) added to them addresses the pair's entries.
    // 3-D: row ri is x-y cell (ri >> 2, ri & 3) of the block, slice s its z plane s:
    //   qx, qy the row's x and y columns, qz the block's 4 z columns (16-byte aligned:
    //   the z table starts at T0 + T1, a multiple of 4).
    // 2-D: row ri is cell (ri >> 3, ri & 7) of a 2 x 8 slice, slice s x rows 2s, 2s + 1:
    //   qx the row's x column in slice 0 (slice s: + 2s), qy its y column.
    const char *pv = (const char *)(s_val + ri);
    const char *qx = (const char *)(s_tab + (D == 3 ? bq0 * 4 + (ri >> 2) : bq0 * 8 + (ri >> 3)));
    const char *qy = (const char *)(s_tab + to1 + (D == 3 ? bq1 * 4 + (ri & 3) : bq1 * 8 + (ri & 7)));
    const char *qz = (const char *)(s_tab + to2 + bq2 * 4);

    if (tid < ttp) s_tab[B * ttp + tid] = 0.f;
    if (tid < 16) s_val[B * ttp + tid] = 0.f;

    // Kahan sums and compensations: set by the first batch (S = batch sums, K = 0), or to
    // zero after the loop when the item has no pairs
    float S[4][4], K[4][4];

    // phase-2 entry of this thread: the 4 consecutive table columns e_i .. e_i + 3 (one
    // axis e_ax, cells e_L .. e_L + 3: the tile extents are multiples of 4) of pairs
    // e_k0, e_k0 + e_kstep, ...; a row holds ttp / 4 column quads (ttp <= 128, so a
    // workgroup of at least 64 threads covers at least two pairs per step)
    const int lgq = lgt - 2;
    const int e_i = (tid & ((ttp >> 2) - 1)) << 2;
    const int e_k0 = tid >> lgq;
    const int e_kstep = (int)blockDim.x >> lgq;
    const int e_ax = e_i < to1 ? 0 : (e_i < to2 ? 1 : 2);
    const int e_L = e_i - (e_ax == 0 ? 0 : (e_ax == 1 ? to1 : to2));
    const int e_T = e_ax == 0 ? T0 : (e_ax == 1 ? T1 : T2);
    const float e_coef = a.coef[e_ax];
    for (int base = item.y; base < item.z; base += B) {
        const int n = min(B, item.z - base);
        if (base != item.y) __syncthreads(); // previous batch fully consumed (uniform)
        // phase 1: per pair, a thread of the first wave loads the coil values and a
        // thread of the second wave (the same thread with one wave) builds the masks and
        // window parameters; each loads the pair's record
        const int pt = split ? (tid & 63) : tid;
        const bool do_vals = tid < n;
        const bool do_msk = split ? (tid >= 64 && pt < n) : tid < n;
        if (do_vals || do_msk) {
            const int q = base + pt;
            const int4 pa = a.pair_a[q];
            const int4 pbi = a.pair_b[q];
            const float4 pb = make_float4(__int_as_float(pbi.x), __int_as_float(pbi.y), __int_as_float(pbi.z),
                                          __int_as_float(pbi.w));
            if (do_vals) {
            const float scale = a.use_dcf ? pb.w : 1.0f;
            float2 v[NCG];
            // coil-last with an even coil count: the pair's coils are contiguous 16-byte pairs
            const bool v4 = NCG >= 2 && a.layout == GN_COIL_LAST && ((a.nc | cg0) & 1) == 0 &&
                            (reinterpret_cast<uintptr_t>(a.in) & 15) == 0;
            if (v4) {
                const float4 *src = reinterpret_cast<const float4 *>(a.in + (int64_t)pa.x * a.nc + cg0);
#pragma unroll
                for (int c2 = 0; c2 < (NCG + 1) / 2; ++c2) {
                    float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (2 * c2 < ncg) t = src[c2];
                    v[2 * c2] = make_float2(t.x, t.y);
                    if (2 * c2 + 1 < NCG) v[2 * c2 + 1] = make_float2(t.z, t.w);
                }
            } else {
#pragma unroll
                for (int c = 0; c < NCG; ++c) {
                    v[c] = make_float2(0.f, 0.f);
                    if (c < ncg) {
                        const int64_t coil = cg0 + c;
                        v[c] = (a.layout == GN_COIL_LAST) ? a.in[(int64_t)pa.x * a.nc + coil] : a.in[coil * a.M + pa.x];
                    }
                }
            }
            float4 *dv = (float4 *)(s_val + pt * ttp);
#pragma unroll
            for (int c2 = 0; c2 < 4; ++c2) {
                float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
                if (2 * c2 < NCG) o = make_float4(v[2 * c2].x * scale, v[2 * c2].y * scale, 0.f, 0.f);
                if (2 * c2 + 1 < NCG) {
                    o.z = v[2 * c2 + 1].x * scale;
                    o.w = v[2 * c2 + 1].y * scale;
                }
                dv[c2] = o;
            }
            }
            if (do_msk) {
            // per axis: the window's tile-local cells (L - l) mod T < c as a bit mask,
            // and the parameters of its weights
            const int Ts[3] = {T0, T1, T2};
            const float ts[3] = {pb.x, pb.y, pb.z};
#pragma unroll
            for (int ax = 0; ax < D; ++ax) {
                const int l = (ax == 0 ? pa.w : ax == 1 ? (pa.w >> 10) : (pa.w >> 20)) & 0x3ff;
                const int c = (pa.z >> (8 * ax)) & 0xff;
                const int j = (pa.y >> (8 * ax)) & 0xff;
                uint64_t m = ((1ull << c) - 1) << l;
                m = (m | (m >> Ts[ax])) & ((1ull << Ts[ax]) - 1);
                s_msk[ax * B + pt] = (int)(uint32_t)m;
                s_par[ax * B + pt] = make_int4(l, c, a.ns - j, __float_as_int(ts[ax]));
            }
            }
        }
        __syncthreads();
        // phase 2: weight tables, one entry per (pair, axis, tile-local cell); a thread
        // keeps one column quad (its axis and 4 cells) and steps over the pairs
        if (e_i < TT) {
            for (int k = e_k0; k < n; k += e_kstep) {
                const int4 par = s_par[e_ax * B + k];
                int rel0 = e_L - par.x; // (L - l) mod T of the quad's first column
                if (rel0 < 0) rel0 += e_T;
                float w[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int rel = rel0 + i;
                    if (i > 0 && rel >= e_T) rel -= e_T;
                    const float u = __int_as_float(par.w) + (float)(par.z - rel);
                    const float g = __builtin_amdgcn_exp2f(-e_coef * u * u);
                    w[i] = rel < par.y ? g : 0.f;
                }
                *(float4 *)(s_tab + (k << lgt) + e_i) = make_float4(w[0], w[1], w[2], w[3]);
            }
        }
        __syncthreads();
        if (!has_block) continue;
        uint32_t m0 = 0, m1 = 0, m2 = 0;
        if (lane < n) {
            m0 = (uint32_t)s_msk[lane];
            m1 = (uint32_t)s_msk[B + lane];
            m2 = D == 3 ? (uint32_t)s_msk[2 * B + lane] : 1u;
        }
        f32x4 acc[4];
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) acc[sb] = f32x4{0.f, 0.f, 0.f, 0.f};
        // phase 3: one list of the pairs whose windows reach the wave's block; per 4
        // listed pairs a lane forms its weight factors once and feeds the block's four
        // slice MFMAs (a pair that misses a slice adds exact zeros there)
        const bool hit = (m0 & bm0) && (m1 & bm1) && (m2 & bm2);
        const unsigned long long hm = __ballot(hit);
        const int cnt = __popcll(hm);
        const int cpad = (cnt + 3) & ~3;
        int *lst = s_list + wave * B;
        const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(hm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0));
        if (hit) lst[pos] = lane << (lgt + 2);
        if (lane >= cnt && lane < cpad) lst[lane] = B << (lgt + 2);
        // the list was written by this wave's own lanes: order the LDS accesses
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int *lst0 = lst + kg;
        for (int t = 0; t < cpad; t += 4) {
            const int off = lst0[t]; // byte offset of the pair's rows
            const float v = *(const float *)(pv + off);
            float w[4];
            if (D == 3) {
                const float wxy = *(const float *)(qx + off) * *(const float *)(qy + off);
                const float4 wz = *(const float4 *)(qz + off);
                w[0] = wxy * wz.x;
                w[1] = wxy * wz.y;
                w[2] = wxy * wz.z;
                w[3] = wxy * wz.w;
            } else {
                const float wy = *(const float *)(qy + off);
#pragma unroll
                for (int sl = 0; sl < 4; ++sl) w[sl] = *(const float *)(qx + off + 8 * sl) * wy;
            }
#pragma unroll
            for (int sl = 0; sl < 4; ++sl) acc[sl] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[sl], v, acc[sl], 0, 0, 0);
        }
        if (base == item.y) {
            // first batch: the Kahan step from zero sums is exact (S = acc, K = 0)
#pragma unroll
            for (int sb = 0; sb < 4; ++sb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    S[sb][v] = acc[sb][v];
                    K[sb][v] = 0.f;
                }
        } else {
#pragma unroll
            for (int sb = 0; sb < 4; ++sb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float y = acc[sb][v] - K[sb][v];
                    const float tsum = S[sb][v] + y;
                    K[sb][v] = (tsum - S[sb][v]) - y;
                    S[sb][v] = tsum;
                }
        }
    }
    // Write-out straight from the registers.  Lane (ri, kg) holds re (ri even) or im
    // (ri odd) of coil ri >> 1 for 16 cells; for each of 4 outer steps i its 4 values
    // o[z] are consecutive cells along the last axis (3-D: z planes sb of x-y cell
    // (kg, i); 2-D: y cells v of slice i).  Neighbouring lanes swap halves (DPP) so
    // that the even lane holds cells z = 0, 1 and the odd lane cells z = 2, 3 as
    // (re, im) pairs: one 16-byte store per lane and step.  An item without pairs (a
    // tile no window reaches, when empty tiles are written) stores the zero sums.
    const int TC = a.TC;
    const int c = ri >> 1;         // coil within the group
    const bool odd = (ri & 1) != 0;
    const bool writer = c < ncg;
    int org[3] = {0, 0, 0}, teff[3] = {1, 1, 1};
    if (item.w < 0) {
        int r = item.x;
        for (int ax = D - 1; ax >= 0; --ax) {
            const int nt = a.ntiles[ax];
            const int qq = r % nt;
            r /= nt;
            org[ax] = qq * a.tile[ax];
            teff[ax] = (int)min((int64_t)a.tile[ax], a.nf[ax] - org[ax]);
        }
    }
    if (item.y >= item.z || !has_block) { // no pairs: the tile's sums are zero
#pragma unroll
        for (int sb = 0; sb < 4; ++sb)
#pragma unroll
            for (int v = 0; v < 4; ++v) S[sb][v] = K[sb][v] = 0.f;
    }
    const bool wide = ((a.G | a.nf[D - 1]) & 1) == 0;
    const int tl = teff[D - 1];
    // this lane's cells of step i: 3-D (L0b, L1b + i, L2), 2-D (L0b + 2 i, L1b); the
    // addresses of step 0 plus i times the step's stride
    const int L0b = D == 3 ? bq0 * 4 + kg : bq0 * 8 + (kg >> 1);
    const int L1b = D == 3 ? bq1 * 4 : bq1 * 8 + 4 * (kg & 1) + (odd ? 2 : 0);
    const int L2 = D == 3 ? bq2 * 4 + (odd ? 2 : 0) : 0;
    const int Ll = D == 3 ? L2 : L1b; // the pair's cell along the last axis
    const bool slot = item.w >= 0;
    float2 *dst0;
    int64_t step;
    bool lane_ok = writer;
    if (slot) {
        dst0 = a.partial + ((int64_t)item.w * a.nc + cg0 + c) * TC + (D == 3 ? (L0b * T1 + L1b) * T2 + L2 : L0b * T1 + L1b);
        step = D == 3 ? T2 : 2 * T1;
    } else {
        lane_ok = lane_ok && Ll < tl && (D == 2 || L0b < teff[0]);
        int64_t gi = (int64_t)(org[0] + L0b) * a.nf[1] + (org[1] + L1b);
        if (D == 3) gi = gi * a.nf[2] + (org[2] + L2);
        dst0 = a.grid + (int64_t)(cg0 + c) * a.G + gi;
        step = D == 3 ? a.nf[2] : 2 * a.nf[1];
    }
    const bool pairw = slot || (wide && Ll + 1 < tl);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float o[4];
#pragma unroll
        for (int z = 0; z < 4; ++z) o[z] = D == 3 ? S[z][i] - K[z][i] : S[i][z] - K[i][z];
        // quad_perm [1, 0, 3, 2]: the value of lane ^ 1
        const float r0 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(odd ? o[0] : o[2]), 0xb1, 0xf, 0xf, false));
        const float r1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(odd ? o[1] : o[3]), 0xb1, 0xf, 0xf, false));
        const float4 w = odd ? make_float4(r0, o[2], r1, o[3]) : make_float4(o[0], r0, o[1], r1);
        if (!lane_ok) continue;
        if (!slot && (D == 3 ? L1b + i >= teff[1] : L0b + 2 * i >= teff[0])) continue;
        float2 *dst = dst0 + i * step;
        if (pairw) {
            *(float4 *)dst = w;
        } else {
            dst[0] = make_float2(w.x, w.y);
            if (Ll + 1 < tl) dst[1] = make_float2(w.z, w.w);
        }
    }
}


template <int D, int NCG>
__global__ __launch_bounds__(64 * kSpreadWavesMax) void k_spread_mfma(SpreadArgs a)
{
    const uint32_t n = gridDim.x * gridDim.y * gridDim.z;
    const uint32_t lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const uint32_t w = xcd_swizzle(lin, n);
    const int bx = (int)(w % gridDim.x);
    const uint32_t yz = w / gridDim.x;
    spread_mfma_body<D, NCG>(a, a.items[bx], (int)(yz % gridDim.y), (int)(yz / gridDim.y));
}

template <int D, int NCG>
__global__ __launch_bounds__(64 * kSpreadWavesMax) void k_spread_mfma_batch(const SpreadArgs *__restrict__ args,
                                                                             const int4 *__restrict__ bitems)
{
    const uint32_t n = gridDim.x * gridDim.y;
    const uint32_t w = xcd_swizzle(blockIdx.x + gridDim.x * blockIdx.y, n);
    const int4 bi = bitems[w % gridDim.x];
    const SpreadArgs &a = args[bi.x]; // read in place: uniform scalar loads
    spread_mfma_body<D, NCG>(a, a.items[bi.y], bi.z, (int)(w / gridDim.x));
}

// Kahan-compensated sum of the partial tiles of each split tile, in slice order;
// one thread per (coil, cell) of a tile (kReduceThreads covers 8 coils of 128 cells).
constexpr int kReduceThreads = 1024;
template <int D>
__device__ __forceinline__ void spread_reduce_body(const SpreadArgs &a, const int4 it)
{
    // it = {tile, slot0, nslots}
    int org[3] = {0, 0, 0}, teff[3] = {1, 1, 1};
    {
        uint32_t r = (uint32_t)it.x; // tile index < 2^31
        for (int ax = D - 1; ax >= 0; --ax) {
            const int q = (int)(r % (uint32_t)a.ntiles[ax]);
            r /= (uint32_t)a.ntiles[ax];
            org[ax] = q * a.tile[ax];
            teff[ax] = (int)min((int64_t)a.tile[ax], a.nf[ax] - org[ax]);
        }
    }
    const int TC = a.TC;
    const int T1 = a.tile[1], T2 = (D == 3) ? a.tile[2] : 1;
    const int ne = (int)(TC * a.nc); // a tile's cells of every coil (< 2^31)
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int64_t coil = e / TC;
        const int cidx = e - (int)coil * TC;
        int x0, x1, x2 = 0;
        if (D == 3) {
            x2 = cidx % T2;
            x1 = (cidx / T2) % T1;
            x0 = cidx / (T2 * T1);
        } else {
            x1 = cidx % T1;
            x0 = cidx / T1;
        }
        if (x0 >= teff[0] || x1 >= teff[1] || (D == 3 && x2 >= teff[2])) continue;
        float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;
        const float2 *src = a.partial + (int64_t)it.y * a.nc * TC + coil * TC + cidx;
        const int64_t sstride = a.nc * TC;
        // slices in groups of kStageUnroll: the loads are issued together, the
        // compensated sum still runs in slice order
        for (int s0 = 0; s0 < it.z; s0 += kStageUnroll) {
            float2 v[kStageUnroll];
#pragma unroll
            for (int u = 0; u < kStageUnroll; ++u)
                v[u] = s0 + u < it.z ? src[(int64_t)(s0 + u) * sstride] : make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < kStageUnroll; ++u) {
                if (s0 + u >= it.z) break;
                const float yr = v[u].x - cr, yi = v[u].y - ci;
                const float tr = sr + yr, ti = si + yi;
                cr = (tr - sr) - yr;
                ci = (ti - si) - yi;
                sr = tr;
                si = ti;
            }
        }
        int64_t gi = (int64_t)(org[0] + x0) * a.nf[1] + (org[1] + x1);
        if (D == 3) gi = gi * a.nf[2] + (org[2] + x2);
        a.grid[coil * a.G + gi] = make_float2(sr - cr, si - ci);
    }
}

template <int D>
__global__ __launch_bounds__(kReduceThreads) void k_spread_reduce(SpreadArgs a, const int4 *__restrict__ ritems)
{
    spread_reduce_body<D>(a, ritems[blockIdx.x]);
}

// bitems[b] = {instance, reduce item}
template <int D>
__global__ __launch_bounds__(kReduceThreads) void k_spread_reduce_batch(const SpreadArgs *__restrict__ args,
                                                             const int4 *__restrict__ bitems)
{
    const int4 bi = bitems[blockIdx.x];
    const SpreadArgs &a = args[bi.x]; // read in place: uniform scalar loads
    spread_reduce_body<D>(a, a.reduce_items[bi.y]);
}

// ---------------------------------------------------------------------------
// Forward gather: one workgroup per (work item, coil pair).  A work item is a
// gather tile and up to kGatherChunk of its samples.  The tile plus its window
// halo is staged in LDS as 16-byte cells holding both coils of the pair, with
// coalesced loads; every thread then interpolates one sample of both coils from
// LDS, one 16-byte LDS read per tap, forming the separable 1-D weights in
// registers.  With an odd coil count the last pair's second coil is zero and
// not written.
// ---------------------------------------------------------------------------

static_assert(kGatherCoils == 2, "the gather stages coil pairs");

template <int D>
__device__ inline void stage_region_pair(float4 *lds, const float2 *__restrict__ g0, const float2 *__restrict__ g1,
                                         bool two, const int64_t *nf, const int *org, int R1, int R2, int RC)
{
    // A thread takes cells rr = tid, tid + NT, ... of the region, rr = (r0 * R1 + r1) * R2 + r2
    // (R2 = 1 in 2-D), and steps its (r0, r1, r2) by NT = (d0 * R1 + d1) * R2 + d2 with
    // carries instead of dividing for every cell.
    const int NT = blockDim.x;
    const int R12 = R1 * R2;
    const int d0 = NT / R12, dr = NT - d0 * R12, d1 = dr / R2, d2 = dr - d1 * R2;
    const int tid = threadIdx.x;
    int r0 = tid / R12;
    int r1 = (tid - r0 * R12) / R2;
    int r2 = tid - r0 * R12 - r1 * R2;
    const int nf0 = (int)nf[0], nf1 = (int)nf[1], nf2 = D == 3 ? (int)nf[2] : 1;
    for (int base = tid; base < RC; base += kStageUnroll * NT) {
        float4 v[kStageUnroll];
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int rr = base + u * NT;
            if (rr < RC) {
                int q0 = org[0] + r0;
                if (q0 >= nf0) q0 -= nf0;
                int q1 = org[1] + r1;
                if (q1 >= nf1) q1 -= nf1;
                int64_t gi = (int64_t)q0 * nf1 + q1;
                if (D == 3) {
                    int q2 = org[2] + r2;
                    if (q2 >= nf2) q2 -= nf2;
                    gi = gi * nf2 + q2;
                }
                const float2 x0 = g0[gi], x1 = two ? g1[gi] : make_float2(0.f, 0.f);
                v[u] = make_float4(x0.x, x0.y, x1.x, x1.y);
            }
            r2 += d2;
            int c = r2 >= R2 ? 1 : 0;
            r2 -= c * R2;
            r1 += d1 + c;
            c = r1 >= R1 ? 1 : 0;
            r1 -= c * R1;
            r0 += d0 + c;
        }
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int rr = base + u * NT;
            if (rr < RC) lds[rr] = v[u];
        }
    }
}

// Row-wise staging pays when whole rows fill at least 7/8 of a wave's lanes (21 or
// 29 cells: 63 or 58 lanes; 25 cells would leave 14 idle) and grid indices fit 31 bits.
__device__ inline bool rows_stage_ok(int RL, int64_t G)
{
    return RL <= 64 && (64 / RL) * RL >= 56 && G < 0x7fffffffll;
}

// Row-wise staging of the same region for rows of at most 64 cells along the last
// (contiguous) axis and grids of fewer than 2^31 cells per coil.  Each wave takes
// 64 / RL whole rows per step (lane = row slot * RL + cell), so a lane keeps one
// last-axis cell for the whole loop (its wrapped grid index is computed once) and
// only the leading-axis coordinates of its row advance, with one carry; grid indices
// are 32-bit.  A wave's lanes write consecutive LDS cells.  PAIR: 16-byte cells
// holding coils g0 and g1 (zeros for a missing second coil); else 8-byte cells of g0.
template <int D, bool PAIR, typename Cell>
__device__ inline void stage_region_rows(Cell *lds, const float2 *__restrict__ g0, const float2 *__restrict__ g1,
                                         bool two, const int64_t *nf, const int *org, int R0, int R1, int R2)
{
    constexpr int U = 4; // rows per lane whose loads are in flight together
    const int RL = D == 3 ? R2 : R1;
    const int nrow = D == 3 ? R0 * R1 : R0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rpw = 64 / RL;
    const int k = lane / RL, c = lane - k * RL;
    if (k >= rpw) return;
    const int step = (int)(blockDim.x >> 6) * rpw;
    const int nf0 = (int)nf[0], nf1 = (int)nf[1], nfl = (int)nf[D - 1];
    int ql = org[D - 1] + c;
    if (ql >= nfl) ql -= nfl;
    int rho = wave * rpw + k;
    int r0, r1 = 0, s0 = step, s1 = 0;
    if (D == 3) {
        r0 = rho / R1;
        r1 = rho - r0 * R1;
        s0 = step / R1;
        s1 = step - s0 * R1;
    } else {
        r0 = rho;
    }
    const uint32_t nf2 = D == 3 ? (uint32_t)nf[2] : 1u;
    for (; rho < nrow; rho += U * step) {
        Cell v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (rho + u * step < nrow) {
                int q0 = org[0] + r0;
                if (q0 >= nf0) q0 -= nf0;
                uint32_t off;
                if (D == 3) {
                    int q1 = org[1] + r1;
                    if (q1 >= nf1) q1 -= nf1;
                    off = ((uint32_t)q0 * (uint32_t)nf1 + (uint32_t)q1) * nf2 + (uint32_t)ql;
                } else {
                    off = (uint32_t)q0 * (uint32_t)nf1 + (uint32_t)ql;
                }
                const float2 x0 = g0[off];
                if constexpr (PAIR) {
                    const float2 x1 = two ? g1[off] : make_float2(0.f, 0.f);
                    v[u] = make_float4(x0.x, x0.y, x1.x, x1.y);
                } else {
                    v[u] = x0;
                }
            }
            if (D == 3) {
                r1 += s1;
                const int cc = r1 >= R1 ? 1 : 0;
                r1 -= cc * R1;
                r0 += s0 + cc;
            } else {
                r0 += step;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (rho + u * step < nrow) lds[(rho + u * step) * RL + c] = v[u];
    }
}

// acc += w * v on both coils of a staged cell
__device__ inline void fma4(float4 &acc, float w, const float4 v)
{
    acc.x += w * v.x;
    acc.y += w * v.y;
    acc.z += w * v.z;
    acc.w += w * v.w;
}

// Lanes of the gather-type kernels when tps lanes share a sample: sample grp of the
// wave, tap share sub; a sample's other lanes are lane + k * stride (k < tps - sub).
// Wide windows (nspread >= 4) space a sample's lanes spw apart, so lanes 0 .. spw-1
// hold consecutive samples and the LDS-bank order of the records (k_bank_order)
// holds for every tps (nspread 5: gather 1.86 -> 1.79 ms); at nspread 3 adjacent
// lanes per sample measured 1% faster.
template <int NS>
__device__ inline void lane_layout(int lane, int tps, int spw, int *grp, int *sub, int *stride)
{
    if (NS >= 4) {
        *sub = lane / spw;
        *grp = lane - *sub * spw;
        *stride = spw;
    } else {
        *grp = lane / tps;
        *sub = lane - *grp * tps;
        *stride = 1;
    }
}

// Tap weights use the hardware exp2 (v_exp_f32) directly.  The exponent is
// coef * u^2 with coef = pi (rat - 1/2) / (rat ns ln 2) and |u| <= ns + 1, so every
// weight is >= 2^-82 for nspread <= 16: far from the denormal range, where the
// hardware exp2 and exp2f agree bit for bit.
template <int D, int NS, bool WRAP>
__device__ __forceinline__ void gather_body(const GatherArgs &a, const int4 item, const int cgi)
{
    constexpr int W = 2 * NS;
    extern __shared__ float4 lds_pair[];
    const int cg0 = cgi * kGatherCoils;
    const bool two = cg0 + 1 < a.nc;

    int org[3] = {0, 0, 0};
    {
        uint32_t r = (uint32_t)item.x; // tile index < 2^31
        for (int ax = D - 1; ax >= 0; --ax) {
            org[ax] = (int)(r % (uint32_t)a.ntiles[ax]) * a.tile[ax];
            r /= (uint32_t)a.ntiles[ax];
        }
    }
    const int R0 = a.region[0], R1 = a.region[1], R2 = (D == 3) ? a.region[2] : 1;
    const int RC = R0 * R1 * R2;

    const float2 *g0 = a.grid + (int64_t)cg0 * a.G;
    if (rows_stage_ok(D == 3 ? R2 : R1, a.G))
        stage_region_rows<D, true>(lds_pair, g0, g0 + a.G, two, a.nf, org, R0, R1, R2);
    else
        stage_region_pair<D>(lds_pair, g0, g0 + a.G, two, a.nf, org, R1, R2, RC);
    __syncthreads();
    // Threads per sample: a work item's samples are spread over all waves of the
    // workgroup.  The tps lanes of a sample take the leading-axis taps t0 = sub,
    // sub + tps, ... (tps divides W); lane sub 0 then adds the partial sums of lanes
    // sub 1 .. tps-1 in that order, so the result does not depend on timing or on the
    // lane layout (lane_layout).
    const int n = item.z - item.y;
    const int nwaves = kGatherThreads / 64;
    int tps = 1;
#pragma unroll
    for (int d = 2; d <= 8; ++d)
        if (W % d == 0 && n <= nwaves * (64 / d)) tps = d;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int spw = 64 / tps;               // samples per wave
    int grp, sub, stride;
    lane_layout<NS>(lane, tps, spw, &grp, &sub, &stride);
    for (int s0 = 0; s0 < n; s0 += nwaves * spw) {
        const int slot = s0 + wave * spw + grp;
        const bool valid = grp < spw && sub < tps && slot < n;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        int orig = 0;
        if (valid) {
            const int j = item.y + slot;
            const int4 rc = a.rec_cell[j];
            const float4 rf = a.rec_frac[j];
            orig = rc.w;
            int l0 = rc.x - org[0];
            if (l0 < 0) l0 += (int)a.nf[0];
            int l1 = rc.y - org[1];
            if (l1 < 0) l1 += (int)a.nf[1];
            int l2 = 0;
            if (D == 3) {
                l2 = rc.z - org[2];
                if (l2 < 0) l2 += (int)a.nf[2];
            }
            float w1[W], w2[D == 3 ? W : 1];
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const float u1 = rf.y + (float)(NS - t);
                w1[t] = __builtin_amdgcn_exp2f(-a.coef[1] * u1 * u1);
                if (D == 3) {
                    const float u2 = rf.z + (float)(NS - t);
                    w2[t] = __builtin_amdgcn_exp2f(-a.coef[2] * u2 * u2);
                }
            }
#pragma unroll 1
            for (int t0 = sub; t0 < W; t0 += tps) {
                const float u0 = rf.x + (float)(NS - t0);
                const float w0 = __builtin_amdgcn_exp2f(-a.coef[0] * u0 * u0);
                int r0 = l0 + t0;
                if (WRAP && r0 >= R0) r0 -= (int)a.nf[0]; // only when the region spans the whole axis
                float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int t1 = 0; t1 < W; ++t1) {
                    int r1 = l1 + t1;
                    if (WRAP && r1 >= R1) r1 -= (int)a.nf[1];
                    if (D == 3) {
                        const int rowb = (r0 * R1 + r1) * R2;
                        float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                        for (int t2 = 0; t2 < W; ++t2) {
                            int r2 = l2 + t2;
                            if (WRAP && r2 >= R2) r2 -= (int)a.nf[2];
                            fma4(q, w2[t2], lds_pair[rowb + r2]);
                        }
                        fma4(p, w1[t1], q);
                    } else {
                        fma4(p, w1[t1], lds_pair[r0 * R1 + r1]);
                    }
                }
                fma4(acc, w0, p);
            }
        }
        // every lane takes part in the exchange (a sample's lanes are in one wave)
        for (int k = 1; k < tps; ++k) {
            const int src = lane + k * stride;
            const float4 o = make_float4(__shfl(acc.x, src), __shfl(acc.y, src), __shfl(acc.z, src),
                                         __shfl(acc.w, src));
            if (sub == 0) {
                acc.x += o.x;
                acc.y += o.y;
                acc.z += o.z;
                acc.w += o.w;
            }
        }
        if (!valid || sub != 0) continue;
        if (a.scale != 1.0f) {
            acc.x *= a.scale;
            acc.y *= a.scale;
            acc.z *= a.scale;
            acc.w *= a.scale;
        }
        const float2 o0 = make_float2(acc.x, acc.y), o1 = make_float2(acc.z, acc.w);
        if (a.layout == GN_COIL_LAST) {
            float2 *o = a.out + (int64_t)orig * a.nc + cg0;
            o[0] = o0;
            if (two) o[1] = o1;
        } else {
            a.out[(int64_t)cg0 * a.M + orig] = o0;
            if (two) a.out[(int64_t)(cg0 + 1) * a.M + orig] = o1;
        }
    }
}

template <int D, int NS, bool WRAP>
__global__ __launch_bounds__(kGatherThreads) void k_gather(GatherArgs a)
{
    const uint32_t n = gridDim.x * gridDim.y;
    const uint32_t w = xcd_swizzle(blockIdx.x + gridDim.x * blockIdx.y, n);
    // coil groups fastest: a tile's coils run side by side and share its sample records in L2
    gather_body<D, NS, WRAP>(a, a.items[w / gridDim.y], (int)(w % gridDim.y));
}

// Batched launch over many plans: bitems[b] = {instance, work item, coil group, 0},
// taken in the XCD-aware order of the plain launch.
template <int D, int NS, bool WRAP>
__global__ __launch_bounds__(kGatherThreads) void k_gather_batch(const GatherArgs *__restrict__ args,
                                                                 const int4 *__restrict__ bitems)
{
    const int4 bi = bitems[xcd_swizzle(blockIdx.x, gridDim.x)];
    const GatherArgs &a = args[bi.x]; // read in place: uniform scalar loads
    gather_body<D, NS, WRAP>(a, a.items[bi.y], bi.z);
}

// ---------------------------------------------------------------------------
// Trajectory gradient: derivative of the type-2 interpolation with respect to
// the sample coordinates, contracted with per-sample complex weights:
//   out[a][j] = sum_c Re( conj(wts[j,c]) * d/dx_a (A image)[j,c] ),
// with d/dx_a w = w * gscale_a * (t + ns - tap), gscale_a = -h_a df / (2 tau),
// the analytic derivative of the Gaussian exp(-(xi - h g)^2 / (4 tau)).
// Same staging as the gather; all coils of a sample are handled by one thread.
// ---------------------------------------------------------------------------

template <int D, int NS, bool WRAP>
__global__ __launch_bounds__(kGatherThreads) void k_traj_grad(GatherArgs a, const float2 *__restrict__ wts,
                                                               float *__restrict__ out, float gs0, float gs1, float gs2)
{
    constexpr int W = 2 * NS;
    extern __shared__ float2 lds_gather[];
    const int4 item = a.items[blockIdx.x];
    int org[3] = {0, 0, 0};
    {
        uint32_t r = (uint32_t)item.x; // tile index < 2^31
        for (int ax = D - 1; ax >= 0; --ax) {
            org[ax] = (int)(r % (uint32_t)a.ntiles[ax]) * a.tile[ax];
            r /= (uint32_t)a.ntiles[ax];
        }
    }
    const int R0 = a.region[0], R1 = a.region[1], R2 = (D == 3) ? a.region[2] : 1;
    const int RC = R0 * R1 * R2;
    // several lanes per sample, as in the gather: the tps lanes of a group take the
    // leading-axis taps t0 = sub, sub + tps, ... and the group's lane 0 adds the
    // partial gradients of the others in lane order
    const int n = item.z - item.y;
    const int nwaves = kGatherThreads / 64;
    int tps = 1;
#pragma unroll
    for (int d = 2; d <= 8; ++d)
        if (W % d == 0 && n <= nwaves * (64 / d)) tps = d;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int spw = 64 / tps;
    int grp, sub, stride;
    lane_layout<NS>(lane, tps, spw, &grp, &sub, &stride);
    for (int s0 = 0; s0 < n; s0 += nwaves * spw) {
        const int slot = s0 + wave * spw + grp;
        const bool valid = grp < spw && sub < tps && slot < n;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f;
        int4 rc = make_int4(0, 0, 0, 0);
        float4 rf = make_float4(0.f, 0.f, 0.f, 0.f);
        int l0 = 0, l1 = 0, l2 = 0;
        float w1[W], u1s[W], w2[D == 3 ? W : 1], u2s[D == 3 ? W : 1];
        if (valid) {
            rc = a.rec_cell[item.y + slot];
            rf = a.rec_frac[item.y + slot];
            l0 = rc.x - org[0];
            if (l0 < 0) l0 += (int)a.nf[0];
            l1 = rc.y - org[1];
            if (l1 < 0) l1 += (int)a.nf[1];
            if (D == 3) {
                l2 = rc.z - org[2];
                if (l2 < 0) l2 += (int)a.nf[2];
            }
        }
#pragma unroll
        for (int t = 0; t < W; ++t) {
            u1s[t] = rf.y + (float)(NS - t);
            w1[t] = __builtin_amdgcn_exp2f(-a.coef[1] * u1s[t] * u1s[t]);
            if (D == 3) {
                u2s[t] = rf.z + (float)(NS - t);
                w2[t] = __builtin_amdgcn_exp2f(-a.coef[2] * u2s[t] * u2s[t]);
            }
        }
        for (int64_t coil = 0; coil < a.nc; ++coil) {
            const float2 *g = a.grid + coil * a.G;
            __syncthreads();
            if (rows_stage_ok(D == 3 ? R2 : R1, a.G))
                stage_region_rows<D, false>(lds_gather, g, g, false, a.nf, org, R0, R1, R2);
            else
                stage_region<D>(lds_gather, g, a.nf, org, R1, R2, RC);
            __syncthreads();
            if (!valid) continue;
            // value sums with plain and derivative weights, separably
            float dxr = 0.f, dxi = 0.f, dyr = 0.f, dyi = 0.f, dzr = 0.f, dzi = 0.f;
#pragma unroll 1
            for (int t0 = sub; t0 < W; t0 += tps) {
                const float u0 = rf.x + (float)(NS - t0);
                const float w0 = __builtin_amdgcn_exp2f(-a.coef[0] * u0 * u0);
                int r0 = l0 + t0;
                if (WRAP && r0 >= R0) r0 -= (int)a.nf[0];
                float pr = 0.f, pi = 0.f, pyr = 0.f, pyi = 0.f, pzr = 0.f, pzi = 0.f;
#pragma unroll
                for (int t1 = 0; t1 < W; ++t1) {
                    int r1 = l1 + t1;
                    if (WRAP && r1 >= R1) r1 -= (int)a.nf[1];
                    float qr = 0.f, qi = 0.f, qzr = 0.f, qzi = 0.f;
                    if (D == 3) {
                        const int rowb = (r0 * R1 + r1) * R2;
#pragma unroll
                        for (int t2 = 0; t2 < W; ++t2) {
                            int r2 = l2 + t2;
                            if (WRAP && r2 >= R2) r2 -= (int)a.nf[2];
                            const float2 gv = lds_gather[rowb + r2];
                            qr += w2[t2] * gv.x;
                            qi += w2[t2] * gv.y;
                            qzr += w2[t2] * u2s[t2] * gv.x;
                            qzi += w2[t2] * u2s[t2] * gv.y;
                        }
                    } else {
                        const float2 gv = lds_gather[r0 * R1 + r1];
                        qr = gv.x;
                        qi = gv.y;
                    }
                    pr += w1[t1] * qr;
                    pi += w1[t1] * qi;
                    pyr += w1[t1] * u1s[t1] * qr;
                    pyi += w1[t1] * u1s[t1] * qi;
                    pzr += w1[t1] * qzr;
                    pzi += w1[t1] * qzi;
                }
                dxr += w0 * u0 * pr;
                dxi += w0 * u0 * pi;
                dyr += w0 * pyr;
                dyi += w0 * pyi;
                dzr += w0 * pzr;
                dzi += w0 * pzi;
            }
            const float2 wv = (a.layout == GN_COIL_LAST) ? wts[(int64_t)rc.w * a.nc + coil] : wts[coil * a.M + rc.w];
            // Re(conj(wv) * d) = wv.x d.x + wv.y d.y
            g0 += wv.x * dxr + wv.y * dxi;
            g1 += wv.x * dyr + wv.y * dyi;
            g2 += wv.x * dzr + wv.y * dzi;
        }
        for (int k = 1; k < tps; ++k) {
            const int src = lane + k * stride;
            const float o0 = __shfl(g0, src), o1 = __shfl(g1, src), o2 = __shfl(g2, src);
            if (sub == 0) {
                g0 += o0;
                g1 += o1;
                g2 += o2;
            }
        }
        if (valid && sub == 0) {
            out[rc.w] = g0 * gs0;
            out[a.M + rc.w] = g1 * gs1;
            if (D == 3) out[2 * a.M + rc.w] = g2 * gs2;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused pre-pass (deapodize + zero-pad + coil transpose) and post-pass
// (crop + deapodize + scale + coil transpose).  Frequency k = n - m//2 sits at
// grid index k mod nf (fft/nufft_func.py:761-764, 728-729).
// ---------------------------------------------------------------------------



template <int D>
__device__ __forceinline__ void pad_body(const PrepostArgs &a, const float2 *__restrict__ img, float2 *__restrict__ grid)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t gidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < a.G; gidx += stride) {
        int64_t i2 = 0, i1, i0;
        if (D == 3) {
            i2 = gidx % a.nf[2];
            int64_t r = gidx / a.nf[2];
            i1 = r % a.nf[1];
            i0 = r / a.nf[1];
        } else {
            i1 = gidx % a.nf[1];
            i0 = gidx / a.nf[1];
        }
        int64_t n0 = i0 + a.modes[0] / 2;
        if (n0 >= a.nf[0]) n0 -= a.nf[0];
        int64_t n1 = i1 + a.modes[1] / 2;
        if (n1 >= a.nf[1]) n1 -= a.nf[1];
        int64_t n2 = 0;
        bool inside = n0 < a.modes[0] && n1 < a.modes[1];
        if (D == 3) {
            n2 = i2 + a.modes[2] / 2;
            if (n2 >= a.nf[2]) n2 -= a.nf[2];
            inside &= n2 < a.modes[2];
        }
        if (!inside) {
            for (int64_t c = 0; c < a.nc; ++c) grid[c * a.G + gidx] = make_float2(0.f, 0.f);
            continue;
        }
        float d = a.scale * a.deap0[n0] * a.deap1[n1];
        int64_t n = n0 * a.modes[1] + n1;
        if (D == 3) {
            d *= a.deap2[n2];
            n = n * a.modes[2] + n2;
        }
        if (a.sens) { // SENSE: coil c's image is x * S_c
            const float2 x = img[n];
            for (int64_t c = 0; c < a.nc; ++c) {
                const float2 s = (a.layout == GN_COIL_LAST) ? a.sens[n * a.nc + c] : a.sens[c * a.N + n];
                const float2 v = make_float2(x.x * s.x - x.y * s.y, x.x * s.y + x.y * s.x);
                grid[c * a.G + gidx] = make_float2(v.x * d, v.y * d);
            }
            continue;
        }
        for (int64_t c = 0; c < a.nc; ++c) {
            const float2 v = (a.layout == GN_COIL_LAST) ? img[n * a.nc + c] : img[c * a.N + n];
            grid[c * a.G + gidx] = make_float2(v.x * d, v.y * d);
        }
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_pad(PrepostArgs a, const float2 *__restrict__ img, float2 *__restrict__ grid)
{
    pad_body<D>(a, img, grid);
}

// batched: blockIdx.y = instance; args[i].src = image, args[i].dst = grid
template <int D>
__global__ __launch_bounds__(256) void k_pad_batch(const PrepostArgs *__restrict__ args)
{
    const PrepostArgs &a = args[blockIdx.y]; // read in place: uniform scalar loads
    pad_body<D>(a, a.src, a.dst);
}

template <int D>
__device__ __forceinline__ void crop_body(const PrepostArgs &a, const float2 *__restrict__ grid, float2 *__restrict__ img)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += stride) {
        int64_t n2 = 0, n1, n0;
        if (D == 3) {
            n2 = n % a.modes[2];
            int64_t r = n / a.modes[2];
            n1 = r % a.modes[1];
            n0 = r / a.modes[1];
        } else {
            n1 = n % a.modes[1];
            n0 = n / a.modes[1];
        }
        int64_t i0 = n0 - a.modes[0] / 2;
        if (i0 < 0) i0 += a.nf[0];
        int64_t i1 = n1 - a.modes[1] / 2;
        if (i1 < 0) i1 += a.nf[1];
        int64_t gi = i0 * a.nf[1] + i1;
        float d = a.scale * a.deap0[n0] * a.deap1[n1];
        if (D == 3) {
            int64_t i2 = n2 - a.modes[2] / 2;
            if (i2 < 0) i2 += a.nf[2];
            gi = gi * a.nf[2] + i2;
            d *= a.deap2[n2];
        }
        if (a.sens) { // SENSE: sum_c conj(S_c) * image_c, coils in order
            float2 acc = make_float2(0.f, 0.f);
            for (int64_t c = 0; c < a.nc; ++c) {
                const float2 v = grid[c * a.G + gi];
                const float2 o = make_float2(v.x * d, v.y * d);
                const float2 s = (a.layout == GN_COIL_LAST) ? a.sens[n * a.nc + c] : a.sens[c * a.N + n];
                acc.x += s.x * o.x + s.y * o.y;
                acc.y += s.x * o.y - s.y * o.x;
            }
            img[n] = acc;
            continue;
        }
        for (int64_t c = 0; c < a.nc; ++c) {
            const float2 v = grid[c * a.G + gi];
            const float2 o = make_float2(v.x * d, v.y * d);
            if (a.layout == GN_COIL_LAST)
                img[n * a.nc + c] = o;
            else
                img[c * a.N + n] = o;
        }
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_crop(PrepostArgs a, const float2 *__restrict__ grid, float2 *__restrict__ img)
{
    crop_body<D>(a, grid, img);
}

// batched: blockIdx.y = instance; args[i].src = grid, args[i].dst = image
template <int D>
__global__ __launch_bounds__(256) void k_crop_batch(const PrepostArgs *__restrict__ args)
{
    const PrepostArgs &a = args[blockIdx.y]; // read in place: uniform scalar loads
    crop_body<D>(a, a.src, a.dst);
}

inline int grid_blocks(int64_t n, int threads, int64_t cap = 1 << 20)
{
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    return (int)std::min<int64_t>(b, cap);
}

template <typename T>
int dev_alloc(T **p, size_t count, const char *what)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(GN_ENOMEM, std::string("hipMalloc failed for ") + what);
    }
    return GN_OK;
}

// Device scratch of set_points: one pool per device, shared by every plan, that only
// grows.  A set_points call holds the pool (its lock) until its last use of the
// scratch has completed on its stream, so plans on other streams and threads never
// see each other's buffers; buffers are requested in the same order on every call,
// and buffer k of a call reuses buffer k of the previous one.
struct ScratchPool {
    std::mutex mu;
    std::vector<void *> ptr;
    std::vector<size_t> cap;
};

ScratchPool &scratch_pool()
{
    static std::mutex mu;
    static std::map<int, ScratchPool *> pools;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> g(mu);
    ScratchPool *&p = pools[dev];
    if (!p) p = new ScratchPool(); // lives until the process ends
    return *p;
}

struct Scratch {
    ScratchPool &pool;
    std::unique_lock<std::mutex> lock;
    hipStream_t stream;
    size_t next = 0;
    explicit Scratch(hipStream_t st) : pool(scratch_pool()), lock(pool.mu), stream(st) {}
    // also on early returns: nothing queued may still use the pool once it is released
    ~Scratch() { (void)hipStreamSynchronize(stream); }
    template <typename T>
    int alloc(T **p, size_t count, const char *what)
    {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        const size_t k = next++;
        if (k >= pool.ptr.size()) {
            pool.ptr.push_back(nullptr);
            pool.cap.push_back(0);
        }
        if (pool.cap[k] < bytes) {
            if (pool.ptr[k]) (void)hipFree(pool.ptr[k]); // waits for the device
            pool.ptr[k] = nullptr;
            pool.cap[k] = 0;
            const size_t want = bytes + bytes / 4;      // headroom for slightly larger calls
            if (hipMalloc(&pool.ptr[k], want) != hipSuccess) {
                pool.ptr[k] = nullptr;
                *p = nullptr;
                return fail(GN_ENOMEM, std::string("hipMalloc failed for ") + what);
            }
            pool.cap[k] = want;
        }
        *p = (T *)pool.ptr[k];
        return GN_OK;
    }
};

// A buffer owned by the plan that only grows (count elements of T, at least one).
template <typename T>
int grow_alloc(T **p, int64_t *cap, size_t count, const char *what)
{
    count = std::max<size_t>(count, 1);
    if (*p && (size_t)*cap >= count) return GN_OK;
    int rc = dev_alloc(p, count, what);
    *cap = rc ? 0 : (int64_t)count;
    return rc;
}

// Exclusive scan of n ints in place; writes the total to *total_dev (device int).
hipError_t device_scan(int *data, int64_t n, int *bsum, hipStream_t st)
{
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(kScanThreads), 0, st, data, n,
                       bsum);
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(kScanThreads), 0, st, bsum, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(kScanThreads), 0, st, data, n,
                       bsum);
    return hipGetLastError();
}

// Stable counting sort of n items by keys[i*stride] in [0,K): writes the sorted item
// indices and the K+1 bucket offsets.  Asynchronous on st; the temporary buffers
// belong to sc, which the caller keeps alive until the stream has consumed them.
int stable_sort_by_key(int64_t n, const int *keys, int stride, int64_t K, int *sorted, int *bin_start,
                       hipStream_t st, Scratch &sc)
{
    int rc;
    int kbits = 1;
    while (kbits < 31 && ((int64_t)1 << kbits) < K) ++kbits;
    const int npass = (kbits + kRadixBits - 1) / kRadixBits;
    // chunks: about 4096 waves over the GPU, 256..4096 items each
    int64_t chunk_len = (n + 4095) / 4096;
    chunk_len = std::min<int64_t>(4096, std::max<int64_t>(256, ((chunk_len + 63) / 64) * 64));
    const int64_t C = std::max<int64_t>(1, (n + chunk_len - 1) / chunk_len);
    const int64_t nKC = (int64_t)kRadixBins * C;
    const int64_t nb = (nKC + kScanTile - 1) / kScanTile;
    int *cnt = nullptr, *bsum = nullptr, *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
    if ((rc = sc.alloc(&cnt, (size_t)nKC, "sort counts")) || (rc = sc.alloc(&bsum, (size_t)nb + 1, "sort scan")) ||
        (rc = sc.alloc(&k0, (size_t)n, "sort keys")) || (rc = sc.alloc(&k1, (size_t)n, "sort keys")) ||
        (rc = sc.alloc(&v0, (size_t)n, "sort values")))
        return rc;
    v1 = sorted;
    const unsigned nblk = (unsigned)((C + kRadixWaves - 1) / kRadixWaves);
    const int *kin = keys;
    const int *vin = nullptr;
    int kstride = stride;
    // ping-pong so that the last pass writes its values into `sorted`
    int *kout[2] = {k0, k1};
    int *vout[2] = {v0, v1};
    int shift = 0;
    hipError_t e = hipSuccess;
    for (int ps = 0; ps < npass && e == hipSuccess; ++ps) {
        const int bits = std::min(kRadixBits, kbits - shift);
        const int bins = 1 << bits;
        const int o = (npass - 1 - ps) % 2 == 0 ? 1 : 0; // the last pass writes slot 1 (sorted)
        const int64_t nbins = (int64_t)bins * C;
        hipLaunchKernelGGL(k_radix_count, dim3(nblk), dim3(64 * kRadixWaves), 0, st, n, kin, kstride, shift, bins,
                           chunk_len, (int)C, cnt);
        e = hipGetLastError();
        if (e == hipSuccess) e = device_scan(cnt, nbins, bsum, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(64 * kRadixWaves), 0, st, n, kin, kstride, vin, shift,
                               bits, chunk_len, (int)C, cnt, kout[o], vout[o]);
            e = hipGetLastError();
        }
        kin = kout[o];
        vin = vout[o];
        kstride = 1;
        shift += bits;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bin_starts_sorted, dim3(grid_blocks(n + 1, 256, INT32_MAX)), dim3(256), 0, st, n, kin, K,
                           bin_start);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_fail(e, "stable sort");
    return GN_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

void default_tiles(int dim, int ns, const int64_t *nf, int *tile, int *gtile)
{
    host::default_tiles(dim, ns, nf, tile, gtile);
}

// Orders the records of one gather work item for conflict-free LDS reads in the
// gather.  Every tap adds the same offset to all lanes' staged-cell addresses, so if
// the lanes of each group of 16 start at cells of distinct residues mod 16 (the
// 16-byte cells map to 16 disjoint groups of four LDS banks), every tap of that group
// is free of bank conflicts.  Samples are placed by (rank within their residue class,
// residue): position = sum_v min(size_v, k) + #{v < r : size_v > k} for rank k and
// residue r.  Ranks are taken in record order with wave ballots, so the permutation is
// deterministic.  Each sample's result does not depend on its lane.
static_assert(kGatherChunk <= kGatherThreads, "one record per thread in the bank ordering");

template <int D>
__global__ __launch_bounds__(kGatherThreads) void k_bank_order(GatherArgs a, int4 *cell, float4 *frac)
{
    constexpr int kWaves = kGatherThreads / 64;
    __shared__ int s_cnt[kWaves][16];
    __shared__ int s_sz[16];
    const int4 item = a.items[blockIdx.x];
    const int n = item.z - item.y;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int org[3] = {0, 0, 0};
    {
        uint32_t r = (uint32_t)item.x; // tile index < 2^31
        for (int ax = D - 1; ax >= 0; --ax) {
            org[ax] = (int)(r % (uint32_t)a.ntiles[ax]) * a.tile[ax];
            r /= (uint32_t)a.ntiles[ax];
        }
    }
    int4 rc = make_int4(0, 0, 0, 0);
    float4 rf = make_float4(0.f, 0.f, 0.f, 0.f);
    int res = 16; // no sample
    if (tid < n) {
        rc = cell[item.y + tid];
        rf = frac[item.y + tid];
        int l0 = rc.x - org[0];
        if (l0 < 0) l0 += (int)a.nf[0];
        int l1 = rc.y - org[1];
        if (l1 < 0) l1 += (int)a.nf[1];
        int addr = l0 * a.region[1] + l1;
        if (D == 3) {
            int l2 = rc.z - org[2];
            if (l2 < 0) l2 += (int)a.nf[2];
            addr = addr * a.region[2] + l2;
        }
        res = addr & 15;
    }
    int rank = 0;
    for (int v = 0; v < 16; ++v) {
        const unsigned long long m = __ballot(res == v);
        if (res == v) rank = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_cnt[wave][v] = __popcll(m);
    }
    __syncthreads();
    if (tid < 16) {
        int sum = 0;
        for (int w = 0; w < kWaves; ++w) {
            const int c = s_cnt[w][tid];
            s_cnt[w][tid] = sum;
            sum += c;
        }
        s_sz[tid] = sum;
    }
    __syncthreads();
    if (tid < n) {
        const int k = s_cnt[wave][res] + rank;
        int pos = 0;
        for (int v = 0; v < 16; ++v) {
            const int sz = s_sz[v];
            pos += min(sz, k) + ((v < res && sz > k) ? 1 : 0);
        }
        cell[item.y + pos] = rc; // every record of the item was read before the barriers
        frac[item.y + pos] = rf;
    }
}

#define GN_HIP(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call);                                                              \
    } while (0)

int bin_and_sort(gn_plan_s *p, int64_t M, const void *x, const void *y, const void *z, int coord_f64, double df,
                 const float *dcf)
{
    Geometry &g = p->geo;
    Records &R = p->rec;
    hipStream_t st = p->stream;
    const int64_t K = g.K, Kg = g.Kg;
    int rc;
    if (M > R.cap) {
        if ((rc = dev_alloc(&R.cell, (size_t)M, "records")) || (rc = dev_alloc(&R.frac, (size_t)M, "records")))
            return rc;
        R.cap = M;
    }
    if ((rc = grow_alloc(&R.bin_start, &R.bin_start_cap, (size_t)Kg + 1, "bin offsets"))) return rc;
    R.npairs = 0;

    Scratch sc(st);
    std::vector<int> h_home(Kg + 1, 0);
    int counts[4] = {0, 0, 0, 0}; // items of tiles with pairs, partial slots, reduce items, empty tiles
    int *tile_pairs = nullptr; // K + 1 offsets of each spread tile's pairs (scratch)

    if (M > 0) {
        int4 *cell = nullptr;
        float4 *frac = nullptr;
        int *sorted = nullptr;
        if ((rc = sc.alloc(&cell, (size_t)M, "sort scratch")) || (rc = sc.alloc(&frac, (size_t)M, "sort scratch")) ||
            (rc = sc.alloc(&sorted, (size_t)M, "sort scratch")))
            return rc;
        PrepArgs pa;
        pa.M = M;
        pa.x[0] = x;
        pa.x[1] = y;
        pa.x[2] = z;
        pa.coord_f64 = coord_f64;
        pa.df = df;
        pa.dim = g.dim;
        for (int a = 0; a < 3; ++a) {
            pa.nf[a] = g.nf[a];
            pa.tile[a] = g.gtile[a];
            pa.ntiles[a] = g.ngtiles[a];
        }
        pa.ns = g.ns;
        pa.cell = cell;
        pa.frac = frac;
        pa.err = p->err_flag;
        GN_HIP(hipMemsetAsync(p->err_flag, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_prep, dim3(grid_blocks(M, 256, INT32_MAX)), dim3(256), 0, st, pa);
        GN_HIP(hipGetLastError());
        // gather-tile sort (stable) -> records in tile order
        if ((rc = stable_sort_by_key(M, &cell[0].w, 4, Kg, sorted, R.bin_start, st, sc))) return rc;
        hipLaunchKernelGGL(k_make_records, dim3(grid_blocks(M, 256, INT32_MAX)), dim3(256), 0, st, M, sorted, cell,
                           frac, dcf, R.cell, R.frac);
        GN_HIP(hipGetLastError());

        // (spread tile, sample) pairs: keys in record order, sorted by tile, records rebuilt in tile order
        TileGeom tg;
        tg.dim = g.dim;
        for (int a = 0; a < 3; ++a) {
            tg.nf[a] = g.nf[a];
            tg.tile[a] = g.tile[a];
            tg.ntiles[a] = g.ntiles[a];
        }
        tg.W = 2 * g.ns;
        int *pcnt = nullptr, *pbsum = nullptr;
        const int64_t nbp = (M + kScanTile - 1) / kScanTile;
        if ((rc = sc.alloc(&pcnt, (size_t)M, "pair counts")) || (rc = sc.alloc(&pbsum, (size_t)nbp + 1, "pair scan")))
            return rc;
        hipLaunchKernelGGL(k_pair_count, dim3(grid_blocks(M, 256, INT32_MAX)), dim3(256), 0, st, M, R.cell, tg, pcnt);
        GN_HIP(hipGetLastError());
        // the pair total is the scan's grand total (pbsum[nb]); one host round trip
        // returns it together with the coordinate error flag
        GN_HIP(device_scan(pcnt, M, pbsum, st));
        int hv[2] = {0, 0};
        GN_HIP(hipMemcpyAsync(&hv[0], pbsum + nbp, sizeof(int), hipMemcpyDeviceToHost, st));
        GN_HIP(hipMemcpyAsync(&hv[1], p->err_flag, sizeof(int), hipMemcpyDeviceToHost, st));
        GN_HIP(hipStreamSynchronize(st));
        if (hv[1]) return fail(GN_EINVAL, "non-finite sample coordinate");
        const int64_t P = (int64_t)(unsigned)hv[0];
        if (P > (int64_t)INT32_MAX) return fail(GN_EINVAL, "too many (tile, sample) pairs");
        int *pkey = nullptr, *psamp = nullptr, *psorted = nullptr, *pstart = nullptr;
        if ((rc = sc.alloc(&pkey, (size_t)P, "pairs")) || (rc = sc.alloc(&psamp, (size_t)P, "pairs")) ||
            (rc = sc.alloc(&psorted, (size_t)P, "pairs")) || (rc = sc.alloc(&pstart, (size_t)K + 1, "pairs")))
            return rc;
        hipLaunchKernelGGL(k_pair_keys, dim3(grid_blocks(M, 256, INT32_MAX)), dim3(256), 0, st, M, R.cell, tg, pcnt,
                           pkey, psamp);
        GN_HIP(hipGetLastError());
        if ((rc = stable_sort_by_key(P, pkey, 1, K, psorted, pstart, st, sc))) return rc;
        if (P > R.pairs_cap) {
            if ((rc = dev_alloc(&R.pair_a, (size_t)P, "pairs")) || (rc = dev_alloc(&R.pair_b, (size_t)P, "pairs")))
                return rc;
            R.pairs_cap = P;
        }
        hipLaunchKernelGGL(k_pair_records, dim3(grid_blocks(P, 256, INT32_MAX)), dim3(256), 0, st, P, psorted, psamp,
                           pcnt, R.cell, R.frac, tg, R.pair_a, R.pair_b);
        GN_HIP(hipGetLastError());
        tile_pairs = pstart;
        R.npairs = P;
    } else {
        GN_HIP(hipMemsetAsync(R.bin_start, 0, sizeof(int) * (Kg + 1), st));
        if ((rc = sc.alloc(&tile_pairs, (size_t)K + 1, "pairs"))) return rc;
        GN_HIP(hipMemsetAsync(tile_pairs, 0, sizeof(int) * (K + 1), st));
    }

    // spread work items, partial slots, reduce items and occupancy masks on the device;
    // their counts come back with the gather bin offsets in one synchronisation
    {
        const int64_t P = R.npairs;
        const int64_t cap_items = K + P / kSpreadChunk + 1, cap_red = std::min<int64_t>(K, P / kSpreadChunk + 1);
        const int64_t ncolt = (int64_t)g.ntiles[1] * g.ntiles[2];
        if ((rc = grow_alloc(&R.spread_items, &R.spread_items_cap, (size_t)cap_items, "work items")) ||
            (rc = grow_alloc(&R.reduce_items, &R.reduce_items_cap, (size_t)cap_red, "work items")) ||
            (rc = grow_alloc(&R.occ_mask, &R.occ_mask_cap, (size_t)ncolt, "tile masks")))
            return rc;
        const int64_t nbk = (K + kScanTile - 1) / kScanTile;
        int *ni = nullptr, *ne = nullptr, *ns = nullptr, *nr = nullptr;
        int *bi = nullptr, *be = nullptr, *bs = nullptr, *br = nullptr;
        if ((rc = sc.alloc(&ni, (size_t)K, "work items")) || (rc = sc.alloc(&ne, (size_t)K, "work items")) ||
            (rc = sc.alloc(&ns, (size_t)K, "work items")) || (rc = sc.alloc(&nr, (size_t)K, "work items")) ||
            (rc = sc.alloc(&bi, (size_t)nbk + 1, "work items")) || (rc = sc.alloc(&be, (size_t)nbk + 1, "work items")) ||
            (rc = sc.alloc(&bs, (size_t)nbk + 1, "work items")) || (rc = sc.alloc(&br, (size_t)nbk + 1, "work items")))
            return rc;
        const dim3 gk(grid_blocks(K, 256, INT32_MAX));
        hipLaunchKernelGGL(k_item_counts, gk, dim3(256), 0, st, K, tile_pairs, kSpreadChunk, ni, ne, ns, nr);
        GN_HIP(hipGetLastError());
        GN_HIP(device_scan(ni, K, bi, st));
        GN_HIP(device_scan(ne, K, be, st));
        GN_HIP(device_scan(ns, K, bs, st));
        GN_HIP(device_scan(nr, K, br, st));
        hipLaunchKernelGGL(k_item_write, gk, dim3(256), 0, st, K, tile_pairs, kSpreadChunk, ni, ne, bi + nbk, ns, nr,
                           R.spread_items, R.reduce_items);
        GN_HIP(hipGetLastError());
        R.mask_k = (g.ntiles[0] + 63) / 64;
        hipLaunchKernelGGL(k_occ_masks, dim3(grid_blocks(ncolt, 256, INT32_MAX)), dim3(256), 0, st, ncolt, g.ntiles[0],
                           R.mask_k, tile_pairs, R.occ_mask);
        GN_HIP(hipGetLastError());
        GN_HIP(hipMemcpyAsync(&counts[0], bi + nbk, sizeof(int), hipMemcpyDeviceToHost, st));
        GN_HIP(hipMemcpyAsync(&counts[1], bs + nbk, sizeof(int), hipMemcpyDeviceToHost, st));
        GN_HIP(hipMemcpyAsync(&counts[2], br + nbk, sizeof(int), hipMemcpyDeviceToHost, st));
        GN_HIP(hipMemcpyAsync(&counts[3], be + nbk, sizeof(int), hipMemcpyDeviceToHost, st));
        GN_HIP(hipMemcpyAsync(h_home.data(), R.bin_start, sizeof(int) * (Kg + 1), hipMemcpyDeviceToHost, st));
        GN_HIP(hipStreamSynchronize(st));
    }

    // work items: every spread tile gets at least one item (empty tiles write zeros),
    // and the occupancy row masks for the first adjoint FFT pass (the spread leaves
    // tiles without pairs unwritten where that pass skips them)
    std::vector<host::Item> gitems;
    host::build_gather_items(Kg, h_home.data(), kGatherChunk, gitems);
    if ((rc = grow_alloc(&R.gather_items, &R.gather_items_cap, gitems.size(), "work items"))) return rc;
    if (!gitems.empty())
        GN_HIP(hipMemcpy(R.gather_items, gitems.data(), sizeof(int4) * gitems.size(), hipMemcpyHostToDevice));
    R.n_spread_items = counts[0] + counts[3];
    R.n_spread_full = counts[0];
    R.n_slots = counts[1];
    R.n_reduce_items = counts[2];
    R.n_gather_items = (int)gitems.size();
    const int slots = counts[1];
    if (R.n_gather_items > 0) {
        const GatherArgs ga = make_gather_args(p, nullptr);
        if (g.dim == 3)
            hipLaunchKernelGGL(k_bank_order<3>, dim3((unsigned)R.n_gather_items), dim3(kGatherThreads), 0, st, ga,
                               R.cell, R.frac);
        else
            hipLaunchKernelGGL(k_bank_order<2>, dim3((unsigned)R.n_gather_items), dim3(kGatherThreads), 0, st, ga,
                               R.cell, R.frac);
        GN_HIP(hipGetLastError());
    }
    const int64_t TC = (int64_t)g.tile[0] * g.tile[1] * g.tile[2];
    const int64_t need = (int64_t)slots * p->nc * TC;
    if (need > R.partial_cap) {
        if ((rc = dev_alloc(&R.partial, (size_t)need, "partial tiles"))) return rc;
        R.partial_cap = need;
    }
    return GN_OK;
}

namespace {

// Raise the dynamic-LDS limit only for launches that need more than the default 64 KiB.
template <typename K>
void lds_limit(K kernel, size_t lds)
{
    if (lds > 64 * 1024) (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

#ifndef GN_SPREAD_MFMA
#define GN_SPREAD_MFMA 1
#endif
// The matrix-core spread covers whole 64-cell blocks: tiles must be multiples of the
// block (4 per axis in 3-D, 8 in 2-D) of at most kSpreadWavesMax blocks, so that one
// workgroup holds the whole tile.
bool mfma_tiles_ok(const SpreadArgs &a, int dim)
{
    if (!GN_SPREAD_MFMA) return false;
    const int q = dim == 3 ? 4 : 8;
    int nb = 1;
    for (int ax = 0; ax < dim; ++ax) {
        if (a.blk[ax] != q || a.tile[ax] % q != 0) return false;
        nb *= a.tile[ax] / q;
    }
    // phase 2 maps one table column per thread: 64 * waves >= TT rounded up to a power of 2
    const int ttp = a.TT <= 16 ? 16 : a.TT <= 32 ? 32 : a.TT <= 64 ? 64 : 128;
    const int waves = std::min(nb, kSpreadWavesMax);
    // one workgroup per (tile, coil group): it writes the whole tile from its LDS
    return nb <= kSpreadWavesMax && a.TT <= 128 && 64 * waves >= ttp && spread_mfma_lds_bytes(a, waves) <= 160 * 1024;
}

inline size_t spread_lds_bytes(const SpreadArgs &a, int ncg)
{
    return (size_t)kSpreadBatch * a.TT * 4 + (size_t)kSpreadBatch * ncg * 8 + (size_t)kSpreadBatch * 16 +
           (size_t)kSpreadBatch * 32;
}

// the dispatch-order arithmetic of the spread and gather kernels is 32-bit
inline bool grid_ok(dim3 g) { return (uint64_t)g.x * g.y * g.z < (1ull << 31); }

template <int D, int NCG>
hipError_t spread_launch(const SpreadArgs &a, int nitems, int64_t nc, hipStream_t st)
{
    const int ng = (int)((nc + NCG - 1) / NCG);
    const int nb = a.nblk[0] * a.nblk[1] * (D == 3 ? a.nblk[2] : 1);
    const int waves = std::min(nb, kSpreadWavesMax);
    const int nbg = (nb + waves - 1) / waves;
    if (nitems == 0) return hipSuccess;
    const dim3 grid((unsigned)nitems, (unsigned)ng, (unsigned)nbg);
    if (!grid_ok(grid)) return hipErrorInvalidConfiguration;
    if (mfma_tiles_ok(a, D)) {
        const size_t lds = spread_mfma_lds_bytes(a, waves);
        lds_limit(k_spread_mfma<D, NCG>, lds);
        hipLaunchKernelGGL((k_spread_mfma<D, NCG>), grid, dim3(64 * waves), lds, st, a);
    } else {
        const size_t lds = spread_lds_bytes(a, NCG);
        lds_limit(k_spread<D, NCG>, lds);
        hipLaunchKernelGGL((k_spread<D, NCG>), grid, dim3(64 * waves), lds, st, a);
    }
    return hipGetLastError();
}

template <int D>
hipError_t spread_dispatch(const SpreadArgs &a, int nitems, int nritems, const int4 *ritems, int64_t nc,
                           hipStream_t st)
{
    hipError_t e;
    if (nc >= 8)
        e = spread_launch<D, 8>(a, nitems, nc, st);
    else if (nc >= 4)
        e = spread_launch<D, 4>(a, nitems, nc, st);
    else if (nc >= 2)
        e = spread_launch<D, 2>(a, nitems, nc, st);
    else
        e = spread_launch<D, 1>(a, nitems, nc, st);
    if (e == hipSuccess && nritems > 0) {
        hipLaunchKernelGGL(k_spread_reduce<D>, dim3((unsigned)nritems), dim3(kReduceThreads), 0, st, a, ritems);
        e = hipGetLastError();
    }
    return e;
}

template <int D, int NS>
hipError_t gather_dispatch(const GatherArgs &a, int nitems, int64_t nc, hipStream_t st)
{
    const size_t lds = (size_t)a.region[0] * a.region[1] * (D == 3 ? a.region[2] : 1) * kGatherCellBytes;
    bool wrap = false;
    for (int ax = 0; ax < D; ++ax) wrap |= a.region[ax] >= a.nf[ax];
    dim3 grid((unsigned)nitems, (unsigned)((nc + kGatherCoils - 1) / kGatherCoils));
    if (!grid_ok(grid)) return hipErrorInvalidConfiguration;
    if (wrap) {
        lds_limit(k_gather<D, NS, true>, lds);
        hipLaunchKernelGGL((k_gather<D, NS, true>), grid, dim3(kGatherThreads), lds, st, a);
    } else {
        lds_limit(k_gather<D, NS, false>, lds);
        hipLaunchKernelGGL((k_gather<D, NS, false>), grid, dim3(kGatherThreads), lds, st, a);
    }
    return hipGetLastError();
}

#define GN_NS_CASES(D, FN, ...)                                                                                        \
    switch (ns) {                                                                                                      \
    case 2: return FN<D, 2>(__VA_ARGS__);                                                                              \
    case 3: return FN<D, 3>(__VA_ARGS__);                                                                              \
    case 4: return FN<D, 4>(__VA_ARGS__);                                                                              \
    case 5: return FN<D, 5>(__VA_ARGS__);                                                                              \
    case 6: return FN<D, 6>(__VA_ARGS__);                                                                              \
    case 7: return FN<D, 7>(__VA_ARGS__);                                                                              \
    case 8: return FN<D, 8>(__VA_ARGS__);                                                                              \
    default: break;                                                                                                    \
    }

#define GN_NS_CASES_2D_EXTRA(FN, ...)                                                                                  \
    switch (ns) {                                                                                                      \
    case 9: return FN<2, 9>(__VA_ARGS__);                                                                              \
    case 10: return FN<2, 10>(__VA_ARGS__);                                                                            \
    case 11: return FN<2, 11>(__VA_ARGS__);                                                                            \
    case 12: return FN<2, 12>(__VA_ARGS__);                                                                            \
    case 13: return FN<2, 13>(__VA_ARGS__);                                                                            \
    case 14: return FN<2, 14>(__VA_ARGS__);                                                                            \
    case 15: return FN<2, 15>(__VA_ARGS__);                                                                            \
    case 16: return FN<2, 16>(__VA_ARGS__);                                                                            \
    default: break;                                                                                                    \
    }

hipError_t spread_any(int dim, int ns, const SpreadArgs &a, int nitems, int nritems, const int4 *ritems, int64_t nc,
                      hipStream_t st)
{
    (void)ns;
    if (dim == 3) return spread_dispatch<3>(a, nitems, nritems, ritems, nc, st);
    return spread_dispatch<2>(a, nitems, nritems, ritems, nc, st);
}

hipError_t gather_any(int dim, int ns, const GatherArgs &a, int nitems, int64_t nc, hipStream_t st)
{
    if (dim == 3) {
        GN_NS_CASES(3, gather_dispatch, a, nitems, nc, st)
    } else {
        GN_NS_CASES(2, gather_dispatch, a, nitems, nc, st)
        GN_NS_CASES_2D_EXTRA(gather_dispatch, a, nitems, nc, st)
    }
    return hipErrorInvalidValue;
}

PrepostArgs make_prepost_impl(const gn_plan_s *p, float scale)
{
    const Geometry &g = p->geo;
    PrepostArgs a;
    a.dim = g.dim;
    for (int ax = 0; ax < 3; ++ax) {
        a.modes[ax] = g.modes[ax];
        a.nf[ax] = g.nf[ax];
    }
    a.G = g.G;
    a.N = g.N;
    a.nc = p->nc;
    a.layout = p->layout;
    a.scale = scale;
    a.deap0 = p->deap;
    a.deap1 = p->deap + g.modes[0];
    a.deap2 = p->deap + g.modes[0] + g.modes[1];
    a.src = nullptr;
    a.dst = nullptr;
    return a;
}

} // namespace

SpreadArgs make_spread_args(const gn_plan_s *p, const float2 *samples, int use_dcf)
{
    const Geometry &g = p->geo;
    SpreadArgs a;
    for (int ax = 0; ax < 3; ++ax) {
        a.nf[ax] = g.nf[ax];
        a.tile[ax] = g.tile[ax];
        a.ntiles[ax] = g.ntiles[ax];
        a.coef[ax] = g.coef[ax];
    }
    for (int ax = 0; ax < 3; ++ax) {
        a.blk[ax] = ax < g.dim ? (g.dim == 3 ? 4 : 8) : 1;
        a.nblk[ax] = (g.tile[ax] + a.blk[ax] - 1) / a.blk[ax];
    }
    a.G = g.G;
    a.M = p->M;
    a.nc = p->nc;
    a.TC = g.tile[0] * g.tile[1] * g.tile[2];
    a.TT = g.tile[0] + g.tile[1] + (g.dim == 3 ? g.tile[2] : 0);
    a.ns = g.ns;
    a.layout = p->layout;
    a.use_dcf = use_dcf && p->has_dcf;
    a.items = p->rec.spread_items;
    a.reduce_items = p->rec.reduce_items;
    a.pair_a = p->rec.pair_a;
    a.pair_b = p->rec.pair_b;
    a.in = samples;
    a.grid = p->grid;
    a.partial = p->rec.partial;
    return a;
}

int launch_spread(gn_plan_s *p, const float2 *samples, int use_dcf, float2 *grid)
{
    const Geometry &g = p->geo;
    SpreadArgs a = make_spread_args(p, samples, use_dcf);
    int nitems = p->rec.n_spread_items;
    if (grid) {
        a.grid = grid; // every tile written (no pass skips cells of this grid)
    } else if (p->fft_custom && p->fft_engine == GN_FFT_AUTO) {
        // the pruned pass C' reads only occupied tiles: empty ones need no zeros.  With
        // one mask bit per tile every empty tile is skipped, so the empty tiles' items
        // (after the others) are not launched at all.
        a.occ = p->rec.occ_mask;
        a.occ_k = p->rec.mask_k;
        if (p->rec.mask_k == 1) nitems = p->rec.n_spread_full;
    }
    hipError_t e = spread_any(g.dim, g.ns, a, nitems, p->rec.n_reduce_items, p->rec.reduce_items, p->nc, p->stream);
    if (e != hipSuccess) return hip_fail(e, "spread launch");
    return GN_OK;
}

GatherArgs make_gather_args(const gn_plan_s *p, float2 *samples)
{
    const Geometry &g = p->geo;
    GatherArgs a;
    for (int ax = 0; ax < 3; ++ax) {
        a.nf[ax] = g.nf[ax];
        a.tile[ax] = g.gtile[ax];
        a.ntiles[ax] = g.ngtiles[ax];
        a.region[ax] = (int)std::min<int64_t>(g.gtile[ax] + 2 * g.ns - 1, g.nf[ax]);
        a.coef[ax] = g.coef[ax];
    }
    a.G = g.G;
    a.M = p->M;
    a.nc = p->nc;
    a.layout = p->layout;
    a.items = p->rec.gather_items;
    a.rec_cell = p->rec.cell;
    a.rec_frac = p->rec.frac;
    a.grid = p->grid;
    a.out = samples;
    return a;
}

int launch_gather(gn_plan_s *p, float2 *samples, const float2 *grid, float scale)
{
    const Geometry &g = p->geo;
    if (p->M == 0 || p->rec.n_gather_items == 0) return GN_OK;
    GatherArgs a = make_gather_args(p, samples);
    if (grid) a.grid = grid;
    a.scale = scale;
    hipError_t e = gather_any(g.dim, g.ns, a, p->rec.n_gather_items, p->nc, p->stream);
    if (e != hipSuccess) return hip_fail(e, "gather launch");
    return GN_OK;
}

namespace {
template <int D, int NS>
hipError_t traj_grad_dispatch(const GatherArgs &a, int nitems, const float2 *wts, float *out, const float *gs,
                              hipStream_t st)
{
    const size_t lds = (size_t)a.region[0] * a.region[1] * (D == 3 ? a.region[2] : 1) * sizeof(float2);
    bool wrap = false;
    for (int ax = 0; ax < D; ++ax) wrap |= a.region[ax] >= a.nf[ax];
    if (wrap) {
        lds_limit(k_traj_grad<D, NS, true>, lds);
        hipLaunchKernelGGL((k_traj_grad<D, NS, true>), dim3((unsigned)nitems), dim3(kGatherThreads), lds, st, a, wts,
                           out, gs[0], gs[1], gs[2]);
    } else {
        lds_limit(k_traj_grad<D, NS, false>, lds);
        hipLaunchKernelGGL((k_traj_grad<D, NS, false>), dim3((unsigned)nitems), dim3(kGatherThreads), lds, st, a,
                           wts, out, gs[0], gs[1], gs[2]);
    }
    return hipGetLastError();
}

hipError_t traj_grad_any(int dim, int ns, const GatherArgs &a, int nitems, const float2 *wts, float *out,
                         const float *gs, hipStream_t st)
{
    if (dim == 3) {
        GN_NS_CASES(3, traj_grad_dispatch, a, nitems, wts, out, gs, st)
    } else {
        GN_NS_CASES(2, traj_grad_dispatch, a, nitems, wts, out, gs, st)
        GN_NS_CASES_2D_EXTRA(traj_grad_dispatch, a, nitems, wts, out, gs, st)
    }
    return hipErrorInvalidValue;
}
} // namespace

int launch_traj_grad(gn_plan_s *p, const float2 *wts, float *out)
{
    const Geometry &g = p->geo;
    if (p->M == 0 || p->rec.n_gather_items == 0) return GN_OK;
    GatherArgs a;
    for (int ax = 0; ax < 3; ++ax) {
        a.nf[ax] = g.nf[ax];
        a.tile[ax] = g.gtile[ax];
        a.ntiles[ax] = g.ngtiles[ax];
        a.region[ax] = (int)std::min<int64_t>(g.gtile[ax] + 2 * g.ns - 1, g.nf[ax]);
        a.coef[ax] = g.coef[ax];
    }
    a.G = g.G;
    a.M = p->M;
    a.nc = p->nc;
    a.layout = p->layout;
    a.items = p->rec.gather_items;
    a.rec_cell = p->rec.cell;
    a.rec_frac = p->rec.frac;
    a.grid = p->grid;
    a.out = nullptr;
    float gs[3] = {0.f, 0.f, 0.f};
    const double pi = 3.141592653589793;
    for (int ax = 0; ax < g.dim; ++ax) gs[ax] = (float)(-(2 * pi / (double)g.nf[ax]) * p->df / (2.0 * p->tau));
    hipError_t e = traj_grad_any(g.dim, g.ns, a, p->rec.n_gather_items, wts, out, gs, p->stream);
    if (e != hipSuccess) return hip_fail(e, "trajectory gradient launch");
    return GN_OK;
}

PrepostArgs make_prepost(const gn_plan_s *p, float scale) { return make_prepost_impl(p, scale); }

int launch_pad(gn_plan_s *p, const float2 *image, float scale, bool sense)
{
    PrepostArgs a = make_prepost(p, scale);
    if (sense) a.sens = p->sens;
    dim3 grid(grid_blocks(p->geo.G, 256, 1 << 16));
    if (p->geo.dim == 3)
        hipLaunchKernelGGL(k_pad<3>, grid, dim3(256), 0, p->stream, a, image, p->grid);
    else
        hipLaunchKernelGGL(k_pad<2>, grid, dim3(256), 0, p->stream, a, image, p->grid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pad launch");
    return GN_OK;
}

int launch_crop(gn_plan_s *p, float2 *image, float scale, bool sense)
{
    PrepostArgs a = make_prepost(p, scale);
    if (sense) a.sens = p->sens;
    dim3 grid(grid_blocks(p->geo.N, 256, 1 << 16));
    if (p->geo.dim == 3)
        hipLaunchKernelGGL(k_crop<3>, grid, dim3(256), 0, p->stream, a, p->grid, image);
    else
        hipLaunchKernelGGL(k_crop<2>, grid, dim3(256), 0, p->stream, a, p->grid, image);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "crop launch");
    return GN_OK;
}

// x[n] = sum_c conj(S[n, c]) * img[n, c] (coils in order), for SENSE combines whose
// coils do not fit one row-pass block
__global__ __launch_bounds__(256) void k_sense_combine(int64_t N, int64_t nc, int layout, const float2 *__restrict__ sens,
                                                       const float2 *__restrict__ img, float2 *__restrict__ x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += stride) {
        float2 acc = make_float2(0.f, 0.f);
        for (int64_t c = 0; c < nc; ++c) {
            const int64_t i = layout == GN_COIL_LAST ? n * nc + c : c * N + n;
            const float2 s = sens[i], o = img[i];
            acc.x += s.x * o.x + s.y * o.y;
            acc.y += s.x * o.y - s.y * o.x;
        }
        x[n] = acc;
    }
}

int launch_sense_combine(gn_plan_s *p, const float2 *coil_images, float2 *image)
{
    hipLaunchKernelGGL(k_sense_combine, dim3(grid_blocks(p->geo.N, 256, 1 << 16)), dim3(256), 0, p->stream, p->geo.N,
                       p->nc, p->layout, p->sens, coil_images, image);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GN_OK : hip_fail(e, "SENSE combine");
}

// ---------------------------------------------------------------------------
// Batched launchers: one launch per stage over the work of many plans.
// ---------------------------------------------------------------------------

void spread_shape(const SpreadArgs &a, int dim, int *waves, int *nbg)
{
    const int nb = a.nblk[0] * a.nblk[1] * (dim == 3 ? a.nblk[2] : 1);
    *waves = std::min(nb, kSpreadWavesMax);
    *nbg = (nb + *waves - 1) / *waves;
}

bool spread_uses_mfma(const SpreadArgs &a, int dim) { return mfma_tiles_ok(a, dim); }

size_t spread_lds(const SpreadArgs &a, int ncg, bool mfma, int waves)
{
    return mfma ? spread_mfma_lds_bytes(a, waves) : spread_lds_bytes(a, ncg);
}

namespace {
template <int D, int NCG>
hipError_t spread_batch_launch(bool mfma, const SpreadArgs *d_args, const int4 *d_items, int n, int waves, int nbg,
                               size_t lds, hipStream_t st)
{
    if (!grid_ok(dim3((unsigned)n, (unsigned)nbg))) return hipErrorInvalidConfiguration;
    if (mfma) {
        lds_limit(k_spread_mfma_batch<D, NCG>, lds);
        hipLaunchKernelGGL((k_spread_mfma_batch<D, NCG>), dim3((unsigned)n, (unsigned)nbg), dim3(64 * waves), lds, st,
                           d_args, d_items);
    } else {
        lds_limit(k_spread_batch<D, NCG>, lds);
        hipLaunchKernelGGL((k_spread_batch<D, NCG>), dim3((unsigned)n, (unsigned)nbg), dim3(64 * waves), lds, st,
                           d_args, d_items);
    }
    return hipGetLastError();
}

template <int D, int NS>
hipError_t gather_batch_dispatch(bool wrap, const GatherArgs *d_args, const int4 *d_items, int n, size_t lds,
                                 hipStream_t st)
{
    if (wrap) {
        lds_limit(k_gather_batch<D, NS, true>, lds);
        hipLaunchKernelGGL((k_gather_batch<D, NS, true>), dim3((unsigned)n), dim3(kGatherThreads), lds, st, d_args,
                           d_items);
    } else {
        lds_limit(k_gather_batch<D, NS, false>, lds);
        hipLaunchKernelGGL((k_gather_batch<D, NS, false>), dim3((unsigned)n), dim3(kGatherThreads), lds, st, d_args,
                           d_items);
    }
    return hipGetLastError();
}
} // namespace

hipError_t launch_spread_batch(int dim, int ncg, bool mfma, const SpreadArgs *d_args, const int4 *d_items, int n,
                               int waves, int nbg, size_t lds, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (dim == 3) {
        switch (ncg) {
        case 8: return spread_batch_launch<3, 8>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        case 4: return spread_batch_launch<3, 4>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        case 2: return spread_batch_launch<3, 2>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        case 1: return spread_batch_launch<3, 1>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        }
    } else {
        switch (ncg) {
        case 8: return spread_batch_launch<2, 8>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        case 4: return spread_batch_launch<2, 4>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        case 2: return spread_batch_launch<2, 2>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        case 1: return spread_batch_launch<2, 1>(mfma, d_args, d_items, n, waves, nbg, lds, st);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_spread_reduce_batch(int dim, const SpreadArgs *d_args, const int4 *d_items, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (dim == 3)
        hipLaunchKernelGGL(k_spread_reduce_batch<3>, dim3((unsigned)n), dim3(kReduceThreads), 0, st, d_args, d_items);
    else
        hipLaunchKernelGGL(k_spread_reduce_batch<2>, dim3((unsigned)n), dim3(kReduceThreads), 0, st, d_args, d_items);
    return hipGetLastError();
}

hipError_t launch_gather_batch(int dim, int ns, bool wrap, const GatherArgs *d_args, const int4 *d_items, int n,
                               size_t lds, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (dim == 3) {
        GN_NS_CASES(3, gather_batch_dispatch, wrap, d_args, d_items, n, lds, st)
    } else {
        GN_NS_CASES(2, gather_batch_dispatch, wrap, d_args, d_items, n, lds, st)
        GN_NS_CASES_2D_EXTRA(gather_batch_dispatch, wrap, d_args, d_items, n, lds, st)
    }
    return hipErrorInvalidValue;
}

hipError_t launch_pad_batch(int dim, const PrepostArgs *d_args, int ninst, int64_t maxG, hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    dim3 grid(grid_blocks(maxG, 256, 1 << 14), (unsigned)ninst);
    if (dim == 3)
        hipLaunchKernelGGL(k_pad_batch<3>, grid, dim3(256), 0, st, d_args);
    else
        hipLaunchKernelGGL(k_pad_batch<2>, grid, dim3(256), 0, st, d_args);
    return hipGetLastError();
}

hipError_t launch_crop_batch(int dim, const PrepostArgs *d_args, int ninst, int64_t maxN, hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    dim3 grid(grid_blocks(maxN, 256, 1 << 14), (unsigned)ninst);
    if (dim == 3)
        hipLaunchKernelGGL(k_crop_batch<3>, grid, dim3(256), 0, st, d_args);
    else
        hipLaunchKernelGGL(k_crop_batch<2>, grid, dim3(256), 0, st, d_args);
    return hipGetLastError();
}

} // namespace gn
