// HIP kernels of the Gaussian-gridding NUFFT for CDNA4 (gfx950).
//
// Data layout in HBM (per plan):
//   grid   : nc oversampled grids, coil-major, C order (nf0, nf1[, nf2]), complex64
//   records: samples sorted by home tile (tile containing the window start cell)
//            cell = {s0,s1,s2,orig}, frac = {t0,t1,t2,dcf}; bin_start[K+1]
//   image / samples at the boundary: coil-last (reference) or coil-major.
//
// Window convention (fft/nufft_func.py:293-295, 536-546): xi = x mod 2pi,
// m = 1 + xi // h, taps g = m - ns + j (j = 0..2ns-1) wrapped mod nf, weight
// exp(-(xi - h g)^2/(4 tau)) = exp2(-coef (t + ns - j)^2) with t = xi/h - m.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "gn_internal.h"

namespace gn {

namespace {

constexpr double kTwoPi = 6.283185307179586; // 2*np.pi in float64

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
        py_divmod(x, kTwoPi, &xi);                 // xi = x % (2 pi)
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
// Stable counting sort of samples by home tile.
//   cnt[key*C + chunk]  : samples of chunk with that key (key-major)
//   exclusive scan      : offsets, key-major so chunk order is preserved
//   scatter             : each chunk places its samples in index order
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_bin_count(int64_t M, const int4 *__restrict__ cell, int64_t chunk_len, int C,
                                                   int *__restrict__ cnt)
{
    const int chunk = blockIdx.x;
    const int64_t beg = (int64_t)chunk * chunk_len;
    const int64_t end = min(M, beg + chunk_len);
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x)
        atomicAdd(&cnt[(int64_t)cell[i].w * C + chunk], 1);
}

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
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int *__restrict__ data, int64_t n,
                                                            const int *__restrict__ bsum)
{
    const int add = bsum[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
        if (base + k < n) data[base + k] += add;
}

__global__ __launch_bounds__(256) void k_bin_starts(const int *__restrict__ offs, int64_t K, int C, int64_t M,
                                                    int *__restrict__ bin_start)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) bin_start[k] = offs[k * C];
    if (k == K) bin_start[K] = (int)M;
}

// One wave per chunk; samples are placed group by group in index order, so
// equal keys keep their original order (stable, deterministic).
__global__ __launch_bounds__(64) void k_bin_scatter(int64_t M, const int4 *__restrict__ cell, int64_t chunk_len, int C,
                                                    int *__restrict__ offs, int *__restrict__ sorted)
{
    __shared__ int sk[64];
    const int lane = threadIdx.x;
    const int chunk = blockIdx.x;
    const int64_t beg = (int64_t)chunk * chunk_len;
    const int64_t end = min(M, beg + chunk_len);
    for (int64_t g = beg; g < end; g += 64) {
        const int64_t i = g + lane;
        const bool valid = i < end;
        const int key = valid ? cell[i].w : -1;
        sk[lane] = key;
        __syncthreads();
        int rank = 0;
        bool last = true;
        for (int l = 0; l < 64; ++l) {
            const int kl = sk[l];
            if (kl == key) {
                if (l < lane) ++rank;
                if (l > lane) last = false;
            }
        }
        int base = 0;
        if (valid) {
            base = offs[(int64_t)key * C + chunk];
            sorted[base + rank] = (int)i;
        }
        __syncthreads();
        if (valid && last) offs[(int64_t)key * C + chunk] = base + rank + 1;
        __syncthreads();
    }
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
// Adjoint spread: output-stationary LDS tiles, one wave per tile and coil.
//
// A wave owns one tile of the grid for one coil.  It visits the samples whose
// home tile is the tile itself or one of its predecessors (those are the only
// samples whose windows can reach it), in sorted order, one sample at a time;
// the lanes take the sample's taps.  Taps inside the tile update the tile's
// Kahan-compensated (sum, compensation) pair in LDS.  Within one sample every
// lane hits a distinct cell and only this wave touches this LDS region, so the
// summation order is fixed: the result is deterministic.  The tile is written
// to HBM once, coalesced along the fastest axis.  No global atomics.
// ---------------------------------------------------------------------------

struct SpreadArgs {
    int64_t nf[3];
    int tile[3];
    int ntiles[3];
    int npred[3];   // candidate home tiles per axis (tile itself + predecessors)
    int64_t K;
    int64_t G;
    int64_t M;
    int64_t nc;
    int layout;
    int use_dcf;
    float coef[3];
    const int *bin_start;
    const int4 *rec_cell;
    const float4 *rec_frac;
    const float2 *in;
    float2 *grid;
};

template <int D>
__device__ inline bool touches_axis(int s, int origin, int teff, int span, int64_t nf)
{
    int v = s - origin;
    if (v < 0) v += (int)nf;
    return v < teff || v + span - 1 >= nf;
}

template <int D, int NS>
__global__ __launch_bounds__(64 * kSpreadWaves) void k_spread(SpreadArgs a)
{
    constexpr int W = 2 * NS;
    constexpr int NT = (D == 3) ? W * W * W : W * W;
    extern __shared__ float4 lds_spread[];

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t tile_id = (int64_t)blockIdx.x * kSpreadWaves + wave;
    const int64_t coil = blockIdx.y;
    if (tile_id >= a.K) return; // wave-uniform; no block barriers below

    const int TC = a.tile[0] * a.tile[1] * (D == 3 ? a.tile[2] : 1);
    float4 *acc = lds_spread + (int64_t)wave * TC;

    // tile coordinates
    int q[3], org[3], teff[3];
    {
        int64_t r = tile_id;
        for (int ax = D - 1; ax >= 0; --ax) {
            q[ax] = (int)(r % a.ntiles[ax]);
            r /= a.ntiles[ax];
        }
        for (int ax = 0; ax < D; ++ax) {
            org[ax] = q[ax] * a.tile[ax];
            teff[ax] = (int)min((int64_t)a.tile[ax], a.nf[ax] - org[ax]);
        }
    }

    for (int c = lane; c < TC; c += 64) acc[c] = make_float4(0.f, 0.f, 0.f, 0.f);

    const int np0 = a.npred[0], np1 = a.npred[1], np2 = (D == 3) ? a.npred[2] : 1;
    for (int p0 = 0; p0 < np0; ++p0)
        for (int p1 = 0; p1 < np1; ++p1)
            for (int p2 = 0; p2 < np2; ++p2) {
                int bq0 = q[0] - p0, bq1 = q[1] - p1;
                if (bq0 < 0) bq0 += a.ntiles[0];
                if (bq1 < 0) bq1 += a.ntiles[1];
                int64_t bin = (int64_t)bq0 * a.ntiles[1] + bq1;
                if (D == 3) {
                    int bq2 = q[2] - p2;
                    if (bq2 < 0) bq2 += a.ntiles[2];
                    bin = bin * a.ntiles[2] + bq2;
                }
                const int bbeg = a.bin_start[bin];
                const int bend = a.bin_start[bin + 1];
                for (int base = bbeg; base < bend; base += 64) {
                    const int j = base + lane;
                    bool touch = false;
                    if (j < bend) {
                        const int4 rc = a.rec_cell[j];
                        touch = touches_axis<D>(rc.x, org[0], teff[0], W, a.nf[0]) &&
                                touches_axis<D>(rc.y, org[1], teff[1], W, a.nf[1]) &&
                                (D == 2 || touches_axis<D>(rc.z, org[2], teff[2], W, a.nf[2]));
                    }
                    unsigned long long mask = __ballot(touch);
                    while (mask) {
                        const int k = __ffsll((long long)mask) - 1;
                        mask &= mask - 1;
                        const int idx = base + k;
                        const int4 rc = a.rec_cell[idx];
                        const float4 rf = a.rec_frac[idx];
                        const int64_t orig = rc.w;
                        float2 v = (a.layout == GN_COIL_LAST) ? a.in[orig * a.nc + coil] : a.in[coil * a.M + orig];
                        if (a.use_dcf) {
                            v.x *= rf.w;
                            v.y *= rf.w;
                        }
                        for (int tap = lane; tap < NT; tap += 64) {
                            int j0, j1, j2 = 0;
                            if (D == 3) {
                                j2 = tap % W;
                                j1 = (tap / W) % W;
                                j0 = tap / (W * W);
                            } else {
                                j1 = tap % W;
                                j0 = tap / W;
                            }
                            int c0 = rc.x + j0;
                            if (c0 >= a.nf[0]) c0 -= (int)a.nf[0];
                            int l0 = c0 - org[0];
                            int c1 = rc.y + j1;
                            if (c1 >= a.nf[1]) c1 -= (int)a.nf[1];
                            int l1 = c1 - org[1];
                            bool in = (l0 >= 0) & (l0 < teff[0]) & (l1 >= 0) & (l1 < teff[1]);
                            int l2 = 0;
                            if (D == 3) {
                                int c2 = rc.z + j2;
                                if (c2 >= a.nf[2]) c2 -= (int)a.nf[2];
                                l2 = c2 - org[2];
                                in &= (l2 >= 0) & (l2 < teff[2]);
                            }
                            if (!in) continue;
                            const float u0 = rf.x + (float)(NS - j0);
                            const float u1 = rf.y + (float)(NS - j1);
                            float w = exp2f(-a.coef[0] * u0 * u0) * exp2f(-a.coef[1] * u1 * u1);
                            if (D == 3) {
                                const float u2 = rf.z + (float)(NS - j2);
                                w *= exp2f(-a.coef[2] * u2 * u2);
                            }
                            const int li = (D == 3) ? (l0 * a.tile[1] + l1) * a.tile[2] + l2 : l0 * a.tile[1] + l1;
                            float4 s = acc[li];
                            // Kahan: (s.x,s.y) running sum, (s.z,s.w) compensation
                            const float yr = v.x * w - s.z;
                            const float yi = v.y * w - s.w;
                            const float tr = s.x + yr;
                            const float ti = s.y + yi;
                            s.z = (tr - s.x) - yr;
                            s.w = (ti - s.y) - yi;
                            s.x = tr;
                            s.y = ti;
                            acc[li] = s;
                        }
                    }
                }
            }

    // write the tile once; consecutive lanes take consecutive cells of a row
    float2 *g = a.grid + coil * a.G;
    for (int c = lane; c < TC; c += 64) {
        int l0, l1, l2 = 0;
        if (D == 3) {
            l2 = c % a.tile[2];
            l1 = (c / a.tile[2]) % a.tile[1];
            l0 = c / (a.tile[2] * a.tile[1]);
        } else {
            l1 = c % a.tile[1];
            l0 = c / a.tile[1];
        }
        if (l0 >= teff[0] || l1 >= teff[1] || (D == 3 && l2 >= teff[2])) continue;
        int64_t gi = (int64_t)(org[0] + l0) * a.nf[1] + (org[1] + l1);
        if (D == 3) gi = gi * a.nf[2] + (org[2] + l2);
        const float4 s = acc[c];
        g[gi] = make_float2(s.x - s.z, s.y - s.w);
    }
}

// ---------------------------------------------------------------------------
// Forward gather: one workgroup per home tile; the tile plus its window halo
// is staged in LDS one coil at a time with coalesced loads, then every thread
// interpolates one sample from LDS with separable weights in registers.
// ---------------------------------------------------------------------------

struct GatherArgs {
    int64_t nf[3];
    int tile[3];
    int ntiles[3];
    int region[3]; // staged cells per axis
    int64_t K;
    int64_t G;
    int64_t M;
    int64_t nc;
    int layout;
    float coef[3];
    const int *bin_start;
    const int4 *rec_cell;
    const float4 *rec_frac;
    const float2 *grid;
    float2 *out;
};

template <int D, int NS>
__global__ __launch_bounds__(kGatherThreads) void k_gather(GatherArgs a)
{
    constexpr int W = 2 * NS;
    extern __shared__ float2 lds_gather[];
    const int64_t tile_id = blockIdx.x;
    const int bbeg = a.bin_start[tile_id];
    const int bend = a.bin_start[tile_id + 1];
    if (bbeg == bend) return;
    const int cg0 = blockIdx.y * kCoilGroup;
    const int ncg = (int)min((int64_t)kCoilGroup, a.nc - cg0);

    int q[3] = {0, 0, 0}, org[3] = {0, 0, 0};
    {
        int64_t r = tile_id;
        for (int ax = D - 1; ax >= 0; --ax) {
            q[ax] = (int)(r % a.ntiles[ax]);
            r /= a.ntiles[ax];
        }
        for (int ax = 0; ax < D; ++ax) org[ax] = q[ax] * a.tile[ax];
    }
    const int R0 = a.region[0], R1 = a.region[1], R2 = (D == 3) ? a.region[2] : 1;
    const int RC = R0 * R1 * R2;

    for (int cbase = bbeg; cbase < bend; cbase += kGatherThreads) {
        const int j = cbase + threadIdx.x;
        const bool valid = j < bend;
        float w0[W], w1[W], w2[W];
        int l0 = 0, l1 = 0, l2 = 0;
        int64_t orig = 0;
        if (valid) {
            const int4 rc = a.rec_cell[j];
            const float4 rf = a.rec_frac[j];
            orig = rc.w;
            l0 = rc.x - org[0];
            if (l0 < 0) l0 += (int)a.nf[0];
            l1 = rc.y - org[1];
            if (l1 < 0) l1 += (int)a.nf[1];
            if (D == 3) {
                l2 = rc.z - org[2];
                if (l2 < 0) l2 += (int)a.nf[2];
            }
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const float u0 = rf.x + (float)(NS - t);
                const float u1 = rf.y + (float)(NS - t);
                w0[t] = exp2f(-a.coef[0] * u0 * u0);
                w1[t] = exp2f(-a.coef[1] * u1 * u1);
                if (D == 3) {
                    const float u2 = rf.z + (float)(NS - t);
                    w2[t] = exp2f(-a.coef[2] * u2 * u2);
                }
            }
        }
        float2 acc[kCoilGroup];
#pragma unroll
        for (int cc = 0; cc < kCoilGroup; ++cc) acc[cc] = make_float2(0.f, 0.f);

#pragma unroll
        for (int cc = 0; cc < kCoilGroup; ++cc) {
            if (cc >= ncg) break;
            const float2 *g = a.grid + (int64_t)(cg0 + cc) * a.G;
            __syncthreads();
            for (int r = threadIdx.x; r < RC; r += kGatherThreads) {
                int r0, r1, r2 = 0;
                if (D == 3) {
                    r2 = r % R2;
                    r1 = (r / R2) % R1;
                    r0 = r / (R2 * R1);
                } else {
                    r1 = r % R1;
                    r0 = r / R1;
                }
                int g0 = org[0] + r0;
                if (g0 >= a.nf[0]) g0 -= (int)a.nf[0];
                int g1 = org[1] + r1;
                if (g1 >= a.nf[1]) g1 -= (int)a.nf[1];
                int64_t gi = (int64_t)g0 * a.nf[1] + g1;
                if (D == 3) {
                    int g2 = org[2] + r2;
                    if (g2 >= a.nf[2]) g2 -= (int)a.nf[2];
                    gi = gi * a.nf[2] + g2;
                }
                lds_gather[r] = g[gi];
            }
            __syncthreads();
            if (valid) {
                float ar = 0.f, ai = 0.f;
#pragma unroll
                for (int t0 = 0; t0 < W; ++t0) {
                    int r0 = l0 + t0;
                    if (r0 >= R0) r0 -= (int)a.nf[0]; // only when the region spans the whole axis
#pragma unroll
                    for (int t1 = 0; t1 < W; ++t1) {
                        int r1 = l1 + t1;
                        if (r1 >= R1) r1 -= (int)a.nf[1];
                        const float w01 = w0[t0] * w1[t1];
                        if (D == 3) {
                            const int rowb = (r0 * R1 + r1) * R2;
                            float pr = 0.f, pi = 0.f;
#pragma unroll
                            for (int t2 = 0; t2 < W; ++t2) {
                                int r2 = l2 + t2;
                                if (r2 >= R2) r2 -= (int)a.nf[2];
                                const float2 gv = lds_gather[rowb + r2];
                                pr += w2[t2] * gv.x;
                                pi += w2[t2] * gv.y;
                            }
                            ar += w01 * pr;
                            ai += w01 * pi;
                        } else {
                            const float2 gv = lds_gather[r0 * R1 + r1];
                            ar += w01 * gv.x;
                            ai += w01 * gv.y;
                        }
                    }
                }
                acc[cc] = make_float2(ar, ai);
            }
        }
        if (valid) {
#pragma unroll
            for (int cc = 0; cc < kCoilGroup; ++cc) {
                if (cc >= ncg) break;
                const int64_t coil = cg0 + cc;
                if (a.layout == GN_COIL_LAST)
                    a.out[orig * a.nc + coil] = acc[cc];
                else
                    a.out[coil * a.M + orig] = acc[cc];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused pre-pass (deapodize + zero-pad + coil transpose) and post-pass
// (crop + deapodize + scale + coil transpose).  Frequency k = n - m//2 sits at
// grid index k mod nf (fft/nufft_func.py:761-764, 728-729).
// ---------------------------------------------------------------------------

struct PrepostArgs {
    int dim;
    int64_t modes[3];
    int64_t nf[3];
    int64_t G;
    int64_t N;
    int64_t nc;
    int layout;
    float scale;
    const float *deap0, *deap1, *deap2;
};

template <int D>
__global__ __launch_bounds__(256) void k_pad(PrepostArgs a, const float2 *__restrict__ img, float2 *__restrict__ grid)
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
        for (int64_t c = 0; c < a.nc; ++c) {
            const float2 v = (a.layout == GN_COIL_LAST) ? img[n * a.nc + c] : img[c * a.N + n];
            grid[c * a.G + gidx] = make_float2(v.x * d, v.y * d);
        }
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_crop(PrepostArgs a, const float2 *__restrict__ grid, float2 *__restrict__ img)
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
        hipFree(*p);
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

} // namespace

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

void default_tiles(int dim, int ns, const int64_t *nf, int *tile)
{
    if (dim == 3) {
        tile[0] = 8;
        tile[1] = 8;
        tile[2] = 16;
    } else {
        tile[0] = 16;
        tile[1] = 32;
        tile[2] = 1;
    }
    for (int a = 0; a < dim; ++a)
        if (tile[a] > nf[a]) tile[a] = (int)nf[a];
    (void)ns;
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
    hipStream_t st = p->stream;
    int rc;
    if (M > p->rec.cap) {
        if ((rc = dev_alloc(&p->rec.cell, (size_t)M, "records"))) return rc;
        if ((rc = dev_alloc(&p->rec.frac, (size_t)M, "records"))) return rc;
        p->rec.cap = M;
    }
    if ((rc = dev_alloc(&p->rec.bin_start, (size_t)g.K + 1, "bin offsets"))) return rc;
    if (M == 0) {
        GN_HIP(hipMemsetAsync(p->rec.bin_start, 0, sizeof(int) * (g.K + 1), st));
        GN_HIP(hipStreamSynchronize(st));
        return GN_OK;
    }

    // scratch: unsorted cell/frac, counts matrix, sorted index, scan block sums
    const int64_t K = g.K;
    int64_t C = std::max<int64_t>(1, std::min<int64_t>((M + 255) / 256, (int64_t)(8 << 20) / std::max<int64_t>(K, 1)));
    int64_t chunk_len = (M + C - 1) / C;
    chunk_len = ((chunk_len + 63) / 64) * 64;
    C = (M + chunk_len - 1) / chunk_len;
    const int64_t nKC = K * C;
    const int64_t nb = (nKC + kScanTile - 1) / kScanTile;

    int4 *cell = nullptr;
    float4 *frac = nullptr;
    int *cnt = nullptr, *sorted = nullptr, *bsum = nullptr;
    auto cleanup = [&]() {
        hipFree(cell);
        hipFree(frac);
        hipFree(cnt);
        hipFree(sorted);
        hipFree(bsum);
    };
    if ((rc = dev_alloc(&cell, (size_t)M, "sort scratch")) || (rc = dev_alloc(&frac, (size_t)M, "sort scratch")) ||
        (rc = dev_alloc(&cnt, (size_t)nKC, "sort scratch")) || (rc = dev_alloc(&sorted, (size_t)M, "sort scratch")) ||
        (rc = dev_alloc(&bsum, (size_t)nb, "sort scratch"))) {
        cleanup();
        return rc;
    }

    hipError_t e = hipSuccess;
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
        pa.tile[a] = g.tile[a];
        pa.ntiles[a] = g.ntiles[a];
    }
    pa.ns = g.ns;
    pa.cell = cell;
    pa.frac = frac;
    pa.err = p->err_flag;
    if (e == hipSuccess) e = hipMemsetAsync(p->err_flag, 0, sizeof(int), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_prep, dim3(grid_blocks(M, 256, INT32_MAX)), dim3(256), 0, st, pa);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, sizeof(int) * nKC, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bin_count, dim3((unsigned)C), dim3(256), 0, st, M, cell, chunk_len, (int)C, cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanThreads), 0, st, cnt, nKC, bsum);
        hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(kScanThreads), 0, st, bsum, nb);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanThreads), 0, st, cnt, nKC, bsum);
        hipLaunchKernelGGL(k_bin_starts, dim3(grid_blocks(K + 1, 256, INT32_MAX)), dim3(256), 0, st, cnt, K, (int)C,
                           M, p->rec.bin_start);
        hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)C), dim3(64), 0, st, M, cell, chunk_len, (int)C, cnt, sorted);
        hipLaunchKernelGGL(k_make_records, dim3(grid_blocks(M, 256, INT32_MAX)), dim3(256), 0, st, M, sorted, cell,
                           frac, dcf, p->rec.cell, p->rec.frac);
        e = hipGetLastError();
    }
    int herr = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, p->err_flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "set_points");
    if (herr) return fail(GN_EINVAL, "non-finite sample coordinate");
    return GN_OK;
}

namespace {

template <int D, int NS>
hipError_t spread_dispatch(const SpreadArgs &a, int64_t nc, hipStream_t st)
{
    const int TC = a.tile[0] * a.tile[1] * (D == 3 ? a.tile[2] : 1);
    const size_t lds = (size_t)kSpreadWaves * TC * sizeof(float4);
    static bool attr_set = false;
    if (!attr_set) {
        hipFuncSetAttribute((const void *)k_spread<D, NS>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    dim3 grid((unsigned)((a.K + kSpreadWaves - 1) / kSpreadWaves), (unsigned)nc);
    hipLaunchKernelGGL((k_spread<D, NS>), grid, dim3(64 * kSpreadWaves), lds, st, a);
    return hipGetLastError();
}

template <int D, int NS>
hipError_t gather_dispatch(const GatherArgs &a, int64_t nc, hipStream_t st)
{
    const size_t lds = (size_t)a.region[0] * a.region[1] * (D == 3 ? a.region[2] : 1) * sizeof(float2);
    static bool attr_set = false;
    if (!attr_set) {
        hipFuncSetAttribute((const void *)k_gather<D, NS>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    dim3 grid((unsigned)a.K, (unsigned)((nc + kCoilGroup - 1) / kCoilGroup));
    hipLaunchKernelGGL((k_gather<D, NS>), grid, dim3(kGatherThreads), lds, st, a);
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

hipError_t spread_any(int dim, int ns, const SpreadArgs &a, int64_t nc, hipStream_t st)
{
    if (dim == 3) {
        GN_NS_CASES(3, spread_dispatch, a, nc, st)
    } else {
        GN_NS_CASES(2, spread_dispatch, a, nc, st)
        GN_NS_CASES_2D_EXTRA(spread_dispatch, a, nc, st)
    }
    return hipErrorInvalidValue;
}

hipError_t gather_any(int dim, int ns, const GatherArgs &a, int64_t nc, hipStream_t st)
{
    if (dim == 3) {
        GN_NS_CASES(3, gather_dispatch, a, nc, st)
    } else {
        GN_NS_CASES(2, gather_dispatch, a, nc, st)
        GN_NS_CASES_2D_EXTRA(gather_dispatch, a, nc, st)
    }
    return hipErrorInvalidValue;
}

PrepostArgs make_prepost(const gn_plan_s *p, float scale)
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
    return a;
}

} // namespace

int launch_spread(gn_plan_s *p, const float2 *samples, int use_dcf)
{
    const Geometry &g = p->geo;
    SpreadArgs a;
    for (int ax = 0; ax < 3; ++ax) {
        a.nf[ax] = g.nf[ax];
        a.tile[ax] = g.tile[ax];
        a.ntiles[ax] = g.ntiles[ax];
        // A window starting in home tile q-p reaches tile q only if p <= 1 + (2ns-2)/T for
        // every tile extent T it may jump over; the last tile of an axis can be short.
        const int64_t last = g.nf[ax] - (int64_t)(g.ntiles[ax] - 1) * g.tile[ax];
        const int tmin = (int)std::min<int64_t>(g.tile[ax], last);
        const int np = 2 + (2 * g.ns - 2) / tmin;
        a.npred[ax] = std::min(np, g.ntiles[ax]);
        a.coef[ax] = g.coef[ax];
    }
    a.K = g.K;
    a.G = g.G;
    a.M = p->M;
    a.nc = p->nc;
    a.layout = p->layout;
    a.use_dcf = use_dcf && p->has_dcf;
    a.bin_start = p->rec.bin_start;
    a.rec_cell = p->rec.cell;
    a.rec_frac = p->rec.frac;
    a.in = samples;
    a.grid = p->grid;
    hipError_t e = spread_any(g.dim, g.ns, a, p->nc, p->stream);
    if (e != hipSuccess) return hip_fail(e, "spread launch");
    return GN_OK;
}

int launch_gather(gn_plan_s *p, float2 *samples)
{
    const Geometry &g = p->geo;
    if (p->M == 0) return GN_OK;
    GatherArgs a;
    for (int ax = 0; ax < 3; ++ax) {
        a.nf[ax] = g.nf[ax];
        a.tile[ax] = g.tile[ax];
        a.ntiles[ax] = g.ntiles[ax];
        a.region[ax] = (int)std::min<int64_t>(g.tile[ax] + 2 * g.ns - 1, g.nf[ax]);
        a.coef[ax] = g.coef[ax];
    }
    a.K = g.K;
    a.G = g.G;
    a.M = p->M;
    a.nc = p->nc;
    a.layout = p->layout;
    a.bin_start = p->rec.bin_start;
    a.rec_cell = p->rec.cell;
    a.rec_frac = p->rec.frac;
    a.grid = p->grid;
    a.out = samples;
    hipError_t e = gather_any(g.dim, g.ns, a, p->nc, p->stream);
    if (e != hipSuccess) return hip_fail(e, "gather launch");
    return GN_OK;
}

int launch_pad(gn_plan_s *p, const float2 *image, float scale)
{
    PrepostArgs a = make_prepost(p, scale);
    dim3 grid(grid_blocks(p->geo.G, 256, 1 << 16));
    if (p->geo.dim == 3)
        hipLaunchKernelGGL(k_pad<3>, grid, dim3(256), 0, p->stream, a, image, p->grid);
    else
        hipLaunchKernelGGL(k_pad<2>, grid, dim3(256), 0, p->stream, a, image, p->grid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pad launch");
    return GN_OK;
}

int launch_crop(gn_plan_s *p, float2 *image, float scale)
{
    PrepostArgs a = make_prepost(p, scale);
    dim3 grid(grid_blocks(p->geo.N, 256, 1 << 16));
    if (p->geo.dim == 3)
        hipLaunchKernelGGL(k_crop<3>, grid, dim3(256), 0, p->stream, a, p->grid, image);
    else
        hipLaunchKernelGGL(k_crop<2>, grid, dim3(256), 0, p->stream, a, p->grid, image);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "crop launch");
    return GN_OK;
}

} // namespace gn
