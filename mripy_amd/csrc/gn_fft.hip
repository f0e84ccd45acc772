// Pruned oversampled-grid FFTs fused with the pre/post passes.
//
// Type 2 pads an m-point image axis into nf cells (frequency k = n - m/2 at cell
// (k mod nf)) and transforms every axis; type 1 transforms and keeps the same m
// cells.  Done axis by axis, every pass touches only the cells that are non-zero
// (type 2) or needed (type 1) on the axes not yet transformed:
//
//   type 2, 3-D: A  z-rows of the m0 x m1 image lines      image -> cells (n0 < m0, n1 < m1, all k2)
//                B  y-columns, m1 rows in -> nf1 out       -> cells (n0 < m0, all k1, all k2)
//                C  x-columns, m0 rows in -> nf0 out       -> every cell
//   type 1, 3-D: C' x-columns, nf0 rows in -> m0 kept       -> cells (n0 < m0, ...)
//                B' y-columns, nf1 rows in -> m1 kept       -> cells (n0 < m0, n1 < m1, ...)
//                A' z-rows, nf2 in -> m2 kept               -> image
//
// (2-D: passes A and C only.)  Every pass works in place in the plan's grid buffer:
// a block reads all rows of its (slab, column block) into LDS before it writes
// any, and no two blocks share an address, so the partial transforms live in the
// leading rows of the grid itself and no scratch buffer is needed.  Deapodization
// and scale are applied in A (type 2) and A' (type 1), which also read / write the
// caller's coil-last or coil-major image.  Each 1-D FFT runs in LDS as an in-place
// mixed-radix (8, 4, 2, 3, 5) decimation-in-time transform: loads scatter the input
// to digit-reversed slots, so the stages need no reordering and the output comes
// out in natural order.
// Twiddles come from an FP64-computed table of e^{-2 pi i k / L}.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gn_internal.h"
#include "gn_host_plan.h"

namespace gn {
namespace {

// Streaming 16-byte accesses: every grid element a pass reads or writes is touched
// once per pass and the next pass runs over the whole grid, so the column passes load
// and store, and pass A' loads its grid lines and stores its image pairs, with the
// nontemporal hint (measured at 128^3 x 8 coils: fwd+adj 2.54 -> 2.39 ms).  Pass A's
// 8-byte grid stores, the spread's tile stores and the gather's sample stores were
// slower with it.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float2 *p)
{
    const f32x4_t t = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t *>(p));
    return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void store_stream(float2 *p, float4 v)
{
    const f32x4_t t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<f32x4_t *>(p));
}

#ifndef GN_FFT_THREADS
#define GN_FFT_THREADS 256
#endif
#ifndef GN_FFT_COL_THREADS
#define GN_FFT_COL_THREADS 512
#endif
#ifndef GN_FFT_CB
#define GN_FFT_CB 16
#endif
constexpr int kFFTThreads = GN_FFT_THREADS;       // row passes
constexpr int kLoadUnroll = 8;                    // global loads in flight per thread
constexpr int kPadLoadUnroll = 4;                 // the same in pass A (a 256-thread block loads
                                                  // 128 x 8 cells in one step)
constexpr int kColLoadUnroll = 4;                 // the same in the column passes (a 512-thread block
                                                  // loads 256 x 8 cell pairs in one step)
constexpr int kFFTColThreads = GN_FFT_COL_THREADS; // column passes
static_assert((GN_FFT_CB & (GN_FFT_CB - 1)) == 0, "column blocks index rows with shifts");


__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by (sign * i)
__device__ inline float2 mul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// Radix-R DFT of v[0..R) in registers (e^{sign 2 pi i / R} roots): v[k] becomes
// output k.
template <int R>
__device__ __forceinline__ void dft_small(float2 *v, float sign)
{
    if (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = mul_si(csub(v[1], v[3]), sign);
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, d13);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, d13);
    } else if (R == 3) {
        const float c1 = -0.5f, s1 = 0.86602540378443864676f * sign;
        const float2 a = cadd(v[1], v[2]), b = csub(v[1], v[2]);
        const float2 m = make_float2(v[0].x + c1 * a.x, v[0].y + c1 * a.y);
        const float2 r = make_float2(-s1 * b.y, s1 * b.x);
        v[0] = cadd(v[0], a);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    } else if (R == 8) {
        // X[m] = E[m] + w^m O[m], X[m + 4] = E[m] - w^m O[m]: E, O the 4-point DFTs of
        // the even and odd inputs, w = e^{sign 2 pi i / 8}
        const float2 s04 = cadd(v[0], v[4]), d04 = csub(v[0], v[4]);
        const float2 s26 = cadd(v[2], v[6]), d26 = mul_si(csub(v[2], v[6]), sign);
        const float2 e0 = cadd(s04, s26), e1 = cadd(d04, d26), e2 = csub(s04, s26), e3 = csub(d04, d26);
        const float2 s15 = cadd(v[1], v[5]), d15 = csub(v[1], v[5]);
        const float2 s37 = cadd(v[3], v[7]), d37 = mul_si(csub(v[3], v[7]), sign);
        const float2 o0 = cadd(s15, s37), o1 = cadd(d15, d37), o2 = csub(s15, s37), o3 = csub(d15, d37);
        const float r2 = 0.70710678118654752440f;
        // w^1 = (1 + sign i) / sqrt 2, w^2 = sign i, w^3 = (-1 + sign i) / sqrt 2
        const float2 t1 = make_float2(r2 * (o1.x - sign * o1.y), r2 * (o1.y + sign * o1.x));
        const float2 t2 = mul_si(o2, sign);
        const float2 t3 = make_float2(r2 * (-o3.x - sign * o3.y), r2 * (-o3.y + sign * o3.x));
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, t1);
        v[5] = csub(e1, t1);
        v[2] = cadd(e2, t2);
        v[6] = csub(e2, t2);
        v[3] = cadd(e3, t3);
        v[7] = csub(e3, t3);
    } else if (R == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f * sign, s2 = 0.58778525229247312917f * sign;
        const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        const float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
        const float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
        // r1 = i (s1 b1 + s2 b2), r2 = i (s2 b1 - s1 b2)
        const float2 r1 = make_float2(-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x);
        const float2 r2 = make_float2(-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x);
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(m1, r1);
        v[4] = csub(m1, r1);
        v[2] = cadd(m2, r2);
        v[3] = csub(m2, r2);
    }
}

// One radix-r butterfly on LDS elements at base + k*qstride (k < r) of a line, after
// twiddling element k by w_g^{t k}: tw[(t k L / g)] (the table in LDS already holds
// the twiddles of this transform's sign).  TW = false for the first stage (g = r, all
// twiddles 1).  Otherwise every element is multiplied, t = 0 included (tw[0] = 1 and
// x * 1 - y * 0 = x exactly), so the lanes of a wave take no per-element branches.
// NL = 2: the same butterfly on two lines (elements base + k*qstride and lstep after
// them), which share the index arithmetic and the twiddle reads.
template <int R, bool TW, int NL>
__device__ __forceinline__ void butterfly(float2 *buf, int base, int qstride, int lstep, const float2 *tw, int twstep,
                                          int t, float sign)
{
    float2 w[R];
    if (TW) {
#pragma unroll
        for (int k = 1; k < R; ++k) w[k] = tw[t * k * twstep];
    }
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        float2 v[R];
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = buf[base + k * qstride + l * lstep];
        if (TW) {
#pragma unroll
            for (int k = 1; k < R; ++k) v[k] = cmul(v[k], w[k]);
        }
        dft_small<R>(v, sign);
#pragma unroll
        for (int k = 0; k < R; ++k) buf[base + k * qstride + l * lstep] = v[k];
    }
}

// floor(n / d) for 0 <= n < 2^22, 1 <= d <= 2^12, with r = 1/d: the float estimate
// is within one of the quotient, and one correction step each way makes it exact.
__device__ inline int fdiv(int n, int d, float r)
{
    int q = (int)((float)n * r);
    q -= (q * d > n) ? 1 : 0;
    q += ((q + 1) * d <= n) ? 1 : 0;
    return q;
}

// In-place FFT of `nlines` lines of length L held in LDS: element i of line j at
// buf[j * lstride + i * estride].  All threads of the block take part.  NL = 2 (nlines
// even): a thread takes the same butterfly on lines j and j + nlines / 2, so that the
// lanes of a wave still read consecutive lines in each of the two reads.
template <int NL>
__device__ __forceinline__ void fft_lines(float2 *buf, int nlines, int lstride, int estride, int L, const Stages &st,
                                          const float2 *tw, float sign)
{
    const int nw = nlines / NL; // line groups
    // fdiv corrects its quotient estimate, so approximate reciprocals suffice
    const float rl = __builtin_amdgcn_rcpf((float)nw);
    // power-of-two line counts and butterfly groups index with shifts
    const int lgl = (nw & (nw - 1)) == 0 ? __builtin_ctz(nw) : -1;
    for (int s = 0; s < st.n; ++s) {
        const int r = st.radix[s];
        const int q = st.q[s];         // butterfly span before this stage
        const int g = r * q;
        const int twstep = st.tws[s];  // L / g
        const int nbl = st.nb[s] * nw; // butterflies of all line groups
        const float rq = __builtin_amdgcn_rcpf((float)q);
        const int lgq = (q & (q - 1)) == 0 ? __builtin_ctz(q) : -1;
        for (int w = threadIdx.x; w < nbl; w += blockDim.x) {
            // line groups fastest: neighbouring threads work on neighbouring lines
            const int bidx = lgl >= 0 ? (w >> lgl) : fdiv(w, nw, rl);
            const int j = w - bidx * nw;
            const int blk = lgq >= 0 ? (bidx >> lgq) : fdiv(bidx, q, rq);
            const int t = bidx - blk * q;
            const int base = j * lstride + (blk * g + t) * estride;
            if (s == 0) {
                switch (r) {
                case 2: butterfly<2, false, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                case 3: butterfly<3, false, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                case 4: butterfly<4, false, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                case 8: butterfly<8, false, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                default: butterfly<5, false, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                }
            } else {
                switch (r) {
                case 2: butterfly<2, true, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                case 3: butterfly<3, true, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                case 4: butterfly<4, true, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                case 8: butterfly<8, true, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                default: butterfly<5, true, NL>(buf, base, q * estride, nw * lstride, tw, twstep, t, sign); break;
                }
            }
        }
        __syncthreads();
    }
}

// Stage the twiddles of a transform of this sign (the table holds e^{-2 pi i k/L}; sign
// -1 keeps it, +1 conjugates it) and the digit-reversal map of a length-n transform in
// LDS.  Both loads of an entry are issued before either is used.
__device__ inline void load_tables(float2 *tw_dst, int *pinv_dst, const float2 *tw, const int *pinv, int n, float sign)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float2 w = tw[i];
        const int q = pinv[i];
        tw_dst[i] = make_float2(w.x, -sign * w.y);
        pinv_dst[i] = q;
    }
}

// ---------------------------------------------------------------------------
// Column pass: FFT along a strided axis.  A block takes CB consecutive columns of
// one slab; rows are the transformed axis.  Input row n sits at FFT position
// (n - m/2) mod L (in_map) or n; output row n reads FFT index (n - m/2) mod L
// (out_map) or n.
// ---------------------------------------------------------------------------


// Row mask word of grid column c (c < 2^31) for the masked leading-axis passes.
__device__ inline unsigned long long col_mask(const ColArgs &a, const unsigned long long *mask, int64_t c)
{
    const uint32_t y = (uint32_t)c / (uint32_t)a.mcz, z = (uint32_t)c - y * (uint32_t)a.mcz;
    return mask[(y / (uint32_t)a.mt1) * (uint32_t)a.mnt2 + z / (uint32_t)a.mt2];
}

// MASK: 0 none, 1 read mask (cells outside it enter the FFT as zeros)
template <int MASK>
__device__ __forceinline__ void fft_cols_body(const ColArgs &a, int64_t colblk, int64_t slab)
{
    extern __shared__ float2 lds_fft[];
    float2 *tw = lds_fft;
    float2 *buf = lds_fft + a.L;
    const int CB = a.CB;
    const int lstride = 1;        // columns adjacent
    const int estride = CB + 1;   // rows padded by one element
    const int64_t col0 = colblk * CB;
    const int ncb = (int)min((int64_t)CB, a.ncols - col0);
    // slab and slab_inner are below 2^31: 32-bit division
    const uint32_t so = (uint32_t)slab / (uint32_t)a.slab_inner, si = (uint32_t)slab - so * (uint32_t)a.slab_inner;
    float2 *base = a.data + so * a.stride_outer + si * a.stride_inner + col0;
    const int half_in = a.rows_in / 2;
    const int cbs = __builtin_ctz(CB); // CB is a power of two
    // two columns per lane (16-byte accesses) when every row segment of the block is 16-byte aligned
    const bool v2 = (((a.ncols | a.stride_outer | a.stride_inner | ncb) & 1) == 0) &&
                    ((reinterpret_cast<uintptr_t>(a.data) & 15) == 0);
    const int ws = v2 ? 1 : 0; // log2 of the columns per lane
    const int cbw = CB >> ws, cbws = cbs - ws;
    const int total_in = a.rows_in * cbw;
    // Row masks of this lane's columns: the lane index fixes the column (blockDim.x is a
    // multiple of the columns per row), so one word per column serves every row.  Both
    // words are fetched (for in-range columns) before the table staging and selected
    // after it, so that the mask and table loads are in flight together.
    const float rmrow = 1.0f / (float)a.mrow;
    const int lgm = (a.mrow & (a.mrow - 1)) == 0 ? __builtin_ctz(a.mrow) : -1; // row tiles of 2^lgm rows
    unsigned long long mk0 = ~0ull, mk1 = ~0ull, mw0 = 0ull, mw1 = 0ull;
    if (MASK) {
        const int j = (threadIdx.x & (cbw - 1)) << ws;
        mw0 = col_mask(a, a.rmask, col0 + min(j, ncb - 1));
        mw1 = col_mask(a, a.rmask, col0 + min(j + 1, ncb - 1));
    }
    int *spinv = (int *)(buf + a.L * estride); // digit-reversal map, staged in LDS
    // kColLoadUnroll independent global loads in flight per thread before the LDS stores.
    // 16- and 8-byte cells take separate loops, and the mask is applied at the LDS
    // stores: a loaded value that is copied or selected right after its load makes the
    // wave wait for that load before it issues the next one.
    float4 v[kColLoadUnroll];
#pragma unroll
    for (int u = 0; u < kColLoadUnroll; ++u) v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t onb = 0; // masked pass: bit 2u (2u + 1) set when column j (j + 1) of step u lies in a written tile
    auto load_chunk = [&](int e0) {
        // elements left from this chunk's first (uniform): steps past the end are
        // skipped on a scalar branch
        const int left = __builtin_amdgcn_readfirstlane(total_in - (e0 - (int)threadIdx.x));
        onb = 0;
        if (v2) {
#pragma unroll
            for (int u = 0; u < kColLoadUnroll; ++u) {
                if (u * (int)blockDim.x >= left) break;
                const int e = e0 + u * blockDim.x;
                const int n = e >> cbws, j = (e & (cbw - 1)) << ws;
                if (e < total_in && j < ncb) {
                    uint32_t on = 3u;
                    if (MASK == 1) { // cells of tiles the spread did not write enter as zeros
                        const int b = lgm >= 0 ? n >> lgm : fdiv(n, a.mrow, rmrow);
                        on = (uint32_t)((mk0 >> b) & 1ull) | ((uint32_t)((mk1 >> b) & 1ull) << 1);
                        onb |= on << (2 * u);
                    }
                    if (on) v[u] = load_stream(base + (int64_t)n * a.ncols + j);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < kColLoadUnroll; ++u) {
                if (u * (int)blockDim.x >= left) break;
                const int e = e0 + u * blockDim.x;
                const int n = e >> cbws, j = e & (cbw - 1);
                if (e < total_in && j < ncb) {
                    uint32_t on = 1u;
                    if (MASK == 1) {
                        const int b = lgm >= 0 ? n >> lgm : fdiv(n, a.mrow, rmrow);
                        on = (uint32_t)((mk0 >> b) & 1ull);
                        onb |= on << (2 * u);
                    }
                    if (on) {
                        const float2 t = base[(int64_t)n * a.ncols + j];
                        v[u].x = t.x;
                        v[u].y = t.y;
                    }
                }
            }
        }
    };
    // Passes with pruned input (the type-2 column passes) issue the first chunk's loads
    // before the tables are staged, so that the grid and table latencies overlap.  The
    // type-1 column passes, which read every row, were slower that way.
    const bool early = MASK == 0 && a.rows_in < a.L;
    if (early) load_chunk(threadIdx.x);
    load_tables(tw, spinv, a.tw, a.pinv, a.L, a.sign);
    __syncthreads();
    if (MASK) {
        const int j = (threadIdx.x & (cbw - 1)) << ws;
        if (j < ncb) mk0 = mw0;
        mk1 = (v2 && j + 1 < ncb) ? mw1 : mk0;
    }
    // pruned input: the FFT positions no input row maps to, [rows_in - half_in,
    // L - half_in) with in_map and [rows_in, L) without, start as zeros; the stores
    // below write every other slot, so no barrier separates the two
    if (a.rows_in < a.L) {
        const int pz = a.in_map ? a.rows_in - half_in : a.rows_in;
        const int nz = (a.L - a.rows_in) << cbs;
        for (int e = threadIdx.x; e < nz; e += blockDim.x) {
            const int r = e >> cbs, j = e & (CB - 1);
            buf[spinv[pz + r] * estride + j] = make_float2(0.f, 0.f);
        }
    }
    for (int e0 = threadIdx.x; e0 < total_in; e0 += kColLoadUnroll * blockDim.x) {
        if (!early || e0 != (int)threadIdx.x) load_chunk(e0);
        const int left = __builtin_amdgcn_readfirstlane(total_in - (e0 - (int)threadIdx.x));
#pragma unroll
        for (int u = 0; u < kColLoadUnroll; ++u) {
            if (u * (int)blockDim.x >= left) break;
            const int e = e0 + u * blockDim.x;
            const int n = e >> cbws, j = (e & (cbw - 1)) << ws;
            if (e < total_in && j < ncb) {
                int p = n;
                if (a.in_map) {
                    p = n - half_in;
                    if (p < 0) p += a.L;
                }
                float2 *dst = buf + spinv[p] * estride + j;
                float4 w = v[u];
                if (MASK == 1) {
                    if (!((onb >> (2 * u)) & 1u)) w.x = w.y = 0.f;
                    if (!((onb >> (2 * u + 1)) & 1u)) w.z = w.w = 0.f;
                }
                dst[0] = make_float2(w.x, w.y);
                if (v2) dst[1] = make_float2(w.z, w.w);
            }
        }
    }
    __syncthreads();
    // columns in pairs: one index computation and one set of twiddle reads serve both
    if ((CB & 1) == 0)
        fft_lines<2>(buf, CB, lstride, estride, a.L, a.st, tw, a.sign);
    else
        fft_lines<1>(buf, CB, lstride, estride, a.L, a.st, tw, a.sign);
    const int half_out = a.rows_out / 2;
    for (int e = threadIdx.x; e < a.rows_out * cbw; e += blockDim.x) {
        const int n = e >> cbws, j = (e & (cbw - 1)) << ws;
        if (j >= ncb) continue;
        int p = n;
        if (a.out_map) {
            p = n - half_out;
            if (p < 0) p += a.L;
        }
        const float2 *src = buf + p * estride + j;
        float2 *dst = base + (int64_t)n * a.ncols + j;
        if (v2) {
            const float2 lo = src[0], hi = src[1];
            store_stream(dst, make_float4(lo.x, lo.y, hi.x, hi.y));
        } else {
            *dst = src[0];
        }
    }
}

template <int MASK>
__global__ __launch_bounds__(kFFTColThreads) void k_fft_cols(ColArgs a)
{
    fft_cols_body<MASK>(a, blockIdx.x, blockIdx.y);
}

// Pass C of a 3-D type 2 (m0 rows in, nf0 rows out at a stride of a whole plane):
// blocks of twice the threads and kWideCB columns, 256-byte row segments.  Measured
// faster for this pass only; the other passes keep kFFTColThreads and CB <= GN_FFT_CB.
constexpr int kFFTColThreadsWide = 2 * kFFTColThreads;
constexpr int kWideCB = 2 * GN_FFT_CB;
__global__ __launch_bounds__(kFFTColThreadsWide) void k_fft_cols_wide(ColArgs a)
{
    fft_cols_body<0>(a, blockIdx.x, blockIdx.y);
}

// Batched column passes: items[b] = {argument index, column block, slab, 0}.
__global__ __launch_bounds__(kFFTColThreads) void k_fft_cols_batch(const ColArgs *__restrict__ args,
                                                                   const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const ColArgs &a = args[it.x]; // read in place: uniform scalar loads
    fft_cols_body<0>(a, it.y, it.z);
}

// ---------------------------------------------------------------------------
// Row passes along the contiguous image axis, fused with the image read (type 2)
// or write (type 1): deapodization, scale, zero-pad / crop, coil transpose.  A
// block takes one image line position (all leading indices but the last axis)
// and a chunk of CL coils; LDS holds CL lines of L (+1) elements.
// ---------------------------------------------------------------------------


__device__ inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Image elements e = (n, c) of a block: coil-last runs over c fastest within a
// point n, coil-major over n fastest within a coil.  The crop pass moves pairs
// (e, e + 1) as one 16-byte access when the run is even and 16-byte aligned
// (img_pairs).  The pad pass keeps 8-byte accesses: pairs measured slower there.
__device__ inline bool img_pairs(const RowArgs &a, const void *img, int64_t c0, int ncl)
{
    return aligned16(img) && (a.layout == GN_COIL_LAST ? ((a.nc | c0 | ncl) & 1) == 0 : (a.m & 1) == 0);
}
// Grid lines are contiguous runs of L per coil; pairs when L, the line offsets and G are even.
__device__ inline bool grid_pairs(const RowArgs &a)
{
    return aligned16(a.grid) && ((a.L | a.s0 | a.G) & 1) == 0;
}

__device__ __forceinline__ void fft_rows_pad_body(const RowArgs &a, int64_t line, int64_t chunk)
{
    extern __shared__ float2 lds_fft[];
    float2 *tw = lds_fft;
    float2 *buf = lds_fft + a.L;
    const int64_t c0 = chunk * a.CL;
    const int ncl = (int)min((int64_t)a.CL, a.nc - c0);
    const int ls = a.L + 1;
    int *spinv = (int *)(buf + a.CL * ls);
    load_tables(tw, spinv, a.tw, a.pinv, a.L, a.sign);
    for (int e = threadIdx.x; e < ncl * ls; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
    __syncthreads();
    const int64_t n0 = (uint32_t)line / (uint32_t)a.m_b, n1 = line - n0 * a.m_b; // line < 2^31
    const float d01 = a.scale * a.deap_a[n0] * (a.deap_b ? a.deap_b[n1] : 1.0f);
    const int half = a.m / 2;
    const float rncl = 1.0f / (float)ncl, rm = 1.0f / (float)a.m, rL = 1.0f / (float)a.L;
    const int total = a.m * ncl;
    // The image values and their deapodization factors of kPadLoadUnroll steps are loaded
    // together; the SENSE product takes a loop of its own, so that no loaded value is
    // used (which makes the wave wait for it) before the last load of the step issues.
    float2 v[kPadLoadUnroll];
    float dl[kPadLoadUnroll];
#pragma unroll
    for (int u = 0; u < kPadLoadUnroll; ++u) {
        v[u] = make_float2(0.f, 0.f);
        dl[u] = 0.f;
    }
    for (int e0 = threadIdx.x; e0 < total; e0 += kPadLoadUnroll * blockDim.x) {
        int nn[kPadLoadUnroll], cc[kPadLoadUnroll];
#pragma unroll
        for (int u = 0; u < kPadLoadUnroll; ++u) {
            const int e = e0 + u * blockDim.x;
            int n, c;
            if (a.layout == GN_COIL_LAST) {
                n = fdiv(e, ncl, rncl);
                c = e - n * ncl;
            } else {
                c = fdiv(e, a.m, rm);
                n = e - c * a.m;
            }
            nn[u] = n;
            cc[u] = c;
        }
        if (a.sens) { // SENSE: coil c's line is x * S_c
#pragma unroll
            for (int u = 0; u < kPadLoadUnroll; ++u) {
                const int e = e0 + u * blockDim.x, n = nn[u], c = cc[u];
                if (e < total) {
                    const float2 x = a.img_in[line * a.m + n];
                    const float2 sc = a.layout == GN_COIL_LAST ? a.sens[(line * a.m + n) * a.nc + c0 + c]
                                                               : a.sens[(c0 + c) * a.nlines * a.m + line * a.m + n];
                    v[u] = cmul(x, sc);
                    dl[u] = a.deap_last[n];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < kPadLoadUnroll; ++u) {
                const int e = e0 + u * blockDim.x, n = nn[u], c = cc[u];
                if (e < total) {
                    const int64_t src = a.layout == GN_COIL_LAST ? (line * a.m + n) * a.nc + c0 + c
                                                                 : (c0 + c) * a.nlines * a.m + line * a.m + n;
                    v[u] = a.img_in[src];
                    dl[u] = a.deap_last[n];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPadLoadUnroll; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e < total) {
                const float d = d01 * dl[u];
                int p = nn[u] - half;
                if (p < 0) p += a.L;
                buf[cc[u] * ls + spinv[p]] = make_float2(v[u].x * d, v[u].y * d);
            }
        }
    }
    __syncthreads();
    fft_lines<1>(buf, ncl, ls, 1, a.L, a.st, tw, a.sign);
    const int64_t off = n0 * a.s0 + n1 * a.L;
    for (int e = threadIdx.x; e < ncl * a.L; e += blockDim.x) {
        const int c = fdiv(e, a.L, rL), k = e - c * a.L;
        a.grid[(c0 + c) * a.G + off + k] = buf[c * ls + k];
    }
}

__device__ __forceinline__ void fft_rows_crop_body(const RowArgs &a, int64_t line, int64_t chunk)
{
    extern __shared__ float2 lds_fft[];
    float2 *tw = lds_fft;
    float2 *buf = lds_fft + a.L;
    const int64_t c0 = chunk * a.CL;
    const int ncl = (int)min((int64_t)a.CL, a.nc - c0);
    const int ls = a.L + 1;
    const int64_t n0 = (uint32_t)line / (uint32_t)a.m_b, n1 = line - n0 * a.m_b; // line < 2^31
    const int64_t off = n0 * a.s0 + n1 * a.L;
    const float rncl = 1.0f / (float)ncl, rm = 1.0f / (float)a.m, rL = 1.0f / (float)a.L;
    int *spinv = (int *)(buf + a.CL * ls);
    load_tables(tw, spinv, a.tw, a.pinv, a.L, a.sign);
    __syncthreads();
    const int gw = grid_pairs(a) ? 1 : 0; // log2 of the elements per lane
    const int ngp = (ncl * a.L) >> gw;
    for (int q0 = threadIdx.x; q0 < ngp; q0 += kLoadUnroll * blockDim.x) {
        float4 v[kLoadUnroll];
#pragma unroll
        for (int u = 0; u < kLoadUnroll; ++u) {
            const int q = q0 + u * blockDim.x, e = q << gw;
            const int c = fdiv(e, a.L, rL), k = e - c * a.L;
            if (q < ngp) {
                const float2 *src = a.grid + (c0 + c) * a.G + off + k;
                if (gw) {
                    v[u] = load_stream(src);
                } else {
                    const float2 t = *src;
                    v[u] = make_float4(t.x, t.y, 0.f, 0.f);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; ++u) {
            const int q = q0 + u * blockDim.x, e = q << gw;
            const int c = fdiv(e, a.L, rL), k = e - c * a.L;
            if (q < ngp) {
                buf[c * ls + spinv[k]] = make_float2(v[u].x, v[u].y);
                if (gw) buf[c * ls + spinv[k + 1]] = make_float2(v[u].z, v[u].w);
            }
        }
    }
    __syncthreads();
    fft_lines<1>(buf, ncl, ls, 1, a.L, a.st, tw, a.sign);
    const float d01 = a.scale * a.deap_a[n0] * (a.deap_b ? a.deap_b[n1] : 1.0f);
    const int half = a.m / 2;
    const bool cl = a.layout == GN_COIL_LAST;
    if (a.sens) { // SENSE: the block holds every coil (ncl == nc): sum_c conj(S_c) * image_c
        for (int n = threadIdx.x; n < a.m; n += blockDim.x) {
            int p = n - half;
            if (p < 0) p += a.L;
            const float d = d01 * a.deap_last[n];
            float2 acc = make_float2(0.f, 0.f);
            for (int c = 0; c < ncl; ++c) {
                const float2 v = buf[c * ls + p];
                const float2 o = make_float2(v.x * d, v.y * d);
                const float2 sc = cl ? a.sens[(line * a.m + n) * a.nc + c] : a.sens[c * a.nlines * a.m + line * a.m + n];
                acc.x += sc.x * o.x + sc.y * o.y;
                acc.y += sc.x * o.y - sc.y * o.x;
            }
            a.img_out[line * a.m + n] = acc;
        }
        return;
    }
    const int iw = img_pairs(a, a.img_out, c0, ncl) ? 1 : 0;
    auto get = [&](int n, int c, float dl) {
        int p = n - half;
        if (p < 0) p += a.L;
        const float d = d01 * dl;
        const float2 v = buf[c * ls + p];
        return make_float2(v.x * d, v.y * d);
    };
    // the deapodization factors of kCropSteps output steps are loaded together
    constexpr int kCropSteps = 2;
    const int nq = (a.m * ncl) >> iw;
    for (int q0 = threadIdx.x; q0 < nq; q0 += kCropSteps * blockDim.x) {
        int nn[kCropSteps], cc[kCropSteps];
        float dlo[kCropSteps], dhi[kCropSteps];
#pragma unroll
        for (int u = 0; u < kCropSteps; ++u) {
            const int q = q0 + u * blockDim.x, e = q << iw;
            int n, c;
            if (cl) {
                n = fdiv(e, ncl, rncl);
                c = e - n * ncl;
            } else {
                c = fdiv(e, a.m, rm);
                n = e - c * a.m;
            }
            nn[u] = n;
            cc[u] = c;
            dlo[u] = dhi[u] = 0.f;
            if (q < nq) {
                dlo[u] = a.deap_last[n];
                if (iw && !cl) dhi[u] = a.deap_last[n + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < kCropSteps; ++u) {
            const int q = q0 + u * blockDim.x;
            if (q >= nq) continue;
            const int n = nn[u], c = cc[u];
            float2 *dst = cl ? a.img_out + (line * a.m + n) * a.nc + c0 + c
                             : a.img_out + (c0 + c) * a.nlines * a.m + line * a.m + n;
            const float2 lo = get(n, c, dlo[u]);
            if (iw) {
                const float2 hi = cl ? get(n, c + 1, dlo[u]) : get(n + 1, c, dhi[u]);
                store_stream(dst, make_float4(lo.x, lo.y, hi.x, hi.y));
            } else {
                *dst = lo;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

using host::fft_radices;

__global__ __launch_bounds__(kFFTThreads) void k_fft_rows_pad(RowArgs a) { fft_rows_pad_body(a, blockIdx.x, blockIdx.y); }
__global__ __launch_bounds__(kFFTThreads) void k_fft_rows_crop(RowArgs a) { fft_rows_crop_body(a, blockIdx.x, blockIdx.y); }

// Batched row passes: items[b] = {instance, image line, coil chunk, 0}.
__global__ __launch_bounds__(kFFTThreads) void k_fft_rows_pad_batch(const RowArgs *__restrict__ args,
                                                                    const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const RowArgs &a = args[it.x]; // read in place: uniform scalar loads
    fft_rows_pad_body(a, it.y, it.z);
}
__global__ __launch_bounds__(kFFTThreads) void k_fft_rows_crop_batch(const RowArgs *__restrict__ args,
                                                                     const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const RowArgs &a = args[it.x]; // read in place: uniform scalar loads
    fft_rows_crop_body(a, it.y, it.z);
}

Stages make_stages(const FFTAxis &ax)
{
    Stages s;
    s.n = ax.nst;
    int q = 1;
    for (int i = 0; i < kFFTMaxStages; ++i) {
        s.radix[i] = i < ax.nst ? ax.radix[i] : 1;
        s.q[i] = q;
        s.tws[i] = ax.L / (q * s.radix[i]);
        s.nb[i] = ax.L / s.radix[i];
        q *= s.radix[i];
    }
    return s;
}


#ifndef GN_FFT_COL_LDS_MAX
#define GN_FFT_COL_LDS_MAX (64 * 1024)
#endif

int pick_cb(int L)
{
    int cb = GN_FFT_CB;
    while (cb > 1 && col_lds(L, cb) > GN_FFT_COL_LDS_MAX) cb /= 2;
    return cb;
}

// Coils per row-pass block: at most 16, fewer while the block's LDS exceeds this cap.
// The row passes are latency-bound per block, so a small cap (more, smaller blocks per
// CU) pays: on the 64-instance 2-D workload (lines of 256-640 points, up to 32 coils)
// caps of 64, 32, 24, 20, 16 KiB gave 8.76, 8.03, 7.92, 7.95, 8.36 ms per batched
// fwd+adj; 24 KiB keeps the 8 coils of a 256-point 3-D line in one block.
#ifndef GN_FFT_ROW_LDS_MAX
#define GN_FFT_ROW_LDS_MAX (24 * 1024)
#endif

int pick_cl(int L, int64_t nc)
{
    int cl = (int)std::min<int64_t>(nc, 16);
    while (cl > 1 && row_lds(L, cl) > GN_FFT_ROW_LDS_MAX) --cl;
    return cl;
}

ColArgs make_col_args(const FFTAxis &ax, float sign, float2 *data, int rows_in, int rows_out, int in_map, int out_map,
                      int64_t ncols, int64_t nslabs, int64_t slab_inner, int64_t stride_outer, int64_t stride_inner)
{
    ColArgs a;
    a.data = data;
    a.L = ax.L;
    a.rows_in = rows_in;
    a.rows_out = rows_out;
    a.in_map = in_map;
    a.out_map = out_map;
    a.ncols = ncols;
    a.slab_inner = slab_inner;
    a.stride_outer = stride_outer;
    a.stride_inner = stride_inner;
    a.nslabs = nslabs;
    a.CB = pick_cb(ax.L);
    a.sign = sign;
    a.st = make_stages(ax);
    a.tw = ax.tw;
    a.pinv = ax.pinv;
    return a;
}

void allow_lds(const void *kernel, size_t lds)
{
    if (lds > 64 * 1024) (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_cols(const ColArgs &a, hipStream_t st)
{
    dim3 grid((unsigned)((a.ncols + a.CB - 1) / a.CB), (unsigned)a.nslabs);
    const size_t lds = col_lds(a.L, a.CB);
    if (a.rmask) {
        allow_lds((const void *)k_fft_cols<1>, lds);
        hipLaunchKernelGGL(k_fft_cols<1>, grid, dim3(kFFTColThreads), lds, st, a);
    } else {
        allow_lds((const void *)k_fft_cols<0>, lds);
        hipLaunchKernelGGL(k_fft_cols<0>, grid, dim3(kFFTColThreads), lds, st, a);
    }
    return hipGetLastError();
}

RowArgs row_args(const gn_plan_s *p, const FFTAxis &ax, float sign, float scale)
{
    const Geometry &g = p->geo;
    const int d = g.dim;
    RowArgs a;
    a.img_in = nullptr;
    a.img_out = nullptr;
    a.grid = p->grid;
    a.G = g.G;
    a.s0 = g.G / g.nf[0]; // nf1 (* nf2)
    a.L = ax.L;
    a.m = (int)g.modes[d - 1];
    a.nlines = g.N / g.modes[d - 1];
    a.nc = p->nc;
    a.CL = pick_cl(ax.L, p->nc);
    a.layout = p->layout;
    a.scale = scale;
    a.sign = sign;
    a.deap_a = p->deap;
    if (d == 3) {
        a.deap_b = p->deap + g.modes[0];
        a.deap_last = p->deap + g.modes[0] + g.modes[1];
        a.m_b = (int)g.modes[1];
    } else {
        a.deap_b = nullptr;
        a.deap_last = p->deap + g.modes[0];
        a.m_b = 1;
    }
    a.st = make_stages(ax);
    a.tw = ax.tw;
    a.pinv = ax.pinv;
    return a;
}

} // namespace

// LDS: twiddles | lines | digit-reversal map
size_t col_lds(int L, int CB)
{
    return (size_t)L * sizeof(float2) + (size_t)L * (CB + 1) * sizeof(float2) + (size_t)L * sizeof(int);
}
size_t row_lds(int L, int CL)
{
    return (size_t)L * sizeof(float2) + (size_t)CL * (L + 1) * sizeof(float2) + (size_t)L * sizeof(int);
}

RowArgs make_row_args(const gn_plan_s *p, float sign, float scale)
{
    return row_args(p, p->fft_ax[p->geo.dim - 1], sign, scale);
}

// The column passes of a pruned transform in execution order (type 2: B (3-D), C;
// type 1: C', B' (3-D)); returns their number.
int pruned_col_args(const gn_plan_s *p, bool type2, float sign, ColArgs *out)
{
    const Geometry &g = p->geo;
    const int64_t nc = p->nc, m0 = g.modes[0], L2 = g.nf[g.dim - 1];
    // B / B': slabs (coil, n0), rows n1, columns k2; C / C': slabs = coils, rows n0,
    // columns = all cells of the other axes
    auto passB = [&](bool t2) {
        return t2 ? make_col_args(p->fft_ax[1], sign, p->grid, (int)g.modes[1], (int)g.nf[1], 1, 0, L2, nc * m0, m0,
                                  g.G, g.nf[1] * L2)
                  : make_col_args(p->fft_ax[1], sign, p->grid, (int)g.nf[1], (int)g.modes[1], 0, 1, L2, nc * m0, m0,
                                  g.G, g.nf[1] * L2);
    };
    auto passC = [&](bool t2) {
        return t2 ? make_col_args(p->fft_ax[0], sign, p->grid, (int)m0, (int)g.nf[0], 1, 0, g.G / g.nf[0], nc, 1, g.G,
                                  0)
                  : make_col_args(p->fft_ax[0], sign, p->grid, (int)g.nf[0], (int)m0, 0, 1, g.G / g.nf[0], nc, 1, g.G,
                                  0);
    };
    int k = 0;
    if (type2) {
        if (g.dim == 3) out[k++] = passB(true);
        out[k++] = passC(true);
    } else {
        out[k++] = passC(false);
        if (g.dim == 3) out[k++] = passB(false);
    }
    return k;
}

hipError_t launch_fft_cols_batch(const ColArgs *d_args, const int4 *d_items, int n, size_t lds, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    allow_lds((const void *)k_fft_cols_batch, lds);
    hipLaunchKernelGGL(k_fft_cols_batch, dim3((unsigned)n), dim3(kFFTColThreads), lds, st, d_args, d_items);
    return hipGetLastError();
}

hipError_t launch_fft_rows_batch(bool pad, const RowArgs *d_args, const int4 *d_items, int n, size_t lds,
                                 hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (pad) {
        allow_lds((const void *)k_fft_rows_pad_batch, lds);
        hipLaunchKernelGGL(k_fft_rows_pad_batch, dim3((unsigned)n), dim3(kFFTThreads), lds, st, d_args, d_items);
    } else {
        allow_lds((const void *)k_fft_rows_crop_batch, lds);
        hipLaunchKernelGGL(k_fft_rows_crop_batch, dim3((unsigned)n), dim3(kFFTThreads), lds, st, d_args, d_items);
    }
    return hipGetLastError();
}

bool fft_length_supported(int64_t L)
{
    if (L < 2 || L > 2048) return false; // LDS: one line of 2048 + twiddles fits in 64 KiB
    return !fft_radices((int)L).empty() && (int)fft_radices((int)L).size() <= kFFTMaxStages;
}

int fft_axis_init(FFTAxis *ax, int L)
{
    std::vector<int> r = fft_radices(L);
    if (r.empty() || (int)r.size() > kFFTMaxStages) return fail(GN_EINVAL, "FFT length not supported");
    ax->L = L;
    ax->nst = (int)r.size();
    for (int i = 0; i < ax->nst; ++i) ax->radix[i] = r[i];
    // digit reversal: slot i holds position rev(i); pinv maps position -> slot
    std::vector<int> pinv;
    std::vector<float> tw;
    host::fft_tables(L, r, pinv, tw);
    if (hipMalloc((void **)&ax->tw, L * sizeof(float2)) != hipSuccess ||
        hipMalloc((void **)&ax->pinv, L * sizeof(int)) != hipSuccess)
        return fail(GN_ENOMEM, "FFT tables");
    hipError_t e = hipMemcpy(ax->tw, tw.data(), L * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ax->pinv, pinv.data(), L * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "FFT tables");
    return GN_OK;
}

void fft_axis_free(FFTAxis *ax)
{
    hipFree(ax->tw);
    hipFree(ax->pinv);
    ax->tw = nullptr;
    ax->pinv = nullptr;
}

// SENSE pass A: coil-less image x -> leading grid rows of x * S_c for every coil.
int pruned_type2_rows_sense(gn_plan_s *p, const float2 *image, float scale, float sign)
{
    const Geometry &g = p->geo;
    const FFTAxis &az = p->fft_ax[g.dim - 1];
    RowArgs ra = row_args(p, az, sign, scale);
    ra.img_in = image;
    ra.sens = p->sens;
    dim3 ga((unsigned)ra.nlines, (unsigned)((p->nc + ra.CL - 1) / ra.CL));
    hipLaunchKernelGGL(k_fft_rows_pad, ga, dim3(kFFTThreads), row_lds(az.L, ra.CL), p->stream, ra);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GN_OK : hip_fail(e, "FFT pass A (SENSE)");
}

// SENSE pass A': every coil of a line in one block, combined with conj(S_c) into x.
// GN_EINVAL when the coils do not fit the LDS of one block (the caller then combines
// separately).
int pruned_type1_rows_sense(gn_plan_s *p, float2 *image, float scale, float sign)
{
    const Geometry &g = p->geo;
    const FFTAxis &az = p->fft_ax[g.dim - 1];
    RowArgs ra = row_args(p, az, sign, scale);
    const size_t lds = row_lds(az.L, (int)p->nc);
    if (p->nc > 1024 || lds > 160 * 1024) return fail(GN_EINVAL, "SENSE combine: coils do not fit one block");
    ra.CL = (int)p->nc;
    ra.img_out = image;
    ra.sens = p->sens;
    allow_lds((const void *)k_fft_rows_crop, lds);
    hipLaunchKernelGGL(k_fft_rows_crop, dim3((unsigned)ra.nlines, 1), dim3(kFFTThreads), lds, p->stream, ra);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GN_OK : hip_fail(e, "FFT pass A' (SENSE)");
}

bool sense_rows_fit(const gn_plan_s *p)
{
    return p->nc <= 1024 && row_lds(p->fft_ax[p->geo.dim - 1].L, (int)p->nc) <= 160 * 1024;
}

// Type 2, pass A: image -> leading grid rows (pad, deapodize, scale, FFT along the last axis).
int pruned_type2_rows(gn_plan_s *p, const float2 *image, float scale, float sign)
{
    const Geometry &g = p->geo;
    const FFTAxis &az = p->fft_ax[g.dim - 1];
    RowArgs ra = row_args(p, az, sign, scale);
    ra.img_in = image;
    dim3 ga((unsigned)ra.nlines, (unsigned)((p->nc + ra.CL - 1) / ra.CL));
    hipLaunchKernelGGL(k_fft_rows_pad, ga, dim3(kFFTThreads), row_lds(az.L, ra.CL), p->stream, ra);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GN_OK : hip_fail(e, "FFT pass A");
}

// Mask geometry of the leading-axis pass: rows per mask bit, and the spread tiling of
// the other axes that maps a grid column to its mask word.
static void set_mask_geometry(const gn_plan_s *p, ColArgs &a)
{
    const Geometry &g = p->geo;
    a.mrow = g.tile[0] * p->rec.mask_k;
    a.mcz = g.dim == 3 ? (int)g.nf[2] : 1;
    a.mt1 = g.tile[1];
    a.mt2 = g.dim == 3 ? g.tile[2] : 1;
    a.mnt2 = g.dim == 3 ? g.ntiles[2] : 1;
}

// Type 2, passes B (3-D) and C: the remaining axes, expanding m rows to nf.
int pruned_type2_cols(gn_plan_s *p, float sign)
{
    ColArgs a[2];
    const int n = pruned_col_args(p, true, sign, a);
    for (int k = 0; k < n; ++k) {
        hipError_t e;
        ColArgs &c = a[k];
        const size_t wide_lds = col_lds(c.L, kWideCB);
        if (k == n - 1 && p->geo.dim == 3 && c.CB == GN_FFT_CB && wide_lds <= 80 * 1024) {
            c.CB = kWideCB; // two blocks per CU
            dim3 grid((unsigned)((c.ncols + c.CB - 1) / c.CB), (unsigned)c.nslabs);
            allow_lds((const void *)k_fft_cols_wide, wide_lds);
            hipLaunchKernelGGL(k_fft_cols_wide, grid, dim3(kFFTColThreadsWide), wide_lds, p->stream, c);
            e = hipGetLastError();
        } else {
            e = launch_cols(c, p->stream);
        }
        if (e != hipSuccess) return hip_fail(e, k + 1 < n ? "FFT pass B" : "FFT pass C");
    }
    return GN_OK;
}

// Type 1, passes C' and B' (3-D): keep m of nf rows on the leading axes (in place).
// Pass C' reads only the spread tiles that hold pairs (occupancy mask); the spread
// does not write the others.
int pruned_type1_cols(gn_plan_s *p, float sign)
{
    ColArgs a[2];
    const int n = pruned_col_args(p, false, sign, a);
    if (p->rec.occ_mask) {
        a[0].rmask = p->rec.occ_mask;
        set_mask_geometry(p, a[0]);
    }
    for (int k = 0; k < n; ++k) {
        hipError_t e = launch_cols(a[k], p->stream);
        if (e != hipSuccess) return hip_fail(e, k == 0 ? "FFT pass C'" : "FFT pass B'");
    }
    return GN_OK;
}

// Type 1, pass A': FFT along the last axis, crop, deapodize, scale -> image.
int pruned_type1_rows(gn_plan_s *p, float2 *image, float scale, float sign)
{
    const Geometry &g = p->geo;
    const FFTAxis &az = p->fft_ax[g.dim - 1];
    RowArgs ra = row_args(p, az, sign, scale);
    ra.img_out = image;
    dim3 ga((unsigned)ra.nlines, (unsigned)((p->nc + ra.CL - 1) / ra.CL));
    hipLaunchKernelGGL(k_fft_rows_crop, ga, dim3(kFFTThreads), row_lds(az.L, ra.CL), p->stream, ra);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GN_OK : hip_fail(e, "FFT pass A'");
}

} // namespace gn
