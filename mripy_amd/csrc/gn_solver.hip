// Device primitives of the iterative reconstructions (include/gaussnufft_solver.h):
// FP64 deterministic reductions and the Chambolle TV prox in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "gaussnufft_solver.h"
#include "gn_internal.h"

namespace gn {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024; // partials of one reduction (fixed upper bound)
constexpr int kRedUnroll = 4;    // elements per thread per grid-stride round

// ---------------------------------------------------------------- element loads
struct C2 {
    double re, im;
};

__device__ inline C2 load(const float *p, int64_t i) { return {(double)p[i], 0.0}; }
__device__ inline C2 load(const float2 *p, int64_t i)
{
    float2 v = p[i];
    return {(double)v.x, (double)v.y};
}
__device__ inline C2 load(const double2 *p, int64_t i)
{
    double2 v = p[i];
    return {v.x, v.y};
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of (a, b) in a fixed order: butterfly within each 64-lane wave, then
// the 4 wave sums in wave order.  Result valid in thread 0.
__device__ inline void block_sum2(double &a, double &b)
{
    __shared__ double sa[kRedThreads / 64], sb[kRedThreads / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0;
        b = 0;
        for (int k = 0; k < kRedThreads / 64; ++k) {
            a += sa[k];
            b += sb[k];
        }
    }
}

// OP 0: sum conj(a) b;  OP 1: sum |a - b|^2 (b may be null)
template <typename T, int OP>
__global__ __launch_bounds__(kRedThreads) void k_reduce_partial(const T *__restrict__ a, const T *__restrict__ b,
                                                                int64_t n, double2 *__restrict__ part)
{
    double sr = 0.0, si = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kRedThreads;
    for (int64_t base = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; base < n; base += stride * kRedUnroll) {
#pragma unroll
        for (int u = 0; u < kRedUnroll; ++u) {
            const int64_t i = base + u * stride;
            if (i < n) {
                C2 x = load(a, i);
                if (OP == 0) {
                    C2 y = load(b, i);
                    sr += x.re * y.re + x.im * y.im;
                    si += x.re * y.im - x.im * y.re;
                } else {
                    if (b) {
                        C2 y = load(b, i);
                        x.re -= y.re;
                        x.im -= y.im;
                    }
                    sr += x.re * x.re + x.im * x.im;
                }
            }
        }
    }
    block_sum2(sr, si);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr, si);
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const double2 *__restrict__ part, int nb, int nout,
                                                              double *__restrict__ out)
{
    double sr = 0.0, si = 0.0;
    for (int k = threadIdx.x; k < nb; k += kRedThreads) {
        sr += part[k].x;
        si += part[k].y;
    }
    block_sum2(sr, si);
    if (threadIdx.x == 0) {
        out[0] = sr;
        if (nout > 1) out[1] = si;
    }
}

int reduce_blocks(int64_t n)
{
    const int64_t per = (int64_t)kRedThreads * kRedUnroll;
    int64_t nb = (n + per - 1) / per;
    if (nb > kRedBlocks) nb = kRedBlocks;
    return (int)nb;
}

template <int OP>
int reduce(int dtype, int64_t n, const void *a, const void *b, void *work, double *out, void *stream)
{
    if (n < 0 || !out || !work || (n > 0 && !a) || (OP == 0 && n > 0 && !b))
        return fail(GN_EINVAL, "reduction: negative size or null pointer");
    hipStream_t s = (hipStream_t)stream;
    double2 *part = (double2 *)work;
    const int nb = reduce_blocks(n);
    if (nb > 0) {
        switch (dtype) {
        case GN_F32:
            k_reduce_partial<float, OP><<<nb, kRedThreads, 0, s>>>((const float *)a, (const float *)b, n, part);
            break;
        case GN_C64:
            k_reduce_partial<float2, OP><<<nb, kRedThreads, 0, s>>>((const float2 *)a, (const float2 *)b, n, part);
            break;
        case GN_C128:
            k_reduce_partial<double2, OP>
                <<<nb, kRedThreads, 0, s>>>((const double2 *)a, (const double2 *)b, n, part);
            break;
        default:
            return fail(GN_EINVAL, "reduction: unknown dtype " + std::to_string(dtype));
        }
    } else if (dtype != GN_F32 && dtype != GN_C64 && dtype != GN_C128) {
        return fail(GN_EINVAL, "reduction: unknown dtype " + std::to_string(dtype));
    }
    k_reduce_final<<<1, kRedThreads, 0, s>>>(part, nb, OP == 0 ? 2 : 1, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "reduction launch");
    return GN_OK;
}

// ---------------------------------------------------------------- TV prox (FP64)
//
// Arrays are C-ordered (s0, s1[, s2], nt): NDIM image axes, then nt trailing
// elements.  grad(x)_a = x[.. i_a+1 ..] - x[.. i_a ..], zero on the last slice of axis a.
// Div(p) = sum_a adjD_a(p_a) with adjD(p)[i] = p[i-1]-p[i] (1 <= i <= s-2),
// adjD(p)[0] = -p[0], adjD(p)[s-1] = p[s-2]: the exact adjoint of grad.

struct TVGeom {
    int s[3];      // image axes (s[2] = 1 for NDIM 2)
    int nt;        // trailing elements per pixel
    int stride[3]; // element strides of the image axes
    int n;         // total elements
};

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <int NDIM> __device__ inline void tv_coords(const TVGeom &g, int idx, int *c)
{
    int r = idx / g.nt;
    if (NDIM == 3) {
        c[2] = r % g.s[2];
        r /= g.s[2];
    }
    c[1] = r % g.s[1];
    c[0] = r / g.s[1];
}

__device__ inline double2 adj_at(const double2 *__restrict__ p, int i, int s, int stride, int idx)
{
    if (i == 0) return make_double2(-p[idx].x, -p[idx].y);
    if (i == s - 1) return p[idx - stride];
    return csub(p[idx - stride], p[idx]);
}

template <typename T> __device__ inline double2 ld2(const T *p, int i);
template <> __device__ inline double2 ld2<float2>(const float2 *p, int i)
{
    float2 v = p[i];
    return make_double2(v.x, v.y);
}
template <> __device__ inline double2 ld2<double2>(const double2 *p, int i) { return p[i]; }
__device__ inline void st2(float2 *p, int i, double2 v) { p[i] = make_float2((float)v.x, (float)v.y); }
__device__ inline void st2(double2 *p, int i, double2 v) { p[i] = v; }

struct TVField {
    double2 *g[3]; // dual field components
    double2 *D;    // Div(G) - y/lambda
};

// D = Div(G) - y/lambda (FINAL: f = y - lambda*Div(G)).  FIRST: G is zero.
template <typename T, int NDIM, bool FIRST, bool FINAL>
__global__ __launch_bounds__(256) void k_tv_div(TVField fld, const T *__restrict__ y, double lambda, TVGeom g,
                                                T *__restrict__ f)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= g.n) return;
    int c[3];
    tv_coords<NDIM>(g, idx, c);
    const double2 yv = ld2<T>(y, idx);
    double2 dv = make_double2(0.0, 0.0);
    if (!FIRST) {
#pragma unroll
        for (int a = 0; a < NDIM; ++a) dv = cadd(dv, adj_at(fld.g[a], c[a], g.s[a], g.stride[a], idx));
    }
    if (FINAL)
        st2(f, idx, make_double2(yv.x - lambda * dv.x, yv.y - lambda * dv.y));
    else
        fld.D[idx] = make_double2(dv.x - yv.x / lambda, dv.y - yv.y / lambda);
}

// principal complex square root (the C99 csqrt branch: Re >= 0, Im takes the sign of b)
__device__ inline double2 csqrt_d(double2 z)
{
    const double a = z.x, b = z.y;
    if (a == 0.0 && b == 0.0) return make_double2(0.0, b);
    const double t = sqrt((fabs(a) + hypot(a, b)) * 0.5);
    if (a >= 0.0) return make_double2(t, b / (2.0 * t));
    return make_double2(fabs(b) / (2.0 * t), copysign(t, b));
}

__device__ inline double2 cdiv_d(double2 a, double2 d)
{
    // Smith's algorithm
    if (fabs(d.x) >= fabs(d.y)) {
        const double r = d.y / d.x, den = d.x + d.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    const double r = d.x / d.y, den = d.x * r + d.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

// G = G - step*grad(D); G = G / max(amp(G), 1).  FIRST: G is zero on entry.
template <int NDIM, bool FIRST, int AMP>
__global__ __launch_bounds__(256) void k_tv_step(TVField fld, double step, TVGeom g)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= g.n) return;
    int c[3];
    tv_coords<NDIM>(g, idx, c);
    const double2 d0 = fld.D[idx];
    double2 v[NDIM];
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
        const double2 dd = (c[a] < g.s[a] - 1) ? csub(fld.D[idx + g.stride[a]], d0) : make_double2(0.0, 0.0);
        const double2 old = FIRST ? make_double2(0.0, 0.0) : fld.g[a][idx];
        v[a] = make_double2(old.x - step * dd.x, old.y - step * dd.y);
    }
    if (AMP == GN_TV_AMP_SQUARE) {
        // s = sum_a v_a^2 with complex squares; d = sqrt(s); numpy maximum(d, 1) keeps d
        // when d >= 1 in lexicographic (real, imag) order or when d is NaN
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < NDIM; ++a)
            s = cadd(s, make_double2(v[a].x * v[a].x - v[a].y * v[a].y, v[a].x * v[a].y + v[a].y * v[a].x));
        const double2 d = csqrt_d(s);
        const bool keep = d.x > 1.0 || (d.x == 1.0 && d.y >= 0.0) || isnan(d.x) || isnan(d.y);
        if (keep) {
#pragma unroll
            for (int a = 0; a < NDIM; ++a) v[a] = cdiv_d(v[a], d);
        }
    } else {
        double m = 0.0;
#pragma unroll
        for (int a = 0; a < NDIM; ++a) m += v[a].x * v[a].x + v[a].y * v[a].y;
        m = sqrt(m);
        if (m > 1.0) {
#pragma unroll
            for (int a = 0; a < NDIM; ++a) v[a] = make_double2(v[a].x / m, v[a].y / m);
        }
    }
#pragma unroll
    for (int a = 0; a < NDIM; ++a) fld.g[a][idx] = v[a];
}

template <int NDIM, bool FIRST>
void tv_step_launch(int nb, hipStream_t s, TVField fld, double step, TVGeom g, int amp)
{
    if (amp == GN_TV_AMP_SQUARE)
        k_tv_step<NDIM, FIRST, GN_TV_AMP_SQUARE><<<nb, 256, 0, s>>>(fld, step, g);
    else
        k_tv_step<NDIM, FIRST, GN_TV_AMP_MODULUS><<<nb, 256, 0, s>>>(fld, step, g);
}

template <typename T, int NDIM>
int tv_prox(TVGeom g, const T *y, double lambda, double step, int niter, int amp, double2 *work, T *f,
            hipStream_t s)
{
    TVField fld;
    for (int a = 0; a < 3; ++a) fld.g[a] = a < NDIM ? work + (int64_t)a * g.n : nullptr;
    fld.D = work + (int64_t)NDIM * g.n;
    const int nb = (g.n + 255) / 256;
    for (int it = 0; it < niter; ++it) {
        if (it == 0) {
            k_tv_div<T, NDIM, true, false><<<nb, 256, 0, s>>>(fld, y, lambda, g, nullptr);
            tv_step_launch<NDIM, true>(nb, s, fld, step, g, amp);
        } else {
            k_tv_div<T, NDIM, false, false><<<nb, 256, 0, s>>>(fld, y, lambda, g, nullptr);
            tv_step_launch<NDIM, false>(nb, s, fld, step, g, amp);
        }
    }
    if (niter > 0)
        k_tv_div<T, NDIM, false, true><<<nb, 256, 0, s>>>(fld, y, lambda, g, f);
    else
        k_tv_div<T, NDIM, true, true><<<nb, 256, 0, s>>>(fld, y, lambda, g, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "TV prox launch");
    return GN_OK;
}

} // namespace
} // namespace gn

using namespace gn;

extern "C" {

int64_t gn_reduce_work_bytes(void) { return (int64_t)kRedBlocks * (int64_t)sizeof(double2); }

int gn_dot_f64(int dtype, int64_t n, const void *a, const void *b, void *work, double *out, void *stream)
{
    return reduce<0>(dtype, n, a, b, work, out, stream);
}

int gn_sqdist_f64(int dtype, int64_t n, const void *a, const void *b, void *work, double *out, void *stream)
{
    return reduce<1>(dtype, n, a, b, work, out, stream);
}

int64_t gn_tv_work_bytes(int ndim, const int64_t *shape, int64_t nt)
{
    if ((ndim != 2 && ndim != 3) || !shape || nt < 0) return 0;
    int64_t n = nt;
    for (int a = 0; a < ndim; ++a) {
        if (shape[a] < 0) return 0;
        n *= shape[a];
    }
    return (int64_t)(ndim + 1) * n * (int64_t)sizeof(double2);
}

int gn_tv_prox(int dtype, int ndim, const int64_t *shape, int64_t nt, const void *y, double lambda, double step,
               int niter, int amp_rule, void *work, void *f, void *stream)
{
    if ((ndim != 2 && ndim != 3) || !shape) return fail(GN_EINVAL, "TV prox: ndim must be 2 or 3");
    int64_t n = nt;
    for (int a = 0; a < ndim; ++a) {
        if (shape[a] < 2)
            return fail(GN_EINVAL, "TV prox: every image axis needs at least 2 samples (axis " + std::to_string(a) +
                                       " has " + std::to_string(shape[a]) + ")");
        n *= shape[a];
    }
    if (nt < 1) return fail(GN_EINVAL, "TV prox: nt must be >= 1");
    if (n >= (int64_t)1 << 31) return fail(GN_EINVAL, "TV prox: array too large (>= 2^31 elements)");
    if (!y || !f || !work) return fail(GN_EINVAL, "TV prox: null pointer");
    if (y == f) return fail(GN_EINVAL, "TV prox: y and f must not alias");
    if (!(lambda > 0.0) || !std::isfinite(lambda)) return fail(GN_EINVAL, "TV prox: lambda must be positive");
    if (niter < 0) return fail(GN_EINVAL, "TV prox: niter < 0");
    if (amp_rule != GN_TV_AMP_SQUARE && amp_rule != GN_TV_AMP_MODULUS)
        return fail(GN_EINVAL, "TV prox: unknown amplitude rule");
    TVGeom g;
    g.s[0] = (int)shape[0];
    g.s[1] = (int)shape[1];
    g.s[2] = ndim == 3 ? (int)shape[2] : 1;
    g.nt = (int)nt;
    g.stride[2] = (int)nt;
    g.stride[1] = ndim == 3 ? g.s[2] * (int)nt : (int)nt;
    g.stride[0] = g.s[1] * g.stride[1];
    g.n = (int)n;
    hipStream_t s = (hipStream_t)stream;
    double2 *w = (double2 *)work;
    if (dtype == GN_C64) {
        if (ndim == 2) return tv_prox<float2, 2>(g, (const float2 *)y, lambda, step, niter, amp_rule, w, (float2 *)f, s);
        return tv_prox<float2, 3>(g, (const float2 *)y, lambda, step, niter, amp_rule, w, (float2 *)f, s);
    }
    if (dtype == GN_C128) {
        if (ndim == 2)
            return tv_prox<double2, 2>(g, (const double2 *)y, lambda, step, niter, amp_rule, w, (double2 *)f, s);
        return tv_prox<double2, 3>(g, (const double2 *)y, lambda, step, niter, amp_rule, w, (double2 *)f, s);
    }
    return fail(GN_EINVAL, "TV prox: dtype must be GN_C64 or GN_C128");
}

} // extern "C"
