// Cartesian complex-to-complex FFTs over the leading axes of a C-order array whose
// trailing dimensions are a batch (coils), interleaved element by element -- the
// layout of the reference's cuFFT wrappers -- with the centring shifts and the
// k-space mask of the masked Cartesian operators fused into one pass before and one
// pass after rocFFT.
//
//   forward: out = [ifftshift] ( FFT( [fftshift] in ) ) * mask       (unnormalized)
//   inverse: out = [ifftshift] ( IFFT( [fftshift] in ) ) / prod(shape)
//
// rocFFT plans are cached per (rank, shape, batch, direction) for the process; the
// shifted copy lives in a grow-only workspace per device.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "gn_internal.h"

namespace gn {
namespace {

struct CPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
};

std::mutex g_cfft_mu;
std::map<std::tuple<int, int, int64_t, int64_t, int64_t, int64_t, int, int>, CPlan> g_cfft_plans;
std::map<int, std::pair<void *, size_t>> g_cfft_ws;

struct ShiftArgs {
    int rank;
    int64_t n[3];      // transformed lengths (leading axes)
    int64_t batch;     // interleaved batch elements
    int64_t total;     // prod(n) * batch
    int pre;           // roll the input by n/2 (fftshift) while copying
    int post;          // take the output from position (k + n/2) mod n (ifftshift)
    float scale;
    const float *mask; // null, prod(n) (broadcast over the batch) or prod(n) * batch entries
    int64_t mask_batch;
};

// dst[e] = src[roll(e)] (* mask[e] * scale): e = (index of the leading axes) * batch + b.
// pre: fftshift (source index m - n/2); post: ifftshift (source index k + n/2).
__global__ __launch_bounds__(256) void k_shift_copy(ShiftArgs a, const float2 *__restrict__ src,
                                                    float2 *__restrict__ dst, int post)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.total; e += stride) {
        const int64_t b = e % a.batch;
        int64_t r = e / a.batch;
        int64_t srcpos = 0, mul = 1;
        int64_t pos = r;
        for (int ax = a.rank - 1; ax >= 0; --ax) {
            const int64_t n = a.n[ax];
            const int64_t i = r % n;
            r /= n;
            int64_t s = i;
            if (post ? a.post : a.pre) {
                s = post ? i + n / 2 : i - n / 2;
                if (s >= n) s -= n;
                if (s < 0) s += n;
            }
            srcpos += s * mul;
            mul *= n;
        }
        float2 v = src[srcpos * a.batch + b];
        if (post) {
            float f = a.scale;
            if (a.mask) f *= a.mask_batch > 1 ? a.mask[pos * a.batch + b] : a.mask[pos];
            v = make_float2(v.x * f, v.y * f);
        }
        dst[e] = v;
    }
}

int cfft_plan(int rank, const int64_t *n, int64_t batch, bool forward, CPlan **out)
{
    static std::once_flag once;
    std::call_once(once, []() { rocfft_setup(); });
    int dev = 0;
    (void)hipGetDevice(&dev);
    auto key = std::make_tuple(dev, rank, n[0], rank > 1 ? n[1] : 1, rank > 2 ? n[2] : 1, batch, forward ? 1 : 0, 0);
    auto it = g_cfft_plans.find(key);
    if (it != g_cfft_plans.end()) {
        *out = &it->second;
        return GN_OK;
    }
    size_t lengths[3], strides[3];
    for (int a = 0; a < rank; ++a) lengths[a] = (size_t)n[rank - 1 - a]; // fastest first
    strides[0] = (size_t)batch;
    for (int a = 1; a < rank; ++a) strides[a] = strides[a - 1] * lengths[a - 1];
    rocfft_plan_description desc = nullptr;
    rocfft_status s = rocfft_plan_description_create(&desc);
    if (s != rocfft_status_success) return fail(GN_EROCFFT, "rocFFT description");
    s = rocfft_plan_description_set_data_layout(desc, rocfft_array_type_complex_interleaved,
                                                rocfft_array_type_complex_interleaved, nullptr, nullptr, rank, strides,
                                                1, rank, strides, 1);
    CPlan cp;
    if (s == rocfft_status_success)
        s = rocfft_plan_create(&cp.plan, rocfft_placement_inplace,
                               forward ? rocfft_transform_type_complex_forward : rocfft_transform_type_complex_inverse,
                               rocfft_precision_single, rank, lengths, (size_t)batch, desc);
    rocfft_plan_description_destroy(desc);
    if (s != rocfft_status_success) return fail(GN_EROCFFT, "rocFFT plan create (Cartesian FFT)");
    size_t wsz = 0;
    rocfft_plan_get_work_buffer_size(cp.plan, &wsz);
    if ((s = rocfft_execution_info_create(&cp.info)) != rocfft_status_success)
        return fail(GN_EROCFFT, "rocFFT execution info");
    if (wsz) {
        if (hipMalloc(&cp.work, wsz) != hipSuccess) return fail(GN_ENOMEM, "hipMalloc failed for the rocFFT work buffer");
        if (rocfft_execution_info_set_work_buffer(cp.info, cp.work, wsz) != rocfft_status_success)
            return fail(GN_EROCFFT, "rocFFT work buffer");
    }
    auto ins = g_cfft_plans.emplace(key, cp);
    *out = &ins.first->second;
    return GN_OK;
}

} // namespace
} // namespace gn

using namespace gn;

extern "C" {

int gn_fft_c2c(int rank, const int64_t *shape, int64_t batch, int direction, int flags, const float *mask,
               int64_t mask_batch, const void *in, void *out, void *stream)
{
    set_error("");
    if (rank < 1 || rank > 3 || !shape || batch < 1 || !in || !out)
        return fail(GN_EINVAL, "gn_fft_c2c: bad arguments");
    if (direction != GN_FFT_FORWARD && direction != GN_FFT_INVERSE) return fail(GN_EINVAL, "gn_fft_c2c: direction");
    if (mask && mask_batch != 1 && mask_batch != batch) return fail(GN_EINVAL, "gn_fft_c2c: mask batch");
    ShiftArgs a;
    a.rank = rank;
    a.total = batch;
    for (int ax = 0; ax < 3; ++ax) a.n[ax] = ax < rank ? shape[ax] : 1;
    for (int ax = 0; ax < rank; ++ax) {
        if (shape[ax] < 1) return fail(GN_EINVAL, "gn_fft_c2c: non-positive length");
        a.total *= shape[ax];
    }
    a.batch = batch;
    a.pre = (flags & GN_FFT_SHIFT_IN) ? 1 : 0;
    a.post = (flags & GN_FFT_SHIFT_OUT) ? 1 : 0;
    const bool fwd = direction == GN_FFT_FORWARD;
    a.scale = fwd ? 1.0f : (float)(1.0 / (double)(a.total / batch));
    a.mask = mask;
    a.mask_batch = mask_batch;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(g_cfft_mu);
    CPlan *cp = nullptr;
    int rc = cfft_plan(rank, a.n, batch, fwd, &cp);
    if (rc) return rc;
    if (rocfft_execution_info_set_stream(cp->info, st) != rocfft_status_success)
        return fail(GN_EROCFFT, "rocFFT set stream");
    int dev = 0;
    (void)hipGetDevice(&dev);
    auto &ws = g_cfft_ws[dev];
    const size_t bytes = (size_t)a.total * sizeof(float2);
    if (ws.second < bytes) {
        if (ws.first) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(ws.first);
        }
        ws.first = nullptr;
        ws.second = 0;
        if (hipMalloc(&ws.first, bytes) != hipSuccess) return fail(GN_ENOMEM, "hipMalloc failed for the FFT workspace");
        ws.second = bytes;
    }
    float2 *w = (float2 *)ws.first;
    const int nb = (int)std::min<int64_t>((a.total + 255) / 256, 1 << 16);
    // input (shifted) into the workspace, transform in place, output (shifted, masked, scaled)
    hipLaunchKernelGGL(k_shift_copy, dim3(nb), dim3(256), 0, st, a, (const float2 *)in, w, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "FFT pre-pass");
    void *buf[1] = {w};
    if (rocfft_execute(cp->plan, buf, nullptr, cp->info) != rocfft_status_success)
        return fail(GN_EROCFFT, "rocFFT execute (Cartesian FFT)");
    hipLaunchKernelGGL(k_shift_copy, dim3(nb), dim3(256), 0, st, a, (const float2 *)w, (float2 *)out, 1);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "FFT post-pass");
    // the workspace is shared by every call on this device: calls must not overlap
    e = hipStreamSynchronize(st);
    return e == hipSuccess ? GN_OK : hip_fail(e, "FFT synchronize");
}

} // extern "C"
