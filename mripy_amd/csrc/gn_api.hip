// C ABI of the Gaussian-gridding NUFFT (see include/gaussnufft.h).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gn_internal.h"

namespace gn {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return GN_EHIP;
}

static int fft_fail(rocfft_status s, const char *what)
{
    g_err = std::string("rocFFT ") + what + " failed with status " + std::to_string((int)s);
    return GN_EROCFFT;
}

static std::once_flag g_rocfft_once;

// Reference parameter rule, fft/nufft_func.py:198-249.
static int params_from_eps(int dim, const int64_t *modes, double eps, int max_nspread, int *rat, int *ns)
{
    if (!(eps > 1e-33 && eps < 1e-1)) {
        char buf[128];
        snprintf(buf, sizeof(buf), "eps = %.0e; must satisfy 1e-33 < eps < 1e-1.", eps);
        return fail(GN_EINVAL, buf);
    }
    const double pi = 3.141592653589793;
    int r = eps > 1e-11 ? 2 : 3;
    int n = (int)(-std::log(eps) / (pi * (r - 1) / (r - 0.5)) + 0.5);
    if (max_nspread > 0) {
        n = std::min(n, max_nspread);
        n = std::max(n, 3);
    }
    *rat = r;
    *ns = n;
    (void)dim;
    (void)modes;
    return GN_OK;
}

static double tau_of(int rat, int ns, int64_t ms)
{
    const double pi = 3.141592653589793;
    const double lam = ns / (rat * (rat - 0.5));
    return pi * lam / ((double)ms * (double)ms);
}

static int setup_tiles(gn_plan_s *p)
{
    Geometry &g = p->geo;
    int tile[3] = {1, 1, 1}, gtile[3] = {1, 1, 1};
    default_tiles(g.dim, g.ns, g.nf, tile, gtile);
    for (int a = 0; a < g.dim; ++a) {
        if (p->tile_override[a] > 0) tile[a] = p->tile_override[a];
        if (p->gtile_override[a] > 0) gtile[a] = p->gtile_override[a];
    }
    // a window must not leave a tile and re-enter it around the periodic axis:
    // otherwise one tile spans the whole axis
    // spread tiles are at most 32 cells per axis (32-bit cell masks in the kernel)
    auto fit = [&](int t, int a, int tmax) {
        t = (int)std::min<int64_t>(std::min(std::max(t, 1), tmax), g.nf[a]);
        if (t + 2 * g.ns - 1 > g.nf[a]) t = g.nf[a] <= tmax ? (int)g.nf[a] : (int)(g.nf[a] - 2 * g.ns + 1);
        return t;
    };
    g.K = 1;
    g.Kg = 1;
    for (int a = 0; a < 3; ++a) {
        if (a >= g.dim) {
            g.tile[a] = g.gtile[a] = 1;
            g.ntiles[a] = g.ngtiles[a] = 1;
            continue;
        }
        g.tile[a] = fit(tile[a], a, 32);
        g.gtile[a] = fit(gtile[a], a, 1 << 20);
        g.ntiles[a] = (int)((g.nf[a] + g.tile[a] - 1) / g.tile[a]);
        g.ngtiles[a] = (int)((g.nf[a] + g.gtile[a] - 1) / g.gtile[a]);
        g.K *= g.ntiles[a];
        g.Kg *= g.ngtiles[a];
    }
    // LDS limits: the spread keeps per-coil planes of up to kCoilGroup coils of a tile,
    // the gather stages tile + halo of a coil pair (the trajectory gradient one coil)
    const int64_t tc = (int64_t)g.tile[0] * g.tile[1] * g.tile[2];
    const int64_t ncg = std::min<int64_t>(p->nc, kCoilGroup);
    int64_t rc = 1;
    for (int a = 0; a < g.dim; ++a) rc *= std::min<int64_t>(g.gtile[a] + 2 * g.ns - 1, g.nf[a]);
    const int64_t stage = 64 * (int64_t)(g.tile[0] + g.tile[1] + g.tile[2]) * 4 + 64 * ncg * 8 + 64 * 48;
    (void)tc;
    if (stage > 160 * 1024 || rc * kGatherCellBytes > 160 * 1024)
        return fail(GN_EINVAL, "tile shape exceeds the 160 KiB LDS budget");

    if (g.K > (int64_t)INT32_MAX / 2 || g.Kg > (int64_t)INT32_MAX / 2) return fail(GN_EINVAL, "too many tiles");
    return GN_OK;
}

static void destroy_fft(gn_plan_s *p)
{
    if (p->fft_fwd) rocfft_plan_destroy(p->fft_fwd);
    if (p->fft_bwd) rocfft_plan_destroy(p->fft_bwd);
    if (p->fft_info) rocfft_execution_info_destroy(p->fft_info);
    if (p->fft_work) hipFree(p->fft_work);
    p->fft_fwd = p->fft_bwd = nullptr;
    p->fft_info = nullptr;
    p->fft_work = nullptr;
}

static int create_fft(gn_plan_s *p)
{
    rocfft_status s;
    std::call_once(g_rocfft_once, []() { rocfft_setup(); });
    const Geometry &g = p->geo;
    size_t lengths[3];
    for (int a = 0; a < g.dim; ++a) lengths[a] = (size_t)g.nf[g.dim - 1 - a]; // fastest first
    rocfft_plan_description desc = nullptr;
    if ((s = rocfft_plan_description_create(&desc)) != rocfft_status_success) return fft_fail(s, "description");
    size_t strides[3];
    strides[0] = 1;
    for (int a = 1; a < g.dim; ++a) strides[a] = strides[a - 1] * lengths[a - 1];
    s = rocfft_plan_description_set_data_layout(desc, rocfft_array_type_complex_interleaved,
                                                rocfft_array_type_complex_interleaved, nullptr, nullptr, g.dim,
                                                strides, (size_t)g.G, g.dim, strides, (size_t)g.G);
    if (s != rocfft_status_success) {
        rocfft_plan_description_destroy(desc);
        return fft_fail(s, "data layout");
    }
    s = rocfft_plan_create(&p->fft_fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                           rocfft_precision_single, g.dim, lengths, (size_t)p->nc, desc);
    if (s == rocfft_status_success)
        s = rocfft_plan_create(&p->fft_bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                               rocfft_precision_single, g.dim, lengths, (size_t)p->nc, desc);
    rocfft_plan_description_destroy(desc);
    if (s != rocfft_status_success) return fft_fail(s, "plan create");
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(p->fft_fwd, &w1);
    rocfft_plan_get_work_buffer_size(p->fft_bwd, &w2);
    p->fft_work_size = std::max(w1, w2);
    if ((s = rocfft_execution_info_create(&p->fft_info)) != rocfft_status_success) return fft_fail(s, "exec info");
    if (p->fft_work_size) {
        if (hipMalloc(&p->fft_work, p->fft_work_size) != hipSuccess)
            return fail(GN_ENOMEM, "hipMalloc failed for the rocFFT work buffer");
        if ((s = rocfft_execution_info_set_work_buffer(p->fft_info, p->fft_work, p->fft_work_size)) !=
            rocfft_status_success)
            return fft_fail(s, "work buffer");
    }
    if ((s = rocfft_execution_info_set_stream(p->fft_info, p->stream)) != rocfft_status_success)
        return fft_fail(s, "set stream");
    return GN_OK;
}

// Stage timer: records an event pair around fn() when timing is enabled.
template <typename F>
static int timed(gn_plan_s *p, int stage, F fn)
{
    if (!p->timing) return fn();
    auto take = [&]() -> hipEvent_t {
        if (!p->event_pool.empty()) {
            hipEvent_t e = p->event_pool.back();
            p->event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    };
    hipEvent_t a = take(), b = take();
    hipEventRecord(a, p->stream);
    int rc = fn();
    hipEventRecord(b, p->stream);
    p->pending.push_back({stage, a, b});
    return rc;
}

// Oversampled-grid FFT in place; forward = e^{-i}, inverse = e^{+i}, both unnormalized.
static int run_fft(gn_plan_s *p, bool forward)
{
    void *buf[1] = {p->grid};
    rocfft_status s = rocfft_execute(forward ? p->fft_fwd : p->fft_bwd, buf, nullptr, p->fft_info);
    if (s != rocfft_status_success) return fft_fail(s, "execute");
    return GN_OK;
}

} // namespace gn

using namespace gn;

extern "C" {

int gn_version(void) { return GN_VERSION; }

const char *gn_last_error(void) { return g_err.c_str(); }

int gn_grid_params(int dim, const int64_t *modes, double eps, int max_nspread, int *rat, int *nspread, int64_t *nf,
                   double *tau)
{
    g_err.clear();
    if ((dim != 2 && dim != 3) || !modes || !rat || !nspread || !nf || !tau)
        return fail(GN_EINVAL, "gn_grid_params: bad arguments");
    int rc = params_from_eps(dim, modes, eps, max_nspread, rat, nspread);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) nf[a] = a < dim ? std::max<int64_t>((int64_t)*rat * modes[a], 2 * *nspread) : 1;
    *tau = tau_of(*rat, *nspread, modes[0]);
    return GN_OK;
}

} // extern "C"

namespace gn {

// Plan set-up shared by gn_plan_create and batch members.  with_grid: allocate the
// oversampled grid and create the rocFFT plans (batch members use the batch's).
int plan_create_impl(int dim, const int64_t *modes, int rat, int nspread, int iflag, int64_t ncoils, int layout,
                     void *stream, bool with_grid, gn_plan_s **plan)
{
    if (!plan) return fail(GN_EINVAL, "plan pointer is null");
    *plan = nullptr;
    if (dim != 2 && dim != 3) return fail(GN_EINVAL, "dim must be 2 or 3");
    if (!modes) return fail(GN_EINVAL, "modes is null");
    if (rat < 2) return fail(GN_EINVAL, "rat must be >= 2");
    const int max_ns = dim == 3 ? 8 : 16;
    if (nspread < 2 || nspread > max_ns)
        return fail(GN_EINVAL, dim == 3 ? "nspread out of the supported range [2, 8] for 3-D plans"
                                        : "nspread out of the supported range [2, 16] for 2-D plans (eps >= ~1e-18)");
    if (ncoils < 1) return fail(GN_EINVAL, "ncoils must be >= 1");
    if (layout != GN_COIL_LAST && layout != GN_COIL_MAJOR) return fail(GN_EINVAL, "unknown layout");
    for (int a = 0; a < dim; ++a)
        if (modes[a] < 1) return fail(GN_EINVAL, "modes must be >= 1");

    gn_plan_s *p = new gn_plan_s();
    Geometry &g = p->geo;
    g.dim = dim;
    g.ns = nspread;
    g.G = 1;
    g.N = 1;
    for (int a = 0; a < 3; ++a) {
        g.modes[a] = a < dim ? modes[a] : 1;
        g.nf[a] = a < dim ? std::max<int64_t>((int64_t)rat * modes[a], 2 * nspread) : 1;
        g.G *= g.nf[a];
        g.N *= g.modes[a];
    }
    p->rat = rat;
    p->iflag = iflag;
    p->tau = tau_of(rat, nspread, g.modes[0]);
    p->nc = ncoils;
    p->layout = layout;
    p->stream = (hipStream_t)stream;
    const double pi = 3.141592653589793;
    for (int a = 0; a < 3; ++a) {
        const double h = 2 * pi / (double)g.nf[a];
        g.coef[a] = (float)(1.4426950408889634 * h * h / (4.0 * p->tau));
    }
    int rc = setup_tiles(p);
    if (rc) {
        delete p;
        return rc;
    }

    // deapodization tables exp(tau k^2), k = n - m//2, per axis (fft/nufft_func.py:735-738)
    std::vector<float> deap;
    for (int a = 0; a < 3; ++a)
        for (int64_t n = 0; n < g.modes[a]; ++n) {
            const double k = (double)(n - g.modes[a] / 2);
            deap.push_back(a < dim ? (float)std::exp(p->tau * k * k) : 1.0f);
        }
    if (hipMalloc((void **)&p->deap, deap.size() * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&p->err_flag, sizeof(int)) != hipSuccess ||
        (with_grid && hipMalloc((void **)&p->grid, (size_t)g.G * ncoils * sizeof(float2)) != hipSuccess)) {
        gn_plan_destroy(p);
        return fail(GN_ENOMEM, "hipMalloc failed for the oversampled grid");
    }
    hipError_t e = hipMemcpy(p->deap, deap.data(), deap.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        gn_plan_destroy(p);
        return hip_fail(e, "deapodization upload");
    }
    if (with_grid) {
        p->fft_custom = true;
        for (int a = 0; a < dim; ++a) p->fft_custom &= fft_length_supported(g.nf[a]);
        for (int a = 0; a < dim && p->fft_custom && !rc; ++a) rc = fft_axis_init(&p->fft_ax[a], (int)g.nf[a]);
        if (!rc && !p->fft_custom) rc = create_fft(p);
        if (rc) {
            std::string keep = g_err;
            gn_plan_destroy(p);
            g_err = keep;
            return rc;
        }
    }
    *plan = p;
    return GN_OK;
}

} // namespace gn

extern "C" {

int gn_plan_create(int dim, const int64_t *modes, int rat, int nspread, int iflag, int64_t ncoils, int layout,
                   void *stream, gn_plan **plan)
{
    g_err.clear();
    return plan_create_impl(dim, modes, rat, nspread, iflag, ncoils, layout, stream, true, plan);
}

int gn_plan_set_stream(gn_plan *p, void *stream)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    p->stream = (hipStream_t)stream;
    if (p->fft_info) {
        rocfft_status s = rocfft_execution_info_set_stream(p->fft_info, p->stream);
        if (s != rocfft_status_success) return fft_fail(s, "set stream");
    }
    return GN_OK;
}

int gn_plan_set_tiles(gn_plan *p, const int *tile)
{
    g_err.clear();
    if (!p || !tile) return fail(GN_EINVAL, "bad arguments");
    p->grid_t2_img = nullptr;
    int old[3], oldg[3];
    memcpy(old, p->tile_override, sizeof(old));
    memcpy(oldg, p->gtile_override, sizeof(oldg));
    for (int a = 0; a < 3; ++a) {
        p->tile_override[a] = tile[a] > 0 ? tile[a] : 0;
        p->gtile_override[a] = tile[3 + a] > 0 ? tile[3 + a] : 0;
    }
    Geometry save = p->geo;
    int rc = setup_tiles(p);
    if (rc) {
        memcpy(p->tile_override, old, sizeof(old));
        memcpy(p->gtile_override, oldg, sizeof(oldg));
        p->geo = save;
        return rc;
    }
    if (p->has_points && (memcmp(save.tile, p->geo.tile, sizeof(save.tile)) != 0 ||
                          memcmp(save.gtile, p->geo.gtile, sizeof(save.gtile)) != 0)) {
        // records are binned for the old tiles: require a new set_points
        p->has_points = false;
    }
    return GN_OK;
}

int gn_plan_set_sense(gn_plan *p, const void *maps)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    p->grid_t2_img = nullptr;
    if (!maps) {
        hipFree(p->sens);
        hipFree(p->sbuf);
        p->sens = p->sbuf = nullptr;
        return GN_OK;
    }
    const size_t bytes = (size_t)(p->geo.N * p->nc) * sizeof(float2);
    if (!p->sens && hipMalloc((void **)&p->sens, bytes) != hipSuccess) {
        p->sens = nullptr;
        return fail(GN_ENOMEM, "hipMalloc failed for the SENSE maps");
    }
    hipError_t e = hipMemcpyAsync(p->sens, maps, bytes, hipMemcpyDeviceToDevice, p->stream);
    if (e != hipSuccess) return hip_fail(e, "SENSE maps copy");
    return GN_OK;
}

int gn_set_points(gn_plan *p, int64_t M, const void *x, const void *y, const void *z, int coord_f64, double df,
                  const float *dcf)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    if (M < 0 || M > (int64_t)INT32_MAX) return fail(GN_EINVAL, "sample count out of range");
    if (M > 0 && (!x || !y || (p->geo.dim == 3 && !z))) return fail(GN_EINVAL, "coordinate pointer is null");
    if (!std::isfinite(df)) return fail(GN_EINVAL, "df must be finite");
    p->has_points = false;
    int rc = bin_and_sort(p, M, x, y, z, coord_f64, df, dcf);
    if (rc) return rc;
    p->M = M;
    p->df = df;
    p->has_dcf = dcf != nullptr;
    p->has_points = true;
    return GN_OK;
}

} // extern "C"

namespace gn {

// gn_execute's body.  ybuf_given: the type-21 intermediate samples (nc * M complex64)
// provided by the caller instead of the plan's lazily allocated buffer.
static int execute_impl(gn_plan *p, int kind, const void *in, void *out, float2 *ybuf_given)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    p->grid_t2_img = nullptr; // set again below after a plain type 2
    if (!p->has_points) return fail(GN_ESTATE, "gn_set_points must be called before gn_execute");
    const bool sense = kind == GN_SENSE_TYPE1 || kind == GN_SENSE_TYPE2 || kind == GN_SENSE_ADJOINT ||
                       kind == GN_SENSE_TYPE21;
    if (kind == GN_GRID_SPREAD || kind == GN_GRID_GATHER) {
        // the gridding step alone on a caller grid (coil-major nc x nf[0] x .. x nf[d-1])
        if (!in || !out) return fail(GN_EINVAL, "null data pointer");
        if (p->M == 0) {
            if (kind == GN_GRID_SPREAD) {
                hipError_t e = hipMemsetAsync(out, 0, (size_t)(p->geo.G * p->nc) * sizeof(float2), p->stream);
                if (e != hipSuccess) return hip_fail(e, "grid clear");
            }
            return GN_OK;
        }
        if (kind == GN_GRID_SPREAD) return launch_spread(p, (const float2 *)in, 0, (float2 *)out);
        return launch_gather(p, (float2 *)out, (const float2 *)in, (float)(1.0 / (double)p->geo.G));
    }
    if (kind != GN_TYPE1 && kind != GN_TYPE2 && kind != GN_ADJOINT && kind != GN_TYPE21 && !sense)
        return fail(GN_EINVAL, "unknown transform kind");
    if (sense && !p->sens) return fail(GN_ESTATE, "gn_plan_set_sense must be called before a SENSE transform");
    if (sense) {
        switch (kind) {
        case GN_SENSE_TYPE1: kind = GN_TYPE1; break;
        case GN_SENSE_TYPE2: kind = GN_TYPE2; break;
        case GN_SENSE_ADJOINT: kind = GN_ADJOINT; break;
        default: kind = GN_TYPE21; break;
        }
    }
    const Geometry &g = p->geo;
    const bool empty_in = (kind == GN_TYPE2 || kind == GN_TYPE21) ? false : p->M == 0;
    if ((!in && !empty_in) || (!out && !(kind == GN_TYPE2 && p->M == 0)))
        return fail(GN_EINVAL, "null data pointer");
    if ((kind == GN_TYPE1 || kind == GN_TYPE21) && p->M == 0)
        return fail(GN_EINVAL, "type-1 transform of zero samples (1/M undefined)");
    const double pi = 3.141592653589793;
    const double c = std::pow(pi / p->tau, g.dim / 2.0);
    // type 2 uses e^{-ik.x} for iflag >= 0; type 1 the opposite sign
    const bool fwd_dir_type2 = p->iflag >= 0;
    int rc = GN_OK;
    // pruned passes (sign of the type-2 FFT: -1 for iflag >= 0) or pad / rocFFT / crop
    const bool custom = p->fft_custom && p->fft_engine == GN_FFT_AUTO;
    if (!custom && !p->fft_fwd) {
        int r = create_fft(p);
        if (r) return r;
    }
    const float sign2 = fwd_dir_type2 ? -1.0f : 1.0f;
    auto type2 = [&](const void *img, void *samples) -> int {
        const float scale = (float)(c / (double)g.G);
        int r;
        if (custom) {
            r = timed(p, GN_STAGE_PAD, [&] {
                return sense ? pruned_type2_rows_sense(p, (const float2 *)img, scale, sign2)
                             : pruned_type2_rows(p, (const float2 *)img, scale, sign2);
            });
            if (!r) r = timed(p, GN_STAGE_FFT2, [&] { return pruned_type2_cols(p, sign2); });
        } else {
            r = timed(p, GN_STAGE_PAD, [&] { return launch_pad(p, (const float2 *)img, scale, sense); });
            if (!r) r = timed(p, GN_STAGE_FFT2, [&] { return run_fft(p, fwd_dir_type2); });
        }
        if (!r) r = timed(p, GN_STAGE_GATHER, [&] { return launch_gather(p, (float2 *)samples); });
        return r;
    };
    auto type1 = [&](const void *samples, void *img, bool dcf, double extra) -> int {
        const float scale = (float)(c * extra / (double)g.G);
        int r = timed(p, GN_STAGE_SPREAD, [&] { return launch_spread(p, (const float2 *)samples, dcf ? 1 : 0); });
        if (custom) {
            if (!r) r = timed(p, GN_STAGE_FFT1, [&] { return pruned_type1_cols(p, -sign2); });
            if (!r && sense && !sense_rows_fit(p)) {
                // too many coils for one block: coil images into the plan's buffer, then combine
                if (!p->sbuf && hipMalloc((void **)&p->sbuf, (size_t)(g.N * p->nc) * sizeof(float2)) != hipSuccess) {
                    p->sbuf = nullptr;
                    return fail(GN_ENOMEM, "hipMalloc failed for the SENSE coil images");
                }
                r = timed(p, GN_STAGE_CROP, [&] {
                    int q = pruned_type1_rows(p, p->sbuf, scale, -sign2);
                    return q ? q : launch_sense_combine(p, p->sbuf, (float2 *)img);
                });
            } else if (!r) {
                r = timed(p, GN_STAGE_CROP, [&] {
                    return sense ? pruned_type1_rows_sense(p, (float2 *)img, scale, -sign2)
                                 : pruned_type1_rows(p, (float2 *)img, scale, -sign2);
                });
            }
        } else {
            if (!r) r = timed(p, GN_STAGE_FFT1, [&] { return run_fft(p, !fwd_dir_type2); });
            if (!r) r = timed(p, GN_STAGE_CROP, [&] { return launch_crop(p, (float2 *)img, scale, sense); });
        }
        return r;
    };
    switch (kind) {
    case GN_TYPE2:
        rc = type2(in, out);
        if (!rc && !sense) p->grid_t2_img = in;
        break;
    case GN_TYPE1:
        rc = type1(in, out, true, 1.0 / (double)p->M);
        break;
    case GN_ADJOINT:
        rc = type1(in, out, false, 1.0);
        break;
    case GN_TYPE21: {
        const int64_t need = p->M * p->nc;
        if (ybuf_given) {
            rc = type2(in, ybuf_given);
            if (!rc) rc = type1(ybuf_given, out, true, 1.0 / (double)p->M);
            break;
        }
        if (need > p->ybuf_cap) {
            if (p->ybuf) hipFree(p->ybuf);
            p->ybuf = nullptr;
            if (hipMalloc((void **)&p->ybuf, (size_t)need * sizeof(float2)) != hipSuccess)
                return fail(GN_ENOMEM, "hipMalloc failed for the type-21 sample buffer");
            p->ybuf_cap = need;
        }
        rc = type2(in, p->ybuf);
        if (!rc) rc = type1(p->ybuf, out, true, 1.0 / (double)p->M);
        break;
    }
    }
    return rc;
}

// Intermediate samples of gn_execute_type21_noalloc: a file-scope device array, one
// copy per GPU, located with hipGetSymbolAddress once per device.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20; // 32 MiB = 4,194,304 complex64 samples x coils
__device__ __align__(256) char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;

static int scratch_address(void **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void *q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q) return hip_fail(e, "hipGetSymbolAddress of the no-allocation scratch");
        it = g_scratch_ptr.emplace(dev, q).first;
    }
    *ptr = it->second;
    return GN_OK;
}

} // namespace gn

extern "C" {

int gn_execute(gn_plan *p, int kind, const void *in, void *out) { return execute_impl(p, kind, in, out, nullptr); }

int gn_execute_type21_noalloc(gn_plan *p, const void *in, void *out)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    if (!p->has_points) return fail(GN_ESTATE, "gn_set_points must be called before gn_execute_type21_noalloc");
    if ((size_t)(p->M * p->nc) * sizeof(float2) > kNoAllocScratchBytes)
        return fail(GN_EINVAL, "gn_execute_type21_noalloc: M * ncoils exceeds the 4,194,304 complex samples of the "
                               "static scratch");
    if (!(p->fft_custom && p->fft_engine == GN_FFT_AUTO))
        return fail(GN_ESTATE, "gn_execute_type21_noalloc needs the pruned FFT passes (grid lengths with prime "
                               "factors 2, 3, 5 and the default engine): rocFFT plans allocate");
    if (p->timing) return fail(GN_ESTATE, "gn_execute_type21_noalloc: stage timers create events; disable them");
    void *ptr = nullptr;
    int rc = scratch_address(&ptr);
    if (rc) return rc;
    return execute_impl(p, GN_TYPE21, in, out, (float2 *)ptr);
}

int gn_traj_grad(gn_plan *p, const void *image, const void *weights, float *grad)
{
    return gn_traj_grad_ex(p, image, weights, grad, 0);
}

int gn_traj_grad_ex(gn_plan *p, const void *image, const void *weights, float *grad, int flags)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    if (!p->has_points) return fail(GN_ESTATE, "gn_set_points must be called before gn_traj_grad");
    if (p->M == 0) return GN_OK;
    if (!image || !weights || !grad) return fail(GN_EINVAL, "null data pointer");
    if (flags & ~GN_TRAJ_REUSE_GRID) return fail(GN_EINVAL, "unknown gn_traj_grad_ex flag");
    const Geometry &g = p->geo;
    const double pi = 3.141592653589793;
    const double c = std::pow(pi / p->tau, g.dim / 2.0);
    const float scale = (float)(c / (double)g.G);
    int rc;
    if ((flags & GN_TRAJ_REUSE_GRID) && p->grid_t2_img == image) {
        rc = GN_OK; // the grid already holds this image's type-2 grid
    } else if (p->fft_custom && p->fft_engine == GN_FFT_AUTO) {
        p->grid_t2_img = nullptr;
        const float sign2 = p->iflag >= 0 ? -1.0f : 1.0f;
        rc = timed(p, GN_STAGE_PAD, [&] { return pruned_type2_rows(p, (const float2 *)image, scale, sign2); });
        if (!rc) rc = timed(p, GN_STAGE_FFT2, [&] { return pruned_type2_cols(p, sign2); });
    } else {
        p->grid_t2_img = nullptr;
        rc = p->fft_fwd ? GN_OK : create_fft(p);
        if (!rc) rc = timed(p, GN_STAGE_PAD, [&] { return launch_pad(p, (const float2 *)image, scale); });
        if (!rc) rc = timed(p, GN_STAGE_FFT2, [&] { return run_fft(p, p->iflag >= 0); });
    }
    if (!rc) rc = launch_traj_grad(p, (const float2 *)weights, grad);
    p->grid_t2_img = rc ? nullptr : image;
    return rc;
}

int gn_plan_info(const gn_plan *p, int64_t *nf, int *nspread, double *tau, int64_t *M, int64_t *ncoils, int *tile)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    for (int a = 0; a < 3; ++a) {
        if (nf) nf[a] = p->geo.nf[a];
        if (tile) {
            tile[a] = p->geo.tile[a];
            tile[3 + a] = p->geo.gtile[a];
        }
    }
    if (nspread) *nspread = p->geo.ns;
    if (tau) *tau = p->tau;
    if (M) *M = p->M;
    if (ncoils) *ncoils = p->nc;
    return GN_OK;
}

int gn_plan_set_timing(gn_plan *p, int enable)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    p->timing = enable != 0;
    return GN_OK;
}

int gn_plan_stage_times(gn_plan *p, double *ms, int64_t *counts, int reset)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    if (!p->pending.empty()) {
        hipError_t e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) return hip_fail(e, "stage timer synchronize");
        for (auto &q : p->pending) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, q.a, q.b) == hipSuccess) {
                p->stage_ms[q.stage] += t;
                p->stage_n[q.stage] += 1;
            }
            p->event_pool.push_back(q.a);
            p->event_pool.push_back(q.b);
        }
        p->pending.clear();
    }
    for (int i = 0; i < GN_STAGE_COUNT; ++i) {
        if (ms) ms[i] = p->stage_ms[i];
        if (counts) counts[i] = p->stage_n[i];
        if (reset) {
            p->stage_ms[i] = 0;
            p->stage_n[i] = 0;
        }
    }
    return GN_OK;
}

int gn_plan_sample_order(const gn_plan *p, int *order)
{
    g_err.clear();
    if (!p || !order) return fail(GN_EINVAL, "bad arguments");
    if (!p->has_points) return fail(GN_ESTATE, "gn_set_points must be called first");
    if (p->M == 0) return GN_OK;
    hipError_t e = hipMemcpy2D(order, sizeof(int), (const char *)p->rec.cell + 3 * sizeof(int), sizeof(int4),
                               sizeof(int), (size_t)p->M, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "sample order copy");
    // The records of a gather tile are stored in LDS-bank order; the counting sort is
    // stable, so its order within a tile is ascending original index.
    std::vector<int> start((size_t)p->geo.Kg + 1);
    e = hipMemcpy(start.data(), p->rec.bin_start, sizeof(int) * start.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "sample order copy");
    for (int64_t k = 0; k < p->geo.Kg; ++k) std::sort(order + start[(size_t)k], order + start[(size_t)k + 1]);
    return GN_OK;
}

int gn_plan_set_fft_engine(gn_plan *p, int engine)
{
    g_err.clear();
    if (!p) return fail(GN_EINVAL, "plan is null");
    p->grid_t2_img = nullptr;
    if (engine != GN_FFT_AUTO && engine != GN_FFT_ROCFFT) return fail(GN_EINVAL, "unknown FFT engine");
    p->fft_engine = engine;
    return GN_OK;
}

int gn_plan_destroy(gn_plan *p)
{
    if (!p) return GN_OK;
    if (!p->pending.empty()) hipStreamSynchronize(p->stream);
    for (auto &q : p->pending) {
        hipEventDestroy(q.a);
        hipEventDestroy(q.b);
    }
    for (auto e : p->event_pool) hipEventDestroy(e);
    destroy_fft(p);
    for (auto &ax : p->fft_ax) fft_axis_free(&ax);
    hipFree(p->rec.cell);
    hipFree(p->rec.frac);
    hipFree(p->rec.bin_start);
    hipFree(p->rec.pair_a);
    hipFree(p->rec.pair_b);
    hipFree(p->rec.spread_items);
    hipFree(p->rec.reduce_items);
    hipFree(p->rec.gather_items);
    hipFree(p->rec.partial);
    hipFree(p->rec.occ_mask);
    hipFree(p->grid);
    hipFree(p->ybuf);
    hipFree(p->sens);
    hipFree(p->sbuf);
    hipFree(p->deap);
    hipFree(p->err_flag);
    delete p;
    return GN_OK;
}

} // extern "C"
