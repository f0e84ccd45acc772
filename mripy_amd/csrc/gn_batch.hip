// Batched uneven NUFFT (include/gaussnufft_batch.h): one launch per stage over
// the work of many plans, graph replay, and the tile autotuner.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "gaussnufft_batch.h"
#include "gn_internal.h"

namespace gn {
namespace {

constexpr int kKinds = 4; // table/graph slot per kind: type2, type1, adjoint, type21

int kind_slot(int kind)
{
    switch (kind) {
    case GN_TYPE2: return 0;
    case GN_TYPE1: return 1;
    case GN_ADJOINT: return 2;
    case GN_TYPE21: return 3;
    }
    return -1;
}

// instances of one oversampled-grid size: adjacent in the batch grid buffer, one
// batched rocFFT execution per direction
struct FFTGroup {
    int64_t nf[3] = {1, 1, 1};
    int64_t G = 1;
    int64_t batch = 0;
    int64_t offset = 0; // float2 elements into the batch grid buffer
    rocfft_plan fwd = nullptr, bwd = nullptr;
};

// one batched kernel launch: work items {instance, item, coil group, 0} of the
// instances that share its launch shape
struct LaunchList {
    int ncg = 0, waves = 0, nbg = 0; // spread
    bool mfma = false;               // spread on the matrix cores
    bool wrap = false;               // gather
    size_t lds = 0;
    std::vector<int4> items;
    int4 *d = nullptr;
};

// one batched pruned-FFT launch: work items of the instances whose pass has the same
// length; the LDS size is the largest any of them needs
struct FFTLaunch {
    size_t lds = 0;
    std::vector<int4> items;
    int4 *d = nullptr;
};

// device argument tables of one kind
struct Tables {
    SpreadArgs *spread = nullptr;
    GatherArgs *gather = nullptr;
    PrepostArgs *pad = nullptr;
    PrepostArgs *crop = nullptr;
    RowArgs *rowp = nullptr; // pruned pass A (pad rows) per instance
    RowArgs *rowc = nullptr; // pruned pass A' (crop rows) per instance
    std::vector<char> last; // host image of what was uploaded
};

int fft_error(rocfft_status s, const char *what)
{
    return fail(GN_EROCFFT, std::string("rocFFT ") + what + " failed with status " + std::to_string((int)s));
}

} // namespace
} // namespace gn

using namespace gn;

struct gn_batch_s {
    int dim = 2, n = 0, rat = 2, ns = 3, iflag = 1, layout = GN_COIL_LAST, coord_f64 = 0;
    double df = 1.0;
    hipStream_t stream = nullptr;  // the caller's
    hipStream_t gstream = nullptr; // graph replays (default priority)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::vector<gn_plan_s *> m;
    std::vector<void *> coords; // owned copies, 3 per instance
    std::vector<float *> dcf;   // owned copies or null
    float2 *grid = nullptr;     // all instances' oversampled grids
    float2 *ybuf = nullptr;     // type-21 intermediate samples
    std::vector<int64_t> yoff;
    std::vector<FFTGroup> groups;
    bool custom = false;               // pruned FFT passes (every grid length supported)
    std::vector<FFTLaunch> frows;      // items {instance, line, coil chunk}
    std::vector<FFTLaunch> fcols[2][2]; // [type 2, type 1][pass]: items {argument, column block, slab}
    ColArgs *d_cols[2] = {nullptr, nullptr}; // [type 2, type 1]: 2 entries per instance
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
    std::vector<LaunchList> spread_lists, gather_lists;
    std::vector<int4> reduce_items;
    int4 *d_reduce = nullptr;
    Tables tab[kKinds];
    int cur_tile[6] = {0, 0, 0, 0, 0, 0}; // last tile request applied to every instance
    int64_t maxG = 0, maxN = 0;
    bool graph = false;
    hipGraphExec_t exec[kKinds] = {nullptr, nullptr, nullptr, nullptr};
};

namespace gn {
namespace {

void free_lists(gn_batch_s *b)
{
    for (auto &l : b->spread_lists) hipFree(l.d);
    for (auto &l : b->gather_lists) hipFree(l.d);
    b->spread_lists.clear();
    b->gather_lists.clear();
    hipFree(b->d_reduce);
    b->d_reduce = nullptr;
    b->reduce_items.clear();
}

void drop_graphs(gn_batch_s *b)
{
    for (auto &e : b->exec) {
        if (e) hipGraphExecDestroy(e);
        e = nullptr;
    }
    for (auto &t : b->tab) t.last.clear();
}

int upload_list(std::vector<int4> &h, int4 **d)
{
    hipFree(*d);
    *d = nullptr;
    if (h.empty()) return GN_OK;
    if (hipMalloc((void **)d, h.size() * sizeof(int4)) != hipSuccess) return fail(GN_ENOMEM, "batch work items");
    hipError_t e = hipMemcpy(*d, h.data(), h.size() * sizeof(int4), hipMemcpyHostToDevice);
    return e == hipSuccess ? GN_OK : hip_fail(e, "batch work items");
}

// Work-item lists of all instances, grouped by launch shape.
int build_lists(gn_batch_s *b)
{
    free_lists(b);
    std::map<std::tuple<int, bool, int, int, size_t>, size_t> skey;
    std::map<std::tuple<int, size_t>, size_t> gkey;
    for (int i = 0; i < b->n; ++i) {
        gn_plan_s *p = b->m[i];
        const Geometry &g = p->geo;
        const SpreadArgs sa = make_spread_args(p, nullptr, 0);
        const int ncg = p->nc >= 8 ? 8 : p->nc >= 4 ? 4 : p->nc >= 2 ? 2 : 1;
        int waves = 0, nbg = 0;
        spread_shape(sa, g.dim, &waves, &nbg);
        const bool mfma = spread_uses_mfma(sa, g.dim);
        const size_t slds = spread_lds(sa, ncg, mfma, waves);
        auto sk = std::make_tuple(ncg, mfma, waves, nbg, slds);
        if (!skey.count(sk)) {
            skey[sk] = b->spread_lists.size();
            LaunchList l;
            l.ncg = ncg;
            l.mfma = mfma;
            l.waves = waves;
            l.nbg = nbg;
            l.lds = slds;
            b->spread_lists.push_back(l);
        }
        LaunchList &sl = b->spread_lists[skey[sk]];
        const int scg = (int)((p->nc + ncg - 1) / ncg);
        for (int k = 0; k < p->rec.n_spread_items; ++k)
            for (int c = 0; c < scg; ++c) sl.items.push_back(make_int4(i, k, c, 0));
        for (int r = 0; r < p->rec.n_reduce_items; ++r) b->reduce_items.push_back(make_int4(i, r, 0, 0));

        const GatherArgs ga = make_gather_args(p, nullptr);
        bool wrap = false;
        size_t cells = 1;
        for (int a = 0; a < g.dim; ++a) {
            wrap |= ga.region[a] >= ga.nf[a];
            cells *= (size_t)ga.region[a];
        }
        auto gk = std::make_tuple((int)wrap, cells * kGatherCellBytes);
        if (!gkey.count(gk)) {
            gkey[gk] = b->gather_lists.size();
            LaunchList l;
            l.wrap = wrap;
            l.lds = cells * kGatherCellBytes;
            b->gather_lists.push_back(l);
        }
        LaunchList &gl = b->gather_lists[gkey[gk]];
        const int gcg = (int)((p->nc + kGatherCoils - 1) / kGatherCoils);
        if (p->M > 0)
            for (int k = 0; k < p->rec.n_gather_items; ++k)
                for (int c = 0; c < gcg; ++c) gl.items.push_back(make_int4(i, k, c, 0));
    }
    int rc;
    for (auto &l : b->spread_lists)
        if ((rc = upload_list(l.items, &l.d))) return rc;
    for (auto &l : b->gather_lists)
        if ((rc = upload_list(l.items, &l.d))) return rc;
    return upload_list(b->reduce_items, &b->d_reduce);
}

// Re-bin every instance from the batch's own copies of the coordinates.
int rebin(gn_batch_s *b)
{
    for (int i = 0; i < b->n; ++i) {
        gn_plan_s *p = b->m[i];
        int rc = gn_set_points(p, p->M, b->coords[3 * i], b->coords[3 * i + 1], b->coords[3 * i + 2], b->coord_f64,
                               b->df, b->dcf[i]);
        if (rc) return rc;
    }
    drop_graphs(b);
    return build_lists(b);
}

int make_groups(gn_batch_s *b)
{
    // instances of equal grid size are adjacent in the grid buffer
    std::map<std::tuple<int64_t, int64_t, int64_t>, std::vector<int>> by;
    for (int i = 0; i < b->n; ++i) {
        const Geometry &g = b->m[i]->geo;
        by[std::make_tuple(g.nf[0], g.nf[1], g.nf[2])].push_back(i);
    }
    int64_t off = 0;
    for (auto &kv : by) {
        FFTGroup fg;
        fg.nf[0] = std::get<0>(kv.first);
        fg.nf[1] = std::get<1>(kv.first);
        fg.nf[2] = std::get<2>(kv.first);
        fg.G = fg.nf[0] * fg.nf[1] * fg.nf[2];
        fg.offset = off;
        for (int i : kv.second) {
            b->m[i]->grid = (float2 *)(intptr_t)(off * (int64_t)sizeof(float2)); // relocated below
            off += fg.G * b->m[i]->nc;
            fg.batch += b->m[i]->nc;
        }
        b->groups.push_back(fg);
    }
    if (hipMalloc((void **)&b->grid, (size_t)off * sizeof(float2)) != hipSuccess)
        return fail(GN_ENOMEM, "hipMalloc failed for the batch grid buffer");
    for (int i = 0; i < b->n; ++i) b->m[i]->grid = b->grid + ((intptr_t)b->m[i]->grid / (intptr_t)sizeof(float2));

    if (b->custom) return GN_OK; // pruned passes: no rocFFT plans
    rocfft_status s;
    size_t wmax = 0;
    for (auto &fg : b->groups) {
        const int d = b->dim;
        size_t lengths[3], strides[3];
        for (int a = 0; a < d; ++a) lengths[a] = (size_t)fg.nf[d - 1 - a];
        strides[0] = 1;
        for (int a = 1; a < d; ++a) strides[a] = strides[a - 1] * lengths[a - 1];
        rocfft_plan_description desc = nullptr;
        if ((s = rocfft_plan_description_create(&desc)) != rocfft_status_success) return fft_error(s, "description");
        s = rocfft_plan_description_set_data_layout(desc, rocfft_array_type_complex_interleaved,
                                                    rocfft_array_type_complex_interleaved, nullptr, nullptr, d,
                                                    strides, (size_t)fg.G, d, strides, (size_t)fg.G);
        if (s == rocfft_status_success)
            s = rocfft_plan_create(&fg.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                                   rocfft_precision_single, d, lengths, (size_t)fg.batch, desc);
        if (s == rocfft_status_success)
            s = rocfft_plan_create(&fg.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                                   rocfft_precision_single, d, lengths, (size_t)fg.batch, desc);
        rocfft_plan_description_destroy(desc);
        if (s != rocfft_status_success) return fft_error(s, "batch plan create");
        size_t w1 = 0, w2 = 0;
        rocfft_plan_get_work_buffer_size(fg.fwd, &w1);
        rocfft_plan_get_work_buffer_size(fg.bwd, &w2);
        wmax = std::max(wmax, std::max(w1, w2));
    }
    if ((s = rocfft_execution_info_create(&b->info)) != rocfft_status_success) return fft_error(s, "exec info");
    if (wmax) {
        if (hipMalloc(&b->work, wmax) != hipSuccess) return fail(GN_ENOMEM, "rocFFT work buffer");
        if ((s = rocfft_execution_info_set_work_buffer(b->info, b->work, wmax)) != rocfft_status_success)
            return fft_error(s, "work buffer");
    }
    return GN_OK;
}

int run_ffts(gn_batch_s *b, bool forward, hipStream_t st)
{
    rocfft_status s = rocfft_execution_info_set_stream(b->info, st);
    if (s != rocfft_status_success) return fft_error(s, "set stream");
    for (auto &fg : b->groups) {
        void *buf[1] = {b->grid + fg.offset};
        s = rocfft_execute(forward ? fg.fwd : fg.bwd, buf, nullptr, b->info);
        if (s != rocfft_status_success) return fft_error(s, "execute");
    }
    return GN_OK;
}

// Work items and column-pass arguments of the batched pruned FFT (geometry only).
int build_fft_lists(gn_batch_s *b)
{
    const float sign2 = b->iflag >= 0 ? -1.0f : 1.0f;
    std::vector<ColArgs> cols[2];
    std::map<int, size_t> rkey;
    std::map<int, size_t> ckey[2][2];
    for (int t = 0; t < 2; ++t) cols[t].resize(2 * (size_t)b->n);
    for (int i = 0; i < b->n; ++i) {
        const gn_plan_s *p = b->m[i];
        const RowArgs ra = make_row_args(p, sign2, 1.0f);
        if (!rkey.count(ra.L)) {
            rkey[ra.L] = b->frows.size();
            b->frows.push_back(FFTLaunch());
        }
        FFTLaunch &rl = b->frows[rkey[ra.L]];
        rl.lds = std::max(rl.lds, row_lds(ra.L, ra.CL)); // coils per block vary with the instance
        const int64_t nch = (p->nc + ra.CL - 1) / ra.CL;
        for (int64_t line = 0; line < ra.nlines; ++line)
            for (int64_t c = 0; c < nch; ++c) rl.items.push_back(make_int4(i, (int)line, (int)c, 0));
        for (int t = 0; t < 2; ++t) {
            const int np = pruned_col_args(p, t == 0, t == 0 ? sign2 : -sign2, &cols[t][2 * (size_t)i]);
            for (int k = 0; k < np; ++k) {
                const ColArgs &a = cols[t][2 * (size_t)i + k];
                if (!ckey[t][k].count(a.L)) {
                    ckey[t][k][a.L] = b->fcols[t][k].size();
                    FFTLaunch l;
                    l.lds = col_lds(a.L, a.CB);
                    b->fcols[t][k].push_back(l);
                }
                FFTLaunch &cl = b->fcols[t][k][ckey[t][k][a.L]];
                // column blocks fastest: neighbouring blocks read neighbouring columns
                const int64_t ncb = (a.ncols + a.CB - 1) / a.CB;
                for (int64_t sl = 0; sl < a.nslabs; ++sl)
                    for (int64_t cb = 0; cb < ncb; ++cb)
                        cl.items.push_back(make_int4(2 * i + k, (int)cb, (int)sl, 0));
            }
        }
    }
    int rc;
    for (auto &l : b->frows)
        if ((rc = upload_list(l.items, &l.d))) return rc;
    for (int t = 0; t < 2; ++t) {
        for (int k = 0; k < 2; ++k)
            for (auto &l : b->fcols[t][k])
                if ((rc = upload_list(l.items, &l.d))) return rc;
        if (hipMalloc((void **)&b->d_cols[t], cols[t].size() * sizeof(ColArgs)) != hipSuccess)
            return fail(GN_ENOMEM, "batch FFT argument tables");
        hipError_t e = hipMemcpy(b->d_cols[t], cols[t].data(), cols[t].size() * sizeof(ColArgs), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "batch FFT argument tables");
    }
    return GN_OK;
}

int launch_fft_cols(gn_batch_s *b, int t, hipStream_t st)
{
    for (int k = 0; k < 2; ++k)
        for (auto &l : b->fcols[t][k]) {
            hipError_t e = launch_fft_cols_batch(b->d_cols[t], l.d, (int)l.items.size(), l.lds, st);
            if (e != hipSuccess) return hip_fail(e, "batch FFT column pass");
        }
    return GN_OK;
}

int launch_fft_rows(gn_batch_s *b, bool pad, const RowArgs *args, hipStream_t st)
{
    for (auto &l : b->frows) {
        hipError_t e = launch_fft_rows_batch(pad, args, l.d, (int)l.items.size(), l.lds, st);
        if (e != hipSuccess) return hip_fail(e, pad ? "batch FFT pass A" : "batch FFT pass A'");
    }
    return GN_OK;
}

// Host argument tables of one kind for these data pointers; uploaded when changed.
int prepare_tables(gn_batch_s *b, int kind, const void *const *in, void *const *out, hipStream_t st)
{
    const int slot = kind_slot(kind);
    Tables &t = b->tab[slot];
    const int n = b->n;
    std::vector<SpreadArgs> sa(n);
    std::vector<GatherArgs> ga(n);
    std::vector<PrepostArgs> pa(n), ca(n);
    std::vector<RowArgs> rp(b->custom ? n : 0), rcp(b->custom ? n : 0);
    const float sign2 = b->iflag >= 0 ? -1.0f : 1.0f;
    const double pi = 3.141592653589793;
    for (int i = 0; i < n; ++i) {
        gn_plan_s *p = b->m[i];
        const Geometry &g = p->geo;
        const double c = std::pow(pi / p->tau, g.dim / 2.0);
        float2 *y = b->ybuf + b->yoff[i];
        const bool t2 = kind == GN_TYPE2 || kind == GN_TYPE21;
        const bool t1 = kind != GN_TYPE2;
        pa[i] = make_prepost(p, (float)(c / (double)g.G));
        pa[i].src = t2 ? (const float2 *)in[i] : nullptr;
        pa[i].dst = p->grid;
        ga[i] = make_gather_args(p, kind == GN_TYPE21 ? y : (t2 ? (float2 *)out[i] : nullptr));
        const float2 *sin = kind == GN_TYPE21 ? y : (t1 ? (const float2 *)in[i] : nullptr);
        sa[i] = make_spread_args(p, sin, kind == GN_ADJOINT ? 0 : 1);
        const double extra = kind == GN_ADJOINT ? 1.0 : (p->M > 0 ? 1.0 / (double)p->M : 0.0);
        ca[i] = make_prepost(p, (float)(c * extra / (double)g.G));
        ca[i].src = p->grid;
        ca[i].dst = t1 ? (float2 *)out[i] : nullptr;
        if (b->custom) {
            rp[i] = make_row_args(p, sign2, pa[i].scale);
            rp[i].img_in = pa[i].src;
            rcp[i] = make_row_args(p, -sign2, ca[i].scale);
            rcp[i].img_out = ca[i].dst;
        }
    }
    std::vector<char> img;
    auto put = [&](const void *ptr, size_t nb) {
        const char *c = (const char *)ptr;
        img.insert(img.end(), c, c + nb);
    };
    put(sa.data(), sa.size() * sizeof(SpreadArgs));
    put(ga.data(), ga.size() * sizeof(GatherArgs));
    put(pa.data(), pa.size() * sizeof(PrepostArgs));
    put(ca.data(), ca.size() * sizeof(PrepostArgs));
    put(rp.data(), rp.size() * sizeof(RowArgs));
    put(rcp.data(), rcp.size() * sizeof(RowArgs));
    if (img == t.last) return GN_OK;
    hipError_t e = hipMemcpyAsync(t.spread, sa.data(), sa.size() * sizeof(SpreadArgs), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t.gather, ga.data(), ga.size() * sizeof(GatherArgs), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t.pad, pa.data(), pa.size() * sizeof(PrepostArgs), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t.crop, ca.data(), ca.size() * sizeof(PrepostArgs), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->custom)
        e = hipMemcpyAsync(t.rowp, rp.data(), rp.size() * sizeof(RowArgs), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->custom)
        e = hipMemcpyAsync(t.rowc, rcp.data(), rcp.size() * sizeof(RowArgs), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st); // the host vectors die with this scope
    if (e != hipSuccess) return hip_fail(e, "batch argument upload");
    t.last.swap(img);
    return GN_OK;
}

// Enqueue the stages of one kind on st (eagerly, or into a graph being captured).
int enqueue(gn_batch_s *b, int kind, hipStream_t st)
{
    const Tables &t = b->tab[kind_slot(kind)];
    const bool fwd_dir_type2 = b->iflag >= 0;
    hipError_t e;
    int rc;
    auto type2 = [&]() -> int {
        if (b->custom) {
            if ((rc = launch_fft_rows(b, true, t.rowp, st))) return rc;
            if ((rc = launch_fft_cols(b, 0, st))) return rc;
        } else {
            if ((e = launch_pad_batch(b->dim, t.pad, b->n, b->maxG, st)) != hipSuccess) return hip_fail(e, "batch pad");
            if ((rc = run_ffts(b, fwd_dir_type2, st))) return rc;
        }
        for (auto &l : b->gather_lists)
            if ((e = launch_gather_batch(b->dim, b->ns, l.wrap, t.gather, l.d, (int)l.items.size(), l.lds, st)) !=
                hipSuccess)
                return hip_fail(e, "batch gather");
        return GN_OK;
    };
    auto type1 = [&]() -> int {
        for (auto &l : b->spread_lists)
            if ((e = launch_spread_batch(b->dim, l.ncg, l.mfma, t.spread, l.d, (int)l.items.size(), l.waves, l.nbg, l.lds,
                                         st)) != hipSuccess)
                return hip_fail(e, "batch spread");
        if ((e = launch_spread_reduce_batch(b->dim, t.spread, b->d_reduce, (int)b->reduce_items.size(), st)) !=
            hipSuccess)
            return hip_fail(e, "batch spread reduce");
        if (b->custom) {
            if ((rc = launch_fft_cols(b, 1, st))) return rc;
            return launch_fft_rows(b, false, t.rowc, st);
        }
        if ((rc = run_ffts(b, !fwd_dir_type2, st))) return rc;
        if ((e = launch_crop_batch(b->dim, t.crop, b->n, b->maxN, st)) != hipSuccess) return hip_fail(e, "batch crop");
        return GN_OK;
    };
    if (kind == GN_TYPE2) return type2();
    if (kind == GN_TYPE21) {
        if ((rc = type2())) return rc;
    }
    return type1();
}

int execute(gn_batch_s *b, int kind, const void *const *in, void *const *out)
{
    int rc = prepare_tables(b, kind, in, out, b->stream);
    if (rc) return rc;
    if (!b->graph) return enqueue(b, kind, b->stream);
    const int slot = kind_slot(kind);
    hipError_t e;
    if (!b->exec[slot]) {
        hipGraph_t g = nullptr;
        if ((e = hipStreamBeginCapture(b->gstream, hipStreamCaptureModeRelaxed)) != hipSuccess)
            return hip_fail(e, "graph capture");
        rc = enqueue(b, kind, b->gstream);
        e = hipStreamEndCapture(b->gstream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "graph capture");
        e = hipGraphInstantiateWithFlags(&b->exec[slot], g, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->exec[slot] = nullptr;
            return hip_fail(e, "graph instantiate");
        }
    }
    if ((e = hipEventRecord(b->ev_in, b->stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(b->gstream, b->ev_in, 0)) != hipSuccess ||
        (e = hipGraphLaunch(b->exec[slot], b->gstream)) != hipSuccess ||
        (e = hipEventRecord(b->ev_out, b->gstream)) != hipSuccess ||
        (e = hipStreamWaitEvent(b->stream, b->ev_out, 0)) != hipSuccess)
        return hip_fail(e, "graph launch");
    return GN_OK;
}

void destroy(gn_batch_s *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->gstream) hipStreamSynchronize(b->gstream);
    drop_graphs(b);
    free_lists(b);
    for (auto &t : b->tab) {
        hipFree(t.spread);
        hipFree(t.gather);
        hipFree(t.pad);
        hipFree(t.crop);
        hipFree(t.rowp);
        hipFree(t.rowc);
    }
    for (auto &l : b->frows) hipFree(l.d);
    for (int t = 0; t < 2; ++t) {
        for (int k = 0; k < 2; ++k)
            for (auto &l : b->fcols[t][k]) hipFree(l.d);
        hipFree(b->d_cols[t]);
    }
    for (auto &fg : b->groups) {
        if (fg.fwd) rocfft_plan_destroy(fg.fwd);
        if (fg.bwd) rocfft_plan_destroy(fg.bwd);
    }
    if (b->info) rocfft_execution_info_destroy(b->info);
    hipFree(b->work);
    for (auto *p : b->m) {
        p->grid = nullptr; // lives in the batch buffer
        gn_plan_destroy(p);
    }
    for (auto *c : b->coords) hipFree(c);
    for (auto *d : b->dcf) hipFree(d);
    hipFree(b->grid);
    hipFree(b->ybuf);
    if (b->ev_in) hipEventDestroy(b->ev_in);
    if (b->ev_out) hipEventDestroy(b->ev_out);
    if (b->gstream) hipStreamDestroy(b->gstream);
    delete b;
}

// candidate tile shapes {spread tile, gather tile} of the autotuner
std::vector<std::vector<int>> candidates(int dim)
{
    std::vector<std::vector<int>> out;
    // spread tiles set the waves per tile (one wave per 64-cell block, up to 8 per
    // workgroup); gather tiles set the bin of samples staged together
    if (dim == 2) {
        const int sp[][2] = {{8, 16}, {16, 8}, {16, 16}, {8, 32}};
        const int ga[][2] = {{24, 24}, {16, 32}, {32, 32}, {32, 64}, {48, 48}};
        for (auto &s : sp)
            for (auto &g : ga) out.push_back({s[0], s[1], 1, g[0], g[1], 1});
    } else {
        const int sp[][3] = {{4, 4, 8}, {4, 8, 8}, {8, 8, 8}, {4, 4, 16}, {8, 8, 16}};
        const int ga[][3] = {{10, 8, 12}, {10, 10, 16}, {12, 12, 12}, {8, 8, 16}, {8, 8, 8}};
        for (auto &s : sp)
            for (auto &g : ga) out.push_back({s[0], s[1], s[2], g[0], g[1], g[2]});
    }
    return out;
}

template <typename Apply, typename Run>
int tune(int dim, int reps, hipStream_t st, Apply apply, Run run, int *tile, double *best_ms)
{
    if (reps < 1) return fail(GN_EINVAL, "autotune: reps must be >= 1");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    double best = 1e300;
    std::vector<int> best_t;
    int rc = GN_OK;
    for (const auto &c : candidates(dim)) {
        if (apply(c.data())) continue; // shape rejected (LDS budget)
        if ((rc = run())) break;       // warm-up
        hipEventRecord(e0, st);
        for (int r = 0; r < reps && !rc; ++r) rc = run();
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        if (rc) break;
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms / reps < best) {
            best = ms / reps;
            best_t = c;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if (best_t.empty()) return fail(GN_EINVAL, "autotune: no candidate tile shape fits");
    if ((rc = apply(best_t.data()))) return rc;
    if (tile) std::copy(best_t.begin(), best_t.end(), tile);
    if (best_ms) *best_ms = best;
    set_error("");
    return GN_OK;
}

} // namespace
} // namespace gn

extern "C" {

int gn_batch_create(int dim, int n, const gn_instance *inst, int rat, int nspread, int iflag, int layout,
                    int coord_f64, double df, void *stream, gn_batch **batch)
{
    set_error("");
    if (!batch) return fail(GN_EINVAL, "batch pointer is null");
    *batch = nullptr;
    if (n < 1 || !inst) return fail(GN_EINVAL, "a batch needs at least one instance");
    if (!std::isfinite(df)) return fail(GN_EINVAL, "df must be finite");
    gn_batch_s *b = new gn_batch_s();
    b->dim = dim;
    b->n = n;
    b->rat = rat;
    b->ns = nspread;
    b->iflag = iflag;
    b->layout = layout;
    b->coord_f64 = coord_f64 ? 1 : 0;
    b->df = df;
    b->stream = (hipStream_t)stream;
    int rc = GN_OK;
    hipError_t e;
    const size_t csz = coord_f64 ? sizeof(double) : sizeof(float);
    int64_t yel = 0;
    for (int i = 0; i < n && !rc; ++i) {
        gn_plan_s *p = nullptr;
        rc = plan_create_impl(dim, inst[i].modes, rat, nspread, iflag, inst[i].ncoils, layout, stream, false, &p);
        if (rc) break;
        b->m.push_back(p);
        const int64_t M = inst[i].M;
        if (M < 0 || M > (int64_t)INT32_MAX) {
            rc = fail(GN_EINVAL, "instance " + std::to_string(i) + ": sample count out of range");
            break;
        }
        p->M = M;
        const void *src[3] = {inst[i].x, inst[i].y, inst[i].z};
        for (int a = 0; a < 3; ++a) {
            void *d = nullptr;
            if (a < dim && M > 0) {
                if (!src[a]) {
                    rc = fail(GN_EINVAL, "instance " + std::to_string(i) + ": coordinate pointer is null");
                    break;
                }
                if (hipMalloc(&d, (size_t)M * csz) != hipSuccess) {
                    rc = fail(GN_ENOMEM, "coordinate copy");
                    break;
                }
                if ((e = hipMemcpy(d, src[a], (size_t)M * csz, hipMemcpyDeviceToDevice)) != hipSuccess) {
                    hipFree(d);
                    rc = hip_fail(e, "coordinate copy");
                    break;
                }
            }
            b->coords.push_back(d);
        }
        if (rc) break;
        float *dd = nullptr;
        if (inst[i].dcf && M > 0) {
            if (hipMalloc((void **)&dd, (size_t)M * sizeof(float)) != hipSuccess) {
                rc = fail(GN_ENOMEM, "dcf copy");
                break;
            }
            if ((e = hipMemcpy(dd, inst[i].dcf, (size_t)M * sizeof(float), hipMemcpyDeviceToDevice)) != hipSuccess)
                rc = hip_fail(e, "dcf copy");
        }
        b->dcf.push_back(dd);
        b->yoff.push_back(yel);
        yel += M * inst[i].ncoils;
        b->maxG = std::max(b->maxG, p->geo.G);
        b->maxN = std::max(b->maxN, p->geo.N);
    }
    if (!rc) {
        // pruned FFT passes when every instance's grid lengths are supported
        b->custom = true;
        for (auto *p : b->m)
            for (int a = 0; a < dim; ++a) b->custom &= fft_length_supported(p->geo.nf[a]);
        for (auto *p : b->m)
            for (int a = 0; a < dim && b->custom && !rc; ++a) rc = fft_axis_init(&p->fft_ax[a], (int)p->geo.nf[a]);
    }
    if (!rc) rc = make_groups(b);
    if (!rc && b->custom) rc = build_fft_lists(b);
    if (!rc && yel > 0 && hipMalloc((void **)&b->ybuf, (size_t)yel * sizeof(float2)) != hipSuccess)
        rc = fail(GN_ENOMEM, "type-21 sample buffer");
    for (int k = 0; k < kKinds && !rc; ++k) {
        Tables &t = b->tab[k];
        if (hipMalloc((void **)&t.spread, n * sizeof(SpreadArgs)) != hipSuccess ||
            hipMalloc((void **)&t.gather, n * sizeof(GatherArgs)) != hipSuccess ||
            hipMalloc((void **)&t.pad, n * sizeof(PrepostArgs)) != hipSuccess ||
            hipMalloc((void **)&t.crop, n * sizeof(PrepostArgs)) != hipSuccess ||
            (b->custom && (hipMalloc((void **)&t.rowp, n * sizeof(RowArgs)) != hipSuccess ||
                           hipMalloc((void **)&t.rowc, n * sizeof(RowArgs)) != hipSuccess)))
            rc = fail(GN_ENOMEM, "batch argument tables");
    }
    if (!rc) rc = rebin(b);
    if (!rc && ((e = hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking)) != hipSuccess || (e = hipEventCreate(&b->ev_in)) != hipSuccess ||
                (e = hipEventCreate(&b->ev_out)) != hipSuccess))
        rc = hip_fail(e, "batch stream");
    if (rc) {
        std::string keep = gn_last_error();
        destroy(b);
        set_error(keep);
        return rc;
    }
    *batch = b;
    return GN_OK;
}

int gn_batch_execute(gn_batch *b, int kind, const void *const *in, void *const *out)
{
    set_error("");
    if (!b) return fail(GN_EINVAL, "batch is null");
    if (kind_slot(kind) < 0) return fail(GN_EINVAL, "unknown transform kind");
    if (!in || !out) return fail(GN_EINVAL, "null pointer array");
    for (int i = 0; i < b->n; ++i) {
        const gn_plan_s *p = b->m[i];
        if ((kind == GN_TYPE1 || kind == GN_TYPE21) && p->M == 0)
            return fail(GN_EINVAL, "instance " + std::to_string(i) + ": type-1 transform of zero samples");
        const bool need_in = !(kind == GN_TYPE1 || kind == GN_ADJOINT) || p->M > 0;
        const bool need_out = !(kind == GN_TYPE2) || p->M > 0;
        if ((need_in && !in[i]) || (need_out && !out[i]))
            return fail(GN_EINVAL, "instance " + std::to_string(i) + ": null data pointer");
    }
    return execute(b, kind, in, out);
}

int gn_batch_set_graph(gn_batch *b, int enable)
{
    set_error("");
    if (!b) return fail(GN_EINVAL, "batch is null");
    if (b->graph && !enable) {
        hipStreamSynchronize(b->gstream);
        drop_graphs(b);
    }
    b->graph = enable != 0;
    return GN_OK;
}

int gn_batch_set_tiles(gn_batch *b, const int *tile)
{
    set_error("");
    if (!b || !tile) return fail(GN_EINVAL, "bad arguments");
    hipStreamSynchronize(b->stream);
    hipStreamSynchronize(b->gstream);
    for (auto *p : b->m) {
        int rc = gn_plan_set_tiles(p, tile);
        if (rc) {
            // keep every instance on the previous shape
            std::string keep = gn_last_error();
            for (auto *q : b->m) gn_plan_set_tiles(q, b->cur_tile);
            int r2 = rebin(b);
            set_error(keep);
            return r2 ? r2 : rc;
        }
    }
    std::copy(tile, tile + 6, b->cur_tile);
    return rebin(b);
}

int gn_batch_autotune(gn_batch *b, int kind, const void *const *in, void *const *out, int reps, int *tile,
                      double *best_ms)
{
    set_error("");
    if (!b) return fail(GN_EINVAL, "batch is null");
    auto apply = [&](const int *t) { return gn_batch_set_tiles(b, t); };
    auto run = [&]() { return gn_batch_execute(b, kind, in, out); };
    return tune(b->dim, reps, b->stream, apply, run, tile, best_ms);
}

int gn_batch_info(const gn_batch *b, int *n, int *tile, int *launches)
{
    if (!b) return fail(GN_EINVAL, "batch is null");
    if (n) *n = b->n;
    if (tile) {
        const Geometry &g = b->m[0]->geo;
        for (int a = 0; a < 3; ++a) {
            tile[a] = g.tile[a];
            tile[3 + a] = g.gtile[a];
        }
    }
    if (launches) {
        if (b->custom) {
            const int rows = (int)b->frows.size();
            const int c2 = (int)(b->fcols[0][0].size() + b->fcols[0][1].size());
            const int c1 = (int)(b->fcols[1][0].size() + b->fcols[1][1].size());
            launches[0] = rows + c2 + (int)b->gather_lists.size();
            launches[1] = (int)b->spread_lists.size() + (b->reduce_items.empty() ? 0 : 1) + c1 + rows;
        } else {
            const int nfft = (int)b->groups.size();
            launches[0] = 1 + nfft + (int)b->gather_lists.size();
            launches[1] = (int)b->spread_lists.size() + (b->reduce_items.empty() ? 0 : 1) + nfft + 1;
        }
    }
    return GN_OK;
}

int gn_batch_destroy(gn_batch *b)
{
    destroy(b);
    return GN_OK;
}

int gn_plan_autotune(gn_plan *p, int kind, int64_t M, const void *x, const void *y, const void *z, int coord_f64,
                     double df, const float *dcf, const void *in, void *out, int reps, int *tile, double *best_ms)
{
    set_error("");
    if (!p) return fail(GN_EINVAL, "plan is null");
    auto apply = [&](const int *t) {
        int rc = gn_plan_set_tiles(p, t);
        if (!rc) rc = gn_set_points(p, M, x, y, z, coord_f64, df, dcf);
        return rc;
    };
    auto run = [&]() { return gn_execute(p, kind, in, out); };
    return tune(p->geo.dim, reps, p->stream, apply, run, tile, best_ms);
}

} // extern "C"
