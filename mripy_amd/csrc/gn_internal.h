// Internal definitions shared by the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "gaussnufft.h"

namespace gn {

constexpr int kMaxDim = 3;
#ifndef GN_GATHER_THREADS
#define GN_GATHER_THREADS 512
#endif
#ifndef GN_GATHER_CHUNK
#define GN_GATHER_CHUNK 512
#endif
constexpr int kGatherThreads = GN_GATHER_THREADS; // threads per gather workgroup
constexpr int kCoilGroup = 8;       // coils accumulated together per pass
// coils per gather work-group: a coil pair, staged together as 16-byte cells so
// that one LDS read serves both coils of a tap
constexpr int kGatherCoils = 2;
constexpr int kGatherCellBytes = 8 * kGatherCoils; // LDS bytes per staged gather cell
#ifndef GN_SPREAD_CHUNK
#define GN_SPREAD_CHUNK 512
#endif
constexpr int kSpreadChunk = GN_SPREAD_CHUNK; // max (tile, sample) pairs per spread work item
constexpr int kGatherChunk = GN_GATHER_CHUNK;  // max samples per gather work item
constexpr int kFFTMaxStages = 12;   // radix stages of one pruned-FFT axis

// One axis of the in-LDS mixed-radix FFT: radix sequence (decimation in time),
// twiddle table e^{-2 pi i k/L} and the digit-reversal map position -> LDS slot.
struct FFTAxis {
    int L = 0;
    int nst = 0;
    int radix[kFFTMaxStages] = {0};
    float2 *tw = nullptr;
    int *pinv = nullptr;
};

// launch arguments of the pruned FFT passes (fft.hip)

// Radix stages of a length-L transform and, per stage, the butterfly span q before it
// (product of the earlier radices), the twiddle step L / (q r) and the butterflies per
// line L / r.
struct Stages {
    int n;
    int radix[kFFTMaxStages];
    int q[kFFTMaxStages];
    int tws[kFFTMaxStages];
    int nb[kFFTMaxStages];
};

struct ColArgs {
    float2 *data; // in place
    int L;
    int rows_in, rows_out;
    int in_map, out_map; // 1: pruned position map with m = rows_in (in) / rows_out (out)
    int64_t ncols;       // contiguous columns; also the row stride
    int64_t slab_inner;  // slab s starts at (s / slab_inner) * stride_outer + (s % slab_inner) * stride_inner
    int64_t stride_outer, stride_inner;
    int64_t nslabs;      // slabs of this pass (the launch's second dimension)
    int CB;
    float sign;
    Stages st;
    const float2 *tw;
    const int *pinv; // FFT position -> LDS slot (digit reversal)
    // Optional row mask over the grid cells of this pass (the first adjoint pass of a
    // plain plan): element (row n, column c) is read only if bit n / mrow of
    // rmask[(y / mt1) * mnt2 + z / mt2] is set, where y = c / mcz, z = c % mcz.
    // Unread elements enter the FFT as zeros.  Null: every element is read.
    const unsigned long long *rmask = nullptr;
    int mrow = 1, mcz = 1, mt1 = 1, mt2 = 1, mnt2 = 1;
};

struct RowArgs {
    const float2 *img_in; // type 2: image
    float2 *img_out;      // type 1: image
    float2 *grid;         // line (c, n0, n1) at c * G + n0 * s0 + n1 * L
    int64_t G, s0;
    int L;
    int m;                // image points on this axis
    int64_t nlines;       // image lines per coil (product of the other modes)
    int64_t nc;
    int CL;               // coils per block
    int layout;
    float scale;
    float sign;
    const float *deap_last; // deapodization along this axis (m)
    const float *deap_a;    // along axis 0 (m0 entries)
    const float *deap_b;    // along axis 1 (3-D only; m1 entries), null in 2-D
    int m_b;                // m1 in 3-D (line index = n0 * m1 + n1), 1 in 2-D
    Stages st;
    const float2 *tw;
    const int *pinv;
    // SENSE maps (same layout as the coil images) or null.  Type 2 (pad): the image is
    // one coil-less image x and coil c's input is x * S_c.  Type 1 (crop): the block
    // holds every coil of its line and writes sum_c conj(S_c) * image_c.
    const float2 *sens = nullptr;
};

// Sample records in tile-sorted order: start cell of the window on each axis
// (already wrapped to [0,nf)), the sample's original index, its offset
// t = xi/h - m in (-1,0] on each axis and its density weight.
struct Records {
    int4 *cell = nullptr;     // {s0, s1, s2, original index}, sorted by gather tile
    float4 *frac = nullptr;   // {t0, t1, t2, dcf}
    int *bin_start = nullptr; // Kg+1 offsets of each gather tile's samples
    int64_t cap = 0;
    int64_t bin_start_cap = 0, spread_items_cap = 0, reduce_items_cap = 0, gather_items_cap = 0, occ_mask_cap = 0;
    // spread: one 32-byte record per (output tile, sample) pair, grouped by tile:
    //   a = {orig, j0 | j1<<8 | j2<<16, c0 | c1<<8 | c2<<16, l0 | l1<<10 | l2<<20}
    //   b = {t0, t1, t2, dcf} (float bits)
    // (j = first window tap inside the tile, c = taps inside, l = tile-local cell of tap j)
    int4 *pair_a = nullptr;
    int4 *pair_b = nullptr;
    int64_t npairs = 0;
    int64_t pairs_cap = 0;
    // work items
    int4 *spread_items = nullptr; // {tile, pbeg, pend, slot or -1}
    int n_spread_items = 0;
    int n_spread_full = 0; // the first n_spread_full items: tiles with pairs; then one per empty tile
    int4 *reduce_items = nullptr; // {tile, slot0, nslots, 0}
    int n_reduce_items = 0;
    int n_slots = 0;
    int4 *gather_items = nullptr; // {gather tile, sbeg, send, 0}
    int n_gather_items = 0;
    float2 *partial = nullptr;    // n_slots * nc * TC partial tiles
    int64_t partial_cap = 0;
    // Occupancy row masks per column of spread tiles (ntiles[1] * ntiles[2] words):
    // bit b is set when one of the leading-axis tiles [b * mask_k, (b + 1) * mask_k)
    // of that column has (tile, sample) pairs.
    unsigned long long *occ_mask = nullptr;
    int mask_k = 1;
};

struct Geometry {
    int dim;
    int64_t modes[3];
    int64_t nf[3];
    int64_t G;  // prod nf
    int64_t N;  // prod modes
    int ns;
    int tile[3];   // spread (output) tiles
    int ntiles[3];
    int64_t K;     // number of spread tiles
    int gtile[3];  // gather (home) tiles
    int ngtiles[3];
    int64_t Kg;    // number of gather tiles
    float coef[3]; // log2(e) * h_a^2/(4 tau): weight = exp2(-coef * (t + ns - j)^2)
};

// ---- kernel arguments (shared by the plan and batch launchers)

struct SpreadArgs {
    int64_t nf[3];
    int tile[3];
    int ntiles[3];
    int blk[3];  // block (cells per wave) per axis
    int nblk[3]; // blocks per tile per axis
    int64_t G;
    int64_t M;
    int64_t nc;
    int TC;
    int TT;      // tile[0] + tile[1] + tile[2]: weight-table entries per pair
    int ns;
    int layout;
    int use_dcf;
    float coef[3];
    const int4 *items;
    const int4 *reduce_items;
    const int4 *pair_a;
    const int4 *pair_b;
    const float2 *in;
    float2 *grid;
    float2 *partial;
    // Occupancy masks of the spread tiles (bit t0 / occ_k of occ[t1 * ntiles[2] + t2]
    // set when some tile of that row group has pairs).  When set, an empty tile whose
    // group bit is clear is not written: the pass that reads the grid next skips it.
    // Null: every empty tile is written with zeros.
    const unsigned long long *occ = nullptr;
    int occ_k = 1;
};

struct GatherArgs {
    int64_t nf[3];
    int tile[3];
    int ntiles[3];
    int region[3]; // staged cells per axis
    int64_t G;
    int64_t M;
    int64_t nc;
    int layout;
    float coef[3];
    const int4 *items;
    const int4 *rec_cell;
    const float4 *rec_frac;
    const float2 *grid;
    float2 *out;
    float scale = 1.0f; // output factor (1 in the transforms; 1/G for the raw gridding kind)
};

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
    const float2 *src; // batched launches: this instance's input
    float2 *dst;       // and output
    // SENSE maps (same layout as the coil images) or null: pad reads one coil-less
    // image x and writes x * S_c per coil; crop writes sum_c conj(S_c) * image_c
    const float2 *sens = nullptr;
};

} // namespace gn

struct gn_plan_s {
    gn::Geometry geo;
    int rat = 2;
    int iflag = 1;
    double tau = 0;
    int64_t nc = 1;
    int layout = GN_COIL_LAST;
    hipStream_t stream = nullptr;
    int tile_override[3] = {0, 0, 0};
    int gtile_override[3] = {0, 0, 0};

    int64_t M = 0;
    double df = 1.0;
    bool has_points = false;
    bool has_dcf = false;
    gn::Records rec;

    float2 *grid = nullptr;   // nc * G oversampled grids, coil-major
    // the image whose (non-SENSE) type-2 grid `grid` holds after the plan's last
    // operation, or null: set by a GN_TYPE2 execution or a trajectory gradient, cleared
    // by every other operation that may write the grid (gn_traj_grad_ex may reuse it)
    const void *grid_t2_img = nullptr;
    float2 *ybuf = nullptr;   // nc * M intermediate samples (type 21)
    float2 *sens = nullptr;   // N * nc SENSE maps in the plan's layout (gn_plan_set_sense)
    float2 *sbuf = nullptr;   // N * nc coil images of a SENSE combine that does not fit one block
    int64_t ybuf_cap = 0;
    float *deap = nullptr;    // concatenated per-axis exp(tau k^2) tables
    int *err_flag = nullptr;  // device flag set by set_points on bad coordinates

    // FFT engine: the pruned in-place passes (gn_fft.hip) when every nf is a product
    // of 2, 3, 5 (fft_custom), else rocFFT; GN_FFT_ROCFFT forces rocFFT
    bool fft_custom = false;
    int fft_engine = GN_FFT_AUTO;
    gn::FFTAxis fft_ax[3];
    rocfft_plan fft_fwd = nullptr;
    rocfft_plan fft_bwd = nullptr;
    rocfft_execution_info fft_info = nullptr;
    void *fft_work = nullptr;
    size_t fft_work_size = 0;

    // optional stage timers
    bool timing = false;
    struct Pending {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double stage_ms[GN_STAGE_COUNT] = {0};
    int64_t stage_n[GN_STAGE_COUNT] = {0};
};

namespace gn {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

// kernels.hip
int bin_and_sort(gn_plan_s *p, int64_t M, const void *x, const void *y, const void *z, int coord_f64, double df,
                 const float *dcf);
// grid: null = the plan's grid (the transforms); else a caller grid, every tile written
int launch_spread(gn_plan_s *p, const float2 *samples, int use_dcf, float2 *grid = nullptr);
int launch_gather(gn_plan_s *p, float2 *samples, const float2 *grid = nullptr, float scale = 1.0f);
int launch_pad(gn_plan_s *p, const float2 *image, float scale, bool sense = false);
int launch_crop(gn_plan_s *p, float2 *image, float scale, bool sense = false);
int launch_traj_grad(gn_plan_s *p, const float2 *wts, float *out);
int launch_sense_combine(gn_plan_s *p, const float2 *coil_images, float2 *image);
void default_tiles(int dim, int ns, const int64_t *nf, int *tile, int *gtile);

// fft.hip
bool fft_length_supported(int64_t L);
int fft_axis_init(FFTAxis *ax, int L);
void fft_axis_free(FFTAxis *ax);
int pruned_type2_rows(gn_plan_s *p, const float2 *image, float scale, float sign);
int pruned_type2_cols(gn_plan_s *p, float sign);
int pruned_type1_cols(gn_plan_s *p, float sign);
int pruned_type1_rows(gn_plan_s *p, float2 *image, float scale, float sign);
// SENSE variants of the row passes (p->sens set): coil-less image in / out
int pruned_type2_rows_sense(gn_plan_s *p, const float2 *image, float scale, float sign);
int pruned_type1_rows_sense(gn_plan_s *p, float2 *image, float scale, float sign);
bool sense_rows_fit(const gn_plan_s *p);
// pieces of the pruned passes for the batched launch (batch.hip)
size_t col_lds(int L, int CB);
size_t row_lds(int L, int CL);
RowArgs make_row_args(const gn_plan_s *p, float sign, float scale);
int pruned_col_args(const gn_plan_s *p, bool type2, float sign, ColArgs *out);
hipError_t launch_fft_cols_batch(const ColArgs *d_args, const int4 *d_items, int n, size_t lds, hipStream_t st);
hipError_t launch_fft_rows_batch(bool pad, const RowArgs *d_args, const int4 *d_items, int n, size_t lds,
                                 hipStream_t st);

// api.hip
int plan_create_impl(int dim, const int64_t *modes, int rat, int nspread, int iflag, int64_t ncoils, int layout,
                     void *stream, bool with_grid, gn_plan_s **plan);

// argument builders and batched launchers (kernels.hip)
SpreadArgs make_spread_args(const gn_plan_s *p, const float2 *samples, int use_dcf);
GatherArgs make_gather_args(const gn_plan_s *p, float2 *samples);
PrepostArgs make_prepost(const gn_plan_s *p, float scale);
void spread_shape(const SpreadArgs &a, int dim, int *waves, int *nbg);
bool spread_uses_mfma(const SpreadArgs &a, int dim);
size_t spread_lds(const SpreadArgs &a, int ncg, bool mfma, int waves);
hipError_t launch_spread_batch(int dim, int ncg, bool mfma, const SpreadArgs *d_args, const int4 *d_items, int n,
                               int waves, int nbg, size_t lds, hipStream_t st);
hipError_t launch_spread_reduce_batch(int dim, const SpreadArgs *d_args, const int4 *d_items, int n, hipStream_t st);
hipError_t launch_gather_batch(int dim, int ns, bool wrap, const GatherArgs *d_args, const int4 *d_items, int n,
                               size_t lds, hipStream_t st);
hipError_t launch_pad_batch(int dim, const PrepostArgs *d_args, int ninst, int64_t maxG, hipStream_t st);
hipError_t launch_crop_batch(int dim, const PrepostArgs *d_args, int ninst, int64_t maxN, hipStream_t st);

} // namespace gn
