/*
 * FP64 C restatement of mripy's Gaussian gridding loop nests -- TEST INFRASTRUCTURE.
 *
 * Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * as the checker for large sizes and as the timed CPU baseline.  Never linked
 * into the product library.
 *
 * Each function restates one numba loop nest of fft/nufft_func.py in the same
 * loop order and arithmetic (complex128, the E1*E2^k*E3 recurrence):
 *   orc_spread2d  <- build_grid_2d1_fast   fft/nufft_func.py:550-582
 *   orc_gather2d  <- build_grid_2d2_fast   fft/nufft_func.py:607-639
 *   orc_spread3d  <- build_grid_3d1_fast   fft/nufft_func.py:863-909 (+ prange over
 *                    coils, build_grid_array_3d1_fast :931-937)
 *   orc_gather3d  <- build_grid_3d2_fast   fft/nufft_func.py:985-1032 (+ :1034-1038)
 *   orc_interp_spread3d <- build_grid_3d21 fft/nufft_func.py:1058-1106 (+ :1110-1114)
 * Grids are coil-last (nf1, nf2[, nf3], nc) like the reference's ftau[..., t]
 * views; samples are (M, nc) interleaved complex.
 *
 * The window index uses CPython/numba float semantics for x % (2 pi) and
 * xi // h (an fmod-based divmod), fft/nufft_func.py:293-295.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.141592653589793;

/* CPython float_divmod: returns floor quotient, writes the modulo. */
static double py_divmod(double vx, double wx, double *modp)
{
    double mod = fmod(vx, wx);
    double div = (vx - mod) / wx;
    double fl;
    if (mod != 0.0) {
        if ((wx < 0) != (mod < 0)) {
            mod += wx;
            div -= 1.0;
        }
    } else {
        mod = copysign(0.0, wx);
    }
    if (div != 0.0) {
        fl = floor(div);
        if (div - fl > 0.5) fl += 1.0;
    } else {
        fl = copysign(0.0, vx / wx);
    }
    if (modp) *modp = mod;
    return fl;
}

static double py_mod(double x, double m)
{
    double r;
    py_divmod(x, m, &r);
    return r;
}

static inline int64_t wrapi(int64_t a, int64_t n)
{
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

/* window start for one axis: xi in [0,2pi), m = 1 + xi // h, delta = xi - h m */
static void axis_window(double x, int nf, int64_t *m, double *delta)
{
    const double two_pi = 2 * PI;
    const double h = two_pi / nf;
    double xi = py_mod(x, two_pi);
    int64_t mm = 1 + (int64_t)py_divmod(xi, h, NULL);
    *m = mm;
    *delta = xi - h * mm;
}

void orc_cell_index(int64_t M, const double *x, int nf, int64_t *m_out, double *delta_out)
{
    for (int64_t i = 0; i < M; ++i) axis_window(x[i], nf, &m_out[i], &delta_out[i]);
}

/* ------------------------------------------------------------------ 2-D */

static void e3_2d(int ns, int nf1, int nf2, double tau, double *E3)
{
    for (int l = 0; l <= ns; ++l)
        for (int j = 0; j <= ns; ++j)
            E3[j * (ns + 1) + l] = exp(-((PI * j / nf1) * (PI * j / nf1) + (PI * l / nf2) * (PI * l / nf2)) / tau);
}

static void spread2d_one(int64_t M, const double *x, const double *y, const double *c, int nc, int t,
                         int nf1, int nf2, int ns, double tau, const double *E3, double *grid)
{
    const int ld = ns + 1;
    for (int64_t i = 0; i < M; ++i) {
        int64_t mx, my;
        double xi, yi;
        axis_window(x[i], nf1, &mx, &xi);
        axis_window(y[i], nf2, &my, &yi);
        double E1 = exp(-0.25 * (xi * xi + yi * yi) / tau);
        double E2x = exp((xi * PI) / (nf1 * tau));
        double E2y = exp((yi * PI) / (nf2 * tau));
        double v0r = c[2 * (i * nc + t)] * E1, v0i = c[2 * (i * nc + t) + 1] * E1;
        double E2mmx = 1;
        for (int a = 0; a < ns; ++a) {
            double E2mmy = 1;
            int64_t xp = wrapi(mx + a, nf1), xm = wrapi(mx - a - 1, nf1);
            for (int b = 0; b < ns; ++b) {
                int64_t yp = wrapi(my + b, nf2), ym = wrapi(my - b - 1, nf2);
                double w;
                double *g;
                w = E2mmx * E2mmy * E3[a * ld + b];
                g = grid + 2 * ((xp * nf2 + yp) * nc + t); g[0] += v0r * w; g[1] += v0i * w;
                w = E2mmx / (E2mmy * E2y) * E3[a * ld + b + 1];
                g = grid + 2 * ((xp * nf2 + ym) * nc + t); g[0] += v0r * w; g[1] += v0i * w;
                w = 1.0 / (E2mmx * E2x) * E2mmy * E3[(a + 1) * ld + b];
                g = grid + 2 * ((xm * nf2 + yp) * nc + t); g[0] += v0r * w; g[1] += v0i * w;
                w = 1.0 / (E2mmx * E2x) / (E2mmy * E2y) * E3[(a + 1) * ld + b + 1];
                g = grid + 2 * ((xm * nf2 + ym) * nc + t); g[0] += v0r * w; g[1] += v0i * w;
                E2mmy *= E2y;
            }
            E2mmx *= E2x;
        }
    }
}

/* 2-D type-1 spread.  The reference 2-D path is single threaded; coils loop serially. */
void orc_spread2d(int64_t M, const double *x, const double *y, const double *c, int nc,
                  int nf1, int nf2, int ns, double tau, double *grid)
{
    double *E3 = (double *)malloc(sizeof(double) * (ns + 1) * (ns + 1));
    e3_2d(ns, nf1, nf2, tau, E3);
    memset(grid, 0, sizeof(double) * 2 * (size_t)nf1 * nf2 * nc);
    for (int t = 0; t < nc; ++t) spread2d_one(M, x, y, c, nc, t, nf1, nf2, ns, tau, E3, grid);
    free(E3);
}

void orc_gather2d(int64_t M, const double *x, const double *y, const double *grid, int nc,
                  int nf1, int nf2, int ns, double tau, double *out)
{
    const int ld = ns + 1;
    double *E3 = (double *)malloc(sizeof(double) * ld * ld);
    e3_2d(ns, nf1, nf2, tau, E3);
    const double invG = 1.0 / ((double)nf1 * nf2);
    for (int t = 0; t < nc; ++t) {
        for (int64_t i = 0; i < M; ++i) {
            int64_t mx, my;
            double xi, yi;
            axis_window(x[i], nf1, &mx, &xi);
            axis_window(y[i], nf2, &my, &yi);
            double E1 = exp(-0.25 * (xi * xi + yi * yi) / tau);
            double E2x = exp((xi * PI) / (nf1 * tau));
            double E2y = exp((yi * PI) / (nf2 * tau));
            double sr = 0, si = 0;
            double E2mmx = 1;
            for (int a = 0; a < ns; ++a) {
                double E2mmy = 1;
                int64_t xp = wrapi(mx + a, nf1), xm = wrapi(mx - a - 1, nf1);
                for (int b = 0; b < ns; ++b) {
                    int64_t yp = wrapi(my + b, nf2), ym = wrapi(my - b - 1, nf2);
                    const double *g;
                    double w;
                    w = E1 * E2mmx * E2mmy * E3[a * ld + b];
                    g = grid + 2 * ((xp * nf2 + yp) * nc + t); sr += g[0] * w; si += g[1] * w;
                    w = E1 * E2mmx / (E2mmy * E2y) * E3[a * ld + b + 1];
                    g = grid + 2 * ((xp * nf2 + ym) * nc + t); sr += g[0] * w; si += g[1] * w;
                    w = E1 / (E2mmx * E2x) * E2mmy * E3[(a + 1) * ld + b];
                    g = grid + 2 * ((xm * nf2 + yp) * nc + t); sr += g[0] * w; si += g[1] * w;
                    w = E1 / (E2mmx * E2x) / (E2mmy * E2y) * E3[(a + 1) * ld + b + 1];
                    g = grid + 2 * ((xm * nf2 + ym) * nc + t); sr += g[0] * w; si += g[1] * w;
                    E2mmy *= E2y;
                }
                E2mmx *= E2x;
            }
            out[2 * (i * nc + t)] = sr * invG;
            out[2 * (i * nc + t) + 1] = si * invG;
        }
    }
    free(E3);
}

/* ------------------------------------------------------------------ 3-D */

static void e3_3d(int ns, int nf1, int nf2, int nf3, double tau, double *E3)
{
    const int ld = ns + 1;
    for (int p = 0; p <= ns; ++p)
        for (int l = 0; l <= ns; ++l)
            for (int j = 0; j <= ns; ++j) {
                double a = PI * j / nf1, b = PI * l / nf2, cc = PI * p / nf3;
                E3[(j * ld + l) * ld + p] = exp(-(a * a + b * b + cc * cc) / tau);
            }
}

#define IDX3(i1, i2, i3) ((((i1) * (int64_t)nf2 + (i2)) * nf3 + (i3)) * nc + t)

static void spread3d_one(int64_t M, const double *x, const double *y, const double *z, const double *c,
                         int nc, int t, int nf1, int nf2, int nf3, int ns, double tau, const double *E3,
                         double *grid)
{
    const int ld = ns + 1;
    for (int64_t i = 0; i < M; ++i) {
        int64_t mx, my, mz;
        double xi, yi, zi;
        axis_window(x[i], nf1, &mx, &xi);
        axis_window(y[i], nf2, &my, &yi);
        axis_window(z[i], nf3, &mz, &zi);
        double E1 = exp(-0.25 * (xi * xi + yi * yi + zi * zi) / tau);
        double E2x = exp((xi * PI) / (nf1 * tau));
        double E2y = exp((yi * PI) / (nf2 * tau));
        double E2z = exp((zi * PI) / (nf3 * tau));
        double v0r = c[2 * (i * nc + t)] * E1, v0i = c[2 * (i * nc + t) + 1] * E1;
        double E2mmx = 1;
        for (int a = 0; a < ns; ++a) {
            int64_t xp = wrapi(mx + a, nf1), xm = wrapi(mx - a - 1, nf1);
            double E2mmy = 1;
            for (int b = 0; b < ns; ++b) {
                int64_t yp = wrapi(my + b, nf2), ym = wrapi(my - b - 1, nf2);
                double E2mmz = 1;
                for (int q = 0; q < ns; ++q) {
                    int64_t zp = wrapi(mz + q, nf3), zm = wrapi(mz - q - 1, nf3);
                    double px = E2mmx, mxw = 1.0 / (E2mmx * E2x);
                    double py = E2mmy, myw = 1.0 / (E2mmy * E2y);
                    double pz = E2mmz, mzw = 1.0 / (E2mmz * E2z);
                    double w;
                    double *g;
#define ACC(I1, I2, I3, WX, WY, WZ, A, B, Q)                                      \
    w = (WX) * (WY) * (WZ) * E3[((A) * ld + (B)) * ld + (Q)];                     \
    g = grid + 2 * IDX3(I1, I2, I3);                                              \
    g[0] += v0r * w;                                                              \
    g[1] += v0i * w;
                    ACC(xp, yp, zp, px, py, pz, a, b, q)
                    ACC(xp, ym, zp, px, myw, pz, a, b + 1, q)
                    ACC(xm, yp, zp, mxw, py, pz, a + 1, b, q)
                    ACC(xm, ym, zp, mxw, myw, pz, a + 1, b + 1, q)
                    ACC(xp, yp, zm, px, py, mzw, a, b, q + 1)
                    ACC(xp, ym, zm, px, myw, mzw, a, b + 1, q + 1)
                    ACC(xm, yp, zm, mxw, py, mzw, a + 1, b, q + 1)
                    ACC(xm, ym, zm, mxw, myw, mzw, a + 1, b + 1, q + 1)
#undef ACC
                    E2mmz *= E2z;
                }
                E2mmy *= E2y;
            }
            E2mmx *= E2x;
        }
    }
}

/* 3-D type-1 spread, parallel over coils (numba prange, fft/nufft_func.py:935). */
void orc_spread3d(int64_t M, const double *x, const double *y, const double *z, const double *c, int nc,
                  int nf1, int nf2, int nf3, int ns, double tau, double *grid, int threads)
{
    const int ld = ns + 1;
    double *E3 = (double *)malloc(sizeof(double) * ld * ld * ld);
    e3_3d(ns, nf1, nf2, nf3, tau, E3);
    memset(grid, 0, sizeof(double) * 2 * (size_t)nf1 * nf2 * nf3 * nc);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int t = 0; t < nc; ++t) spread3d_one(M, x, y, z, c, nc, t, nf1, nf2, nf3, ns, tau, E3, grid);
    free(E3);
}

void orc_gather3d(int64_t M, const double *x, const double *y, const double *z, const double *grid, int nc,
                  int nf1, int nf2, int nf3, int ns, double tau, double *out, int threads)
{
    const int ld = ns + 1;
    double *E3 = (double *)malloc(sizeof(double) * ld * ld * ld);
    e3_3d(ns, nf1, nf2, nf3, tau, E3);
    const double invG = 1.0 / ((double)nf1 * nf2 * nf3);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int t = 0; t < nc; ++t) {
        for (int64_t i = 0; i < M; ++i) {
            int64_t mx, my, mz;
            double xi, yi, zi;
            axis_window(x[i], nf1, &mx, &xi);
            axis_window(y[i], nf2, &my, &yi);
            axis_window(z[i], nf3, &mz, &zi);
            double E1 = exp(-0.25 * (xi * xi + yi * yi + zi * zi) / tau);
            double E2x = exp((xi * PI) / (nf1 * tau));
            double E2y = exp((yi * PI) / (nf2 * tau));
            double E2z = exp((zi * PI) / (nf3 * tau));
            double sr = 0, si = 0;
            double E2mmx = 1;
            for (int a = 0; a < ns; ++a) {
                int64_t xp = wrapi(mx + a, nf1), xm = wrapi(mx - a - 1, nf1);
                double E2mmy = 1;
                for (int b = 0; b < ns; ++b) {
                    int64_t yp = wrapi(my + b, nf2), ym = wrapi(my - b - 1, nf2);
                    double E2mmz = 1;
                    for (int q = 0; q < ns; ++q) {
                        int64_t zp = wrapi(mz + q, nf3), zm = wrapi(mz - q - 1, nf3);
                        double px = E2mmx, mxw = 1.0 / (E2mmx * E2x);
                        double py = E2mmy, myw = 1.0 / (E2mmy * E2y);
                        double pz = E2mmz, mzw = 1.0 / (E2mmz * E2z);
                        double w;
                        const double *g;
#define GAT(I1, I2, I3, WX, WY, WZ, A, B, Q)                                      \
    w = E1 * (WX) * (WY) * (WZ) * E3[((A) * ld + (B)) * ld + (Q)];                \
    g = grid + 2 * IDX3(I1, I2, I3);                                              \
    sr += g[0] * w;                                                               \
    si += g[1] * w;
                        GAT(xp, yp, zp, px, py, pz, a, b, q)
                        GAT(xp, ym, zp, px, myw, pz, a, b + 1, q)
                        GAT(xm, yp, zp, mxw, py, pz, a + 1, b, q)
                        GAT(xm, ym, zp, mxw, myw, pz, a + 1, b + 1, q)
                        GAT(xp, yp, zm, px, py, mzw, a, b, q + 1)
                        GAT(xp, ym, zm, px, myw, mzw, a, b + 1, q + 1)
                        GAT(xm, yp, zm, mxw, py, mzw, a + 1, b, q + 1)
                        GAT(xm, ym, zm, mxw, myw, mzw, a + 1, b + 1, q + 1)
#undef GAT
                        E2mmz *= E2z;
                    }
                    E2mmy *= E2y;
                }
                E2mmx *= E2x;
            }
            out[2 * (i * nc + t)] = sr * invG;
            out[2 * (i * nc + t) + 1] = si * invG;
        }
    }
    free(E3);
}

/* Per-sample gather * dcf / G, then spread with the same direct weights
 * (build_grid_3d21, fft/nufft_func.py:1058-1106), parallel over coils. */
void orc_interp_spread3d(int64_t M, const double *x, const double *y, const double *z, const double *dcf,
                         const double *grid_in, int nc, int nf1, int nf2, int nf3, int ns, double tau,
                         double *grid_out, int threads)
{
    const double two_pi = 2 * PI;
    const double hx = two_pi / nf1, hy = two_pi / nf2, hz = two_pi / nf3;
    const double invG = 1.0 / ((double)nf1 * nf2 * nf3);
    const int w2 = 2 * ns;
    memset(grid_out, 0, sizeof(double) * 2 * (size_t)nf1 * nf2 * nf3 * nc);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int t = 0; t < nc; ++t) {
        double *Em = (double *)malloc(sizeof(double) * w2 * w2 * w2);
        for (int64_t i = 0; i < M; ++i) {
            double xi = py_mod(x[i], two_pi), yi = py_mod(y[i], two_pi), zi = py_mod(z[i], two_pi);
            int64_t m1 = 1 + (int64_t)py_divmod(xi, hx, NULL);
            int64_t m2 = 1 + (int64_t)py_divmod(yi, hy, NULL);
            int64_t m3 = 1 + (int64_t)py_divmod(zi, hz, NULL);
            double sr = 0, si = 0;
            for (int a = -ns; a < ns; ++a)
                for (int b = -ns; b < ns; ++b)
                    for (int q = -ns; q < ns; ++q) {
                        double dx = xi - hx * (m1 + a), dy = yi - hy * (m2 + b), dz = zi - hz * (m3 + q);
                        double w = exp(-0.25 * (dx * dx + dy * dy + dz * dz) / tau);
                        Em[((a + ns) * w2 + (b + ns)) * w2 + (q + ns)] = w;
                        const double *g = grid_in + 2 * IDX3(wrapi(m1 + a, nf1), wrapi(m2 + b, nf2), wrapi(m3 + q, nf3));
                        sr += g[0] * w;
                        si += g[1] * w;
                    }
            double d = (dcf ? dcf[i] : 1.0) * invG;
            sr *= d;
            si *= d;
            for (int a = -ns; a < ns; ++a)
                for (int b = -ns; b < ns; ++b)
                    for (int q = -ns; q < ns; ++q) {
                        double w = Em[((a + ns) * w2 + (b + ns)) * w2 + (q + ns)];
                        double *g = grid_out + 2 * IDX3(wrapi(m1 + a, nf1), wrapi(m2 + b, nf2), wrapi(m3 + q, nf3));
                        g[0] += sr * w;
                        g[1] += si * w;
                    }
        }
        free(Em);
    }
}
#undef IDX3

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
