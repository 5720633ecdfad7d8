/*
 * curafft.h -- C-ABI of libcurafft.so, the MI355X (gfx950) W-stacking imager.
 *
 * The first eight functions are drop-in replacements for the reference's
 * exported symbols (RapidsAtHKUST/cuGridder include/cugridder.h:20-46,
 * defined at src/astro/cugridder.cu:540-571) with the same names, argument
 * order and meaning.  Complex arrays are passed as interleaved (re, im)
 * scalars, which is the memory layout of std::complex<PCS> in the reference
 * and of numpy complex128/complex64 used by its ctypes binding
 * (python/curagridder/cursl.py:29-72).
 *
 * Conventions (reference): uvw is host [nrow][3] (u, v, w in metres); vis is
 * [nrow]; wgt is [nrow] (NULL in the _1 variants); dirty is [nxdirty][nydirty]
 * row-major, dirty[i*ny + j] <-> l = (i - nx/2)*fov/nx, m = (j - ny/2)*fov/ny;
 * fov in degrees; freq in Hz; sign multiplies w (the NIFTY-style binding
 * passes -1).  ms2dirty computes
 *   dirty(l,m) = 1/n sum_j wgt_j vis_j exp(2 pi i f/c (u l + v m + sign w (n-1)))
 * and dirty2ms its adjoint, to relative error ~epsilon.
 *
 * Return value: 0 ok, 1 epsilon was clamped to the precision floor (as the
 * reference's setup_conv_opts), 2 invalid argument, 3 unsupported size,
 * 4 HIP runtime error, 5 device allocation failed.  The library never exits
 * the process.
 */
#ifndef CURAFFT_H
#define CURAFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference entry points, double precision (complex128) ------------- */

/* replaces ms2dirty_2, src/astro/cugridder.cu:541-547 (MS2DIRTY_2) */
int ms2dirty_2(int nrow, int nxdirty, int nydirty, double fov, double freq,
               double *uvw, double *vis, double *wgt, double *dirty,
               double epsilon, double sigma, int sign);
/* replaces ms2dirty_1, src/astro/cugridder.cu:549-555 (no weights) */
int ms2dirty_1(int nrow, int nxdirty, int nydirty, double fov, double freq,
               double *uvw, double *vis, double *dirty, double epsilon,
               double sigma, int sign);
/* replaces dirty2ms_1, src/astro/cugridder.cu:557-563 */
int dirty2ms_1(int nrow, int nxdirty, int nydirty, double fov, double freq,
               double *uvw, double *vis, double *dirty, double epsilon,
               double sigma, int sign);
/* replaces dirty2ms_2, src/astro/cugridder.cu:565-571 */
int dirty2ms_2(int nrow, int nxdirty, int nydirty, double fov, double freq,
               double *uvw, double *vis, double *wgt, double *dirty,
               double epsilon, double sigma, int sign);

/* ---- reference entry points, single precision (complex64, FP32 uvw) ---- */

/* replaces ms2dirtyf_2 (the -DSINGLE build of cugridder.cu:541-547) */
int ms2dirtyf_2(int nrow, int nxdirty, int nydirty, float fov, float freq,
                float *uvw, float *vis, float *wgt, float *dirty,
                float epsilon, float sigma, int sign);
/* replaces ms2dirtyf_1 (-DSINGLE cugridder.cu:549-555) */
int ms2dirtyf_1(int nrow, int nxdirty, int nydirty, float fov, float freq,
                float *uvw, float *vis, float *dirty, float epsilon,
                float sigma, int sign);
/* replaces dirty2msf_1 (-DSINGLE cugridder.cu:557-563) */
int dirty2msf_1(int nrow, int nxdirty, int nydirty, float fov, float freq,
                float *uvw, float *vis, float *dirty, float epsilon,
                float sigma, int sign);
/* replaces dirty2msf_2 (-DSINGLE cugridder.cu:565-571) */
int dirty2msf_2(int nrow, int nxdirty, int nydirty, float fov, float freq,
                float *uvw, float *vis, float *wgt, float *dirty,
                float epsilon, float sigma, int sign);

/* ---- additions with no reference counterpart ---------------------------- */

/* Mixed precision: FP64 uvw (coordinate scaling in FP64), complex64 vis and
 * image, FP32 weights (NULL allowed).  Same semantics as ms2dirty_2 /
 * dirty2ms_2; the float ABI above forces FP32 uvw. */
int ms2dirty_mixed(int nrow, int nxdirty, int nydirty, double fov, double freq,
                   const double *uvw, const float *vis, const float *wgt,
                   float *dirty, double epsilon, double sigma, int sign);
int dirty2ms_mixed(int nrow, int nxdirty, int nydirty, double fov, double freq,
                   const double *uvw, float *vis, const float *wgt,
                   const float *dirty, double epsilon, double sigma, int sign);

/* 2-D NUFFT (the reference reaches these only through its test drivers,
 * test/nufft_2d_1_test.cu:126-258 and test/nufft_2d_2_test.cu:140-315).
 * Points x, y in [-pi, pi); modes fk[N2][N1], k1 = j - N1/2 (fast), k2 =
 * i - N2/2.  type 1: fk = sum_j c_j e^{+i(k1 x_j + k2 y_j)};
 * type 2: c_j = sum_k fk e^{-i(k1 x_j + k2 y_j)}. */
int nufft2d1(int M, const double *x, const double *y, const double *c, int N1,
             int N2, double *fk, double epsilon, double sigma);
int nufft2d2(int M, const double *x, const double *y, double *c, int N1,
             int N2, const double *fk, double epsilon, double sigma);
int nufft2d1f(int M, const float *x, const float *y, const float *c, int N1,
              int N2, float *fk, float epsilon, float sigma);
int nufft2d2f(int M, const float *x, const float *y, float *c, int N1, int N2,
              const float *fk, float epsilon, float sigma);

/* Several channels (no counterpart in the reference C-ABI: its binding passes
 * freq[0] only, python/curagridder/cursl.py:103).  freq[nchan] (Hz, host),
 * vis and wgt are [nrow][nchan] (NIFTY layout); every (row, channel) is one
 * visibility at uvw * freq[c] / c and the dirty image sums all of them.
 * wgt may be NULL.  Double: complex128 vis/dirty; mixed: complex64 vis/dirty
 * with FP32 wgt. */
int ms2dirty_mc_2(int nrow, int nchan, int nxdirty, int nydirty, double fov,
                  const double *freq, const double *uvw, const double *vis,
                  const double *wgt, double *dirty, double epsilon, double sigma,
                  int sign);
int dirty2ms_mc_2(int nrow, int nchan, int nxdirty, int nydirty, double fov,
                  const double *freq, const double *uvw, double *vis,
                  const double *wgt, const double *dirty, double epsilon,
                  double sigma, int sign);
int ms2dirty_mc_mixed(int nrow, int nchan, int nxdirty, int nydirty, double fov,
                      const double *freq, const double *uvw, const float *vis,
                      const float *wgt, float *dirty, double epsilon,
                      double sigma, int sign);
int dirty2ms_mc_mixed(int nrow, int nchan, int nxdirty, int nydirty, double fov,
                      const double *freq, const double *uvw, float *vis,
                      const float *wgt, const float *dirty, double epsilon,
                      double sigma, int sign);

/* NIFTY geometry (no counterpart in the reference C-ABI: its binding
 * ignores rad_pix_y and do_wstacking, python/curagridder/cursl.py:78-99).
 * pix_x, pix_y are the pixel sizes in radians along nxdirty and nydirty
 * (the reference derives pix_y = pix_x * nxdirty / nydirty); do_wstacking
 * 0 asks for plain 2-D gridding: w is ignored and there is no 1/n,
 *   dirty(l,m) = sum_j wgt_j vis_j exp(2 pi i f/c (u l + v m)),
 * one grid plane, as NIFTY's do_wstacking=False.  Host arrays; freq[nchan],
 * vis and wgt [nrow][nchan] (wgt may be NULL).  prec 0: FP64 uvw, complex64
 * vis/dirty, FP32 wgt; 1: FP64 uvw, complex128, FP64 wgt; 2: FP32 uvw,
 * complex64, FP32 wgt. */
int ms2dirty_px(int prec, int nrow, int nchan, int nxdirty, int nydirty,
                double pix_x, double pix_y, const double *freq, const void *uvw,
                const void *vis, const void *wgt, void *dirty, double epsilon,
                double sigma, int sign, int do_wstacking);
int dirty2ms_px(int prec, int nrow, int nchan, int nxdirty, int nydirty,
                double pix_x, double pix_y, const double *freq, const void *uvw,
                void *vis, const void *wgt, const void *dirty, double epsilon,
                double sigma, int sign, int do_wstacking);

/* Contexts: a reusable device workspace bound to a HIP stream (NULL = the
 * null stream).  Device-pointer calls below run the same pipeline as the
 * host entry points on inputs already in HBM. */
int cg_ctx_create(void **ctx, void *stream);
int cg_ctx_destroy(void *ctx);
/* options: "timing" (0/1, per-stage HIP events), "lds_budget" (degridding
 * staging bytes), "grid_lds_budget" (gridding accumulator bytes),
 * "tile_grid", "tile_degrid" (preferred tile edge, power of two <= 32),
 * "tile_grid2d" (the same for 2-D NUFFT type 1, <= 64; default 32),
 * "tile_degrid2d" (the same for 2-D NUFFT type 2; default 32),
 * "mfma2d" (1: complex64 2-D NUFFT type 1 grids 16 x 16 quadrants as MFMA
 * products when the tiling allows; 0: LDS-atomic gridding),
 * "dense_min" (points per gridding bin from which the complex64 gridder
 * uses MFMA products over 2 x 2 start cells; 0 disables),
 * "degrid_dense_min" (points per degridding tile from which the complex64
 * degridder uses MFMA products; 0 disables; default 2048),
 * "tap_poly" (1: the complex64 MFMA kernels take kernel values from per-tap
 * polynomials fitted at plan time instead of the exp2 form; a learned
 * kernel always does), "row_mask" (1: the FFT
 * skips grid rows that no point reaches; 0: transforms every row),
 * "prerec" (1: complex64 records are built in the caller's order and
 * gathered as aligned slots; 0: one gather pass; results are identical),
 * "band_fft" (1: complex64 column passes run band-limited when the grid
 * rows points reach fit in 1024 / 2048 rows, see cg_ctx_last_band),
 * "compact_keys" (1: sort keys number only the tiles inside the u / v
 * bands: the same permutation with fewer radix passes),
 * "grid_limit_mb" (the largest W-stacking grid buffer in MiB, 0 = 90% of
 * the free device memory; a larger grid is processed in slabs of w-planes
 * through one buffer, the out-of-core flow of the reference's mem_limit
 * pipeline, src/astro/ra_exec.cu:292-389, 520-613; see cg_ctx_last_slab;
 * the host entry points read it from the environment, CG_GRID_LIMIT_MB),
 * "stop_after" (test hook, see below) */
int cg_ctx_set_option(void *ctx, const char *name, int value);
/* per-stage times summed over calls, out[16] = {ms2dirty[7], dirty2ms[7],
 * dense, dense2} in ms, stages: plan, sort, records, grid/degrid, fft rows,
 * fft columns, image (epilogue/prologue); dense / dense2 = the MFMA
 * gridding / degridding kernel's part of the ms2dirty / dirty2ms grid
 * stage; calls[2] */
int cg_ctx_stage_times(void *ctx, double *out, long long *calls);
int cg_ctx_reset_times(void *ctx);
/* parameters of the last call: nf1 nf2 nf3 kw beta es_c pix_x pix_y
 * i_center i_half_width o_center o_half_width gamma h tile_x tile_y */
int cg_ctx_last_plan(void *ctx, double *out);
/* row band of the last W-stacking call: out[0] = first grid row (nf2 axis)
 * that points can reach, out[1] = number of rows, cyclic (0: no band, the
 * column passes run as the four-step split).  With a band of at most 1024
 * or 2048 rows the complex64 column passes run band-limited: the input is
 * read and the spectrum written once, no intermediate in HBM. */
int cg_ctx_last_band(void *ctx, int *out);
/* out[0] = w-planes per slab of the last call's grid buffer (0: the whole
 * grid at once), out[1] = the number of slabs */
int cg_ctx_last_slab(void *ctx, int *out);

/* prec 0: complex64 vis/image, FP32 gridding and FFT; prec 1: complex128.
 * uvw_prec, wgt_prec: 0 float, 1 double (prec 1 requires both 1).  Supported
 * combinations: (0,1,0) mixed, (0,0,0) float, (1,1,1) double. */
int cg_ms2dirty_dev(void *ctx, int prec, int nrow, int nxdirty, int nydirty,
                    double fov, double freq, const void *d_uvw, int uvw_prec,
                    const void *d_vis, const void *d_wgt, int wgt_prec,
                    void *d_dirty, double epsilon, double sigma, int sign);
int cg_dirty2ms_dev(void *ctx, int prec, int nrow, int nxdirty, int nydirty,
                    double fov, double freq, const void *d_uvw, int uvw_prec,
                    void *d_vis, const void *d_wgt, int wgt_prec,
                    const void *d_dirty, double epsilon, double sigma, int sign);
/* the same with nchan channels: freq[nchan] on the host, d_vis and d_wgt
 * [nrow][nchan] on the device */
int cg_ms2dirty_mc_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty,
                       int nydirty, double fov, const double *freq,
                       const void *d_uvw, int uvw_prec, const void *d_vis,
                       const void *d_wgt, int wgt_prec, void *d_dirty,
                       double epsilon, double sigma, int sign);
int cg_dirty2ms_mc_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty,
                       int nydirty, double fov, const double *freq,
                       const void *d_uvw, int uvw_prec, void *d_vis,
                       const void *d_wgt, int wgt_prec, const void *d_dirty,
                       double epsilon, double sigma, int sign);
/* the same with the NIFTY geometry of ms2dirty_px / dirty2ms_px */
int cg_ms2dirty_px_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty,
                       int nydirty, double pix_x, double pix_y,
                       const double *freq, const void *d_uvw, int uvw_prec,
                       const void *d_vis, const void *d_wgt, int wgt_prec,
                       void *d_dirty, double epsilon, double sigma, int sign,
                       int do_wstacking);
int cg_dirty2ms_px_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty,
                       int nydirty, double pix_x, double pix_y,
                       const double *freq, const void *d_uvw, int uvw_prec,
                       void *d_vis, const void *d_wgt, int wgt_prec,
                       const void *d_dirty, double epsilon, double sigma,
                       int sign, int do_wstacking);
/* Batches of independent channels (no counterpart in the reference, which
 * images one channel per call): every item has its own rows, image size,
 * field of view and frequency; complex64 vis/image, FP64 uvw, FP32 wgt (may
 * be NULL), all on the device.  cg_batch_create plans every item (flags:
 * CG_BATCH_AUTOTUNE picks tile sizes and the dense-gridding threshold per
 * image size and visibility count from measured runs; CG_BATCH_GRAPH
 * captures the whole batch into one HIP graph), runs it once, and keeps a
 * private workspace; cg_batch_run replays it on the context's stream
 * (direction 0: vis -> dirty, 1: dirty -> vis).  The device buffers must
 * stay valid while the batch exists.  Each item's plan is fixed from the uvw
 * contents at creation: the w-range (planes), the u and v ranges (the grid
 * rows and columns the FFT passes keep, and the rectangle of tiles the sort
 * keys number).  A replay must therefore see the same uvw; vis, wgt and
 * dirty contents may change between replays.  Points outside the planned
 * ranges would be gridded into the wrong tiles: rebuild the batch when uvw
 * changes. */
typedef struct {
    int nrow, nxdirty, nydirty;
    double fov, freq;
    const void *d_uvw;
    void *d_vis;
    const void *d_wgt;
    void *d_dirty;
} cg_batch_item;
#define CG_BATCH_AUTOTUNE 1
#define CG_BATCH_GRAPH 2
int cg_batch_create(void **batch, void *ctx, int direction, int n,
                    const cg_batch_item *items, double epsilon, double sigma,
                    int sign, int flags);
int cg_batch_run(void *batch);
/* out[5 * i ...] = tile_grid, tile_degrid, dense threshold in use, tuned ms
 * per run, nf3 of item i; returns 1 when the batch runs as a graph */
int cg_batch_info(void *batch, double *out);
int cg_batch_destroy(void *batch);

/* type 1 reads d_c and writes d_fk; type 2 reads d_fk and writes d_c */
int cg_nufft2d_dev(void *ctx, int type, int prec, int M, const void *d_x,
                   const void *d_y, void *d_c, int N1, int N2, void *d_fk,
                   double epsilon, double sigma);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cg_nufft2d_dev without a context, on `stream`, whose
 * workspace is a fixed 64 MiB device array inside this library (one copy per
 * GPU, held until the process ends).  It calls no hipMalloc / hipFree.  A
 * call whose workspace does not fit returns 3 before launching anything
 * (complex64 at 1024 x 1024 modes fits about 0.5 M points; use
 * cg_nufft2d_dev beyond that).  Calls that overlap on one GPU race on the
 * array; the array keeps the last call's contents. */
int cg_nufft2d_noalloc(void *stream, int type, int prec, int M,
                       const void *d_x, const void *d_y, void *d_c, int N1,
                       int N2, void *d_fk, double epsilon, double sigma);

/* Host-only planner (no device calls): out[16] = nf1 nf2 nf3 kw beta es_c
 * pix_x pix_y i_center i_half_width o_center o_half_width gamma h f_over_c
 * ier, given the range [wmin, wmax] of the scaled w = w * 2 pi f/c * sign. */
int cg_plan_query(int nrow, int nxdirty, int nydirty, double fov, double freq,
                  double wmin, double wmax, double epsilon, double sigma,
                  int sign, int single, double *out);

/* Test hooks.  Device pointers into the context workspace from the last call
 * (valid until the next call): the w-plane grid and its cell count; the
 * sorted keys (uint32: tile << n[2] | start cell inside the tile), row
 * permutation (int32), tile starts (int32) with n[0] = rows, n[1] = tiles.  Option "stop_after" (1: after gridding or
 * the dirty2ms prologue, 2: after the plane FFTs) stops a call early. */
int cg_ctx_grid(void *ctx, void **ptr, long long *cells);
/* host only: per-tap polynomial coefficients of the complex64 kernel values
 * psi(d0 + t) / e^beta in x = 2 d0 + kw - 1 (out[t * 12 + q], q <= deg <
 * 12), for the ES kernel (kq NULL) or the learned one (ndeg + 1 coefficients
 * of its exponent's polynomial in s = 2 c d^2 - 1) */
int cg_tap_poly(int kw, double beta, double c, const float *kq, int ndeg, int deg,
                float *out);
/* Order of the spectrum's column axis in the grid after the forward column
 * FFT: returns 1 and n = na * nb when frequency ka + na * kb is stored at
 * row ka * nb + kb ("split order"), 0 when rows are in natural order.  The
 * dirty2ms prologue writes the same order, and only the fast-axis columns
 * holding image pixels. */
int cg_fft_col_split(int n, int *na, int *nb);
int cg_ctx_sorted(void *ctx, void **keys, void **perm, void **tstart,
                  long long *n);
/* In-place unnormalised 2-D FFT of batch planes [batch][n2][n1], exponent
 * sign dir (+1/-1); prec 0 complex64, 1 complex128. */
int cg_fft2d_dev(void *ctx, int prec, void *d_data, int batch, int n2, int n1,
                 int dir);
/* Stable sort of m (uint32 key, int32 value) pairs by the low key_bits bits,
 * in place on the context stream. */
int cg_sort_pairs_dev(void *ctx, void *d_keys, void *d_vals, int m,
                      int key_bits);

/* Device-to-device copy of bytes on a HIP stream (test hook). */
int cg_memcpy_dev(void *dst, const void *src, void *stream, size_t bytes);

int cg_version(void);

/* ---- Learned gridding kernel.  No counterpart in the reference (which has
 * only the ES kernel, src/fourier/conv_interp_invoker.cu:49-119); these are
 * for callers who trade a network-defined kernel of smaller support for
 * fewer taps or a lower error.  The kernel is
 *   psi(d) = exp(beta (sqrt(1 - x^2) - 1) + y(s)), x = 2 d / kw, s = 2 x^2 - 1,
 *   |d| < kw/2, y = W3 . tanh(W2 tanh(W1 s + b1) + b2) + b3 (1 -> 32 -> 32 -> 1,
 *   FP32), beta the ES kernel's for sigma = 2 (conv_interp_invoker.cu:96-104),
 * so W3 = b3 = 0 is the ES kernel; the parameter vector is W1[32] b1[32] W2[32][32] b2[32] W3[32] b3
 * (cg_learn_nparams() = 1153 floats; row o of W2 feeds hidden unit o).  Its
 * correction factors use the Gauss-Legendre rule the reference applies to
 * the ES kernel (src/fourier/deconv.cu:80-94).  All arrays are host arrays;
 * the calls are synchronous on the default stream. */
int cg_learn_nparams(void);
/* Gridding error functional on samples (u[i], theta[i]) (fractional grid
 * offset, image mode in radians per grid cell):
 *   loss = mean |sum_t psi(d_t) e^{-i theta d_t} / psihat(theta) - 1|^2,
 *   d_t = ceil(u - kw/2) + t - u, psihat the quadrature transform of psi,
 * and its gradient grad[1153] with respect to the parameters. */
int cg_learn_grad(int kw, const float *params, const double *u,
                  const double *theta, int nsamp, float *grad, double *loss);
/* Adam training on the GPU (lr, betas 0.9 / 0.999), minibatches of `batch`
 * samples taken in order.  mode 1: minimise the functional above on (a, b) =
 * (u, theta); mode 0: least squares of y(s(d)) to targets, (a, b) = (d,
 * log psi(d)).  params is read and overwritten; losses[steps] may be NULL. */
int cg_learn_train(int kw, int mode, const double *a, const double *b, int n,
                   int batch, int steps, double lr, float *params,
                   double *losses);
/* Network output y at n inputs s (inference; without the ES exponent). */
int cg_learn_eval(const float *params, const float *s, int n, float *y);
/* Tabulated form the gridders use: coef[k] (k <= deg <= 15), the
 * coefficient of s^k in the Chebyshev interpolant of y(s) on [-1, 1]
 * (the gridders evaluate exp(beta sqrt(1 - x^2) + sum_k coef[k] s^k));
 * qpsi[p] = psi at the p = floor(2 + 1.5 kw) quadrature nodes. */
int cg_learn_tabulate(int kw, int deg, const float *params, float *coef,
                      double *qpsi);
/* Installs a learned kernel of support kw in a context, certified by the
 * caller at relative error eps (eps <= 0 removes it).  Later complex64 /
 * mixed-precision calls on the context whose epsilon is >= eps and whose ES
 * kernel is not narrower use it (narrowest first); option "use_learned" 0
 * turns this off.  complex128 calls and the reference entry points always
 * use the ES kernel.  Batches copy the context's kernels when created. */
int cg_ctx_set_kernel(void *ctx, int kw, int deg, const float *params,
                      double eps);

#ifdef __cplusplus
}
#endif

#endif /* CURAFFT_H */
