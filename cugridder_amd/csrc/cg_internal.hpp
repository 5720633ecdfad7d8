// Launchers shared between the HIP translation units of libcurafft.so.
#pragma once
#include "cg_common.hpp"
#include "cg_plan.hpp"

namespace cg {

// Geometry needed by the coordinate kernels (all FP64: the uvw scaling to grid coordinates).
struct CoordArgs {
    double scale_uv;   // 2 pi f/c (one channel)
    double scale_w;    // 2 pi f/c * sign
    // several channels: point j is row j / nchan at channel j % nchan, with
    // scale_uv = chan_scale[channel] (device array of 2 pi f_c/c)
    int nchan;
    const double *chan_scale;
    double pix_x, pix_y;
    double i_center, gamma, o_center;
    int nf1, nf2, nf3, kw;
    int tx, ty, ntx, nty;
    int cshift;        // key = tile << cshift | cell of the start inside the tile
    int pirange;       // 1: coordinates already in [-pi,pi) radians (2-D NUFFT)
    int has_w;         // 0: 2-D (no w axis)
    // compact tile numbering (cnbx > 0): every point's tile lies in the
    // rectangle of tile rows cby0 .. cby0 + cnby - 1 and tile columns cbx0 ..
    // cbx0 + cnbx - 1 (no wrap), and the key numbers the rectangle's tiles
    // row-major, which sorts exactly as the full numbering with fewer bits
    int cby0, cnby, cbx0, cnbx;
};
// the key map's tile number of grid tile (bx, by), and of the first
// rectangle tile at or after grid tile t in row-major order
struct KeyRect {
    int ntx, cby0, cnby, cbx0, cnbx;
    __host__ __device__ int64_t lower(int64_t t, int64_t ntiles) const
    {
        if (!cnbx) return t;
        if (t >= ntiles) return (int64_t)cnby * cnbx;
        const int by = (int)(t / ntx), bx = (int)(t - (int64_t)by * ntx);
        if (by < cby0) return 0;
        if (by >= cby0 + cnby) return (int64_t)cnby * cnbx;
        const int ry = by - cby0;
        if (bx < cbx0) return (int64_t)ry * cnbx;
        if (bx >= cbx0 + cnbx) return (int64_t)(ry + 1) * cnbx;
        return (int64_t)ry * cnbx + (bx - cbx0);
    }
};
inline KeyRect key_rect(const CoordArgs &a) { return KeyRect{a.ntx, a.cby0, a.cnby, a.cbx0, a.cnbx}; }

CoordArgs make_coord_args(const WPlan &p, const Tiling &t);
// fill the tiling fields of a; the start cell joins the sort key when the
// key still fits 32 bits
void set_coord_tiling(CoordArgs &a, const Tiling &t);
// bits of the sort key
int sort_key_bits(const CoordArgs &a, int64_t ntiles);

// ---- prep (cg_prep.hip)
// min/max of w * scale_w (FP64) over nrow rows of uvw [nrow][3], and
// uv[4] = {umin, umax, vmin, vmax} when uv is given; partial buffer needs
// 6150 doubles.  Results copied to host (synchronises).
template <typename TU>
int launch_wminmax(const TU *uvw, int64_t nrow, double scale_w, double *d_part,
                   double &wmin, double &wmax, hipStream_t s, double *uv = nullptr);
// keys[j] = (tile of the footprint start of row j) << cshift | start cell
// inside the tile; vals[j] = j (vals may be null)
template <typename TU>
int launch_keys(const TU *uvw, int64_t nrow, const CoordArgs &a, uint32_t *keys,
                int32_t *vals, hipStream_t s);
// 2-D NUFFT variant: separate coordinate arrays in [-pi,pi)
template <typename TU>
int launch_keys_xy(const TU *x, const TU *y, int64_t m, const CoordArgs &a,
                   uint32_t *keys, int32_t *vals, hipStream_t s);
// p[0..n) = v (a kernel, so that it can sit in a captured graph)
int launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s);
// mask[y] (y < nf2) = 1 for the grid rows and mask[nf2 + x] (x < nf1) = 1
// for the grid columns that points of non-empty tiles reach; other rows and
// columns are zero after gridding and unread by degridding
int launch_row_mask(const int32_t *tile_start, const Tiling &t, int kw, int nf1, int nf2,
                    uint8_t *mask, hipStream_t s);
// tile_start[t] = first sorted position with key >> shift >= t, t in [0, ntiles]
int launch_tile_bounds(const uint32_t *sorted_keys, int64_t m, int64_t ntiles, int shift,
                       const KeyRect &kr,
                       int32_t *tile_start, hipStream_t s);

// Sorted point records: geo = {d1, d2, d3, packed(lx | ly<<8 | z0<<16)} and a
// complex value (ms2dirty: vis*wgt*e^{i o_c w}; dirty2ms: wgt*e^{-i o_c w}).
template <typename T> struct Geo;
template <> struct Geo<float> { using type = float4; };
template <> struct Geo<double> { using type = double4; };
template <typename T> using geo_t = typename Geo<T>::type;

// sorted records geo / val of the points; with rec (32 B per point of
// scratch, complex64 only) in two coalescing passes, else one gather pass
template <typename T, typename TU, typename TV, typename TW>
int launch_records(const TU *uvw, const TV *vis, const TW *wgt, const int32_t *perm,
                   int64_t nrow, const CoordArgs &a, int adjoint, geo_t<T> *geo,
                   cx_t<T> *val, float4 *rec, hipStream_t s);
template <typename T, typename TU, typename TV>
int launch_records_xy(const TU *x, const TU *y, const TV *c, const int32_t *perm,
                      int64_t m, const CoordArgs &a, geo_t<T> *geo, cx_t<T> *val,
                      hipStream_t s);
// complex64: the key pass (launch_keys) and the caller-order record slots of
// launch_records' two-pass path in one pass over uvw; launch_gather_records
// is then the records' second pass
template <typename TU, typename TV, typename TW>
int launch_keys_prerecords(const TU *uvw, const TV *vis, const TW *wgt, int64_t nrow,
                           const CoordArgs &a, int adjoint, uint32_t *keys, int32_t *vals,
                           float4 *rec, hipStream_t s);
int launch_gather_records(const float4 *rec, const int32_t *perm, int64_t nrow, float4 *geo,
                          float2 *val, hipStream_t s);
// complex64 2-D: keys, identity vals and the caller-order record slots in one
// pass (rec holds 32 B per point with c, 16 B without), then the gather
int launch_keys_prerecords_xy(const float *x, const float *y, const float2 *c, int64_t m,
                              const CoordArgs &a, uint32_t *keys, int32_t *vals, float4 *rec,
                              hipStream_t s);
int launch_gather_records_xy(const float4 *rec, const int32_t *perm, int64_t m, int has_c,
                             float4 *geo, float2 *val, hipStream_t s);

// ---- sort (cg_sort.hip): stable LSD radix sort of (key, val) pairs by the
// low `key_bits` bits.  tmp_keys/tmp_vals have m entries; hist needs
// radix_hist_size(m) ints.  Result ends in keys/vals (copy_back), or, with
// copy_back = 0, in keys/vals after an even and tmp after an odd number of
// passes (radix_passes): a caller that starts an odd sort in its tmp
// buffers gets the result in its keys without a copy.  identity_vals: vals
// are not read, the first pass takes value i for item i (the key kernels
// then need not write them).
int64_t radix_hist_size(int64_t m);
int radix_passes(int key_bits);
int launch_radix_sort(uint32_t *keys, int32_t *vals, uint32_t *tmp_keys,
                      int32_t *tmp_vals, int64_t m, int key_bits, int32_t *hist,
                      int32_t *scan_tmp, hipStream_t s, int copy_back = 1, int identity_vals = 0);
int64_t scan_tmp_size(int64_t n);
// exclusive scan of n ints in place
int launch_exclusive_scan(int32_t *data, int64_t n, int32_t *scan_tmp, hipStream_t s);

// learned kernel sizes
constexpr int LEARN_NPARAMS = 1153;  // 1 -> 32 -> 32 -> 1 tanh MLP
constexpr int LEARN_MAXKW = 16;
constexpr int LEARN_MAXQ = 32;       // half-line quadrature nodes, floor(2 + 1.5 kw)
constexpr int LEARN_MAXDEG = 15;     // degree of the network's polynomial in s

// ---- gridding / degridding (cg_grid.hip)
struct GridArgs {
    int m;  // number of points (rows)
    int nf1, nf2, nf3;
    int tx, ty, ntx, nty, sbx, sby, pc, nchunk;
    int cshift;     // sort key = tile << cshift | start cell inside the tile
    int dense_min;  // > 0: bins with at least this many points go to the dense path
    double beta, es_c;
    // tiles whose rows all have rowmask 0, or whose columns all have
    // colmask 0, are not written (null: every tile is written); the sparse
    // gridder then collects the others in tile_list (ntiles + 1 ints)
    const uint8_t *rowmask, *colmask;
    int32_t *tile_list;
    // learned kernel (learned != 0): the value at offset d is
    // exp(beta sqrt(1 - es_c d^2) + sum_k kq[k] s^k), s = 2 es_c d^2 - 1,
    // instead of the ES kernel's exp(beta sqrt(1 - es_c d^2))
    int learned;
    float kq[LEARN_MAXDEG + 1];
    // FP32 kernel constants (set with beta, es_c): -beta log2(e), es_c, and
    // e^beta, e^(2 beta), e^(3 beta) that scale values / sums of products of
    // values divided by e^beta back
    float f_nb, f_ec, f_e1, f_e2;
    double e3;
    KeyRect kr;  // the sort key's tile numbering
    // 1: complex64 one-plane gridding may run as per-quadrant MFMA products
    // (grid2d_mfma_supported decides); 0: always the LDS-atomic kernel
    int mfma2d;
    // FP32 kernel values as per-tap polynomials (tap_polynomials, kw <= 8):
    // tp_on != 0 selects them over the exp2 form in the complex64 kernels
    int tp_on;
    float tp[8 * kTapStride];
    // w-plane slab held by the grid buffer: planes zb .. zb + zn - 1 of the
    // nf3 (periodic) planes sit at buffer planes 0 .. zn - 1 (zn = 0: the
    // whole grid, zb = 0).  Gridding writes only the slab; degridding adds
    // only the slab's taps, and with accumulate != 0 adds its result to
    // the output instead of storing it.
    int zb, zn, accumulate;
};
// planes held by the grid buffer
__host__ __device__ inline int slab_planes(const GridArgs &g) { return g.zn > 0 ? g.zn : g.nf3; }
bool grid2d_mfma_supported(const GridArgs &g, int kw);
template <typename T>
int launch_grid(const geo_t<T> *geo, const cx_t<T> *val, const int32_t *tile_start,
                cx_t<T> *fw, const GridArgs &g, int kw, hipStream_t s);
// list (ntiles + 2 ints of scratch, or nullptr): run over the listed
// non-empty tiles below g.dense_min instead of one workgroup per tile
template <typename T, typename TV>
int launch_degrid(const geo_t<T> *geo, const cx_t<T> *fac, const int32_t *perm,
                  const int32_t *tile_start, const cx_t<T> *fw, TV *out,
                  const GridArgs &g, int kw, int32_t *list, hipStream_t s);
int grid_lds_bytes(const GridArgs &g, int elem);
// Dense bins (fp32, kw <= 8, nf3 <= 32, cshift > 0): the sparse kernel
// launched with g.dense_min > 0 leaves them out; this adds them to fw with
// per-start-cell MFMA products.  list needs dense_list_ints(ntiles, g.m)
// ints of scratch.  The dense kernels' work units are chunks of at most
// dense_chunk(g.m) sorted points of one bin.
int dense_chunk(int64_t m);
int64_t dense_list_ints(int64_t ntiles, int64_t m);
bool grid_dense_supported(const GridArgs &g, int kw, int elem);
int launch_grid_dense(const geo_t<float> *geo, const cx_t<float> *val, const uint32_t *keys,
                      const int32_t *tile_start, int32_t *list, cx_t<float> *fw, const GridArgs &g,
                      int kw, hipStream_t s);
int degrid_lds_bytes(const GridArgs &g, int kw, int elem);
// Dense bins of the complex64 degridder (kw <= 8, nf3 <= 32, start cell in
// the sort key, whole grid staged, no slabs): k_degrid launched with
// g.dense_min > 0 leaves the bins with at least that many points to this
// one, which interpolates them on MFMA.  list needs dense_list_ints ints.
bool degrid_dense_supported(const GridArgs &g, int kw);
int launch_degrid_dense(const float4 *geo, const float2 *fac, const int32_t *perm,
                        const uint32_t *keys, const int32_t *tile_start, int32_t *list,
                        const float2 *fw, float2 *out, const GridArgs &g, int kw, hipStream_t s);

// ---- FFT (cg_fft.hip): in-place unnormalised 2-D FFT of `batch` planes
// [batch][n2][n1], exponent sign dir (+1 / -1), natural order in and out.
// tw1/tw2 are device tables of e^{+2 pi i t/n} (length n1 / n2) in
// precision T; tmp (batch*n2*n1 elements) is needed when n2 is split.
template <typename T>
int launch_fft2d(cx_t<T> *data, int batch, int n2, int n1, int dir, const cx_t<T> *tw1,
                 const cx_t<T> *tw2, cx_t<T> *tmp, hipStream_t s);
// One pass: rows (fast axis) or columns.  Only fast-axis indices k < live_a
// or k >= live_b matter: the column pass transforms only those columns, a
// dir > 0 row pass stores only those outputs and a dir < 0 row pass reads
// only those inputs.  A dir > 0 column pass leaves a split column axis in
// split order (frequency ka + na kb at row ka nb + kb, see fft_col_split);
// a dir < 0 column pass expects that order.
// rowmask (nf2 + nf1 bytes, may be null): rows with mask 0 are zero on
// input to a dir > 0 pass and not needed on output of a dir < 0 pass, so
// they may be skipped (the row pass and the natural-order column step use
// it); the columns x with rowmask[nf2 + x] = 0 likewise for the row pass.
// Frequencies k along n2 with spec_a <= k < spec_b are dead in the spectrum
// (not needed after a dir > 0 column pass, zero before a dir < 0 one); the
// split-order column step skips them when fft_spectrum_pruned(n2).
// A column pass given a band (band_len > 0: rows band_s .. band_s + band_len
// - 1, cyclic, hold every non-zero input row of a dir > 0 pass / every row a
// dir < 0 pass must produce; rows outside it are neither read nor written)
// runs as one band-limited pass when the band is short enough (complex64).
template <typename T>
int launch_fft_pass(cx_t<T> *data, int batch, int n2, int n1, int dir, bool rows,
                    const cx_t<T> *tw, int live_a, int live_b, const uint8_t *rowmask,
                    int spec_a, int spec_b, hipStream_t s, int band_s = 0, int band_len = 0);
// true when the column pass honours the dead spectrum rows (then a dir < 0
// column pass does not read them and they need not be written); it then
// also honours the row mask (a dir > 0 pass does not read masked rows)
template <typename T>
bool fft_spectrum_pruned(int n2);
// true when the row pass honours the row and column masks (a dir > 0 pass
// reads neither masked rows nor masked columns)
template <typename T>
bool fft_rows_masked(int n1);
// n = na * nb when columns of length n are transformed by the split
bool fft_col_split(int n, int &na, int &nb);
template <typename T>
bool fft_supported(int n);
// complex128 rows too long for one workgroup, transformed with a radix-2
// step around two half-length rows: their twiddle table (tw1) holds the
// table of n followed by the table of n / 2
template <typename T>
bool fft_row_long(int n);
template <typename T>
int launch_twiddles(cx_t<T> *tw, int n, hipStream_t s);

// ---- W-stacking per-pixel stages (cg_wstack.hip), FP64 phases & factors
// Gauss-Legendre quadrature of the ES kernel's Fourier transform
struct Quad {
    int nq;
    double xq[32], gq[32];
};
constexpr int kChebMax = 96;
Quad make_quad(const ConvOpts &o);
// phi[k] = 2 sum_i g_i cos(pi k (x_i - N + 1)/(N - 1)), N = nf/2 + 1
// (src/fourier/deconv.cu:12-26), k in [0, nf/2]
int launch_fseries_grid(const Quad &q, int nf, double *phi, hipStream_t s);

struct WArgs {
    int N1, N2, nf1, nf2, nf3;
    int cna, cnb;  // split of the column axis (0: natural order)
    int skip_dead; // 1: the prologue leaves spectrum rows outside the image unwritten
    // w-plane slab of the grid buffer (planes zb .. zb + zn - 1 at buffer
    // planes 0 .. zn - 1; zn = 0: every plane); the epilogue adds its part
    // of the w-DFT to the image when accumulate != 0
    int zb, zn, accumulate;
    double pix_x, pix_y, o_center, i_center, gamma, h;
    Quad qd;
    // psihat(z') over the image's z' range [cheb_a, cheb_b] as a Chebyshev
    // series of cheb_n terms (0: evaluate the quadrature sum per pixel)
    int cheb_n;
    double cheb_a, cheb_b;
    double cheb[kChebMax];
};
WArgs make_wargs(const WPlan &p);
// z' and psihat(z') on the (N2/2+1) x (N1/2+1) quarter image
int launch_quarter_table(const WArgs &w, double *zq, double *phi3, hipStream_t s);
// ms2dirty epilogue: w-DFT over planes + deconvolution + w-correction + 1/n
template <typename T, typename TV>
int launch_ms2dirty_epilogue(const cx_t<T> *fw, const double *zq, const double *phi3,
                             const double *phi1, const double *phi2, TV *dirty,
                             const WArgs &w, hipStream_t s);
// dirty2ms prologue: 1/n, w-correction, deconvolution and w-IDFT onto planes
// (writes every cell of the live columns, zero outside the image footprint)
template <typename T, typename TV>
int launch_dirty2ms_prologue(const TV *dirty, const double *zq, const double *phi3,
                             const double *phi1, const double *phi2, cx_t<T> *fw,
                             const WArgs &w, hipStream_t s);
// 2-D NUFFT deconvolution (type 1: grid -> modes; type 2: modes -> grid)
template <typename T, typename TV>
int launch_deconv2d(cx_t<T> *fw, TV *fk, const double *phi1, const double *phi2,
                    int N1, int N2, int nf1, int nf2, int type, hipStream_t s);

// ---- learned kernel (cg_learn.hip)
struct LearnQuad {
    int kw, nq;
    double beta;  // ES exponent the network corrects
    double xq[LEARN_MAXQ], gq[LEARN_MAXQ];
};
double learn_beta(int kw);
LearnQuad learn_quad(int kw);
// tabulated form of a trained network: y as a polynomial in s (coef[k] of
// s^k, zero-padded to LEARN_MAXDEG) and the kernel exp(beta (sqrt(1 - x^2)
// - 1) + y) at the quadrature nodes qx (FP64)
struct LearnedKernel {
    int kw = 0, deg = 0, nq = 0;
    double beta = 0;
    float coef[LEARN_MAXDEG + 1];
    double qx[LEARN_MAXQ], qg[LEARN_MAXQ], qpsi[LEARN_MAXQ];
};
void cheb_to_monomial(int deg, double *zc, double *M);
int learn_grad(int kw, const float *params, const double *u, const double *th, int nsamp,
               float *grad, double *loss, hipStream_t s);
int learn_train(int kw, int mode, const double *a, const double *b, int n, int batch, int steps,
                double lr, float *params, double *losses, hipStream_t s);
int learn_eval(const float *params, const float *s_in, int n, float *y, hipStream_t s);
int learn_kernel_build(int kw, int deg, const float *params, LearnedKernel &K, hipStream_t s);

}  // namespace cg
