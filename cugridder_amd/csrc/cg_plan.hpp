// Host-side planner for improved W-stacking (pure C++, no device calls).
//
// Restates the reference's planning: gridder_setting / setup_gridder_plan
// (src/astro/cugridder.cu:29-210), setup_conv_opts
// (src/fourier/conv_interp_invoker.cu:49-119), get_num_cells (:30-47),
// set_nhg_w (src/utils.cu:670-709) and next235beven
// (src/common_utils.cu:76-101), and adds the MI355X tiling choice.
#pragma once
#include <cstdint>
#include <vector>

namespace cg {

int next235beven(int n, int b);

struct ConvOpts {
    int kw = 0;
    double beta = 0, es_c = 0, halfwidth = 0;
    // tabulated (learned) kernel: its values at the kernel_quadrature nodes
    // replace the ES formula when tab_nq equals the node count
    int tab_nq = 0;
    double tab_psi[32] = {};
};
// returns ier: 0 ok, 1 eps clamped, 2 invalid sigma
int setup_conv_opts(double eps, double sigma, bool single, ConvOpts &o);
int num_cells(int ms, double sigma, int kw);
void set_nhg_w(double S, double X, int kw, double sigma, int &nf, double &h,
               double &gamma);

// Gauss-Legendre rule on [-1,1], ascending (Newton on P_n); weights
// renormalised to sum to 2 as contrib/legendre_rule_fast.c:82-90 does.
void gauss_legendre(int n, double *x, double *w);

// quadrature nodes x_i and weights g_i = (kw/2) w_i psi(x_i), i < p =
// floor(2 + 1.5 kw), for the kernel Fourier transform
// (src/fourier/deconv.cu:80-94).
int kernel_quadrature(const ConvOpts &o, double *xq, double *gq);

// Correction factors on the regular grid, src/fourier/deconv.cu:12-26:
// out[k] = 2 sum g_i cos(pi k (x_i - N + 1)/(N - 1)), N = nf/2+1.
void fseries_grid(const ConvOpts &o, int nf, double *out);

// Per-tap polynomials of the FP32 kernel values psi(d) / e^beta: for tap t
// of a point whose first tap sits at offset d0 in [-kw/2, -kw/2 + 1),
//   psi(d0 + t) / e^beta = sum_q out[t * kTapStride + q] x^q,  x = 2 d0 + kw - 1,
// fitted in FP64 by Chebyshev interpolation at deg + 1 nodes and returned as
// monomial coefficients (their absolute sums stay near 1, so FP32 Horner
// evaluation is stable).  psi is the ES kernel exp(beta sqrt(1 - c d^2)), or
// with kq (ndeg + 1 coefficients) the learned kernel, whose exponent adds
// sum_k kq[k] s^k, s = 2 c d^2 - 1.  Fills kw * kTapStride floats.
constexpr int kTapStride = 12;
void tap_polynomials(int kw, double beta, double c, const float *kq, int ndeg, int deg,
                     float *out);

struct Tiling {
    int tx = 0, ty = 0;      // tile extent along nf1 (fast) and nf2
    int ntx = 0, nty = 0;    // tiles per axis
    int sbx = 0, sby = 0;    // source-bin reach (bins per axis feeding a tile)
    int pc = 0;              // planes per gridding chunk
    int nchunk = 0;          // plane chunks
    int64_t ntiles() const { return (int64_t)ntx * nty; }
};

struct WPlan {
    // call parameters
    int64_t nrow = 0;
    int nx = 0, ny = 0;       // image rows (x, u) and columns (y, v)
    double fov = 0, freq = 0, eps = 0, sigma = 2.0;
    int sign = -1;
    bool single = false;      // float ABI: EPSILON of the float build
    // derived (reference names)
    int N1 = 0, N2 = 0;       // N1 = ny (fast image axis), N2 = nx
    int nf1 = 0, nf2 = 0, nf3 = 0;
    ConvOpts conv;
    double f_over_c = 0, pix_x = 0, pix_y = 0;
    double i_center = 0, i_half_width = 0, o_center = 0, o_half_width = 0;
    double gamma = 0, h = 0;
    int ier = 0;
    Tiling tile;
    // grid rows (nf2 axis) the points' footprints can reach: band_len rows
    // from band_s, cyclic (band_len = 0: unknown or the whole axis); the
    // same for the grid columns (nf1 axis)
    int band_s = 0, band_len = 0, cband_s = 0, cband_len = 0;
    // 1: improved W-stacking (w-planes, w-correction, 1/n); 0: plain 2-D
    // gridding that ignores w (one plane, no n), NIFTY's do_wstacking=False
    int wstack = 1;
    int64_t grid_cells() const { return (int64_t)nf1 * nf2 * nf3; }
};

// Image geometry beyond the reference's (fov, nx, ny): pix_y > 0 sets the
// pixel size along ny (radians) instead of fov / ny, as NIFTY's rad_pix_y;
// wstack = 0 asks for plain 2-D gridding (w ignored, no 1/n).
struct PlanGeom {
    double pix_y = 0;
    int wstack = 1;
};

// Sets p.band_s / p.band_len from the range [ulo, uhi] of u * 2 pi f / c
// over all points (before multiplying by pix_x), and p.cband_s / p.cband_len
// from that of v (pix_y, nf1): the footprint rows (columns) of the lowest
// and highest grid coordinate, plus one of margin on each side.  No band on
// an axis whose range wraps.
void plan_bands(WPlan &p, double ulo, double uhi, double vlo, double vhi);

// Fills every field from the call parameters and the w range (min/max of
// the scaled w = w*f/c*sign*2pi over all rows).  Returns ier (>=2 is fatal).
// conv, when given, replaces the kernel setup_conv_opts picks for eps (a
// learned kernel of smaller support); every size follows its kw.
int make_wplan(WPlan &p, int64_t nrow, int nxdirty, int nydirty, double fov,
               double freq, double eps, double sigma, int sign, bool single,
               double wmin_scaled, double wmax_scaled, const ConvOpts *conv = nullptr,
               const PlanGeom *geom = nullptr);

// Chooses tile extents and plane chunking for an LDS budget (bytes) and
// complex element size; `nplanes` is the grid depth and `halo` the extra
// cells per axis a workgroup keeps in LDS (kw-1 when degridding, 0 when
// gridding).
Tiling choose_tiling(int nf1, int nf2, int nplanes, int kw, int elem_bytes,
                     int lds_budget, int tile_pref, int halo);

}  // namespace cg
