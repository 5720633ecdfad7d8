// Host-side planner; see cg_plan.hpp for the reference functions restated.
#include "cg_plan.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace cg {

static constexpr double PI = 3.14159265358979323846264338327950288;

int next235beven(int n, int b)
{
    // src/common_utils.cu:76-101
    if (n <= 2) return 2;
    if (n % 2 == 1) n += 1;
    int nplus = n - 2;
    int numdiv = 2;
    while (numdiv > 1 || nplus % b != 0) {
        nplus += 2;
        numdiv = nplus;
        while (numdiv % 2 == 0) numdiv /= 2;
        while (numdiv % 3 == 0) numdiv /= 3;
        while (numdiv % 5 == 0) numdiv /= 5;
    }
    return nplus;
}

int setup_conv_opts(double eps, double sigma, bool single, ConvOpts &o)
{
    // src/fourier/conv_interp_invoker.cu:49-119 (kerevalmeth = 0)
    if (!(sigma > 1.0)) return 2;
    int ier = 0;
    const double eps_min = single ? 6e-08 : 1.1e-16;  // include/utils.h:34-38
    if (eps < eps_min) {
        eps = eps_min;
        ier = 1;
    }
    int kw = (int)std::ceil(-std::log10(eps / 10.0));
    if (sigma != 2.0) kw = (int)std::ceil(-std::log(eps) / (PI * std::sqrt(1.0 - 1.0 / sigma)));
    kw = std::max(2, kw);
    kw = std::min(kw, 16);  // MAX_KERNEL_WIDTH, include/curafft_plan.h:14
    double bo = 2.30;
    if (kw == 2) bo = 2.20;
    if (kw == 3) bo = 2.26;
    if (kw == 4) bo = 2.38;
    if (sigma != 2.0) bo = 0.97 * PI * (1.0 - 1.0 / (2.0 * sigma));
    o.kw = kw;
    o.beta = bo * kw;
    o.es_c = 4.0 / (double)(kw * kw);
    o.halfwidth = kw / 2.0;
    return ier;
}

int num_cells(int ms, double sigma, int kw)
{
    // src/fourier/conv_interp_invoker.cu:30-47
    int nf = (int)(sigma * ms);
    if (nf < 2 * kw) nf = 2 * kw;
    return next235beven(nf, 1);
}

void set_nhg_w(double S, double X, int kw, double sigma, int &nf, double &h,
               double &gamma)
{
    // src/utils.cu:670-709
    int nss = kw + 1;
    double Xsafe = X, Ssafe = S;
    if (X == 0.0) {
        if (S == 0.0) {
            Xsafe = 1.0;
            Ssafe = 1.0;
        } else {
            Xsafe = std::max(Xsafe, 1.0 / S);
        }
    } else {
        Ssafe = std::max(Ssafe, 1.0 / X);
    }
    double nfd = 2.0 * sigma * Ssafe * Xsafe / PI + nss;
    if (!std::isfinite(nfd)) nfd = 0.0;
    int n = (int)nfd;
    n = n % 2 + n;
    if (n < 2 * kw) n = 2 * kw;
    nf = n;
    h = 2.0 * PI / n;
    gamma = (double)n / (2.0 * sigma * Ssafe);
}

void gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) {
                p0 = 1.0;
                p1 = z;
                for (int k = 2; k <= n; k++) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1.0);
                break;
            }
        }
        double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[i] = wi;
        w[n - 1 - i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += w[i];
    for (int i = 0; i < n; i++) w[i] = 2.0 * w[i] / s;
}

int kernel_quadrature(const ConvOpts &o, double *xq, double *gq)
{
    double alpha = o.kw / 2.0;
    int p = (int)(2 + 3.0 * alpha);
    std::vector<double> x(2 * p), w(2 * p);
    gauss_legendre(2 * p, x.data(), w.data());
    for (int n = 0; n < p; n++) {
        double xn = x[n] * alpha;
        double phi = 0.0;
        if (o.tab_nq == p) phi = o.tab_psi[n];
        else if (std::fabs(xn) <= o.halfwidth) phi = std::exp(o.beta * std::sqrt(1.0 - o.es_c * xn * xn));
        xq[n] = xn;
        gq[n] = alpha * w[n] * phi;
    }
    return p;
}

void tap_polynomials(int kw, double beta, double c, const float *kq, int ndeg, int deg,
                     float *out)
{
    const int n = deg + 1;
    std::vector<double> xs(n), f(n), a(n), T(n * n, 0.0), mono(n);
    // T[k][q]: monomial coefficients of the Chebyshev polynomial T_k
    T[0] = 1.0;
    if (n > 1) T[1 * n + 1] = 1.0;
    for (int k = 2; k < n; k++)
        for (int q = 0; q < n; q++)
            T[k * n + q] = (q > 0 ? 2.0 * T[(k - 1) * n + q - 1] : 0.0) - T[(k - 2) * n + q];
    for (int i = 0; i < n; i++) xs[i] = std::cos(PI * (i + 0.5) / n);
    for (int t = 0; t < kw; t++) {
        for (int i = 0; i < n; i++) {
            const double d = (xs[i] + 1.0) / 2.0 - kw / 2.0 + t;
            const double u = c * d * d;
            double e = beta * (std::sqrt(std::max(1.0 - u, 0.0)) - 1.0);
            if (kq) {
                const double sv = 2.0 * u - 1.0;
                double y = 0.0;
                for (int k = ndeg; k >= 0; k--) y = y * sv + kq[k];
                e += y;
            }
            f[i] = std::fabs(d) < kw / 2.0 ? std::exp(e) : 0.0;
        }
        // Chebyshev coefficients from the values at the nodes
        for (int k = 0; k < n; k++) {
            double sacc = 0.0;
            for (int i = 0; i < n; i++) sacc += f[i] * std::cos(PI * k * (i + 0.5) / n);
            a[k] = (k == 0 ? 1.0 : 2.0) * sacc / n;
        }
        for (int q = 0; q < n; q++) {
            double m = 0.0;
            for (int k = 0; k < n; k++) m += a[k] * T[k * n + q];
            mono[q] = m;
        }
        for (int q = 0; q < kTapStride; q++) out[t * kTapStride + q] = q < n ? (float)mono[q] : 0.0f;
    }
}

void fseries_grid(const ConvOpts &o, int nf, double *out)
{
    double xq[64], gq[64];
    int p = kernel_quadrature(o, xq, gq);
    int N = nf / 2 + 1;
    for (int k = 0; k < N; k++) {
        double s = 0.0;
        for (int i = 0; i < p; i++) s += gq[i] * std::cos((xq[i] - N + 1) / (double)(N - 1) * PI * k);
        out[k] = 2.0 * s;
    }
}

// footprint cells of grid coordinates (x pix / 2 pi + 1/2) nf, x in [xlo, xhi]
static void axis_band(int nf, double pix, int kw, double xlo, double xhi, int &s, int &len)
{
    s = len = 0;
    const double a = xlo * pix, b = xhi * pix;
    // valid while x pix stays inside [-pi, pi) (the shift-rescale's wrap)
    if (!(a <= b) || !(a > -PI + 1e-9) || !(b < PI - 1e-9)) return;
    const double glo = (a * (0.5 / PI) + 0.5) * nf, ghi = (b * (0.5 / PI) + 0.5) * nf;
    const int64_t y0 = (int64_t)std::ceil(glo - 0.5 * kw) - 1;
    const int64_t y1 = (int64_t)std::ceil(ghi - 0.5 * kw) + kw;  // one past the last cell, + 1 margin
    const int64_t n = y1 - y0 + 1;
    if (n <= 0 || n >= nf) return;
    s = (int)(((y0 % nf) + nf) % nf);
    len = (int)n;
}

void plan_bands(WPlan &p, double ulo, double uhi, double vlo, double vhi)
{
    axis_band(p.nf2, p.pix_x, p.conv.kw, ulo, uhi, p.band_s, p.band_len);
    axis_band(p.nf1, p.pix_y, p.conv.kw, vlo, vhi, p.cband_s, p.cband_len);
}

Tiling choose_tiling(int nf1, int nf2, int nplanes, int kw, int elem_bytes,
                     int lds_budget, int tile_pref, int halo)
{
    // Tiles must divide the grid (so no tile is partial); pick the largest
    // power of two <= tile_pref that divides each axis.
    auto pick = [&](int nf) {
        int t = tile_pref;
        while (t > 1 && nf % t != 0) t >>= 1;
        return std::max(t, 1);
    };
    Tiling T;
    T.tx = pick(nf1);
    T.ty = pick(nf2);
    // a staged plane with a halo (degridding) has an odd row stride in LDS
    auto plane_bytes = [&]() {
        const int64_t w = halo ? ((T.tx + halo) | 1) : T.tx;
        return w * (T.ty + halo) * elem_bytes;
    };
    // shrink the tile while even one plane does not fit the LDS budget
    while (plane_bytes() > lds_budget && (T.tx > 2 || T.ty > 2)) {
        if (T.tx >= T.ty && T.tx > 2) T.tx >>= 1;
        else T.ty >>= 1;
    }
    T.ntx = nf1 / T.tx;
    T.nty = nf2 / T.ty;
    // a point whose footprint starts in bin b touches tiles b .. b+sb-1
    T.sbx = 1 + (kw - 1 + T.tx - 1) / T.tx;
    T.sby = 1 + (kw - 1 + T.ty - 1) / T.ty;
    T.sbx = std::min(T.sbx, T.ntx);
    T.sby = std::min(T.sby, T.nty);
    int64_t per_plane = plane_bytes();
    int pc = (int)std::max<int64_t>(1, lds_budget / per_plane);
    pc = std::min(pc, nplanes);
    T.nchunk = (nplanes + pc - 1) / pc;
    pc = (nplanes + T.nchunk - 1) / T.nchunk;  // balance chunks
    T.pc = pc;
    return T;
}

int make_wplan(WPlan &p, int64_t nrow, int nxdirty, int nydirty, double fov,
               double freq, double eps, double sigma, int sign, bool single,
               double wmin, double wmax, const ConvOpts *conv, const PlanGeom *geom)
{
    if (nrow < 0 || nxdirty <= 0 || nydirty <= 0 || !(fov > 0) || !(freq > 0))
        return 2;
    p = WPlan();
    p.nrow = nrow;
    p.nx = nxdirty;
    p.ny = nydirty;
    p.fov = fov;
    p.freq = freq;
    p.eps = eps;
    p.sigma = sigma;
    p.sign = sign;
    p.single = single;
    p.N1 = nydirty;
    p.N2 = nxdirty;
    int ier = setup_conv_opts(eps, sigma, single, p.conv);
    if (ier >= 2) return ier;
    if (conv) p.conv = *conv;
    p.ier = ier;
    p.f_over_c = freq / 299792458.0;
    // src/astro/cugridder.cu:48-76
    p.pix_x = fov / 180.0 * PI / (double)p.N2;
    p.pix_y = fov / 180.0 * PI / (double)p.N1;
    if (geom && geom->pix_y > 0) p.pix_y = geom->pix_y;
    p.nf1 = num_cells(p.N1, sigma, p.conv.kw);
    p.nf2 = num_cells(p.N2, sigma, p.conv.kw);
    if (p.nf1 > 65536 || p.nf2 > 65536) return 3;
    if (geom && !geom->wstack) {
        // one plane, no w-terms: the w-phase, w-correction and 1/n vanish
        p.wstack = 0;
        p.nf3 = 1;
        p.gamma = 1.0;
        p.h = 0.0;
        return ier;
    }
    double l_min = -0.5 * p.pix_x * p.N2;
    double m_min = -0.5 * p.pix_y * p.N1;
    double temp = 1.0 - l_min * l_min - m_min * m_min;
    double n_lm = temp <= 0 ? 0 : std::sqrt(temp);
    if (nrow == 0) {
        wmin = 0;
        wmax = 0;
    }
    p.i_center = (wmax + wmin) / 2.0;
    p.i_half_width = (wmax - wmin) / 2.0;
    double o_min = n_lm - 1.0;
    p.o_center = o_min / 2.0;
    p.o_half_width = std::fabs(o_min / 2.0);
    set_nhg_w(p.o_half_width, p.i_half_width, p.conv.kw, sigma, p.nf3, p.h, p.gamma);
    if (p.nf3 > 65535) return 3;
    return ier;
}

}  // namespace cg
