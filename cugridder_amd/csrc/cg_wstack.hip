// Per-pixel stages of improved W-stacking, with FP64 kept for n-1, the
// w-phase and the correction factors.
//
// ms2dirty epilogue fuses, per image pixel, the reference's
//   w_term_dft      (src/fourier/cuft.cu:597-645): sum_p F_p e^{i z'(p-P/2)}
//   deconv_2d type1 (src/fourier/deconv.cu:250-294): / (phi1 phi2)
//   w_term_deconv   (src/fourier/deconv.cu:401-432): * e^{i(n-1-o_c) i_c} / phi3
//   div_n_lm        (src/astro/ra_exec.cu:23-37):   / n
// and the dirty2ms prologue fuses their adjoints (div_n_lm, w_term_deconv
// with flag -1, deconv_2d type 2, w_term_idft cuft.cu:647-689) and writes
// every cell of the live columns of every plane (zero outside the image
// footprint); the other columns are never read by the inverse FFT, so the
// grid needs no memset.  Both index the spectrum's column axis in the order
// the column FFT leaves it (split order for long columns).
#include <vector>

#include "cg_internal.hpp"

namespace cg {

// psihat(z') = sum_i g_i 2 cos(x_i z') (the quadrature of the kernel's
// Fourier transform, deconv.cu:49-65), FP64 on the host
static double psihat_at(const Quad &q, double zp)
{
    double s = 0.0;
    for (int i = 0; i < q.nq; i++) s += q.gq[i] * 2.0 * std::cos(q.xq[i] * zp);
    return s;
}

// Chebyshev series of psihat(z') over the z' range of the image's pixels:
// z = n - 1 falls monotonically with l^2 + m^2 inside the horizon, so the
// range is z' at the centre and at the quarter image's far corner.  The
// shortest series (16 .. kChebMax terms) whose last three coefficients are
// below 1e-14 of the coefficient sum and which agrees with the direct sum
// to 1e-14 at 17 points of the range is kept; none (images reaching the
// horizon, or a series that does not converge) leaves cheb_n = 0.  Host
// cost: a few hundred cosines per plan (the DCT's cosines are a table).
static constexpr int kChebSizes[] = {16, 24, 32, 48, 64, kChebMax};

static const std::vector<double> &cheb_cos_table(int n)
{
    static const std::vector<std::vector<double>> tabs = [] {
        std::vector<std::vector<double>> t;
        for (int m : kChebSizes) {
            std::vector<double> c((size_t)m * m);
            for (int k = 0; k < m; k++)
                for (int j = 0; j < m; j++) c[(size_t)k * m + j] = std::cos(kPi * k * (j + 0.5) / m);
            t.push_back(std::move(c));
        }
        return t;
    }();
    int i = 0;
    while (kChebSizes[i] != n) i++;
    return tabs[i];
}

static double clenshaw(const double *c, int n, double x)
{
    double b1 = 0.0, b2 = 0.0;
    for (int k = n - 1; k >= 1; k--) {
        const double t = 2.0 * x * b1 - b2 + c[k];
        b2 = b1;
        b1 = t;
    }
    return x * b1 - b2 + c[0];
}

static void fit_psihat(WArgs &w)
{
    w.cheb_n = 0;
    const double lx = (w.N2 / 2) * w.pix_x, my = (w.N1 / 2) * w.pix_y;
    const double r2 = lx * lx + my * my;
    if (!(r2 < 1.0)) return;
    const double zfar = -r2 / (std::sqrt(1.0 - r2) + 1.0);
    const double z0 = (0.0 - w.o_center) * w.gamma * w.h, z1 = (zfar - w.o_center) * w.gamma * w.h;
    const double a = std::min(z0, z1), b = std::max(z0, z1);
    if (!(b - a > 1e-9 * std::max(1.0, std::fabs(a) + std::fabs(b)))) return;
    for (int n : kChebSizes) {
        const std::vector<double> &ct = cheb_cos_table(n);
        double f[kChebMax], c[kChebMax];
        for (int j = 0; j < n; j++) f[j] = psihat_at(w.qd, a + (ct[(size_t)n + j] + 1.0) * 0.5 * (b - a));
        double csum = 0.0;
        for (int k = 0; k < n; k++) {
            double s = 0.0;
            for (int j = 0; j < n; j++) s += f[j] * ct[(size_t)k * n + j];
            c[k] = 2.0 * s / n;
            csum += std::fabs(c[k]);
        }
        c[0] *= 0.5;
        const double tail = std::max(std::fabs(c[n - 1]), std::max(std::fabs(c[n - 2]), std::fabs(c[n - 3])));
        if (!(tail <= 1e-14 * csum)) continue;
        double err = 0.0, scale = 1.0;
        for (int t = 0; t <= 16; t++) {
            const double zp = a + (b - a) * t / 16.0;
            const double ref = psihat_at(w.qd, zp);
            scale = std::max(scale, std::fabs(ref));
            err = std::max(err, std::fabs(clenshaw(c, n, (2.0 * zp - (a + b)) / (b - a)) - ref));
        }
        if (err <= 1e-14 * scale) {
            w.cheb_n = n;
            w.cheb_a = a;
            w.cheb_b = b;
            for (int k = 0; k < n; k++) w.cheb[k] = c[k];
        }
        return;
    }
}

WArgs make_wargs(const WPlan &p)
{
    WArgs w{};
    w.N1 = p.N1;
    w.N2 = p.N2;
    w.nf1 = p.nf1;
    w.nf2 = p.nf2;
    w.nf3 = p.nf3;
    w.pix_x = p.pix_x;
    w.pix_y = p.pix_y;
    w.o_center = p.o_center;
    w.i_center = p.i_center;
    w.gamma = p.gamma;
    w.h = p.h;
    w.qd = make_quad(p.conv);
    fit_psihat(w);
    if (!(p.nf2 > 1 && fft_col_split(p.nf2, w.cna, w.cnb))) w.cna = w.cnb = 0;
    w.skip_dead = fft_spectrum_pruned<float>(p.nf2) ? 1 : 0;
    return w;
}

Quad make_quad(const ConvOpts &o)
{
    Quad q{};
    q.nq = kernel_quadrature(o, q.xq, q.gq);
    return q;
}

__global__ __launch_bounds__(256) void k_fseries_grid(Quad qd, int nf, double *phi)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int N = nf / 2 + 1;
    if (k >= N) return;
    double s = 0.0;
    for (int i = 0; i < qd.nq; i++) s += qd.gq[i] * cos((qd.xq[i] - N + 1) / (double)(N - 1) * kPi * k);
    phi[k] = 2.0 * s;
}

int launch_fseries_grid(const Quad &q, int nf, double *phi, hipStream_t s)
{
    const int N = nf / 2 + 1;
    hipLaunchKernelGGL(k_fseries_grid, dim3((N + 255) / 256), dim3(256), 0, s, q, nf, phi);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// k_generation (src/astro/precomp.cu:215-225) + pre_stage_2 on z
// (cuft.cu:102-105) + fourier_series_appro at z' (deconv.cu:49-65)
__global__ __launch_bounds__(256) void k_quarter(WArgs w, double *zq, double *phi3)
{
    const int W = w.N1 / 2 + 1, H = w.N2 / 2 + 1;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= W * H) return;
    const int r = q / W, c = q - r * W;
    const double l = r * w.pix_x, m = c * w.pix_y;
    const double t = 1.0 - l * l - m * m;
    // n-1, written without cancellation for t >= 0
    const double z = t < 0 ? sqrt(-t) - 1.0 : -(l * l + m * m) / (sqrt(t) + 1.0);
    const double zp = (z - w.o_center) * w.gamma * w.h;
    double s = 0.0;
    if (w.cheb_n > 0) {
        // Clenshaw recurrence of the host-fitted series
        const double x = (2.0 * zp - (w.cheb_a + w.cheb_b)) / (w.cheb_b - w.cheb_a);
        double b1 = 0.0, b2 = 0.0;
        for (int k = w.cheb_n - 1; k >= 1; k--) {
            const double t2 = 2.0 * x * b1 - b2 + w.cheb[k];
            b2 = b1;
            b1 = t2;
        }
        s = x * b1 - b2 + w.cheb[0];
    } else {
        for (int i = 0; i < w.qd.nq; i++) s += w.qd.gq[i] * 2.0 * cos(w.qd.xq[i] * zp);
    }
    zq[q] = zp;
    phi3[q] = s;
}

int launch_quarter_table(const WArgs &w, double *zq, double *phi3, hipStream_t s)
{
    int nq = (w.N1 / 2 + 1) * (w.N2 / 2 + 1);
    hipLaunchKernelGGL(k_quarter, dim3((nq + 255) / 256), dim3(256), 0, s, w, zq, phi3);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
template <typename TV> __device__ __forceinline__ TV from_d2(double2 v);
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 v) { return v; }

// Spectrum row holding column-axis frequency k, and the inverse map: the
// column FFT leaves a split column (na > 0) with frequency ka + na kb at row
// ka nb + kb.
__device__ __forceinline__ int spec_row(int k, int na, int nb) { return na ? (k % na) * nb + k / na : k; }
__device__ __forceinline__ int spec_freq(int r, int na, int nb) { return na ? r / nb + na * (r % nb) : r; }

// live column cc in [0, N) of the fast axis -> grid column w1 and image
// column: the first N - N/2 live columns are w1 = cc (image col cc + N/2),
// the rest sit at the top of the axis (image col cc - (N - N/2))
__device__ __forceinline__ void live_col(int cc, int N, int nf, int &w1, int &col)
{
    const int A = N - N / 2;
    if (cc < A) {
        w1 = cc;
        col = cc + N / 2;
    } else {
        w1 = cc - A + nf - N / 2;
        col = cc - A;
    }
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// n-1 and n at image offsets (l, m); n is sqrt(t) (NaN beyond the horizon,
// as div_n_lm); the w-correction phase uses sqrt(|t|)-1 as w_term_deconv.
__device__ __forceinline__ void pixel_n(double l, double m, double &nm1_phase, double &n)
{
    const double r2 = l * l + m * m;
    const double t = 1.0 - r2;
    n = sqrt(t);
    nm1_phase = t < 0 ? sqrt(-t) - 1.0 : -r2 / (n + 1.0);
}

template <typename T, typename TV>
__global__ __launch_bounds__(256) void k_ms2dirty_epi(const cx_t<T> *__restrict__ fw,
                                                      const double *__restrict__ zq,
                                                      const double *__restrict__ phi3,
                                                      const double *__restrict__ phi1,
                                                      const double *__restrict__ phi2,
                                                      TV *__restrict__ dirty, WArgs w)
{
    const int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (idx >= (int64_t)w.N1 * w.N2) return;
    const int row = (int)(idx / w.N1), col = (int)(idx - (int64_t)row * w.N1);
    const int dc = col - w.N1 / 2, dr = row - w.N2 / 2;
    const int w1 = dc >= 0 ? dc : w.nf1 + dc;
    const int w2 = dr >= 0 ? dr : w.nf2 + dr;
    const int q = abs(dc) + abs(dr) * (w.N1 / 2 + 1);
    const size_t plane = (size_t)w.nf1 * w.nf2;
    const cx_t<T> *src = fw + (size_t)spec_row(w2, w.cna, w.cnb) * w.nf1 + w1;
    const double zp = zq[q];
    // e^{i z' (p - P/2)} over the slab's planes p, FP64 recurrence
    const int zn = w.zn > 0 ? w.zn : w.nf3;
    double sv, cv;
    sincos(zp, &sv, &cv);
    const double2 step = make_double2(cv, sv);
    sincos(zp * (w.zb - w.nf3 / 2), &sv, &cv);
    double2 ph = make_double2(cv, sv);
    double2 acc = make_double2(0.0, 0.0);
    auto add_plane = [&](cx_t<T> v) {
        const double2 f = to_d2(v);
        acc.x += f.x * ph.x - f.y * ph.y;
        acc.y += f.x * ph.y + f.y * ph.x;
        ph = zmul(ph, step);
    };
    // the planes in batches of eight loads in flight, then their sums in
    // plane order (a load waited for per plane leaves the stage bound by
    // latency, not bandwidth)
    int p = 0;
    for (; p + 8 <= zn; p += 8) {
        cx_t<T> fb[8];
#pragma unroll
        for (int k = 0; k < 8; k++) fb[k] = src[plane * (p + k)];
#pragma unroll
        for (int k = 0; k < 8; k++) add_plane(fb[k]);
    }
    for (; p < zn; p++) add_plane(src[plane * p]);
    const double l = dr * w.pix_x, m = dc * w.pix_y;
    double nm1, n;
    pixel_n(l, m, nm1, n);
    sincos((nm1 - w.o_center) * w.i_center, &sv, &cv);
    const double scale = 1.0 / (phi1[abs(dc)] * phi2[abs(dr)] * phi3[q] * n);
    acc = zmul(acc, make_double2(cv, sv));
    double2 r = make_double2(acc.x * scale, acc.y * scale);
    if (w.accumulate) {
        const double2 prev = to_d2(dirty[idx]);
        r.x += prev.x;
        r.y += prev.y;
    }
    dirty[idx] = from_d2<TV>(r);
}

template <typename T, typename TV>
int launch_ms2dirty_epilogue(const cx_t<T> *fw, const double *zq, const double *phi3,
                             const double *phi1, const double *phi2, TV *dirty, const WArgs &w,
                             hipStream_t s)
{
    int64_t n = (int64_t)w.N1 * w.N2;
    hipLaunchKernelGGL((k_ms2dirty_epi<T, TV>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                       fw, zq, phi3, phi1, phi2, dirty, w);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// grid cell -> image index along one axis (inverse of w = (k - N/2) mod nf),
// -1 when the cell lies outside the image footprint
__device__ __forceinline__ int cell_to_img(int wv, int N, int nf)
{
    if (wv < N - N / 2) return wv + N / 2;
    if (wv >= nf - N / 2) return wv - nf + N / 2;
    return -1;
}

template <typename T, typename TV>
__global__ __launch_bounds__(256) void k_dirty2ms_pro(const TV *__restrict__ dirty,
                                                      const double *__restrict__ zq,
                                                      const double *__restrict__ phi3,
                                                      const double *__restrict__ phi1,
                                                      const double *__restrict__ phi2,
                                                      cx_t<T> *__restrict__ fw, WArgs w)
{
    const int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x;
    const size_t plane = (size_t)w.nf1 * w.nf2;
    if (idx >= (int64_t)w.nf2 * w.N1) return;
    const int r = (int)(idx / w.N1), cc = (int)(idx - (int64_t)r * w.N1);
    int w1, col;
    live_col(cc, w.N1, w.nf1, w1, col);
    const int row = cell_to_img(spec_freq(r, w.cna, w.cnb), w.N2, w.nf2);
    cx_t<T> *dst = fw + (size_t)r * w.nf1 + w1;
    const int zn = w.zn > 0 ? w.zn : w.nf3;
    if (row < 0) {
        // a frequency outside the image rows: zero, unless the column FFT
        // treats it as zero without reading it
        if (!w.skip_dead)
            for (int p = 0; p < zn; p++) dst[plane * p] = mkc((T)0, (T)0);
        return;
    }
    const int dc = col - w.N1 / 2, dr = row - w.N2 / 2;
    const int q = abs(dc) + abs(dr) * (w.N1 / 2 + 1);
    const double l = dr * w.pix_x, m = dc * w.pix_y;
    double nm1, n;
    pixel_n(l, m, nm1, n);
    double2 f = to_d2(dirty[(size_t)row * w.N1 + col]);
    double sv, cv;
    sincos(-(nm1 - w.o_center) * w.i_center, &sv, &cv);
    const double scale = 1.0 / (n * phi3[q] * phi1[abs(dc)] * phi2[abs(dr)]);
    f = zmul(f, make_double2(cv * scale, sv * scale));
    const double zp = zq[q];
    sincos(-zp, &sv, &cv);
    const double2 step = make_double2(cv, sv);
    sincos(-zp * (w.zb - w.nf3 / 2), &sv, &cv);
    double2 ph = make_double2(cv, sv);
    for (int p = 0; p < zn; p++) {
        const double2 v = zmul(f, ph);
        store_nt(dst + plane * p, mkc((T)v.x, (T)v.y));
        ph = zmul(ph, step);
    }
}

template <typename T, typename TV>
int launch_dirty2ms_prologue(const TV *dirty, const double *zq, const double *phi3,
                             const double *phi1, const double *phi2, cx_t<T> *fw, const WArgs &w,
                             hipStream_t s)
{
    int64_t n = (int64_t)w.nf2 * w.N1;
    hipLaunchKernelGGL((k_dirty2ms_pro<T, TV>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                       dirty, zq, phi3, phi1, phi2, fw, w);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// 2-D NUFFT deconvolution, deconv_2d (src/fourier/deconv.cu:250-294), CMCL
// mode order.  Type 2 writes every cell of the live columns (zero outside
// the modes); the inverse row FFT never reads the other columns.
template <typename T, typename TV>
__global__ __launch_bounds__(256) void k_deconv2d_t1(const cx_t<T> *fw, TV *fk, const double *phi1,
                                                     const double *phi2, int N1, int N2, int nf1,
                                                     int nf2, int na, int nb)
{
    const int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (idx >= (int64_t)N1 * N2) return;
    const int k2 = (int)(idx / N1), k1 = (int)(idx - (int64_t)k2 * N1);
    const int d1 = k1 - N1 / 2, d2 = k2 - N2 / 2;
    const int w1 = d1 >= 0 ? d1 : nf1 + d1, w2 = d2 >= 0 ? d2 : nf2 + d2;
    const double2 f = to_d2(fw[(size_t)spec_row(w2, na, nb) * nf1 + w1]);
    const double s = 1.0 / (phi1[abs(d1)] * phi2[abs(d2)]);
    fk[idx] = from_d2<TV>(make_double2(f.x * s, f.y * s));
}

template <typename T, typename TV>
__global__ __launch_bounds__(256) void k_deconv2d_t2(cx_t<T> *fw, const TV *fk, const double *phi1,
                                                     const double *phi2, int N1, int N2, int nf1,
                                                     int nf2, int na, int nb)
{
    const int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (idx >= (int64_t)nf2 * N1) return;
    const int r = (int)(idx / N1), cc = (int)(idx - (int64_t)r * N1);
    int w1, k1;
    live_col(cc, N1, nf1, w1, k1);
    const int k2 = cell_to_img(spec_freq(r, na, nb), N2, nf2);
    const size_t cell = (size_t)r * nf1 + w1;
    if (k2 < 0) {
        fw[cell] = mkc((T)0, (T)0);
        return;
    }
    const int d1 = k1 - N1 / 2, d2 = k2 - N2 / 2;
    const double2 f = to_d2(fk[(size_t)k2 * N1 + k1]);
    const double s = 1.0 / (phi1[abs(d1)] * phi2[abs(d2)]);
    fw[cell] = mkc((T)(f.x * s), (T)(f.y * s));
}

template <typename T, typename TV>
int launch_deconv2d(cx_t<T> *fw, TV *fk, const double *phi1, const double *phi2, int N1, int N2,
                    int nf1, int nf2, int type, hipStream_t s)
{
    int na = 0, nb = 0;
    if (!(nf2 > 1 && fft_col_split(nf2, na, nb))) na = nb = 0;
    if (type == 1) {
        int64_t n = (int64_t)N1 * N2;
        hipLaunchKernelGGL((k_deconv2d_t1<T, TV>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                           fw, fk, phi1, phi2, N1, N2, nf1, nf2, na, nb);
    } else {
        int64_t n = (int64_t)nf2 * N1;
        hipLaunchKernelGGL((k_deconv2d_t2<T, TV>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                           fw, fk, phi1, phi2, N1, N2, nf1, nf2, na, nb);
    }
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template int launch_ms2dirty_epilogue<float, float2>(const float2 *, const double *, const double *, const double *, const double *, float2 *, const WArgs &, hipStream_t);
template int launch_ms2dirty_epilogue<double, double2>(const double2 *, const double *, const double *, const double *, const double *, double2 *, const WArgs &, hipStream_t);
template int launch_dirty2ms_prologue<float, float2>(const float2 *, const double *, const double *, const double *, const double *, float2 *, const WArgs &, hipStream_t);
template int launch_dirty2ms_prologue<double, double2>(const double2 *, const double *, const double *, const double *, const double *, double2 *, const WArgs &, hipStream_t);
template int launch_deconv2d<float, float2>(float2 *, float2 *, const double *, const double *, int, int, int, int, int, hipStream_t);
template int launch_deconv2d<double, double2>(double2 *, double2 *, const double *, const double *, int, int, int, int, int, hipStream_t);

}  // namespace cg
