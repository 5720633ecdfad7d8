/*
 * cugrid_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the improved W-stacking imaging pair ms2dirty /
 * dirty2ms of RapidsAtHKUST/cuGridder, plus an FP64 direct-DFT oracle and the
 * 2-D NUFFT type 1/2 used by the reference's nufft_2d tests.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker / CPU baseline.  The product path
 * (libcurafft.so) never links or calls it.
 *
 * Every function names the reference file:line it restates.  Parity pins
 * (what each function is checked against) are listed in oracle/README.md.
 *
 * Conventions (reference): image dirty[row*N1 + col], row in [0,N2=nx) <-> l
 * (u, pixelsize_x), col in [0,N1=ny) <-> m (v, pixelsize_y); the grid is
 * fw[p][b2][b1] with the fast axis b1 (nf1) driven by v and b2 (nf2) by u
 * (src/astro/cugridder.cu:191).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define ORC_PI 3.14159265358979323846264338327950288
#define ORC_C 299792458.0 /* include/utils.h:42 SPEEDOFLIGHT */
#define ORC_MAX_KW 16     /* include/curafft_plan.h:14 MAX_KERNEL_WIDTH */

/* ------------------------------------------------------------------------ */
/* Planner pieces                                                            */
/* ------------------------------------------------------------------------ */

/* src/common_utils.cu:76-101 next235beven: smallest even n' >= n whose only
 * prime factors are 2,3,5 and which is a multiple of b. */
int orc_next235beven(int n, int b)
{
    if (n <= 2)
        return 2;
    if (n % 2 == 1)
        n += 1;
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

/* Gauss-Legendre nodes/weights on [-1,1], ascending, by Newton iteration on
 * P_n.  The reference computes the same rule with the Glaser-Liu-Rokhlin
 * method (contrib/legendre_rule_fast.c:15-93) and renormalises the weights to
 * sum to 2 (:82-90); the renormalisation is kept here. */
void orc_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(ORC_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p0 = 1.0; p1 = z; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) {
                /* one more derivative evaluation at the converged root */
                p0 = 1.0; p1 = z;
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
    if (n % 2 == 1)
        x[n / 2] = 0.0;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += w[i];
    for (int i = 0; i < n; i++) w[i] = 2.0 * w[i] / s;
}

/* src/fourier/conv_interp_invoker.cu:49-119 setup_conv_opts (kerevalmeth=0).
 * single!=0 selects the float build's EPSILON (include/utils.h:34-38).
 * Returns ier (0 ok, 1 eps clamped, 2 bad sigma). */
int orc_setup_conv(double eps, double sigma, int single, int *kw_out,
                   double *beta_out, double *c_out)
{
    if (sigma != 2.0 && sigma <= 1.0)
        return 2;
    int ier = 0;
    double epsilon_min = single ? 6e-08 : 1.1e-16;
    if (eps < epsilon_min) {
        eps = epsilon_min;
        ier = 1;
    }
    int kw = (int)ceil(-log10(eps / 10.0));
    if (sigma != 2.0)
        kw = (int)ceil(-log(eps) / (ORC_PI * sqrt(1.0 - 1.0 / sigma)));
    if (kw < 2) kw = 2;
    if (kw > ORC_MAX_KW) kw = ORC_MAX_KW;
    double betaoverns = 2.30;
    if (kw == 2) betaoverns = 2.20;
    if (kw == 3) betaoverns = 2.26;
    if (kw == 4) betaoverns = 2.38;
    if (sigma != 2.0)
        betaoverns = 0.97 * ORC_PI * (1.0 - 1.0 / (2.0 * sigma));
    *kw_out = kw;
    *beta_out = betaoverns * kw;
    *c_out = 4.0 / (double)(kw * kw);
    return ier;
}

/* src/fourier/conv_interp_invoker.cu:30-47 get_num_cells */
int orc_num_cells(int ms, double sigma, int kw)
{
    int nf = (int)(sigma * ms);
    if (nf < 2 * kw) nf = 2 * kw;
    return orc_next235beven(nf, 1);
}

/* src/utils.cu:670-709 set_nhg_w (type-3 w-plane count nf, spacing h,
 * rescale gamma).  S = output (n-1) half width, X = input w half width. */
void orc_set_nhg_w(double S, double X, int kw, double sigma, int *nf,
                   double *h, double *gam)
{
    int nss = kw + 1;
    double Xsafe = X, Ssafe = S;
    if (X == 0.0) {
        if (S == 0.0) { Xsafe = 1.0; Ssafe = 1.0; }
        else Xsafe = fmax(Xsafe, 1.0 / S);
    } else {
        Ssafe = fmax(Ssafe, 1.0 / X);
    }
    double nfd = 2.0 * sigma * Ssafe * Xsafe / ORC_PI + nss;
    if (!isfinite(nfd)) nfd = 0.0;
    int n = (int)nfd;
    n = n % 2 + n;
    if (n < 2 * kw) n = 2 * kw;
    *nf = n;
    *h = 2.0 * ORC_PI / n;
    *gam = (double)n / (2.0 * sigma * Ssafe);
}

/* Quadrature g_i = (kw/2) w_i psi(x_i) on the first p = floor(2+1.5 kw)
 * nodes (negative half) of the 2p-point rule, psi unnormalised
 * (src/fourier/deconv.cu:80-94,126-141,173-190). */
static int orc_quad(int kw, double beta, double c, double *xq, double *gq)
{
    double alpha = kw / 2.0;
    int p = (int)(2 + 3.0 * alpha);
    double x[2 * 128], w[2 * 128];
    orc_gauss_legendre(2 * p, x, w);
    for (int n = 0; n < p; n++) {
        double xn = x[n] * alpha;
        double phi = 0.0;
        if (fabs(xn) <= kw / 2.0)
            phi = exp(beta * sqrt(1.0 - c * xn * xn));
        xq[n] = xn;
        gq[n] = alpha * w[n] * phi;
    }
    return p;
}

/* src/fourier/deconv.cu:12-26 (3-arg fourier_series_appro) via the invoker
 * 67-110 with N = nf/2+1: out[k] = 2 sum_i g_i cos(pi k (x_i-N+1)/(N-1)),
 * i.e. (-1)^k psihat(2 pi k/nf). */
void orc_fseries_grid(int kw, double beta, double c, int nf, double *out)
{
    double xq[128], gq[128];
    int p = orc_quad(kw, beta, c, xq, gq);
    int N = nf / 2 + 1;
    for (int k = 0; k < N; k++) {
        double s = 0.0;
        for (int i = 0; i < p; i++)
            s += gq[i] * cos((xq[i] - N + 1) / (double)(N - 1) * ORC_PI * k);
        out[k] = 2.0 * s;
    }
}

/* src/fourier/deconv.cu:49-65 (k-arg fourier_series_appro) via the invoker
 * 159-206: out[q] = sum_i 2 g_i cos(x_i k_q)  (psihat at arbitrary k). */
void orc_fseries_at(int kw, double beta, double c, const double *k, int n,
                    double *out)
{
    double xq[128], gq[128];
    int p = orc_quad(kw, beta, c, xq, gq);
    for (int q = 0; q < n; q++) {
        double s = 0.0;
        for (int i = 0; i < p; i++)
            s += gq[i] * 2.0 * cos(xq[i] * k[q]);
        out[q] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix Stockham FFT (unnormalised, e^{sign 2 pi i jk/n})             */
/* Replaces cuFFT C2C/Z2Z (src/fourier/cuft.cu:20-44,192-212).               */
/* ------------------------------------------------------------------------ */

static int orc_factor(int n, int *rad)
{
    int nr = 0;
    while (n % 4 == 0) { rad[nr++] = 4; n /= 4; }
    while (n % 2 == 0) { rad[nr++] = 2; n /= 2; }
    while (n % 3 == 0) { rad[nr++] = 3; n /= 3; }
    while (n % 5 == 0) { rad[nr++] = 5; n /= 5; }
    for (int f = 7; n > 1; f += 2)
        while (n % f == 0) { rad[nr++] = f; n /= f; }
    return nr;
}

/* in-place 1-D FFT of x[0..n-1] (contiguous); work has n entries */
static void orc_fft_contig(int n, cplx *x, cplx *work, int sign,
                           const cplx *tw /* exp(sign 2 pi i t/n) */)
{
    int rad[64];
    int nr = orc_factor(n, rad);
    cplx *src = x, *dst = work;
    int p = 1;
    for (int s = 0; s < nr; s++) {
        int R = rad[s];
        int T = n / R;
        int twstep = n / (p * R);
        for (int i = 0; i < T; i++) {
            int k = i % p;
            cplx u[64], v[64];
            for (int r = 0; r < R; r++) {
                long t = (long)r * k * twstep % n;
                u[r] = src[i + r * T] * tw[t];
            }
            for (int q = 0; q < R; q++) {
                cplx acc = 0;
                for (int r = 0; r < R; r++)
                    acc += u[r] * tw[(long)r * q % R * (n / R)];
                v[q] = acc;
            }
            int j = (i - k) * R + k;
            for (int r = 0; r < R; r++)
                dst[j + r * p] = v[r];
        }
        p *= R;
        cplx *t = src; src = dst; dst = t;
    }
    if (src != x)
        memcpy(x, src, sizeof(cplx) * n);
}

static cplx *orc_twiddles(int n, int sign)
{
    cplx *tw = (cplx *)malloc(sizeof(cplx) * n);
    for (int t = 0; t < n; t++) {
        double a = 2.0 * ORC_PI * (double)t / n;
        tw[t] = cos(a) + I * (sign * sin(a));
    }
    return tw;
}

/* 1-D FFT of n points at stride 'stride' (exported for tests) */
void orc_fft1d(int n, double *xri, int sign)
{
    cplx *x = (cplx *)xri;
    cplx *work = (cplx *)malloc(sizeof(cplx) * n);
    cplx *tw = orc_twiddles(n, sign);
    orc_fft_contig(n, x, work, sign, tw);
    free(work);
    free(tw);
}

/* 2-D FFT of 'batch' planes [batch][n2][n1] (row-major, fast n1). */
void orc_fft2d(int batch, int n2, int n1, double *data, int sign)
{
    cplx *d = (cplx *)data;
    cplx *tw1 = orc_twiddles(n1, sign);
    cplx *tw2 = orc_twiddles(n2, sign);
    for (int b = 0; b < batch; b++) {
        cplx *pl = d + (size_t)b * n1 * n2;
#pragma omp parallel
        {
            cplx *work = (cplx *)malloc(sizeof(cplx) * (n1 > n2 ? n1 : n2) * 9);
#pragma omp for schedule(static)
            for (int r = 0; r < n2; r++)
                orc_fft_contig(n1, pl + (size_t)r * n1, work, sign, tw1);
            /* columns, 8 at a time through a contiguous buffer */
            cplx *col = work + (n1 > n2 ? n1 : n2);
#pragma omp for schedule(static)
            for (int c0 = 0; c0 < n1; c0 += 8) {
                int nc = n1 - c0 < 8 ? n1 - c0 : 8;
                for (int cc = 0; cc < nc; cc++) {
                    cplx *cb = col + (size_t)cc * n2;
                    for (int r = 0; r < n2; r++)
                        cb[r] = pl[(size_t)r * n1 + c0 + cc];
                }
                for (int cc = 0; cc < nc; cc++)
                    orc_fft_contig(n2, col + (size_t)cc * n2, work, sign, tw2);
                for (int cc = 0; cc < nc; cc++) {
                    cplx *cb = col + (size_t)cc * n2;
                    for (int r = 0; r < n2; r++)
                        pl[(size_t)r * n1 + c0 + cc] = cb[r];
                }
            }
            free(work);
        }
    }
    free(tw1);
    free(tw2);
}

/* ------------------------------------------------------------------------ */
/* ES kernel, rescale, spreading and interpolation                           */
/* ------------------------------------------------------------------------ */

/* include/conv.h:24 SHIFT_RESCALE: p=0 folds x into [-pi,pi) first. */
static inline double orc_shift_rescale(double x, int N, int pirange)
{
    const double twopi = 6.28318530717958648;
    double y = x;
    if (!pirange) {
        double t = x - floor(x / twopi) * twopi;
        y = t - (t >= ORC_PI ? twopi : 0.0);
    }
    return (y * 0.159154943091895336 + 0.5) * N;
}

/* src/fourier/conv.cu:41-48 val_kernel_vec: ker[i] = psi(x+i), 0 outside */
static inline void orc_kervec(double *ker, double x, int kw, double c,
                              double beta)
{
    for (int i = 0; i < kw; i++) {
        double d = x + i;
        ker[i] = (fabs(d) >= kw / 2.0) ? 0.0 : exp(beta * sqrt(1.0 - c * d * d));
    }
}

static inline int orc_wrap(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

/* Periodic 3-D spreading, src/fourier/conv.cu:166-244 (conv_3d_nputsdriven;
 * the output-driven method-6 kernel conv.cu:522-627 computes the same sum).
 * x->nf1 (fast), y->nf2, z->nf3.  fw must be zeroed by the caller. */
void orc_spread3d(int M, const double *x, const double *y, const double *z,
                  const double *c /* complex */, int nf1, int nf2, int nf3,
                  int kw, double beta, double es_c, int pirange, double *fwri)
{
    cplx *fw = (cplx *)fwri;
    const cplx *cc = (const cplx *)c;
    for (int j = 0; j < M; j++) {
        double g1 = orc_shift_rescale(x[j], nf1, pirange);
        double g2 = orc_shift_rescale(y[j], nf2, pirange);
        double g3 = orc_shift_rescale(z[j], nf3, pirange);
        int x0 = (int)ceil(g1 - kw / 2.0);
        int y0 = (int)ceil(g2 - kw / 2.0);
        int z0 = (int)ceil(g3 - kw / 2.0);
        double k1[ORC_MAX_KW], k2[ORC_MAX_KW], k3[ORC_MAX_KW];
        orc_kervec(k1, x0 - g1, kw, es_c, beta);
        orc_kervec(k2, y0 - g2, kw, es_c, beta);
        orc_kervec(k3, z0 - g3, kw, es_c, beta);
        for (int dz = 0; dz < kw; dz++) {
            int iz = orc_wrap(z0 + dz, nf3);
            for (int dy = 0; dy < kw; dy++) {
                int iy = orc_wrap(y0 + dy, nf2);
                double kyz = k2[dy] * k3[dz];
                cplx *row = fw + ((size_t)iz * nf2 + iy) * nf1;
                for (int dx = 0; dx < kw; dx++) {
                    int ix = orc_wrap(x0 + dx, nf1);
                    row[ix] += cc[j] * (k1[dx] * kyz);
                }
            }
        }
    }
}

/* Periodic 3-D interpolation, src/fourier/interp.cu:132-210 */
void orc_interp3d(int M, const double *x, const double *y, const double *z,
                  double *cout /* complex */, int nf1, int nf2, int nf3,
                  int kw, double beta, double es_c, int pirange,
                  const double *fwri)
{
    const cplx *fw = (const cplx *)fwri;
    cplx *co = (cplx *)cout;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < M; j++) {
        double g1 = orc_shift_rescale(x[j], nf1, pirange);
        double g2 = orc_shift_rescale(y[j], nf2, pirange);
        double g3 = orc_shift_rescale(z[j], nf3, pirange);
        int x0 = (int)ceil(g1 - kw / 2.0);
        int y0 = (int)ceil(g2 - kw / 2.0);
        int z0 = (int)ceil(g3 - kw / 2.0);
        double k1[ORC_MAX_KW], k2[ORC_MAX_KW], k3[ORC_MAX_KW];
        orc_kervec(k1, x0 - g1, kw, es_c, beta);
        orc_kervec(k2, y0 - g2, kw, es_c, beta);
        orc_kervec(k3, z0 - g3, kw, es_c, beta);
        cplx acc = 0;
        for (int dz = 0; dz < kw; dz++) {
            int iz = orc_wrap(z0 + dz, nf3);
            for (int dy = 0; dy < kw; dy++) {
                int iy = orc_wrap(y0 + dy, nf2);
                double kyz = k2[dy] * k3[dz];
                const cplx *row = fw + ((size_t)iz * nf2 + iy) * nf1;
                for (int dx = 0; dx < kw; dx++) {
                    int ix = orc_wrap(x0 + dx, nf1);
                    acc += row[ix] * (k1[dx] * kyz);
                }
            }
        }
        co[j] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* W-stacking plan (src/astro/cugridder.cu:29-210)                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nrow, N1, N2; /* N1 = nydirty (fast image axis), N2 = nxdirty */
    int nf1, nf2, nf3, kw, ier;
    double beta, es_c, sigma, fov, freq, f_over_c;
    double pix_x, pix_y;
    double i_center, i_half_width, o_center, o_half_width, gamma, h;
    int sign;
} orc_plan;

/* Pixel size along ny (radians) for the plans that follow, when > 0: the
 * reference derives it as fov/ny (cugridder.cu:48-52); NIFTY's rad_pix_y
 * sets it separately.  0 restores the reference's derivation. */
static double orc_pix_y_override = 0.0;
void orc_set_pix_y(double pix_y) { orc_pix_y_override = pix_y > 0 ? pix_y : 0.0; }

/* Builds the plan exactly as gridder_setting/setup_gridder_plan do,
 * with the w range taken from the scaled w (w*f/c*sign*2pi). */
int orc_make_plan(int nrow, int nxdirty, int nydirty, double fov, double freq,
                  const double *uvw, double eps, double sigma, int sign,
                  int single, orc_plan *P)
{
    memset(P, 0, sizeof(*P));
    P->nrow = nrow;
    P->N1 = nydirty;
    P->N2 = nxdirty;
    P->sign = sign;
    P->fov = fov;
    P->freq = freq;
    P->sigma = sigma;
    P->f_over_c = freq / ORC_C;
    P->ier = orc_setup_conv(eps, sigma, single, &P->kw, &P->beta, &P->es_c);
    if (P->ier == 2)
        return 2;
    P->pix_x = fov / 180.0 * ORC_PI / (double)P->N2;
    P->pix_y = fov / 180.0 * ORC_PI / (double)P->N1;
    if (orc_pix_y_override > 0) P->pix_y = orc_pix_y_override;
    double l_min = -0.5 * P->pix_x * P->N2;
    double m_min = -0.5 * P->pix_y * P->N1;
    double temp = 1.0 - l_min * l_min - m_min * m_min;
    double n_lm = temp <= 0 ? 0 : sqrt(temp);
    double wmax = -INFINITY, wmin = INFINITY;
    for (int j = 0; j < nrow; j++) {
        double w = uvw[3 * j + 2] * P->f_over_c * sign * 2.0 * ORC_PI;
        if (w > wmax) wmax = w;
        if (w < wmin) wmin = w;
    }
    if (nrow == 0) { wmax = 0; wmin = 0; }
    P->i_center = (wmax + wmin) / 2.0;
    P->i_half_width = (wmax - wmin) / 2.0;
    double o_min = n_lm - 1.0;
    P->o_center = o_min / 2.0;
    P->o_half_width = fabs(o_min / 2.0);
    orc_set_nhg_w(P->o_half_width, P->i_half_width, P->kw, sigma, &P->nf3,
                  &P->h, &P->gamma);
    P->nf1 = orc_num_cells(P->N1, sigma, P->kw);
    P->nf2 = orc_num_cells(P->N2, sigma, P->kw);
    return P->ier;
}

/* z = n-1 on the (N2/2+1) x (N1/2+1) quarter image,
 * src/astro/precomp.cu:215-225 k_generation (sqrt(-t)-1 beyond horizon). */
static void orc_k_generation(const orc_plan *P, double *z)
{
    int W = P->N1 / 2 + 1;
    int H = P->N2 / 2 + 1;
    for (int r = 0; r < H; r++)
        for (int c = 0; c < W; c++) {
            double t = 1.0 - pow(r * P->pix_x, 2) - pow(c * P->pix_y, 2);
            z[r * W + c] = t < 0 ? sqrt(-t) - 1 : sqrt(t) - 1;
        }
}

/* scaled coordinates: x (grid axis 1) = v*2pi f/c*pix_y, y (axis 2) =
 * u*2pi f/c*pix_x, z = (w*2pi f/c*sign - i_center)/gamma; also returns the
 * scaled w (before pre_stage_2). precomp.cu:48-80, cugridder.cu:191,201-202,
 * cuft.cu:82-103 */
static void orc_coords(const orc_plan *P, const double *uvw, double *x,
                       double *y, double *z, double *wt)
{
    for (int j = 0; j < P->nrow; j++) {
        double u = uvw[3 * j] * P->f_over_c * 2.0 * ORC_PI;
        double v = uvw[3 * j + 1] * P->f_over_c * 2.0 * ORC_PI;
        double w = uvw[3 * j + 2] * P->f_over_c * P->sign * 2.0 * ORC_PI;
        y[j] = u * P->pix_x;
        x[j] = v * P->pix_y;
        wt[j] = w;
        z[j] = (w - P->i_center) / P->gamma;
    }
}

/* ms2dirty restatement: src/astro/ra_exec.cu:95-176 (cura_prestage, iflag=1)
 * then :184-468 (exec_vis2dirty, flow 1, mem_limit=0).  dirty out is
 * [N2][N1] complex.  wgt may be NULL.  Returns ier. */
int orc_ms2dirty(int nrow, int nxdirty, int nydirty, double fov, double freq,
                 const double *uvw, const double *vis, const double *wgt,
                 double *dirty, double eps, double sigma, int sign, int single)
{
    orc_plan P;
    int ier = orc_make_plan(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma,
                            sign, single, &P);
    if (ier == 2)
        return ier;
    int M = nrow, N1 = P.N1, N2 = P.N2;
    int nf1 = P.nf1, nf2 = P.nf2, nf3 = P.nf3;
    double *x = malloc(sizeof(double) * (M + 1));
    double *y = malloc(sizeof(double) * (M + 1));
    double *z = malloc(sizeof(double) * (M + 1));
    double *wt = malloc(sizeof(double) * (M + 1));
    cplx *c = malloc(sizeof(cplx) * (M + 1));
    orc_coords(&P, uvw, x, y, z, wt);
    const cplx *vv = (const cplx *)vis;
    for (int j = 0; j < M; j++) {
        c[j] = vv[j];
        if (wgt) c[j] *= wgt[j];                       /* ra_exec.cu:103-111 */
        if (P.o_center != 0)                            /* cuft.cu:46-80 */
            c[j] *= cexp(I * (P.o_center * wt[j]));
    }
    int W = N1 / 2 + 1, H = N2 / 2 + 1;
    double *zq = malloc(sizeof(double) * W * H);
    orc_k_generation(&P, zq);
    for (int q = 0; q < W * H; q++)                      /* cuft.cu:102-105 */
        zq[q] = (zq[q] - P.o_center) * P.gamma * P.h;
    double *phi1 = malloc(sizeof(double) * (nf1 / 2 + 1));
    double *phi2 = malloc(sizeof(double) * (nf2 / 2 + 1));
    double *phi3 = malloc(sizeof(double) * W * H);
    orc_fseries_grid(P.kw, P.beta, P.es_c, nf1, phi1);
    orc_fseries_grid(P.kw, P.beta, P.es_c, nf2, phi2);
    orc_fseries_at(P.kw, P.beta, P.es_c, zq, W * H, phi3);
    size_t C = (size_t)nf1 * nf2 * nf3;
    cplx *fw = calloc(C, sizeof(cplx));
    orc_spread3d(M, x, y, z, (const double *)c, nf1, nf2, nf3, P.kw, P.beta,
                 P.es_c, 0, (double *)fw);
    orc_fft2d(nf3, nf2, nf1, (double *)fw, +1);         /* cuft.cu:192-212 */
    cplx *out = (cplx *)dirty;
    size_t plane = (size_t)nf1 * nf2;
#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < N1 * N2; idx++) {
        int row = idx / N1, col = idx % N1;
        int w1 = col >= N1 / 2 ? col - N1 / 2 : nf1 + col - N1 / 2;
        int w2 = row >= N2 / 2 ? row - N2 / 2 : nf2 + row - N2 / 2;
        size_t ifw = (size_t)w1 + (size_t)w2 * nf1;
        int iz = abs(col - N1 / 2) + abs(row - N2 / 2) * W;
        /* w_term_dft, cuft.cu:597-645: plane p carries frequency p-nf3/2 */
        cplx acc = 0;
        for (int p = 0; p < nf3; p++)
            acc += fw[ifw + plane * p] * cexp(I * (zq[iz] * (p - nf3 / 2)));
        /* deconv_2d type 1, deconv.cu:250-294 (CMCL) */
        double kv = phi1[abs(col - N1 / 2)] * phi2[abs(row - N2 / 2)];
        acc /= kv;
        /* w_term_deconv flag=+1, deconv.cu:401-432 */
        double l = (row - N2 / 2) * P.pix_x, m = (col - N1 / 2) * P.pix_y;
        double tmp = 1.0 - l * l - m * m;
        double nm1 = tmp < 0 ? sqrt(-tmp) - 1 : sqrt(tmp) - 1;
        double ph = (nm1 - P.o_center) * P.i_center;
        acc = acc * cexp(I * ph) / phi3[iz];
        /* div_n_lm, ra_exec.cu:23-37 */
        acc /= sqrt(tmp);
        out[idx] = acc;
    }
    free(x); free(y); free(z); free(wt); free(c); free(zq);
    free(phi1); free(phi2); free(phi3); free(fw);
    return ier;
}

/* dirty2ms restatement: ra_exec.cu:95-176 (iflag=-1) then :470-680. */
int orc_dirty2ms(int nrow, int nxdirty, int nydirty, double fov, double freq,
                 const double *uvw, double *vis, const double *wgt,
                 const double *dirty, double eps, double sigma, int sign,
                 int single)
{
    orc_plan P;
    int ier = orc_make_plan(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma,
                            sign, single, &P);
    if (ier == 2)
        return ier;
    int M = nrow, N1 = P.N1, N2 = P.N2;
    int nf1 = P.nf1, nf2 = P.nf2, nf3 = P.nf3;
    double *x = malloc(sizeof(double) * (M + 1));
    double *y = malloc(sizeof(double) * (M + 1));
    double *z = malloc(sizeof(double) * (M + 1));
    double *wt = malloc(sizeof(double) * (M + 1));
    orc_coords(&P, uvw, x, y, z, wt);
    int W = N1 / 2 + 1, H = N2 / 2 + 1;
    double *zq = malloc(sizeof(double) * W * H);
    orc_k_generation(&P, zq);
    for (int q = 0; q < W * H; q++)
        zq[q] = (zq[q] - P.o_center) * P.gamma * P.h;
    double *phi1 = malloc(sizeof(double) * (nf1 / 2 + 1));
    double *phi2 = malloc(sizeof(double) * (nf2 / 2 + 1));
    double *phi3 = malloc(sizeof(double) * W * H);
    orc_fseries_grid(P.kw, P.beta, P.es_c, nf1, phi1);
    orc_fseries_grid(P.kw, P.beta, P.es_c, nf2, phi2);
    orc_fseries_at(P.kw, P.beta, P.es_c, zq, W * H, phi3);
    size_t C = (size_t)nf1 * nf2 * nf3;
    size_t plane = (size_t)nf1 * nf2;
    cplx *fw = calloc(C, sizeof(cplx));
    const cplx *in = (const cplx *)dirty;
#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < N1 * N2; idx++) {
        int row = idx / N1, col = idx % N1;
        double l = (row - N2 / 2) * P.pix_x, m = (col - N1 / 2) * P.pix_y;
        double tmp = 1.0 - l * l - m * m;
        cplx f = in[idx] / sqrt(tmp);                    /* div_n_lm */
        double nm1 = tmp < 0 ? sqrt(-tmp) - 1 : sqrt(tmp) - 1;
        int iz = abs(col - N1 / 2) + abs(row - N2 / 2) * W;
        f = f * cexp(I * (-(nm1 - P.o_center) * P.i_center)) / phi3[iz];
        double kv = phi1[abs(col - N1 / 2)] * phi2[abs(row - N2 / 2)];
        f /= kv;                                         /* deconv_2d type 2 */
        int w1 = col >= N1 / 2 ? col - N1 / 2 : nf1 + col - N1 / 2;
        int w2 = row >= N2 / 2 ? row - N2 / 2 : nf2 + row - N2 / 2;
        size_t ifw = (size_t)w1 + (size_t)w2 * nf1;
        /* w_term_idft flag=-1, cuft.cu:647-689 */
        for (int p = 0; p < nf3; p++)
            fw[ifw + plane * p] = f * cexp(I * (-zq[iz] * (p - nf3 / 2)));
    }
    orc_fft2d(nf3, nf2, nf1, (double *)fw, -1);
    cplx *c = (cplx *)vis;
    orc_interp3d(M, x, y, z, (double *)c, nf1, nf2, nf3, P.kw, P.beta, P.es_c,
                 0, (const double *)fw);
    for (int j = 0; j < M; j++) {
        if (wgt) c[j] *= wgt[j];                          /* utils.cu:661-668 */
        double ph = -(z[j] * P.gamma + P.i_center) * P.o_center;
        c[j] *= cexp(I * ph);                             /* ra_exec.cu:55-74 */
    }
    free(x); free(y); free(z); free(wt); free(zq);
    free(phi1); free(phi2); free(phi3); free(fw);
    return ier;
}

/* Stage exports for kernel-level checks: the ms2dirty grid right after
 * spreading (before the plane FFTs) and the dirty2ms grid right after the
 * w-IDFT (before the FFTs), exactly as inside orc_ms2dirty / orc_dirty2ms. */
int orc_ms2dirty_grid(int nrow, int nxdirty, int nydirty, double fov,
                      double freq, const double *uvw, const double *vis,
                      const double *wgt, double eps, double sigma, int sign,
                      int single, double *fwout)
{
    orc_plan P;
    int ier = orc_make_plan(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma,
                            sign, single, &P);
    if (ier == 2) return ier;
    int M = nrow;
    double *x = malloc(sizeof(double) * (M + 1));
    double *y = malloc(sizeof(double) * (M + 1));
    double *z = malloc(sizeof(double) * (M + 1));
    double *wt = malloc(sizeof(double) * (M + 1));
    cplx *c = malloc(sizeof(cplx) * (M + 1));
    orc_coords(&P, uvw, x, y, z, wt);
    const cplx *vv = (const cplx *)vis;
    for (int j = 0; j < M; j++) {
        c[j] = vv[j];
        if (wgt) c[j] *= wgt[j];
        if (P.o_center != 0) c[j] *= cexp(I * (P.o_center * wt[j]));
    }
    memset(fwout, 0, sizeof(cplx) * (size_t)P.nf1 * P.nf2 * P.nf3);
    orc_spread3d(M, x, y, z, (const double *)c, P.nf1, P.nf2, P.nf3, P.kw, P.beta,
                 P.es_c, 0, fwout);
    free(x); free(y); free(z); free(wt); free(c);
    return ier;
}

int orc_dirty2ms_planes(int nrow, int nxdirty, int nydirty, double fov,
                        double freq, const double *uvw, const double *dirty,
                        double eps, double sigma, int sign, int single,
                        double *fwout)
{
    orc_plan P;
    int ier = orc_make_plan(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma,
                            sign, single, &P);
    if (ier == 2) return ier;
    int N1 = P.N1, N2 = P.N2, nf1 = P.nf1, nf2 = P.nf2, nf3 = P.nf3;
    int W = N1 / 2 + 1, H = N2 / 2 + 1;
    double *zq = malloc(sizeof(double) * W * H);
    orc_k_generation(&P, zq);
    for (int q = 0; q < W * H; q++) zq[q] = (zq[q] - P.o_center) * P.gamma * P.h;
    double *phi1 = malloc(sizeof(double) * (nf1 / 2 + 1));
    double *phi2 = malloc(sizeof(double) * (nf2 / 2 + 1));
    double *phi3 = malloc(sizeof(double) * W * H);
    orc_fseries_grid(P.kw, P.beta, P.es_c, nf1, phi1);
    orc_fseries_grid(P.kw, P.beta, P.es_c, nf2, phi2);
    orc_fseries_at(P.kw, P.beta, P.es_c, zq, W * H, phi3);
    size_t plane = (size_t)nf1 * nf2;
    cplx *fw = (cplx *)fwout;
    memset(fw, 0, sizeof(cplx) * plane * nf3);
    const cplx *in = (const cplx *)dirty;
    for (int idx = 0; idx < N1 * N2; idx++) {
        int row = idx / N1, col = idx % N1;
        double l = (row - N2 / 2) * P.pix_x, m = (col - N1 / 2) * P.pix_y;
        double tmp = 1.0 - l * l - m * m;
        cplx f = in[idx] / sqrt(tmp);
        double nm1 = tmp < 0 ? sqrt(-tmp) - 1 : sqrt(tmp) - 1;
        int iz = abs(col - N1 / 2) + abs(row - N2 / 2) * W;
        f = f * cexp(I * (-(nm1 - P.o_center) * P.i_center)) / phi3[iz];
        f /= phi1[abs(col - N1 / 2)] * phi2[abs(row - N2 / 2)];
        int w1 = col >= N1 / 2 ? col - N1 / 2 : nf1 + col - N1 / 2;
        int w2 = row >= N2 / 2 ? row - N2 / 2 : nf2 + row - N2 / 2;
        size_t ifw = (size_t)w1 + (size_t)w2 * nf1;
        for (int p = 0; p < nf3; p++)
            fw[ifw + plane * p] = f * cexp(I * (-zq[iz] * (p - nf3 / 2)));
    }
    free(zq); free(phi1); free(phi2); free(phi3);
    return ier;
}

/* ------------------------------------------------------------------------ */
/* FP64 direct DFT oracle (python/curagridder/test_wst.py:15-32 formula with  */
/* the sign convention of the C-ABI; w_s_degridder_test.cu:249-263 adjoint). */
/* dirty(l,m) = 1/n sum_j wgt_j vis_j exp(2 pi i f/c (u l + v m + s w (n-1))) */
/* ------------------------------------------------------------------------ */

static inline double orc_nm1(double l, double m)
{
    double r2 = l * l + m * m;
    return -r2 / (sqrt(1.0 - r2) + 1.0);
}

/* pixels: list of npix flat indices (row*N1+col); out[npix] complex */
void orc_dft_ms2dirty(int nrow, int nxdirty, int nydirty, double pix_x,
                      double pix_y, double freq, const double *uvw,
                      const double *vis, const double *wgt, int sign,
                      int npix, const int64_t *pixels, double *out)
{
    const cplx *vv = (const cplx *)vis;
    cplx *o = (cplx *)out;
    double foc = 2.0 * ORC_PI * freq / ORC_C;
    int N1 = nydirty, N2 = nxdirty;
#pragma omp parallel for schedule(dynamic, 16)
    for (int q = 0; q < npix; q++) {
        int64_t idx = pixels[q];
        int row = (int)(idx / N1), col = (int)(idx % N1);
        double l = (row - N2 / 2) * pix_x, m = (col - N1 / 2) * pix_y;
        double nm1 = orc_nm1(l, m);
        cplx acc = 0;
        for (int j = 0; j < nrow; j++) {
            double ph = foc * (uvw[3 * j] * l + uvw[3 * j + 1] * m +
                               sign * uvw[3 * j + 2] * nm1);
            cplx cj = vv[j] * (wgt ? wgt[j] : 1.0);
            acc += cj * (cos(ph) + I * sin(ph));
        }
        o[q] = acc / (nm1 + 1.0);
    }
}

/* rows: list of nsel visibility indices; out[nsel] complex */
void orc_dft_dirty2ms(int nrow, int nxdirty, int nydirty, double pix_x,
                      double pix_y, double freq, const double *uvw,
                      const double *dirty, const double *wgt, int sign,
                      int nsel, const int64_t *rows, double *out)
{
    (void)nrow;
    const cplx *d = (const cplx *)dirty;
    cplx *o = (cplx *)out;
    double foc = 2.0 * ORC_PI * freq / ORC_C;
    int N1 = nydirty, N2 = nxdirty;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nsel; q++) {
        int64_t j = rows[q];
        cplx acc = 0;
        for (int row = 0; row < N2; row++) {
            double l = (row - N2 / 2) * pix_x;
            for (int col = 0; col < N1; col++) {
                double m = (col - N1 / 2) * pix_y;
                double nm1 = orc_nm1(l, m);
                double ph = -foc * (uvw[3 * j] * l + uvw[3 * j + 1] * m +
                                    sign * uvw[3 * j + 2] * nm1);
                acc += d[(size_t)row * N1 + col] / (nm1 + 1.0) *
                       (cos(ph) + I * sin(ph));
            }
        }
        o[q] = acc * (wgt ? wgt[j] : 1.0);
    }
}

/* the same sum for an image that is zero except at npx pixels (flat indices
 * pixels[], complex values vals[]): cost nsel * npx, so that images of
 * 8192^2 pixels can be checked on sampled rows */
void orc_dft_dirty2ms_sparse(int nxdirty, int nydirty, double pix_x,
                             double pix_y, double freq, const double *uvw,
                             int npx, const int64_t *pixels, const double *vals,
                             const double *wgt, int sign, int nsel,
                             const int64_t *rows, double *out)
{
    const cplx *d = (const cplx *)vals;
    cplx *o = (cplx *)out;
    double foc = 2.0 * ORC_PI * freq / ORC_C;
    int N1 = nydirty, N2 = nxdirty;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nsel; q++) {
        int64_t j = rows[q];
        cplx acc = 0;
        for (int k = 0; k < npx; k++) {
            int row = (int)(pixels[k] / N1), col = (int)(pixels[k] % N1);
            double l = (row - N2 / 2) * pix_x, m = (col - N1 / 2) * pix_y;
            double nm1 = orc_nm1(l, m);
            double ph = -foc * (uvw[3 * j] * l + uvw[3 * j + 1] * m +
                                sign * uvw[3 * j + 2] * nm1);
            acc += d[k] / (nm1 + 1.0) * (cos(ph) + I * sin(ph));
        }
        o[q] = acc * (wgt ? wgt[j] : 1.0);
    }
}

/* ------------------------------------------------------------------------ */
/* 2-D NUFFT type 1 / type 2 (test/nufft_2d_1_test.cu:126-258,              */
/* test/nufft_2d_2_test.cu:140-315), points in [-pi,pi), pirange=1.           */
/* type 1: fk[i*N1+j] = sum_k c_k exp(+i((j-N1/2) x_k + (i-N2/2) y_k))        */
/* type 2: c_k = sum fk[i*N1+j] exp(-i((j-N1/2) x_k + (i-N2/2) y_k))          */
/* ------------------------------------------------------------------------ */

int orc_nufft2d1(int M, const double *xk, const double *yk, const double *c,
                 int N1, int N2, double eps, double sigma, int single,
                 double *fk)
{
    int kw;
    double beta, es_c;
    int ier = orc_setup_conv(eps, sigma, single, &kw, &beta, &es_c);
    if (ier == 2) return ier;
    int nf1 = orc_num_cells(N1, sigma, kw), nf2 = orc_num_cells(N2, sigma, kw);
    double *zz = calloc(M + 1, sizeof(double));
    cplx *fw = calloc((size_t)nf1 * nf2, sizeof(cplx));
    /* with nf3 = 1 every point sits on plane 0 with kernel psi(0)=e^beta in z;
     * spread in 2-D directly instead (conv.cu:110-164) */
    const cplx *cc = (const cplx *)c;
    for (int j = 0; j < M; j++) {
        double g1 = orc_shift_rescale(xk[j], nf1, 1);
        double g2 = orc_shift_rescale(yk[j], nf2, 1);
        int x0 = (int)ceil(g1 - kw / 2.0), y0 = (int)ceil(g2 - kw / 2.0);
        double k1[ORC_MAX_KW], k2[ORC_MAX_KW];
        orc_kervec(k1, x0 - g1, kw, es_c, beta);
        orc_kervec(k2, y0 - g2, kw, es_c, beta);
        for (int dy = 0; dy < kw; dy++) {
            cplx *row = fw + (size_t)orc_wrap(y0 + dy, nf2) * nf1;
            for (int dx = 0; dx < kw; dx++)
                row[orc_wrap(x0 + dx, nf1)] += cc[j] * (k1[dx] * k2[dy]);
        }
    }
    orc_fft2d(1, nf2, nf1, (double *)fw, +1);
    double *phi1 = malloc(sizeof(double) * (nf1 / 2 + 1));
    double *phi2 = malloc(sizeof(double) * (nf2 / 2 + 1));
    orc_fseries_grid(kw, beta, es_c, nf1, phi1);
    orc_fseries_grid(kw, beta, es_c, nf2, phi2);
    cplx *o = (cplx *)fk;
    for (int i = 0; i < N2; i++)
        for (int j = 0; j < N1; j++) {
            int w1 = j >= N1 / 2 ? j - N1 / 2 : nf1 + j - N1 / 2;
            int w2 = i >= N2 / 2 ? i - N2 / 2 : nf2 + i - N2 / 2;
            o[(size_t)i * N1 + j] = fw[(size_t)w2 * nf1 + w1] /
                                    (phi1[abs(j - N1 / 2)] * phi2[abs(i - N2 / 2)]);
        }
    free(zz); free(fw); free(phi1); free(phi2);
    return ier;
}

int orc_nufft2d2(int M, const double *xk, const double *yk, double *c,
                 int N1, int N2, double eps, double sigma, int single,
                 const double *fk)
{
    int kw;
    double beta, es_c;
    int ier = orc_setup_conv(eps, sigma, single, &kw, &beta, &es_c);
    if (ier == 2) return ier;
    int nf1 = orc_num_cells(N1, sigma, kw), nf2 = orc_num_cells(N2, sigma, kw);
    cplx *fw = calloc((size_t)nf1 * nf2, sizeof(cplx));
    double *phi1 = malloc(sizeof(double) * (nf1 / 2 + 1));
    double *phi2 = malloc(sizeof(double) * (nf2 / 2 + 1));
    orc_fseries_grid(kw, beta, es_c, nf1, phi1);
    orc_fseries_grid(kw, beta, es_c, nf2, phi2);
    const cplx *in = (const cplx *)fk;
    for (int i = 0; i < N2; i++)
        for (int j = 0; j < N1; j++) {
            int w1 = j >= N1 / 2 ? j - N1 / 2 : nf1 + j - N1 / 2;
            int w2 = i >= N2 / 2 ? i - N2 / 2 : nf2 + i - N2 / 2;
            fw[(size_t)w2 * nf1 + w1] = in[(size_t)i * N1 + j] /
                                       (phi1[abs(j - N1 / 2)] * phi2[abs(i - N2 / 2)]);
        }
    orc_fft2d(1, nf2, nf1, (double *)fw, -1);
    cplx *co = (cplx *)c;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < M; j++) {
        double g1 = orc_shift_rescale(xk[j], nf1, 1);
        double g2 = orc_shift_rescale(yk[j], nf2, 1);
        int x0 = (int)ceil(g1 - kw / 2.0), y0 = (int)ceil(g2 - kw / 2.0);
        double k1[ORC_MAX_KW], k2[ORC_MAX_KW];
        orc_kervec(k1, x0 - g1, kw, es_c, beta);
        orc_kervec(k2, y0 - g2, kw, es_c, beta);
        cplx acc = 0;
        for (int dy = 0; dy < kw; dy++) {
            const cplx *row = fw + (size_t)orc_wrap(y0 + dy, nf2) * nf1;
            for (int dx = 0; dx < kw; dx++)
                acc += row[orc_wrap(x0 + dx, nf1)] * (k1[dx] * k2[dy]);
        }
        co[j] = acc;
    }
    free(fw); free(phi1); free(phi2);
    return ier;
}

/* direct sums for the 2-D NUFFT on a subset (modes or points) */
void orc_dft2d1(int M, const double *xk, const double *yk, const double *c,
                int N1, int N2, int nsel, const int64_t *modes, double *out)
{
    const cplx *cc = (const cplx *)c;
    cplx *o = (cplx *)out;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nsel; q++) {
        int i = (int)(modes[q] / N1), j = (int)(modes[q] % N1);
        cplx acc = 0;
        for (int k = 0; k < M; k++) {
            double ph = (j - N1 / 2) * xk[k] + (i - N2 / 2) * yk[k];
            acc += cc[k] * (cos(ph) + I * sin(ph));
        }
        o[q] = acc;
    }
}

void orc_dft2d2(const double *xk, const double *yk, int N1, int N2,
                const double *fk, int nsel, const int64_t *pts, double *out)
{
    const cplx *f = (const cplx *)fk;
    cplx *o = (cplx *)out;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nsel; q++) {
        int64_t k = pts[q];
        cplx acc = 0;
        for (int i = 0; i < N2; i++)
            for (int j = 0; j < N1; j++) {
                double ph = -((j - N1 / 2) * xk[k] + (i - N2 / 2) * yk[k]);
                acc += f[(size_t)i * N1 + j] * (cos(ph) + I * sin(ph));
            }
        o[q] = acc;
    }
}

/* plan export for host-logic tests: fills out[16] doubles */
int orc_plan_export(int nrow, int nxdirty, int nydirty, double fov,
                    double freq, const double *uvw, double eps, double sigma,
                    int sign, int single, double *out)
{
    orc_plan P;
    int ier = orc_make_plan(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma,
                            sign, single, &P);
    out[0] = P.nf1; out[1] = P.nf2; out[2] = P.nf3; out[3] = P.kw;
    out[4] = P.beta; out[5] = P.es_c; out[6] = P.pix_x; out[7] = P.pix_y;
    out[8] = P.i_center; out[9] = P.i_half_width; out[10] = P.o_center;
    out[11] = P.o_half_width; out[12] = P.gamma; out[13] = P.h;
    out[14] = P.f_over_c; out[15] = ier;
    return ier;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
