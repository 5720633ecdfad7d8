// Learned gridding kernel: a small MLP per support width stands in for the
// ES kernel taps and, through the Gauss-Legendre rule the reference uses for
// the ES kernel's Fourier transform (src/fourier/deconv.cu:12-65,
// src/fourier/deconv.cu:80-94), for the correction factors.
//
//   psi(d) = exp(beta (sqrt(1 - x^2) - 1) + y(s)),  x = 2 d / kw,  |d| < kw/2,
//   s = 2 x^2 - 1 (psi is even),  y = W3 . tanh(W2 tanh(W1 s + b1) + b2) + b3,
//   1 -> 32 -> 32 -> 1, FP32 weights.
// The first term is the ES kernel's exponent with the reference's beta for
// sigma = 2 (src/fourier/conv_interp_invoker.cu:96-104): the network learns
// a correction to it, and W3 = b3 = 0 is the ES kernel itself.
//
// Training data is the gridding error the method itself produces: gridding a
// unit visibility at fractional grid offset u with support kw and evaluating
// the corrected output at image mode theta (radians per grid cell) returns
// F(u, theta) times the exact direct-DFT value, with
//   F = sum_t psi(d_t) e^{-i theta d_t} / psihat(theta),  d_t = ceil(u - kw/2) + t - u,
//   psihat(theta) = 2 sum_q g_q psi(x_q) cos(theta x_q)   (half-line nodes x_q).
// The loss is mean |F - 1|^2 over samples (u, theta); per axis the error of
// a 2-D / 3-D gridding is the sum of these to first order.
//
// Kernels: k_mlp_eval (inference), k_learn_samples (per sample: forward at
// the taps, F, loss, dL/dy at the taps, dL/dpsi at the quadrature nodes),
// k_learn_fit (least squares to a target log-kernel, the warm start),
// k_mlp_backward (parameter gradients of a list of (s, dL/dy), reduced per
// block in LDS), k_adam, k_learn_tabulate (y as a polynomial in s for the
// gridders, from its values at Chebyshev points).
#include <cmath>
#include <vector>

#include "cg_internal.hpp"

namespace cg {

constexpr int kH = 32;
constexpr int kW1 = 0, kB1 = 32, kW2 = 64, kB2 = 64 + 1024, kW3 = kB2 + 32, kB3 = kW3 + 32;
constexpr int kNParams = kB3 + 1;  // 1153
static_assert(kNParams == LEARN_NPARAMS, "parameter layout");

__device__ __forceinline__ void mlp_forward(const float *__restrict__ p, float s, float *h1,
                                            float *h2, float &y)
{
#pragma unroll
    for (int j = 0; j < kH; j++) h1[j] = tanhf(p[kW1 + j] * s + p[kB1 + j]);
#pragma unroll 4
    for (int o = 0; o < kH; o++) {
        float z = p[kB2 + o];
#pragma unroll
        for (int j = 0; j < kH; j++) z += p[kW2 + o * kH + j] * h1[j];
        h2[o] = tanhf(z);
    }
    float z = p[kB3];
#pragma unroll
    for (int o = 0; o < kH; o++) z += p[kW3 + o] * h2[o];
    y = z;
}

__device__ __forceinline__ float mlp_input(double d, int kw)
{
    const double x = 2.0 * d / kw;
    return (float)(2.0 * x * x - 1.0);
}

// ES exponent at offset d (|d| <= kw/2)
__device__ __forceinline__ double es_prior(double d, int kw, double beta)
{
    const double x = 2.0 * d / kw;
    return beta * (sqrt(fmax(0.0, 1.0 - x * x)) - 1.0);
}

// the same from the network input s = 2 x^2 - 1
__device__ __forceinline__ double es_prior_s(float s, double beta)
{
    return beta * (sqrt(fmax(0.0, 0.5 * (1.0 - (double)s))) - 1.0);
}

// y at n inputs s[i]
__global__ __launch_bounds__(256) void k_mlp_eval(const float *__restrict__ params,
                                                  const float *__restrict__ s, int n,
                                                  float *__restrict__ y)
{
    __shared__ float p[kNParams];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) p[k] = params[k];
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float h1[kH], h2[kH], yy;
    mlp_forward(p, s[i], h1, h2, yy);
    y[i] = yy;
}

// psi at the quadrature nodes: psiq[q] = exp(y(s(x_q)))
__global__ __launch_bounds__(64) void k_node_psi(const float *__restrict__ params, LearnQuad lq,
                                                 double *__restrict__ psiq)
{
    __shared__ float p[kNParams];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) p[k] = params[k];
    __syncthreads();
    const int q = threadIdx.x;
    if (q >= lq.nq) return;
    float h1[kH], h2[kH], y;
    mlp_forward(p, mlp_input(lq.xq[q], lq.kw), h1, h2, y);
    psiq[q] = exp(es_prior(lq.xq[q], lq.kw, lq.beta) + (double)y);
}

// One thread per sample (u, theta).  Writes the MLP input s and dL/dy of
// every tap (kw per sample), adds dL/dpsi_q into gnode and the loss into
// loss, all scaled by 1 / nsamp.
__global__ __launch_bounds__(256) void k_learn_samples(const float *__restrict__ params,
                                                       const double *__restrict__ su,
                                                       const double *__restrict__ st, int nsamp,
                                                       LearnQuad lq,
                                                       const double *__restrict__ psiq,
                                                       float *__restrict__ s_out,
                                                       float *__restrict__ dyt,
                                                       double *__restrict__ gnode,
                                                       double *__restrict__ loss)
{
    __shared__ float p[kNParams];
    __shared__ double s_gnode[LEARN_MAXQ];
    __shared__ double s_loss;
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) p[k] = params[k];
    for (int k = threadIdx.x; k < LEARN_MAXQ; k += blockDim.x) s_gnode[k] = 0.0;
    if (threadIdx.x == 0) s_loss = 0.0;
    __syncthreads();
    const int kw = lq.kw, nq = lq.nq;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nsamp) {
        const double u = su[i], th = st[i];
        double ph = 0.0;
        for (int q = 0; q < nq; q++) ph += 2.0 * lq.gq[q] * psiq[q] * cos(th * lq.xq[q]);
        double ar = 0.0, ai = 0.0;
        double psi[LEARN_MAXKW], cs[LEARN_MAXKW], sn[LEARN_MAXKW];
        const double d0 = ceil(u - 0.5 * kw) - u;
        for (int t = 0; t < kw; t++) {
            const double d = d0 + t;
            const float s = mlp_input(d, kw);
            float h1[kH], h2[kH], y;
            mlp_forward(p, s, h1, h2, y);
            const double v = fabs(d) < 0.5 * kw ? exp(es_prior(d, kw, lq.beta) + (double)y) : 0.0;
            psi[t] = v;
            sincos(th * d, &sn[t], &cs[t]);
            ar += v * cs[t];
            ai -= v * sn[t];
            s_out[(size_t)i * kw + t] = s;
        }
        const double inv = 1.0 / nsamp;
        const double er = ar / ph - 1.0, ei = ai / ph;  // E = F - 1
        atomicAdd(&s_loss, (er * er + ei * ei) * inv);
        // dL/dpsi_t = 2 Re(conj(E) e^{-i th d_t}) / ph; dL/dy_t = dL/dpsi_t psi_t
        for (int t = 0; t < kw; t++) {
            const double g = 2.0 * (er * cs[t] - ei * sn[t]) / ph * inv;
            dyt[(size_t)i * kw + t] = (float)(g * psi[t]);
        }
        // dL/dpsihat = -2 Re(conj(E) F) / ph; dpsihat/dpsi_q = 2 g_q cos(th x_q)
        const double gph = -2.0 * (er * (er + 1.0) + ei * ei) / ph * inv;
        for (int q = 0; q < nq; q++)
            atomicAdd(&s_gnode[q], gph * 2.0 * lq.gq[q] * cos(th * lq.xq[q]));
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nq; k += blockDim.x) atomicAdd(&gnode[k], s_gnode[k]);
    if (threadIdx.x == 0) atomicAdd(loss, s_loss);
}

// node gradients dL/dpsi_q -> (s_q, dL/dy_q) pairs for k_mlp_backward
__global__ __launch_bounds__(64) void k_node_dy(LearnQuad lq, const double *__restrict__ gnode,
                                                const double *__restrict__ psiq,
                                                float *__restrict__ s_out, float *__restrict__ dy)
{
    const int q = threadIdx.x;
    if (q >= lq.nq) return;
    s_out[q] = mlp_input(lq.xq[q], lq.kw);
    dy[q] = (float)(gnode[q] * psiq[q]);
}

// least squares to a target log-kernel: loss = mean (log psi(s_i) - target_i)^2
__global__ __launch_bounds__(256) void k_learn_fit(const float *__restrict__ params,
                                                   const float *__restrict__ s,
                                                   const float *__restrict__ target, int n,
                                                   double beta, float *__restrict__ dy,
                                                   double *__restrict__ loss)
{
    __shared__ float p[kNParams];
    __shared__ double s_loss;
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) p[k] = params[k];
    if (threadIdx.x == 0) s_loss = 0.0;
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        float h1[kH], h2[kH], y;
        mlp_forward(p, s[i], h1, h2, y);
        const double r = es_prior_s(s[i], beta) + (double)y - (double)target[i];
        atomicAdd(&s_loss, r * r / n);
        dy[i] = (float)(2.0 * r / n);
    }
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, s_loss);
}

// Parameter gradients of sum_e dy_e y(s_e).  Each thread back-propagates one
// evaluation; the block reduces its 256 evaluations in LDS (outer products
// for W2, column sums for the rest) and adds one partial per parameter.
__global__ __launch_bounds__(256) void k_mlp_backward(const float *__restrict__ params,
                                                      const float *__restrict__ s_in,
                                                      const float *__restrict__ dy, int n,
                                                      float *__restrict__ grad)
{
    __shared__ float p[kNParams];
    __shared__ float s_a[256][kH + 1];
    __shared__ float s_b[256][kH + 1];
    __shared__ float s_c[256][kH + 1];
    __shared__ float s_g[256];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) p[k] = params[k];
    __syncthreads();
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int r = threadIdx.x;
    float h1[kH], h2[kH], d2[kH];
    float g = 0.f, se = 0.f;
    if (e < n) {
        se = s_in[e];
        g = dy[e];
        float y;
        mlp_forward(p, se, h1, h2, y);
    } else {
#pragma unroll
        for (int j = 0; j < kH; j++) h1[j] = h2[j] = 0.f;
    }
#pragma unroll
    for (int o = 0; o < kH; o++) d2[o] = g * p[kW3 + o] * (1.f - h2[o] * h2[o]);
#pragma unroll
    for (int j = 0; j < kH; j++) {
        s_a[r][j] = h1[j];
        s_b[r][j] = d2[j];
        s_c[r][j] = g * h2[j];
    }
    s_g[r] = g;
    __syncthreads();
    // dW2[o][j] = sum_r dz2[r][o] h1[r][j]
    for (int k = threadIdx.x; k < kH * kH; k += blockDim.x) {
        const int o = k / kH, j = k % kH;
        float acc = 0.f;
        for (int rr = 0; rr < 256; rr++) acc += s_b[rr][o] * s_a[rr][j];
        atomicAdd(&grad[kW2 + k], acc);
    }
    if (threadIdx.x < kH) {
        const int o = threadIdx.x;
        float a3 = 0.f, a2 = 0.f;
        for (int rr = 0; rr < 256; rr++) {
            a3 += s_c[rr][o];
            a2 += s_b[rr][o];
        }
        atomicAdd(&grad[kW3 + o], a3);
        atomicAdd(&grad[kB2 + o], a2);
    } else if (threadIdx.x == kH) {
        float a = 0.f;
        for (int rr = 0; rr < 256; rr++) a += s_g[rr];
        atomicAdd(&grad[kB3], a);
    }
    __syncthreads();
    // dz1[j] = (sum_o W2[o][j] dz2[o]) (1 - h1[j]^2)
#pragma unroll
    for (int j = 0; j < kH; j++) {
        float a = 0.f;
#pragma unroll
        for (int o = 0; o < kH; o++) a += p[kW2 + o * kH + j] * d2[o];
        const float d1 = a * (1.f - h1[j] * h1[j]);
        s_a[r][j] = d1;
        s_c[r][j] = d1 * se;
    }
    __syncthreads();
    if (threadIdx.x < kH) {
        const int j = threadIdx.x;
        float ab = 0.f, aw = 0.f;
        for (int rr = 0; rr < 256; rr++) {
            ab += s_a[rr][j];
            aw += s_c[rr][j];
        }
        atomicAdd(&grad[kB1 + j], ab);
        atomicAdd(&grad[kW1 + j], aw);
    }
}

// Adam (Kingma & Ba 2015) step with bias corrections c1 = 1/(1-b1^t),
// c2 = 1/(1-b2^t); clears the gradient for the next step
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ params, float *__restrict__ grad,
                                              float *__restrict__ m, float *__restrict__ v, int n,
                                              float lr, float c1, float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    grad[i] = 0.f;
}

__global__ __launch_bounds__(256) void k_fill_f32(float *__restrict__ p, int n, float v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(64) void k_fill_f64(double *__restrict__ p, int n)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Tabulated form for the gridders: the network output y as a polynomial in
// s on [-1, 1], y(s) ~ sum_k coef[k] s^k, from its values at the deg+1
// Chebyshev points sc (FP64 values mapped by M, row k: coefficient k).
// One block.
__global__ __launch_bounds__(64) void k_learn_tabulate(const float *__restrict__ params, int deg,
                                                       const double *__restrict__ sc,
                                                       const double *__restrict__ M,
                                                       float *__restrict__ coef)
{
    __shared__ float p[kNParams];
    __shared__ double vals[LEARN_MAXDEG + 1];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) p[k] = params[k];
    __syncthreads();
    const int np = deg + 1;
    for (int j = threadIdx.x; j < np; j += blockDim.x) {
        float h1[kH], h2[kH], y;
        mlp_forward(p, (float)sc[j], h1, h2, y);
        vals[j] = (double)y;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < np; c += blockDim.x) {
        double a = 0.0;
        for (int j = 0; j < np; j++) a += M[c * np + j] * vals[j];
        coef[c] = (float)a;
    }
}

double learn_beta(int kw)
{
    // src/fourier/conv_interp_invoker.cu:96-104 at sigma = 2
    const double bo = kw == 2 ? 2.20 : kw == 3 ? 2.26 : kw == 4 ? 2.38 : 2.30;
    return bo * kw;
}

LearnQuad learn_quad(int kw)
{
    // half-line Gauss-Legendre rule of p = floor(2 + 1.5 kw) nodes, as the
    // reference integrates the ES kernel (src/fourier/deconv.cu:80-94)
    LearnQuad q{};
    q.kw = kw;
    q.beta = learn_beta(kw);
    const int p = (int)(2 + 1.5 * kw);
    std::vector<double> x(2 * p), w(2 * p);
    gauss_legendre(2 * p, x.data(), w.data());
    q.nq = p;
    for (int i = 0; i < p; i++) {
        q.xq[i] = 0.5 * kw * x[i];
        q.gq[i] = 0.5 * kw * w[i];
    }
    return q;
}

void cheb_to_monomial(int deg, double *zc, double *M)
{
    // Chebyshev coefficients by the discrete cosine sum at the points
    // z_j = cos(pi (j + 1/2) / n), then T_k -> z^m by the recurrence
    // T_k = 2 z T_{k-1} - T_{k-2}; FP64 on the host
    const int n = deg + 1;
    for (int j = 0; j < n; j++) zc[j] = std::cos(M_PI * (j + 0.5) / n);
    std::vector<double> C(n * n), T(n * n, 0.0);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            C[k * n + j] = (k == 0 ? 1.0 : 2.0) / n * std::cos(M_PI * k * (j + 0.5) / n);
    T[0] = 1.0;
    if (n > 1) T[1 * n + 1] = 1.0;
    for (int k = 2; k < n; k++)
        for (int m = 0; m < n; m++)
            T[k * n + m] = (m > 0 ? 2.0 * T[(k - 1) * n + m - 1] : 0.0) - T[(k - 2) * n + m];
    for (int m = 0; m < n; m++)
        for (int j = 0; j < n; j++) {
            double a = 0.0;
            for (int k = 0; k < n; k++) a += T[k * n + m] * C[k * n + j];
            M[m * n + j] = a;
        }
}

namespace {

// device workspace of one training / tabulation call, freed on return
struct LearnWs {
    std::vector<void *> bufs;
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    double *u = nullptr, *th = nullptr;
    float *s_fit = nullptr, *target = nullptr;
    float *s_in = nullptr, *dy = nullptr, *sq = nullptr, *dyq = nullptr;
    double *psiq = nullptr, *gnode = nullptr, *loss = nullptr;
    ~LearnWs()
    {
        for (void *p : bufs) (void)hipFree(p);
    }
    template <typename X>
    int get(X *&p, size_t count)
    {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count * sizeof(X), 256)) != hipSuccess) {
            (void)hipGetLastError();
            return CG_ERR_ALLOC;
        }
        bufs.push_back(q);
        p = (X *)q;
        return CG_OK;
    }
    int init(int mode, int n, int nevals)
    {
        CG_TRY(get(params, kNParams));
        CG_TRY(get(grad, kNParams));
        CG_TRY(get(m, kNParams));
        CG_TRY(get(v, kNParams));
        if (mode == 0) {
            CG_TRY(get(s_fit, n));
            CG_TRY(get(target, n));
        } else {
            CG_TRY(get(u, n));
            CG_TRY(get(th, n));
        }
        CG_TRY(get(s_in, nevals));
        CG_TRY(get(dy, nevals));
        CG_TRY(get(sq, LEARN_MAXQ));
        CG_TRY(get(dyq, LEARN_MAXQ));
        CG_TRY(get(psiq, LEARN_MAXQ));
        CG_TRY(get(gnode, LEARN_MAXQ));
        CG_TRY(get(loss, 1));
        return CG_OK;
    }
};

// gradient of the error functional on samples [off, off + cnt), added to
// ws.grad; the loss goes to ws.loss
void learn_grad_launch(LearnWs &ws, const LearnQuad &lq, int off, int cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, s, ws.gnode, LEARN_MAXQ);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, s, ws.loss, 1);
    hipLaunchKernelGGL(k_node_psi, dim3(1), dim3(64), 0, s, (const float *)ws.params, lq, ws.psiq);
    hipLaunchKernelGGL(k_learn_samples, dim3((unsigned)ceil_div(cnt, 256)), dim3(256), 0, s,
                       (const float *)ws.params, (const double *)ws.u + off,
                       (const double *)ws.th + off, cnt, lq, (const double *)ws.psiq, ws.s_in,
                       ws.dy, ws.gnode, ws.loss);
    const int ne = cnt * lq.kw;
    hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0, s,
                       (const float *)ws.params, (const float *)ws.s_in, (const float *)ws.dy, ne,
                       ws.grad);
    hipLaunchKernelGGL(k_node_dy, dim3(1), dim3(64), 0, s, lq, (const double *)ws.gnode,
                       (const double *)ws.psiq, ws.sq, ws.dyq);
    hipLaunchKernelGGL(k_mlp_backward, dim3(1), dim3(256), 0, s, (const float *)ws.params,
                       (const float *)ws.sq, (const float *)ws.dyq, lq.nq, ws.grad);
}

}  // namespace

int learn_grad(int kw, const float *params, const double *u, const double *th, int nsamp,
               float *grad, double *loss, hipStream_t s)
{
    if (kw < 2 || kw > LEARN_MAXKW || nsamp <= 0 || !params || !u || !th || !grad || !loss)
        return CG_ERR_ARG;
    const LearnQuad lq = learn_quad(kw);
    LearnWs ws;
    CG_TRY(ws.init(1, nsamp, nsamp * kw));
    CG_HIP(hipMemcpyAsync(ws.params, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(ws.u, u, sizeof(double) * nsamp, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(ws.th, th, sizeof(double) * nsamp, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_f32, dim3((kNParams + 255) / 256), dim3(256), 0, s, ws.grad,
                       kNParams, 0.f);
    learn_grad_launch(ws, lq, 0, nsamp, s);
    CG_HIP(hipGetLastError());
    CG_HIP(hipMemcpyAsync(grad, ws.grad, sizeof(float) * kNParams, hipMemcpyDeviceToHost, s));
    CG_HIP(hipMemcpyAsync(loss, ws.loss, sizeof(double), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    return CG_OK;
}

int learn_train(int kw, int mode, const double *a, const double *b, int n, int batch, int steps,
                double lr, float *params, double *losses, hipStream_t s)
{
    if (kw < 2 || kw > LEARN_MAXKW || (mode != 0 && mode != 1) || n <= 0 || steps < 0 ||
        !params || !a || !b || !(lr > 0))
        return CG_ERR_ARG;
    batch = std::max(1, std::min(batch, n));
    const LearnQuad lq = learn_quad(kw);
    LearnWs ws;
    CG_TRY(ws.init(mode, n, mode == 0 ? batch : batch * kw));
    CG_HIP(hipMemcpyAsync(ws.params, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, s));
    std::vector<float> sf, tf;
    if (mode == 1) {
        CG_HIP(hipMemcpyAsync(ws.u, a, sizeof(double) * n, hipMemcpyHostToDevice, s));
        CG_HIP(hipMemcpyAsync(ws.th, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
    } else {
        // mode 0: a = offsets d, b = target log-kernel values
        sf.resize(n);
        tf.resize(n);
        for (int i = 0; i < n; i++) {
            const double x = 2.0 * a[i] / kw;
            sf[i] = (float)(2.0 * x * x - 1.0);
            tf[i] = (float)b[i];
        }
        CG_HIP(hipMemcpyAsync(ws.s_fit, sf.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
        CG_HIP(hipMemcpyAsync(ws.target, tf.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    const int nb = (kNParams + 255) / 256;
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, s, ws.grad, kNParams, 0.f);
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, s, ws.m, kNParams, 0.f);
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, s, ws.v, kNParams, 0.f);
    std::vector<double> dloss(std::max(steps, 1));
    int off = 0;
    for (int it = 0; it < steps; it++) {
        const int cnt = std::min(batch, n - off);
        if (mode == 0) {
            hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, s, ws.loss, 1);
            hipLaunchKernelGGL(k_learn_fit, dim3((unsigned)ceil_div(cnt, 256)), dim3(256), 0, s,
                               (const float *)ws.params, (const float *)ws.s_fit + off,
                               (const float *)ws.target + off, cnt, lq.beta, ws.dy, ws.loss);
            hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)ceil_div(cnt, 256)), dim3(256), 0, s,
                               (const float *)ws.params, (const float *)ws.s_fit + off,
                               (const float *)ws.dy, cnt, ws.grad);
        } else {
            learn_grad_launch(ws, lq, off, cnt, s);
        }
        if (losses)
            CG_HIP(hipMemcpyAsync(&dloss[it], ws.loss, sizeof(double), hipMemcpyDeviceToHost, s));
        const double c1 = 1.0 / (1.0 - std::pow(0.9, it + 1));
        const double c2 = 1.0 / (1.0 - std::pow(0.999, it + 1));
        hipLaunchKernelGGL(k_adam, dim3(nb), dim3(256), 0, s, ws.params, ws.grad, ws.m, ws.v,
                           kNParams, (float)lr, (float)c1, (float)c2);
        CG_HIP(hipGetLastError());
        off += cnt;
        if (off >= n) off = 0;
    }
    CG_HIP(hipMemcpyAsync(params, ws.params, sizeof(float) * kNParams, hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    if (losses)
        for (int it = 0; it < steps; it++) losses[it] = dloss[it];
    return CG_OK;
}

int learn_eval(const float *params, const float *s_in, int n, float *y, hipStream_t s)
{
    if (!params || !s_in || !y || n <= 0) return CG_ERR_ARG;
    LearnWs ws;
    float *dp, *dx, *dy;
    CG_TRY(ws.get(dp, kNParams));
    CG_TRY(ws.get(dx, n));
    CG_TRY(ws.get(dy, n));
    CG_HIP(hipMemcpyAsync(dp, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(dx, s_in, sizeof(float) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                       (const float *)dp, (const float *)dx, n, dy);
    CG_HIP(hipGetLastError());
    CG_HIP(hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    return CG_OK;
}

int learn_kernel_build(int kw, int deg, const float *params, LearnedKernel &K, hipStream_t s)
{
    if (kw < 2 || kw > LEARN_MAXKW || deg < 1 || deg > LEARN_MAXDEG || !params) return CG_ERR_ARG;
    const int np = deg + 1;
    std::vector<double> sc(np), M(np * np);
    cheb_to_monomial(deg, sc.data(), M.data());
    const LearnQuad lq = learn_quad(kw);
    LearnWs ws;
    float *dp, *dcoef;
    double *dsc, *dM, *dpsiq;
    CG_TRY(ws.get(dp, kNParams));
    CG_TRY(ws.get(dcoef, np));
    CG_TRY(ws.get(dsc, np));
    CG_TRY(ws.get(dM, np * np));
    CG_TRY(ws.get(dpsiq, LEARN_MAXQ));
    CG_HIP(hipMemcpyAsync(dp, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(dsc, sc.data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(dM, M.data(), sizeof(double) * np * np, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_learn_tabulate, dim3(1), dim3(64), 0, s, (const float *)dp, deg,
                       (const double *)dsc, (const double *)dM, dcoef);
    hipLaunchKernelGGL(k_node_psi, dim3(1), dim3(64), 0, s, (const float *)dp, lq, dpsiq);
    CG_HIP(hipGetLastError());
    K = LearnedKernel{};
    K.kw = kw;
    K.deg = deg;
    K.nq = lq.nq;
    K.beta = lq.beta;
    float coef[LEARN_MAXDEG + 1] = {};
    CG_HIP(hipMemcpyAsync(coef, dcoef, sizeof(float) * np, hipMemcpyDeviceToHost, s));
    CG_HIP(hipMemcpyAsync(K.qpsi, dpsiq, sizeof(double) * lq.nq, hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    for (int k = 0; k <= LEARN_MAXDEG; k++) K.coef[k] = k < np ? coef[k] : 0.f;
    for (int q = 0; q < lq.nq; q++) {
        K.qx[q] = lq.xq[q];
        K.qg[q] = lq.gq[q];
    }
    return CG_OK;
}

}  // namespace cg
