// Hand-written batched 2-D FFT of the w-planes (replaces the reference's
// cuFFT C2C/Z2Z plans, src/fourier/cuft.cu:20-44 and :192-212).
//
// Unnormalised transform y[k] = sum_j x[j] exp(dir 2 pi i j k / n), applied
// along the fast axis (rows) then the slow axis (columns) of each plane.
// Each workgroup loads whole transforms into LDS (rows: a run of contiguous
// rows; columns: a block of adjacent columns, so every global access is a
// contiguous segment), runs a mixed-radix Stockham autosort (radix 8, 4, 2,
// 3, 5 in any order) in place -- every pass reads its butterflies into
// registers, barriers, then writes -- and stores the transforms back.
// Twiddles come from an FP64-computed table e^{+2 pi i t/n}.
#include "cg_internal.hpp"

namespace cg {

// max elements per thread: 16 complex64 or 8 complex128 (128 VGPRs at 1024
// threads), so one workgroup holds up to 16384 / 8192 elements = 128 KB LDS
template <typename T>
constexpr int fft_ept() { return sizeof(T) == 4 ? 16 : 8; }

template <typename T>
__device__ __forceinline__ cx_t<T> cmul(cx_t<T> a, cx_t<T> b)
{
    return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cx_t<T> cadd(cx_t<T> a, cx_t<T> b) { return mkc(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx_t<T> csub(cx_t<T> a, cx_t<T> b) { return mkc(a.x - b.x, a.y - b.y); }
// multiply by dir*i
template <typename T, int DIR>
__device__ __forceinline__ cx_t<T> cmul_i(cx_t<T> a)
{
    return DIR > 0 ? mkc(-a.y, a.x) : mkc(a.y, -a.x);
}

template <typename T, int R, int DIR>
__device__ __forceinline__ void dft_small(cx_t<T> *a)
{
    if constexpr (R == 2) {
        cx_t<T> t = a[0];
        a[0] = cadd<T>(t, a[1]);
        a[1] = csub<T>(t, a[1]);
    } else if constexpr (R == 4) {
        cx_t<T> t0 = cadd<T>(a[0], a[2]), t1 = csub<T>(a[0], a[2]);
        cx_t<T> t2 = cadd<T>(a[1], a[3]), t3 = cmul_i<T, DIR>(csub<T>(a[1], a[3]));
        a[0] = cadd<T>(t0, t2);
        a[2] = csub<T>(t0, t2);
        a[1] = cadd<T>(t1, t3);
        a[3] = csub<T>(t1, t3);
    } else if constexpr (R == 8) {
        const T r2 = (T)0.70710678118654752440084436210485;
        cx_t<T> e[4] = {a[0], a[2], a[4], a[6]};
        cx_t<T> o[4] = {a[1], a[3], a[5], a[7]};
        dft_small<T, 4, DIR>(e);
        dft_small<T, 4, DIR>(o);
        // W8^k o_k, W8 = (1 + dir i)/sqrt2
        cx_t<T> w1 = mkc(r2 * (o[1].x - DIR * o[1].y), r2 * (o[1].y + DIR * o[1].x));
        cx_t<T> w2 = cmul_i<T, DIR>(o[2]);
        cx_t<T> w3 = mkc(r2 * (-o[3].x - DIR * o[3].y), r2 * (-o[3].y + DIR * o[3].x));
        a[0] = cadd<T>(e[0], o[0]);
        a[4] = csub<T>(e[0], o[0]);
        a[1] = cadd<T>(e[1], w1);
        a[5] = csub<T>(e[1], w1);
        a[2] = cadd<T>(e[2], w2);
        a[6] = csub<T>(e[2], w2);
        a[3] = cadd<T>(e[3], w3);
        a[7] = csub<T>(e[3], w3);
    } else if constexpr (R == 3) {
        const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
        cx_t<T> t = cadd<T>(a[1], a[2]);
        cx_t<T> d = csub<T>(a[1], a[2]);
        cx_t<T> m = mkc(a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y);
        cx_t<T> e = cmul_i<T, DIR>(mkc(h * d.x, h * d.y));
        a[0] = cadd<T>(a[0], t);
        a[1] = cadd<T>(m, e);
        a[2] = csub<T>(m, e);
    } else if constexpr (R == 5) {
        const T c1 = (T)0.30901699437494742410229341718282;    // cos(2pi/5)
        const T c2 = (T)-0.80901699437494742410229341718282;   // cos(4pi/5)
        const T s1 = (T)0.95105651629515357211643933337938;    // sin(2pi/5)
        const T s2 = (T)0.58778525229247312916870595463907;    // sin(4pi/5)
        cx_t<T> t1 = cadd<T>(a[1], a[4]), t2 = cadd<T>(a[2], a[3]);
        cx_t<T> d1 = csub<T>(a[1], a[4]), d2 = csub<T>(a[2], a[3]);
        cx_t<T> b1 = mkc(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
        cx_t<T> b2 = mkc(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
        cx_t<T> e1 = cmul_i<T, DIR>(mkc(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
        cx_t<T> e2 = cmul_i<T, DIR>(mkc(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
        a[0] = mkc(a[0].x + t1.x + t2.x, a[0].y + t1.y + t2.y);
        a[1] = cadd<T>(b1, e1);
        a[4] = csub<T>(b1, e1);
        a[2] = cadd<T>(b2, e2);
        a[3] = csub<T>(b2, e2);
    }
}

struct FftShape {
    int n;        // transform length
    int nb;       // transforms per workgroup
    int lt, le;   // LDS strides: element (b, e) at b*lt + e*le
    int nrad;
    int rad[16];
};

// one Stockham pass of radix R over all nb transforms held in LDS
template <typename T, int R, int DIR>
__device__ __forceinline__ void stockham_pass(cx_t<T> *buf, const FftShape &S, int p,
                                              const cx_t<T> *__restrict__ tw)
{
    constexpr int K = (fft_ept<T>() + R - 1) / R;
    const int Tn = S.n / R;
    const int nbf = S.nb * Tn;
    const int twstep = S.n / (p * R);
    cx_t<T> u[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * blockDim.x;
        if (q < nbf) {
            const int b = q / Tn, i = q - b * Tn;
#pragma unroll
            for (int r = 0; r < R; r++) u[k][r] = buf[b * S.lt + (i + r * Tn) * S.le];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * blockDim.x;
        if (q < nbf) {
            const int b = q / Tn, i = q - b * Tn;
            const int kk = i % p;
#pragma unroll
            for (int r = 1; r < R; r++) {
                cx_t<T> w = tw[r * kk * twstep];
                if (DIR < 0) w.y = -w.y;
                u[k][r] = cmul<T>(u[k][r], w);
            }
            dft_small<T, R, DIR>(u[k]);
            const int j = (i - kk) * R + kk;
#pragma unroll
            for (int r = 0; r < R; r++) buf[b * S.lt + (j + r * p) * S.le] = u[k][r];
        }
    }
    __syncthreads();
}

template <typename T, int DIR>
__device__ void lds_fft(cx_t<T> *buf, const FftShape &S, const cx_t<T> *tw)
{
    int p = 1;
    for (int s = 0; s < S.nrad; s++) {
        const int R = S.rad[s];
        switch (R) {
        case 8: stockham_pass<T, 8, DIR>(buf, S, p, tw); break;
        case 4: stockham_pass<T, 4, DIR>(buf, S, p, tw); break;
        case 2: stockham_pass<T, 2, DIR>(buf, S, p, tw); break;
        case 3: stockham_pass<T, 3, DIR>(buf, S, p, tw); break;
        case 5: stockham_pass<T, 5, DIR>(buf, S, p, tw); break;
        default: break;
        }
        p *= R;
    }
}

// rows: transforms along the contiguous axis; blockIdx.x = row group,
// blockIdx.y = plane
template <typename T, int DIR>
__global__ __launch_bounds__(1024) void k_fft_rows(cx_t<T> *data, int n2, int n1, FftShape S,
                                                   const cx_t<T> *__restrict__ tw)
{
    extern __shared__ unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    cx_t<T> *pl = data + (size_t)blockIdx.y * n2 * n1;
    const int row0 = blockIdx.x * S.nb;
    const int nrow = min(S.nb, n2 - row0);
    const int E = S.nb * n1;
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int b = k / n1;
        buf[k] = b < nrow ? pl[(size_t)row0 * n1 + k] : mkc((T)0, (T)0);
    }
    __syncthreads();
    lds_fft<T, DIR>(buf, S, tw);
    for (int k = threadIdx.x; k < nrow * n1; k += blockDim.x) pl[(size_t)row0 * n1 + k] = buf[k];
}

// columns: blockIdx.x = column group of nb adjacent columns
template <typename T, int DIR>
__global__ __launch_bounds__(1024) void k_fft_cols(cx_t<T> *data, int n2, int n1, FftShape S,
                                                   const cx_t<T> *__restrict__ tw)
{
    extern __shared__ unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    cx_t<T> *pl = data + (size_t)blockIdx.y * n2 * n1;
    const int c0 = blockIdx.x * S.nb;
    const int ncol = min(S.nb, n1 - c0);
    const int E = S.nb * n2;
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int e = k / S.nb, b = k - e * S.nb;
        buf[k] = b < ncol ? pl[(size_t)e * n1 + c0 + b] : mkc((T)0, (T)0);
    }
    __syncthreads();
    lds_fft<T, DIR>(buf, S, tw);
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int e = k / S.nb, b = k - e * S.nb;
        if (b < ncol) pl[(size_t)e * n1 + c0 + b] = buf[k];
    }
}

static int factor_radices(int n, int *rad)
{
    int nr = 0;
    while (n % 8 == 0) { rad[nr++] = 8; n /= 8; }
    while (n % 4 == 0) { rad[nr++] = 4; n /= 4; }
    while (n % 2 == 0) { rad[nr++] = 2; n /= 2; }
    while (n % 3 == 0) { rad[nr++] = 3; n /= 3; }
    while (n % 5 == 0) { rad[nr++] = 5; n /= 5; }
    return n == 1 ? nr : -1;
}

template <typename T>
bool fft_supported(int n)
{
    int rad[32];
    if (n < 1 || factor_radices(n, rad) < 0) return false;
    return (int64_t)n * (int64_t)sizeof(cx_t<T>) <= 131072 && n <= 1024 * fft_ept<T>();
}

static int round64(int x) { return (x + 63) / 64 * 64; }

template <typename T, int DIR>
static int fft_axis(cx_t<T> *data, int batch, int n2, int n1, bool rows, const cx_t<T> *tw,
                    hipStream_t s)
{
    const int n = rows ? n1 : n2;
    const int other = rows ? n2 : n1;
    FftShape S{};
    S.n = n;
    S.nrad = factor_radices(n, S.rad);
    if (S.nrad < 0) return CG_ERR_SIZE;
    const int elem = (int)sizeof(cx_t<T>);
    // elements per workgroup: up to 4096 (32 KB fp32) for short transforms,
    // a whole transform when it is longer
    int target = std::max(n, 4096);
    int nb = std::max(1, target / n);
    nb = std::min(nb, other);
    while (nb > 1 && (int64_t)nb * n * elem > 65536) nb--;
    const int E = nb * n;
    if (E > 1024 * fft_ept<T>() || (int64_t)E * elem > 131072) return CG_ERR_SIZE;
    int threads = round64((E + fft_ept<T>() - 1) / fft_ept<T>());
    threads = std::min(std::max(threads, 64), 1024);
    S.nb = nb;
    if (rows) {
        S.lt = n;
        S.le = 1;
    } else {
        S.lt = 1;
        S.le = nb;
    }
    size_t lds = (size_t)E * elem;
    dim3 grid((unsigned)((other + nb - 1) / nb), (unsigned)batch);
    if (rows) {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute((const void *)k_fft_rows<T, DIR>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_fft_rows<T, DIR>), grid, dim3(threads), lds, s, data, n2, n1, S, tw);
    } else {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute((const void *)k_fft_cols<T, DIR>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_fft_cols<T, DIR>), grid, dim3(threads), lds, s, data, n2, n1, S, tw);
    }
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T>
int launch_fft2d(cx_t<T> *data, int batch, int n2, int n1, int dir, const cx_t<T> *tw1,
                 const cx_t<T> *tw2, hipStream_t s)
{
    if (batch <= 0) return CG_OK;
    int r;
    if (dir > 0) {
        if ((r = fft_axis<T, 1>(data, batch, n2, n1, true, tw1, s))) return r;
        if (n2 > 1 && (r = fft_axis<T, 1>(data, batch, n2, n1, false, tw2, s))) return r;
    } else {
        if ((r = fft_axis<T, -1>(data, batch, n2, n1, true, tw1, s))) return r;
        if (n2 > 1 && (r = fft_axis<T, -1>(data, batch, n2, n1, false, tw2, s))) return r;
    }
    return CG_OK;
}

template <typename T>
int launch_fft_pass(cx_t<T> *data, int batch, int n2, int n1, int dir, bool rows,
                    const cx_t<T> *tw, hipStream_t s)
{
    if (batch <= 0 || (!rows && n2 <= 1)) return CG_OK;
    return dir > 0 ? fft_axis<T, 1>(data, batch, n2, n1, rows, tw, s)
                   : fft_axis<T, -1>(data, batch, n2, n1, rows, tw, s);
}

template <typename T>
__global__ void k_twiddles(cx_t<T> *tw, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double sv, cv;
    sincospi(2.0 * (double)t / (double)n, &sv, &cv);
    tw[t] = mkc((T)cv, (T)sv);
}

template <typename T>
int launch_twiddles(cx_t<T> *tw, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_twiddles<T>, dim3((n + 255) / 256), dim3(256), 0, s, tw, n);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template int launch_fft2d<float>(float2 *, int, int, int, int, const float2 *, const float2 *, hipStream_t);
template int launch_fft2d<double>(double2 *, int, int, int, int, const double2 *, const double2 *, hipStream_t);
template int launch_fft_pass<float>(float2 *, int, int, int, int, bool, const float2 *, hipStream_t);
template int launch_fft_pass<double>(double2 *, int, int, int, int, bool, const double2 *, hipStream_t);
template bool fft_supported<float>(int);
template bool fft_supported<double>(int);
template int launch_twiddles<float>(float2 *, int, hipStream_t);
template int launch_twiddles<double>(double2 *, int, hipStream_t);

}  // namespace cg
