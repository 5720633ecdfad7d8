// Hand-written batched 2-D FFT of the w-planes (replaces the reference's
// cuFFT C2C/Z2Z plans, src/fourier/cuft.cu:20-44 and :192-212).
//
// Unnormalised transform y[k] = sum_j x[j] exp(dir 2 pi i j k / n) along the
// fast axis (rows, length n1) and the slow axis (columns, length n2) of each
// [n2][n1] plane.  Every transform runs as a mixed-radix Stockham autosort
// (radix 8, 4, 2, 3, 5) in LDS: each pass reads its butterflies into
// registers, barriers, then writes them back in place.
//
// Rows: persistent workgroups walk groups of whole rows (one 16384-point
// fp32 row is 132 KB of padded LDS).
//
// Columns: a short column (n2 <= 128, or a length with no usable split) is
// transformed whole by a workgroup holding a block of adjacent columns.  A
// long column uses the in-place four-step split n2 = na * nb on blocks of 64
// (fp32) / 32 (fp64) adjacent columns, i.e. 512-byte row segments: step A
// transforms the na-point sub-columns of stride nb and applies the twiddles
// w^(b ka); step B transforms the nb-point contiguous sub-columns.  Without a
// transposition the spectrum of a split column lands in "split order":
// frequency k = ka + na kb sits at row ka nb + kb.  The forward (dir > 0)
// column pass produces that order and the inverse (dir < 0) consumes it; the
// per-pixel kernels that read or write the spectrum index through it, and
// launch_fft2d restores natural order for callers that need it.
//
// Pruning: only the fast-axis indices k < live.a or k >= live.b carry image
// pixels.  The column pass transforms only those columns.  A forward row pass
// stores only those outputs; an inverse row pass reads only those inputs
// (the others are zero by construction).
//
// Twiddles come from FP64-computed tables e^{+2 pi i t/n}; the sub-
// transforms of a split read the length-n2 table with a stride.
#include <cstdlib>

#include "cg_internal.hpp"

namespace cg {

// max elements per thread: 16 complex64 or 8 complex128
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

// workgroup barrier that orders LDS only: outstanding global loads (the
// row prefetch) are not waited for
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
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
    } else if constexpr (R == 16) {
        // 16 = 4 x 4: X[k1 + 4 k2] = sum_n2 w16^(n2 k1) [sum_n1 x[4 n1 + n2] w4^(n1 k1)] w4^(n2 k2)
        const T c1 = (T)0.92387953251128675612818318939679;  // cos(pi/8)
        const T s1 = (T)0.38268343236508977172845998403040;  // sin(pi/8)
        const T r2 = (T)0.70710678118654752440084436210485;
        cx_t<T> y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            cx_t<T> t[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
            dft_small<T, 4, DIR>(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = t[k1];
        }
        // w16^(DIR m) for m = n2 k1 in {1, 2, 3, 4, 6, 9}
        auto tw = [&](cx_t<T> v, T wr, T wi) { return mkc(v.x * wr - v.y * wi, v.x * wi + v.y * wr); };
        const T d = (T)DIR;
        y[1][1] = tw(y[1][1], c1, d * s1);
        y[1][2] = tw(y[1][2], r2, d * r2);
        y[1][3] = tw(y[1][3], s1, d * c1);
        y[2][1] = tw(y[2][1], r2, d * r2);
        y[2][2] = cmul_i<T, DIR>(y[2][2]);
        y[2][3] = tw(y[2][3], -r2, d * r2);
        y[3][1] = tw(y[3][1], s1, d * c1);
        y[3][2] = tw(y[3][2], -r2, d * r2);
        y[3][3] = tw(y[3][3], -c1, -d * s1);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            cx_t<T> t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            dft_small<T, 4, DIR>(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) a[k1 + 4 * k2] = t[k2];
        }
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
    int lt, le;   // LDS position of element e of transform b: b*lt + e*le
    int bfast;    // 1: consecutive threads take consecutive transforms (nb = 2^nbsh)
    int nbsh;
    int pad;      // > 0: one padding slot after every 2^pad LDS elements
    int tws;      // twiddle table stride (table length n * tws)
    int nrad;
    int rad[16];
};

// fast-axis indices that carry image pixels: k < a or k >= b
struct Live {
    int a, b;
    __host__ __device__ bool operator()(int k) const { return k < a || k >= b; }
    __host__ __device__ int count(int n) const { return a + (n - b); }
    // k-th live index
    __host__ __device__ int at(int k) const { return k < a ? k : k - a + b; }
};

__device__ __forceinline__ int lds_pos(const FftShape &S, int x) { return S.pad ? x + (x >> S.pad) : x; }

// one Stockham pass of radix R over all nb transforms held in LDS
template <typename T, int R, int DIR>
__device__ __forceinline__ void stockham_pass(cx_t<T> *buf, const FftShape &S, int p,
                                              const cx_t<T> *__restrict__ tw)
{
    constexpr int K = (fft_ept<T>() + R - 1) / R;
    const int Tn = S.n / R;
    const int nbf = S.nb * Tn;
    const int twstep = S.n / (p * R) * S.tws;
    cx_t<T> u[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * blockDim.x;
        if (q < nbf) {
            int b, i;
            if (S.bfast) {
                b = q & (S.nb - 1);
                i = q >> S.nbsh;
            } else {
                b = q / Tn;
                i = q - b * Tn;
            }
#pragma unroll
            for (int r = 0; r < R; r++) u[k][r] = buf[lds_pos(S, b * S.lt + (i + r * Tn) * S.le)];
        }
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * blockDim.x;
        if (q < nbf) {
            int b, i;
            if (S.bfast) {
                b = q & (S.nb - 1);
                i = q >> S.nbsh;
            } else {
                b = q / Tn;
                i = q - b * Tn;
            }
            const int kk = i % p;
#pragma unroll
            for (int r = 1; r < R; r++) {
                cx_t<T> w = tw[(unsigned)(r * kk * twstep)];
                if (DIR < 0) w.y = -w.y;
                u[k][r] = cmul<T>(u[k][r], w);
            }
            dft_small<T, R, DIR>(u[k]);
            const int j = (i - kk) * R + kk;
#pragma unroll
            for (int r = 0; r < R; r++) buf[lds_pos(S, b * S.lt + (j + r * p) * S.le)] = u[k][r];
        }
    }
    lds_barrier();
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

// Rows.  Group g holds rows [g nb, g nb + nb) of the nrows rows of all planes
// (rows are contiguous across planes).  DIR > 0 stores only live outputs;
// DIR < 0 reads only live inputs.
template <typename T, int DIR>
__global__ __launch_bounds__(1024) void k_fft_rows(cx_t<T> *data, int64_t nrows, int n1,
                                                   FftShape S, const cx_t<T> *__restrict__ tw,
                                                   Live live)
{
    constexpr int EPT = fft_ept<T>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    const int E = S.nb * n1;
    const int64_t ngroups = (nrows + S.nb - 1) / S.nb;
    cx_t<T> pre[EPT];
    // elements of group grp that exist (the last group may be short)
    auto count = [&](int64_t grp) -> int {
        return (int)min((int64_t)E, (nrows - grp * S.nb) * n1);
    };
    auto load = [&](int64_t grp) {
        const cx_t<T> *src = data + grp * (int64_t)E;
        const int lim = count(grp);
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * blockDim.x;
            const int e = S.nb == 1 ? k : k % n1;
            pre[u] = (k < lim && (DIR > 0 || live(e))) ? src[k] : mkc((T)0, (T)0);
        }
    };
    for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        load(g);
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * blockDim.x;
            if (k < E) buf[lds_pos(S, k)] = pre[u];
        }
        lds_barrier();
        lds_fft<T, DIR>(buf, S, tw);
        cx_t<T> *dst = data + g * (int64_t)E;
        const int lim = count(g);
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * blockDim.x;
            const int e = S.nb == 1 ? k : k % n1;
            if (k < lim && (DIR < 0 || live(e))) dst[k] = buf[lds_pos(S, k)];
        }
        lds_barrier();
    }
}

// Columns, whole transform per workgroup: blockIdx.x = block of S.nb live
// columns, blockIdx.y = plane.  Natural order in and out.
template <typename T, int DIR>
__global__ __launch_bounds__(1024) void k_fft_cols(cx_t<T> *data, int n2, int n1, FftShape S,
                                                   const cx_t<T> *__restrict__ tw, Live live)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    cx_t<T> *pl = data + (size_t)blockIdx.y * n2 * n1;
    const int nlive = live.count(n1);
    const int c0 = blockIdx.x * S.nb;
    const int E = S.nb * n2;
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int e = k >> S.nbsh, b = k & (S.nb - 1);
        const int c = c0 + b;
        buf[k] = c < nlive ? pl[(size_t)e * n1 + live.at(c)] : mkc((T)0, (T)0);
    }
    lds_barrier();
    lds_fft<T, DIR>(buf, S, tw);
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int e = k >> S.nbsh, b = k & (S.nb - 1);
        const int c = c0 + b;
        if (c < nlive) pl[(size_t)e * n1 + live.at(c)] = buf[k];
    }
}

// One step of the in-place split of long columns.  Sub-transform j
// (blockIdx.y) of length L = S.n reads and writes rows j*sj + e*se, e < L;
// twmode 1 multiplies output e by w^(j e) after the transform, twmode 2
// multiplies input e by w^(j e) before it, w = e^{dir 2 pi i/n2} (tw2 is the
// length-n2 table).  blockIdx.x = block of S.nb live columns, blockIdx.z =
// plane.
template <typename T, int DIR>
__global__ __launch_bounds__(1024) void k_fft_cols_step(cx_t<T> *data, int n2, int n1, FftShape S,
                                                        int se, int sj, int twmode,
                                                        const cx_t<T> *__restrict__ tw2, Live live)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    const int j = blockIdx.y;
    cx_t<T> *pl = data + (size_t)blockIdx.z * n2 * n1;
    const int nlive = live.count(n1);
    const int c0 = blockIdx.x * S.nb;
    const int E = S.nb * S.n;
    const size_t row0 = (size_t)j * sj;
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int e = k >> S.nbsh, b = k & (S.nb - 1);
        const int c = c0 + b;
        cx_t<T> v = c < nlive ? pl[(row0 + (size_t)e * se) * n1 + live.at(c)] : mkc((T)0, (T)0);
        if (twmode == 2) {
            cx_t<T> w = tw2[j * e];
            if (DIR < 0) w.y = -w.y;
            v = cmul<T>(v, w);
        }
        buf[k] = v;
    }
    lds_barrier();
    lds_fft<T, DIR>(buf, S, tw2);
    for (int k = threadIdx.x; k < E; k += blockDim.x) {
        const int e = k >> S.nbsh, b = k & (S.nb - 1);
        const int c = c0 + b;
        if (c >= nlive) continue;
        cx_t<T> v = buf[k];
        if (twmode == 1) {
            cx_t<T> w = tw2[j * e];
            if (DIR < 0) w.y = -w.y;
            v = cmul<T>(v, w);
        }
        pl[(row0 + (size_t)e * se) * n1 + live.at(c)] = v;
    }
}

// ------------------------------------------------------------ compile-time plans
//
// Power-of-two lengths get kernels whose radix plan, strides and thread
// counts are compile-time constants: radix-16 passes for fp32 (one 16-point
// butterfly per thread per pass) or radix-8 passes for fp64, then one
// smaller pass.  Twiddles come from LDS.  A length-n table is split in two
// levels, w^t = H[t / B] * L[t % B] with B = 2^s, so one complex multiply
// replaces a global load.

template <typename T>
struct TwTab {
    const cx_t<T> *H, *L;
    int sh;  // B = 1 << sh
    __device__ __forceinline__ cx_t<T> get(int t) const
    {
        const cx_t<T> h = H[t >> sh], l = L[t & ((1 << sh) - 1)];
        return cmul<T>(h, l);
    }
};

// load the two-level table of e^{2 pi i t / n}, t < n, from the full table
// tw (length n, any n) into LDS: H has ceil(n / B) entries, L has B
template <typename T>
__device__ __forceinline__ TwTab<T> load_twtab(cx_t<T> *lds, const cx_t<T> *__restrict__ tw, int n,
                                               int sh)
{
    const int B = 1 << sh, nh = (n + B - 1) >> sh;
    for (int t = threadIdx.x; t < B; t += blockDim.x) lds[t] = t < n ? tw[t] : mkc((T)1, (T)0);
    for (int h = threadIdx.x; h < nh; h += blockDim.x) lds[B + h] = tw[h << sh];
    return TwTab<T>{lds + B, lds, sh};
}

template <int NB, bool BFAST, int N, bool PAD>
__device__ __forceinline__ int ct_pos(int b, int e)
{
    if constexpr (BFAST) return b + e * NB;
    const int x = b * N + e;
    return PAD ? x + (x >> 5) : x;
}

// one Stockham pass of radix R over NB transforms of length N held in LDS,
// NT threads; twiddle w_N^t read as tab.get(t * ts) (tab may describe a
// longer transform whose length is ts * N)
template <typename T, int DIR, int N, int R, int P, int NB, bool BFAST, bool PAD, int NT>
__device__ __forceinline__ void ct_pass(cx_t<T> *buf, const TwTab<T> &tab, int ts)
{
    constexpr int TN = N / R;
    constexpr int NBF = NB * TN;
    constexpr int K = (NBF + NT - 1) / NT;
    constexpr int TWS = N / (P * R);
    cx_t<T> u[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * NT;
        if (NBF % NT == 0 || q < NBF) {
            const int b = BFAST ? q % NB : q / TN;
            const int i = BFAST ? q / NB : q % TN;
#pragma unroll
            for (int r = 0; r < R; r++) u[k][r] = buf[ct_pos<NB, BFAST, N, PAD>(b, i + r * TN)];
        }
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * NT;
        if (NBF % NT == 0 || q < NBF) {
            const int b = BFAST ? q % NB : q / TN;
            const int i = BFAST ? q / NB : q % TN;
            const int kk = i % P;
            if constexpr (P > 1) {
                // W^r for r = 4 r1 + r2 as (W^4)^r1 W^r2: two table lookups
                cx_t<T> w1 = tab.get(kk * TWS * ts);
                if (DIR < 0) w1.y = -w1.y;
                const cx_t<T> w2 = cmul<T>(w1, w1);
                const cx_t<T> pw[4] = {mkc((T)1, (T)0), w1, w2, cmul<T>(w2, w1)};
                if constexpr (R <= 4) {
#pragma unroll
                    for (int r = 1; r < R; r++) u[k][r] = cmul<T>(u[k][r], pw[r]);
                } else {
                    cx_t<T> w4 = tab.get(4 * kk * TWS * ts);
                    if (DIR < 0) w4.y = -w4.y;
                    const cx_t<T> w8 = cmul<T>(w4, w4);
                    const cx_t<T> qw[4] = {mkc((T)1, (T)0), w4, w8, cmul<T>(w8, w4)};
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        cx_t<T> v = u[k][r];
                        if (r & 3) v = cmul<T>(v, pw[r & 3]);
                        if (r >> 2) v = cmul<T>(v, qw[r >> 2]);
                        u[k][r] = v;
                    }
                }
            }
            dft_small<T, R, DIR>(u[k]);
            const int j = (i - kk) * R + kk;
#pragma unroll
            for (int r = 0; r < R; r++) buf[ct_pos<NB, BFAST, N, PAD>(b, j + r * P)] = u[k][r];
        }
    }
    lds_barrier();
}

// all passes: radix RMAX while possible, then the remainder
template <typename T, int DIR, int N, int P, int NB, bool BFAST, bool PAD, int NT, int RMAX>
__device__ __forceinline__ void ct_fft(cx_t<T> *buf, const TwTab<T> &tab, int ts)
{
    if constexpr (P < N) {
        constexpr int R = (N / P) >= RMAX ? RMAX : (N / P);
        ct_pass<T, DIR, N, R, P, NB, BFAST, PAD, NT>(buf, tab, ts);
        ct_fft<T, DIR, N, P * R, NB, BFAST, PAD, NT, RMAX>(buf, tab, ts);
    }
}

template <typename T>
constexpr int ct_rmax() { return sizeof(T) == 4 ? 16 : 8; }

// one row of N held in padded LDS (NB = 1): every pass but the last, then
// the last pass from LDS into registers, its outputs x = i + r P handed to
// out(x, value) instead of written back (P = N / R is a multiple of NT, so
// the thread's outputs are its own columns x = threadIdx.x + u NT)
template <typename T, int DIR, int N, int P, int NT, int RMAX, typename F>
__device__ __forceinline__ void ct_fft_row_out(cx_t<T> *buf, const TwTab<T> &tab, F out)
{
    constexpr int R = (N / P) >= RMAX ? RMAX : (N / P);
    if constexpr (P * R < N) {
        ct_pass<T, DIR, N, R, P, 1, false, true, NT>(buf, tab, 1);
        ct_fft_row_out<T, DIR, N, P * R, NT, RMAX>(buf, tab, out);
    } else {
        static_assert(P > 1 && P % NT == 0, "last pass layout");
        constexpr int K = P / NT;
        cx_t<T> u[K][R];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = (int)threadIdx.x + k * NT;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int x = i + r * P;
                u[k][r] = buf[x + (x >> 5)];
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = (int)threadIdx.x + k * NT;
            // W^r for r = 4 r1 + r2 as (W^4)^r1 W^r2, W = w_N^i
            cx_t<T> w1 = tab.get(i);
            if (DIR < 0) w1.y = -w1.y;
            const cx_t<T> w2 = cmul<T>(w1, w1);
            const cx_t<T> pw[4] = {mkc((T)1, (T)0), w1, w2, cmul<T>(w2, w1)};
            if constexpr (R <= 4) {
#pragma unroll
                for (int r = 1; r < R; r++) u[k][r] = cmul<T>(u[k][r], pw[r]);
            } else {
                cx_t<T> w4 = tab.get(4 * i);
                if (DIR < 0) w4.y = -w4.y;
                const cx_t<T> w8 = cmul<T>(w4, w4);
                const cx_t<T> qw[4] = {mkc((T)1, (T)0), w4, w8, cmul<T>(w8, w4)};
#pragma unroll
                for (int r = 1; r < R; r++) {
                    cx_t<T> v = u[k][r];
                    if (r & 3) v = cmul<T>(v, pw[r & 3]);
                    if (r >> 2) v = cmul<T>(v, qw[r >> 2]);
                    u[k][r] = v;
                }
            }
            dft_small<T, R, DIR>(u[k]);
#pragma unroll
            for (int r = 0; r < R; r++) out(i + r * P, u[k][r]);
        }
    }
}

// rows of N = 2^LOGN points, one row per workgroup, N / EPT threads,
// persistent over the nrows rows of all planes; rows no point reaches are
// skipped.  PF: the next live row is loaded into registers while the current
// one is transformed; it is found 64 candidate rows at a time (each lane of
// a wave reads one mask byte and a ballot names the first live row, the same
// in every wave).
template <typename T, int DIR, int LOGN, bool PF>
__global__ __launch_bounds__((1 << LOGN) / fft_ept<T>()) void k_fft_rows_ct(cx_t<T> *data, int64_t nrows,
                                                      const cx_t<T> *__restrict__ tw, Live live,
                                                      const uint8_t *__restrict__ rowmask, int n2)
{
    constexpr int N = 1 << LOGN;
    constexpr int EPT = fft_ept<T>();
    constexpr int NT = N / EPT;
    constexpr int SH = LOGN / 2;
    constexpr int E = N + (N >> 5);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    const TwTab<T> tab = load_twtab<T>(buf + E, tw, N, SH);
    // rowmask[n2 + x] marks the columns points reach: the others are zero on
    // input (dir > 0; they need not have been written) or unread on output
    // (dir < 0; they are not stored).  EPT <= 32 columns per thread.
    // this thread's columns k = threadIdx.x + u NT are the same in every row:
    // their mask bits are read once
    uint32_t cbits = 0xffffffffu;
    if (rowmask) {
        cbits = 0;
#pragma unroll
        for (int u = 0; u < EPT; u++)
            if (rowmask[n2 + threadIdx.x + u * NT]) cbits |= 1u << u;
    }
    // live bits of the workgroup's 64 rows g + j gridDim.x, j < 64 (a masked
    // row is zero for dir > 0 and unread for dir < 0).  Rows are numbered in
    // 32 bits (the host launches the PF kernel for fewer than 2^30 rows); the
    // lane and n2 are made opaque so that no per-lane value (the lane offset,
    // the reciprocal of n2) is hoisted and kept live across the transform.
    auto probe = [&](int64_t g) -> uint64_t {
        int lane = threadIdx.x & 63, m = n2;
        asm volatile("" : "+v"(lane), "+v"(m));
        const int c = (int)g + lane * (int)gridDim.x;
        return __ballot(c < (int)nrows && rowmask[(unsigned)c % (unsigned)m] != 0);
    };
    auto load = [&](int64_t g, cx_t<T> *dst) {
        const cx_t<T> *src = data + g * (int64_t)N;
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * NT;
            const bool ld = DIR > 0 ? ((cbits >> u) & 1u) != 0 : live(k);
            dst[u] = ld ? src[k] : mkc((T)0, (T)0);
        }
    };
    // the first Stockham pass (radix EPT, no twiddles) reads elements
    // threadIdx.x + u NT, the ones this thread loaded: it runs in registers
    // and writes its outputs EPT threadIdx.x + r to LDS
    static_assert(EPT == ct_rmax<T>(), "first pass radix");
    auto first_pass = [&](cx_t<T> *v) {
        dft_small<T, EPT, DIR>(v);
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int x = (int)threadIdx.x * EPT + r;
            buf[x + (x >> 5)] = v[r];
        }
    };
    // the remaining passes; the last one stores from registers (output x is
    // column threadIdx.x + u NT of this thread, u = x / NT)
    auto transform_store = [&](int64_t g) {
        lds_barrier();
        cx_t<T> *dst = data + g * (int64_t)N;
        ct_fft_row_out<T, DIR, N, EPT, NT, ct_rmax<T>()>(buf, tab, [&](int x, cx_t<T> v) {
            const bool st = DIR < 0 ? ((cbits >> (x / NT)) & 1u) != 0 : live(x);
            if (st) dst[x] = v;
        });
        lds_barrier();
    };
    if constexpr (PF) {
        // the first live row at or after g in steps of gridDim.x (nrows if none)
        auto next_live = [&](int64_t g) -> int64_t {
            if (!rowmask) return g;
            for (; g < nrows; g += 64 * (int64_t)gridDim.x) {
                const uint64_t on = probe(g);
                if (on) return g + (int64_t)__builtin_ctzll(on) * gridDim.x;
            }
            return nrows;
        };
        cx_t<T> pre[EPT];
        int64_t g = next_live(blockIdx.x);
        if (g < nrows) load(g, pre);
        while (g < nrows) {
            first_pass(pre);
            const int64_t gn = next_live(g + gridDim.x);
            if (gn < nrows) load(gn, pre);
            transform_store(g);
            g = gn;
        }
    } else {
        for (int64_t g = blockIdx.x; g < nrows; g += gridDim.x) {
            if (rowmask && !rowmask[g % n2]) continue;
            cx_t<T> v[EPT];
            load(g, v);
            first_pass(v);
            transform_store(g);
        }
    }
}

// One step of the split column FFT with sub-transform length L = 2^LOGL on
// CB adjacent live columns (see k_fft_cols_step; tw2 is the length-n2 table)
// as a persistent kernel: a workgroup walks tiles (column
// block x fastest, then sub-transform j, then plane) and loads the next
// tile into registers while the current one is transformed in LDS and
// stored, so the loads in flight are not limited to what fits in LDS.
template <typename T, int DIR, int LOGL, int CB>
__global__ __launch_bounds__(CB * (1 << LOGL) / fft_ept<T>()) void k_fft_cols_step_pf(
    cx_t<T> *data, int n2, int n1, int se, int sj, int nsub, int batch, int twmode,
    const cx_t<T> *__restrict__ tw2, Live live, const uint8_t *__restrict__ rowmask, int mode_in,
    int mode_out, int spec_a, int spec_b)
{
    constexpr int L = 1 << LOGL;
    constexpr int EPT = fft_ept<T>();
    constexpr int E = CB * L;
    constexpr int NT = E / EPT;
    constexpr int SH2 = 7;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *buf = reinterpret_cast<cx_t<T> *>(smem_raw);
    const TwTab<T> tab = load_twtab<T>(buf + E, tw2, n2, SH2);
    lds_barrier();
    const int nlive = live.count(n1);
    const int nx = (nlive + CB - 1) / CB;
    const int64_t ntile = (int64_t)nx * nsub * batch;
    const int tstride = n2 >> LOGL;
    auto dead = [&](int j, size_t row, int e) -> bool {
        if (sj == 1) return rowmask && !rowmask[row];
        const int k = j + nsub * e;
        return k >= spec_a && k < spec_b;
    };
    cx_t<T> pre[EPT];
    auto load = [&](int64_t t) {
        const int x = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % nsub);
        const cx_t<T> *pl = data + (size_t)(r / nsub) * n2 * n1;
        const size_t row0 = (size_t)j * sj;
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * NT;
            const int e = k / CB, b = k % CB;
            const int c = x * CB + b;
            const size_t row = row0 + (size_t)e * se;
            const bool rd = c < nlive && !(mode_in && dead(j, row, e));
            pre[u] = rd ? pl[row * n1 + live.at(c)] : mkc((T)0, (T)0);
        }
    };
    int64_t t = blockIdx.x;
    if (t < ntile) load(t);
    for (; t < ntile; t += gridDim.x) {
        const int x = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % nsub);
        cx_t<T> *pl = data + (size_t)(r / nsub) * n2 * n1;
        const size_t row0 = (size_t)j * sj;
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * NT;
            cx_t<T> v = pre[u];
            if (twmode == 2) {
                cx_t<T> w = tab.get(j * (k / CB));
                if (DIR < 0) w.y = -w.y;
                v = cmul<T>(v, w);
            }
            buf[k] = v;
        }
        if (t + gridDim.x < ntile) load(t + gridDim.x);
        lds_barrier();
        ct_fft<T, DIR, L, 1, CB, true, false, NT, ct_rmax<T>()>(buf, tab, tstride);
#pragma unroll
        for (int u = 0; u < EPT; u++) {
            const int k = threadIdx.x + u * NT;
            const int e = k / CB, b = k % CB;
            const int c = x * CB + b;
            if (c >= nlive) continue;
            const size_t row = row0 + (size_t)e * se;
            if (mode_out && dead(j, row, e)) continue;
            cx_t<T> v = buf[k];
            if (twmode == 1) {
                cx_t<T> w = tab.get(j * e);
                if (DIR < 0) w.y = -w.y;
                v = cmul<T>(v, w);
            }
            pl[row * n1 + live.at(c)] = v;
        }
        lds_barrier();
    }
}

// Two-pass form of the persistent step for fp32 and L = 16 R2 (R2 = 2..16):
// thread t holds column b = t % CB at rows e0 + (L/16) u, u < 16
// (e0 = t / CB), which are exactly the inputs of the first radix-16
// Stockham butterfly (P = 1, no twiddles).  That butterfly runs in
// registers on the loaded tile and goes to LDS once; the second (last)
// pass, radix R2 with P = 16, reads it back and its outputs go straight
// from registers to HBM: one LDS write and one read per element.
template <int DIR, int LOGL, int CB>
__global__ __launch_bounds__(CB * (1 << LOGL) / 16, CB % 64 == 0 ? 4 : 2) void k_fft_cols_step_r2(
    float2 *data, int n2, int n1, int se, int sj, int nsub, int batch, int twmode,
    const float2 *__restrict__ tw2, Live live, const uint8_t *__restrict__ rowmask, int mode_in,
    int mode_out, int spec_a, int spec_b)
{
    using T = float;
    constexpr int L = 1 << LOGL;
    constexpr int R2 = L / 16;
    constexpr int E = CB * L;
    constexpr int NT = E / 16;
    constexpr int K2 = 16 * CB / NT;  // last-pass butterflies per thread
    constexpr int SH2 = 7;
    static_assert(R2 >= 2 && R2 <= 16 && NT % CB == 0 && (16 * CB) % NT == 0, "shape");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *buf = reinterpret_cast<float2 *>(smem_raw);
    const TwTab<T> tab = load_twtab<T>(buf + E, tw2, n2, SH2);
    lds_barrier();
    const int nlive = live.count(n1);
    const int nx = (nlive + CB - 1) / CB;
    const int64_t ntile = (int64_t)nx * nsub * batch;
    const int ts = n2 >> LOGL;
    const int b0 = threadIdx.x % CB, e0 = threadIdx.x / CB;
    auto dead = [&](int j, size_t row, int e) -> bool {
        if (sj == 1) return rowmask && !rowmask[row];
        const int k = j + nsub * e;
        return k >= spec_a && k < spec_b;
    };
    // with 64 columns per tile a wavefront covers one row per load, so the
    // row (and its mask byte) is wave-uniform and stays in scalar registers
    const int e0w = CB % 64 == 0 ? __builtin_amdgcn_readfirstlane(e0) : e0;
    float2 pre[16];
    auto load = [&](int64_t t) {
        const int x = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % nsub);
        const float2 *pl = data + (size_t)(r / nsub) * n2 * n1;
        const size_t row0 = (size_t)j * sj;
        const int c = x * CB + b0;
        const bool cok = c < nlive;
        const int lc = live.at(c);
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = e0w + u * R2;
            const size_t row = row0 + (size_t)e * se;
            const bool rd = cok && !(mode_in && dead(j, row, e));
            pre[u] = rd ? pl[row * n1 + lc] : make_float2(0.f, 0.f);
        }
    };
    int64_t t = blockIdx.x;
    if (t < ntile) load(t);
    for (; t < ntile; t += gridDim.x) {
        const int x = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % nsub);
        float2 *pl = data + (size_t)(r / nsub) * n2 * n1;
        const size_t row0 = (size_t)j * sj;
        float2 u1[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            float2 v = pre[u];
            if (twmode == 2) {
                float2 w = tab.get(j * (e0 + u * R2));
                if (DIR < 0) w.y = -w.y;
                v = cmul<T>(v, w);
            }
            u1[u] = v;
        }
        if (t + gridDim.x < ntile) load(t + gridDim.x);
        // first pass: butterfly i = e0, outputs at 16 e0 + r
        dft_small<T, 16, DIR>(u1);
#pragma unroll
        for (int rr = 0; rr < 16; rr++) buf[(16 * e0 + rr) * CB + b0] = u1[rr];
        lds_barrier();
        // last pass: butterfly i < 16 of column b, inputs i + 16 r, twiddle
        // w_L^(i r), outputs at rows i + 16 r
#pragma unroll
        for (int k = 0; k < K2; k++) {
            const int q = threadIdx.x + k * NT;
            const int b = q % CB;
            const int i = CB % 64 == 0 ? __builtin_amdgcn_readfirstlane(q / CB) : q / CB;
            float2 v[R2];
#pragma unroll
            for (int rr = 0; rr < R2; rr++) v[rr] = buf[(i + 16 * rr) * CB + b];
            float2 w1 = tab.get(i * ts);
            if (DIR < 0) w1.y = -w1.y;
            const float2 w2 = cmul<T>(w1, w1);
            const float2 pw[4] = {make_float2(1.f, 0.f), w1, w2, cmul<T>(w2, w1)};
            if constexpr (R2 <= 4) {
#pragma unroll
                for (int rr = 1; rr < R2; rr++) v[rr] = cmul<T>(v[rr], pw[rr]);
            } else {
                float2 w4 = tab.get(4 * i * ts);
                if (DIR < 0) w4.y = -w4.y;
                const float2 w8 = cmul<T>(w4, w4);
                const float2 qw[4] = {make_float2(1.f, 0.f), w4, w8, cmul<T>(w8, w4)};
#pragma unroll
                for (int rr = 1; rr < R2; rr++) {
                    float2 z = v[rr];
                    if (rr & 3) z = cmul<T>(z, pw[rr & 3]);
                    if (rr >> 2) z = cmul<T>(z, qw[rr >> 2]);
                    v[rr] = z;
                }
            }
            dft_small<T, R2, DIR>(v);
            const int c = x * CB + b;
            if (c < nlive) {
#pragma unroll
                for (int rr = 0; rr < R2; rr++) {
                    const int e = i + 16 * rr;
                    const size_t row = row0 + (size_t)e * se;
                    if (mode_out && dead(j, row, e)) continue;
                    float2 o = v[rr];
                    if (twmode == 1) {
                        float2 w = tab.get(j * e);
                        if (DIR < 0) w.y = -w.y;
                        o = cmul<T>(o, w);
                    }
                    pl[row * n1 + live.at(c)] = o;
                }
            }
        }
        lds_barrier();
    }
}

// Band-limited columns, fp32.  When the non-zero rows of the input (forward)
// or the rows wanted from the output (inverse) lie in one cyclic band of
// B <= M rows (M = 1024 or 2048), a length-n2 column transform with
// P = n2 / M is P transforms of length M:
//   forward  X[ka + P kb] = FFT_M over r of  y_ka[r] = x[n_r] w^(n_r ka)
//   inverse  x[n] = w^(n ka) summed over ka of IFFT_M(X[ka + P kb] over kb)[n mod M]
// (w = e^(dir 2 pi i / n2), n_r the band row of residue r).  A workgroup
// holds CB adjacent live columns (16 for M = 1024, 8 for M = 2048, so the
// transform buffer is 128 KB of LDS, next to a 16 KB table of the length-M
// roots the passes use); thread (e, b) keeps the band rows of
// residues e + S j (S = 1024 / CB, j < 16) of column b in registers.  Those
// residues are the inputs of the first radix-16 Stockham butterfly, which
// therefore runs in registers; the other two passes run in LDS.  The
// spectrum is read / written in the split order of the four-step column
// pass (frequency k = ka + na kb at row ka nb + kb), live frequencies only.
// No intermediate spectrum goes through HBM: the input is read once and the
// output written once.
constexpr int kBandTwSh = 7;  // two-level twiddle table: 128 low + n2 / 128 high entries

// one Stockham pass of radix R (stage p = PP) over the CB columns of length
// M in LDS, 1024 threads; mt[t] = e^{2 pi i t / M} (an LDS table)
template <int R, int PP, int DIR, int M, int CB>
__device__ __forceinline__ void band_pass(float2 *buf, const float2 *mt)
{
    constexpr int Tn = M / R, K = CB * Tn / 1024, LCB = CB == 16 ? 4 : 3;
    static_assert(K >= 1 && CB * Tn % 1024 == 0 && (1 << LCB) == CB, "shape");
    // addresses are recomputed per call rather than held across the caller's
    // loop (hidden from loop-invariant code motion)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    float2 u[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = tid + 1024 * k, b = q & (CB - 1), i = q >> LCB;
#pragma unroll
        for (int r = 0; r < R; r++) u[k][r] = buf[(i + r * Tn) * CB + b];
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = tid + 1024 * k, b = q & (CB - 1), i = q >> LCB;
        const int kk = i % PP;
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = mt[r * kk * (M / (PP * R))];
            if (DIR < 0) w.y = -w.y;
            u[k][r] = cmul<float>(u[k][r], w);
        }
        dft_small<float, R, DIR>(u[k]);
        const int j = (i - kk) * R + kk;
#pragma unroll
        for (int r = 0; r < R; r++) buf[(j + r * PP) * CB + b] = u[k][r];
    }
    lds_barrier();
}

// the last Stockham pass (radix R, stage PP = M / R), handing the outputs
// of each butterfly to out(k, j, u) instead of writing them back to LDS:
// u[r] (r < R) is element j + PP r of the transform, slot k R + r of the
// thread (k its butterfly, j = its butterfly index i); the column of a
// thread's butterflies is its own (1024 is a multiple of CB).  The caller
// may rewrite buf after this returns (every read precedes its barrier).
template <int R, int PP, int DIR, int M, int CB, typename F>
__device__ __forceinline__ void band_pass_out(const float2 *buf, const float2 *mt, F &&out)
{
    constexpr int Tn = M / R, K = CB * Tn / 1024, LCB = CB == 16 ? 4 : 3;
    static_assert(K >= 1 && CB * Tn % 1024 == 0 && (1 << LCB) == CB, "shape");
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    float2 u[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = tid + 1024 * k, b = q & (CB - 1), i = q >> LCB;
#pragma unroll
        for (int r = 0; r < R; r++) u[k][r] = buf[(i + r * Tn) * CB + b];
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = tid + 1024 * k, i = q >> LCB;
        const int kk = i % PP;
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = mt[r * kk * (M / (PP * R))];
            if (DIR < 0) w.y = -w.y;
            u[k][r] = cmul<float>(u[k][r], w);
        }
        dft_small<float, R, DIR>(u[k]);
        out(k, (i - kk) * R + kk, u[k]);
    }
}

template <int DIR, int M>
__global__ __launch_bounds__(1024) void k_fft_cols_band(float2 *data, int n2, int n1, int np,
                                                        int bs, int blen, Live live,
                                                        const uint8_t *__restrict__ rowmask,
                                                        int nash, int nb, int spec_a, int spec_b,
                                                        const float2 *__restrict__ tw2)
{
    using T = float;
    constexpr int CB = M == 1024 ? 16 : 8, S = 1024 / CB, LCB = CB == 16 ? 4 : 3;
    static_assert(M == 1024 || M == 2048, "band length");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *buf = reinterpret_cast<float2 *>(smem_raw);
    const int nlive = live.count(n1);
    // column block: workgroups w and w + 8 (the same XCD when workgroups are
    // dealt to the 8 XCDs round-robin) take adjacent blocks, so the two
    // halves of a 128-byte row segment (8 columns each at M = 2048) meet in
    // one L2
    const int w = blockIdx.x;
    const int cb = M == 2048 ? 2 * (8 * (w >> 4) + (w & 7)) + ((w >> 3) & 1) : w;
    if (cb * CB >= nlive) return;  // padding workgroups (uniform, before any barrier)
    const TwTab<T> tab = load_twtab<T>(buf + M * CB + M, tw2, n2, kBandTwSh);
    // the length-M roots of the passes, one LDS read per twiddle
    float2 *mt = buf + M * CB;
    const int tws = n2 / M;
    for (int t = threadIdx.x; t < M; t += blockDim.x) mt[t] = tw2[t * tws];
    const int tid = threadIdx.x, b = tid & (CB - 1), e = tid >> LCB;
    const int c = cb * CB + b;
    const bool cok = c < nlive;
    const int lc = cok ? live.at(c) : 0;
    float2 *pl = data + (size_t)blockIdx.y * n2 * n1;
    const int mask = n2 - 1;
    // split-order row of frequency k (na = 2^nash)
    auto srow = [&](int k) { return (k & ((1 << nash) - 1)) * nb + (k >> nash); };
    auto twid = [&](int n, int ka) {
        float2 w = tab.get((n * ka) & mask);
        if (DIR < 0) w.y = -w.y;
        return w;
    };
    // spectrum addresses: a uniform base per output r (per input m) plus one
    // 32-bit byte offset per lane (a plane is < 4 GiB, checked at launch), so
    // that the loads and stores take the scalar-base form; dead frequencies
    // k are those with (unsigned)(k - spec_a) < dw
    char *const plb = reinterpret_cast<char *>(pl);
    const unsigned dw = spec_b > spec_a ? (unsigned)(spec_b - spec_a) : 0u;
    auto lane_off = [&](int k) {
        return ((uint32_t)srow(k) * (uint32_t)n1 + (uint32_t)lc) * (uint32_t)sizeof(float2);
    };
    // residue e + S j of this thread: band offset o = (e + S j - bs) mod M,
    // row (bs + o) mod n2 when o < blen
    float2 xb[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int o = (e + S * j - bs) & (M - 1), n = (bs + o) & mask;
        xb[j] = make_float2(0.f, 0.f);
        if (DIR > 0 && o < blen && cok && (!rowmask || rowmask[n])) xb[j] = pl[(size_t)n * n1 + lc];
    }
    // inverse: input m of sub-transform ka is frequency k0 + m F (k0 = ka +
    // np e, F = np S, a multiple of the split factor na), at split-order row
    // srow(k0) + m (F / na); the inputs of ka + 1 are loaded while ka is
    // transformed (the loop's barriers would otherwise wait on them)
    float2 vn[16];
    auto load_inputs = [&](int ka, int eo) {
        const int k0 = ka + np * eo, F = np * S;
        const uint32_t boff = lane_off(k0);
        const size_t mstride = (size_t)(F >> nash) * n1 * sizeof(float2);
        const int d0 = k0 - spec_a;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            vn[m] = make_float2(0.f, 0.f);
            if (cok && (unsigned)(d0 + m * F) >= dw)
                vn[m] = *reinterpret_cast<const float2 *>(plb + m * mstride + boff);
        }
    };
    if constexpr (DIR < 0) load_inputs(0, e);
    lds_barrier();  // twiddle table
    for (int ka = 0; ka < np; ka++) {
        // recomputing the band rows every iteration is cheaper than keeping
        // their indices live across the loop: hide e from loop-invariant
        // code motion
        int eo = e;
        asm volatile("" : "+v"(eo));
        // v[j]: element e + S j of the length-M input, the inputs of the
        // first radix-16 Stockham butterfly (no twiddles), run in registers
        float2 v[16];
        if (DIR > 0) {
#pragma unroll
            for (int j = 0; j < 16; j++) v[j] = cmul<T>(xb[j], twid(bs + ((eo + S * j - bs) & (M - 1)), ka));
        } else {
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = vn[m];
            if (ka + 1 < np) load_inputs(ka + 1, eo);
        }
        dft_small<T, 16, DIR>(v);
#pragma unroll
        for (int r = 0; r < 16; r++) buf[(16 * eo + r) * CB + b] = v[r];
        lds_barrier();
        band_pass<16, 16, DIR, M, CB>(buf, mt);
        if constexpr (DIR > 0) {
            // the last pass stores its live outputs straight to the spectrum.
            // Output r of a butterfly is frequency k0 + np 256 r; np 256 =
            // n2 / 8 (M = 2048) or n2 / 4 (M = 1024) is a multiple of the
            // split factor na (powers of two, na <= 256 <= n2 / 8), so its
            // split-order row is srow(k0) + r np 256 / na
            constexpr int R = M == 1024 ? 4 : 8;
            const int T = np * 256;
            const size_t rstride = (size_t)(T >> nash) * n1 * sizeof(float2);
            auto store = [&](int, int j, const float2 *u) {
                const int k0 = ka + np * j;
                const uint32_t boff = lane_off(k0);
                const int d0 = k0 - spec_a;
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (cok && (unsigned)(d0 + r * T) >= dw)
                        store_nt(reinterpret_cast<float2 *>(plb + r * rstride + boff), u[r]);
            };
            if constexpr (M == 1024) band_pass_out<4, 256, DIR, M, CB>(buf, mt, store);
            else band_pass_out<8, 256, DIR, M, CB>(buf, mt, store);
            continue;  // buf was read before the pass's barrier
        } else {
            // the last pass's outputs (residue i + 256 r of butterfly i, slot
            // k R + r) are accumulated where they are computed: this thread
            // owns the band rows of those residues
            constexpr int R = M == 1024 ? 4 : 8;
            auto acc = [&](int kb, int j, const float2 *u) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int rr = j + 256 * r;
                    const float2 z = cmul<T>(u[r], twid(bs + ((rr - bs) & (M - 1)), ka));
                    xb[kb * R + r].x += z.x;
                    xb[kb * R + r].y += z.y;
                }
            };
            if constexpr (M == 1024) band_pass_out<4, 256, DIR, M, CB>(buf, mt, acc);
            else band_pass_out<8, 256, DIR, M, CB>(buf, mt, acc);
        }
    }
    if constexpr (DIR < 0) {
        // rows of the residues owned in the last pass: i + 256 r for the
        // thread's butterflies i = (tid + 1024 k) / CB
        constexpr int R = M == 1024 ? 4 : 8, K = 16 / R;
        if (cok) {
#pragma unroll
            for (int k = 0; k < K; k++)
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int rr = ((tid + 1024 * k) >> LCB) + 256 * r;
                    const int o = (rr - bs) & (M - 1);
                    if (o < blen) pl[(size_t)((bs + o) & mask) * n1 + lc] = xb[k * R + r];
                }
        }
    }
}

// split order <-> natural order of the column axis, out of place:
// to_natural: dst[k] = src[pos(k)], else dst[pos(k)] = src[k]
template <typename T>
__global__ __launch_bounds__(256) void k_col_permute(const cx_t<T> *__restrict__ src,
                                                     cx_t<T> *__restrict__ dst, int64_t total,
                                                     int n2, int n1, int na, int nb, int to_natural)
{
    const int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (idx >= total) return;
    const int64_t rowg = idx / n1;
    const int c = (int)(idx - rowg * n1);
    const int64_t plane = rowg / n2;
    const int k = (int)(rowg - plane * n2);
    const int pos = (k % na) * nb + k / na;
    const int64_t base = plane * n2;
    if (to_natural) dst[idx] = src[(base + pos) * n1 + c];
    else dst[(base + pos) * n1 + c] = src[idx];
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

static int round64(int x) { return (x + 63) / 64 * 64; }
static int ilog2(int x)
{
    int l = 0;
    while ((1 << (l + 1)) <= x) l++;
    return l;
}

// columns up to this length are transformed whole; longer ones are split
constexpr int kColWhole = 128;
constexpr int kColSubMax = 256;
constexpr int kLdsMax = 163840;

bool fft_col_split(int n, int &na, int &nb)
{
    na = nb = 0;
    if (n <= kColWhole) return false;
    int best = 0;
    for (int a = 2; a <= kColSubMax; a++) {
        if (n % a) continue;
        const int b = n / a;
        if (b > kColSubMax) continue;
        if (best == 0 || std::abs(a - b) < std::abs(best - n / best)) best = a;
    }
    if (!best) return false;
    na = best;
    nb = n / best;
    return true;
}

// tuning switches read from the environment
static int fft_knob(const char *name, int dflt)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// rows: row padded in LDS by one slot per 32 elements (bank spread of the
// strided Stockham writes)
template <typename T>
static int row_lds_bytes(int n1, int nb)
{
    const int E = nb * n1;
    return (E + (E >> 5)) * (int)sizeof(cx_t<T>);
}

// a row transform held whole in one workgroup's LDS
template <typename T>
static bool fft_row_whole(int n)
{
    int rad[32];
    if (n < 1 || factor_radices(n, rad) < 0) return false;
    return n <= 1024 * fft_ept<T>() && row_lds_bytes<T>(n, 1) <= kLdsMax;
}

template <typename T>
bool fft_row_long(int n)
{
    // complex128 rows beyond one workgroup (n > 8192): one radix-2 step in
    // HBM around two half-length whole-row transforms; the columns of such
    // a grid go through the four-step split
    int na, nb, rad[32];
    return sizeof(T) == 8 && !fft_row_whole<T>(n) && n % 2 == 0 && n <= 16384 &&
           fft_row_whole<T>(n / 2) && fft_col_split(n, na, nb) && factor_radices(na, rad) >= 0 &&
           factor_radices(nb, rad) >= 0;
}

template <typename T>
bool fft_supported(int n)
{
    return fft_row_whole<T>(n) || fft_row_long<T>(n);
}

static int log2_exact(int n)
{
    if (n <= 0 || (n & (n - 1))) return -1;
    return ilog2(n);
}

template <typename T, int DIR, int LOGN, bool PF>
static int rows_ct_launch(cx_t<T> *data, int64_t nrows, const cx_t<T> *tw, Live live,
                          const uint8_t *rowmask, int n2, size_t lds, hipStream_t s)
{
    constexpr int NT = (1 << LOGN) / fft_ept<T>();
    if (lds > 65536)
        CG_HIP(hipFuncSetAttribute((const void *)k_fft_rows_ct<T, DIR, LOGN, PF>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = std::max<int>(1, (int)(kLdsMax / lds));
    per_cu = std::min(per_cu, std::max(1, 2048 / NT));
    const int64_t grid = std::min<int64_t>(nrows, (int64_t)256 * per_cu);
    hipLaunchKernelGGL((k_fft_rows_ct<T, DIR, LOGN, PF>), dim3((unsigned)grid), dim3(NT), lds, s, data,
                       nrows, tw, live, rowmask, n2);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// compile-time row kernel for n1 = 2^LOGN (returns 1 if LOGN is not covered)
template <typename T, int DIR, int LOGN>
static int rows_ct(cx_t<T> *data, int64_t nrows, const cx_t<T> *tw, Live live,
                   const uint8_t *rowmask, int n2, hipStream_t s)
{
    constexpr int N = 1 << LOGN;
    constexpr int NT = N / fft_ept<T>();
    if constexpr (NT < 64 || NT > 1024) {
        return 1;
    } else {
        constexpr int SH = LOGN / 2;
        const size_t lds = (size_t)(N + (N >> 5) + (1 << SH) + (N >> SH)) * sizeof(cx_t<T>);
        if (lds > (size_t)kLdsMax) return 1;
        // the prefetched row takes 2 EPT more VGPRs per thread: only where
        // the workgroup is at most 512 threads (up to 256 VGPRs each)
        constexpr bool pf_fits = NT <= 512;
        if (pf_fits && fft_knob("CG_FFT_ROWPF", 1) && nrows < (int64_t)1 << 30)
            return rows_ct_launch<T, DIR, LOGN, pf_fits>(data, nrows, tw, live, rowmask, n2, lds, s);
        return rows_ct_launch<T, DIR, LOGN, false>(data, nrows, tw, live, rowmask, n2, lds, s);
    }
}

template <typename T, int DIR>
static int rows_ct_dispatch(cx_t<T> *data, int64_t nrows, int n1, const cx_t<T> *tw, Live live,
                            const uint8_t *m, int n2, hipStream_t s)
{
    switch (log2_exact(n1)) {
    case 9: return rows_ct<T, DIR, 9>(data, nrows, tw, live, m, n2, s);
    case 10: return rows_ct<T, DIR, 10>(data, nrows, tw, live, m, n2, s);
    case 11: return rows_ct<T, DIR, 11>(data, nrows, tw, live, m, n2, s);
    case 12: return rows_ct<T, DIR, 12>(data, nrows, tw, live, m, n2, s);
    case 13: return rows_ct<T, DIR, 13>(data, nrows, tw, live, m, n2, s);
    case 14: return rows_ct<T, DIR, 14>(data, nrows, tw, live, m, n2, s);
    default: return 1;
    }
}

// compile-time column step for sub-transforms of L = 2^LOGL (1 if not covered)
struct DeadRows {
    const uint8_t *rowmask;  // natural-order rows
    int spec_a, spec_b;      // live split-order frequencies: k < spec_a or k >= spec_b
    int mode_in, mode_out;   // skip loads of / stores to dead rows
};

template <typename T, int DIR, int LOGL, int CB>
static int cols_step_ct_cb(cx_t<T> *data, int batch, int n2, int n1, int se, int sj, int nsub,
                           int twmode, const cx_t<T> *tw2, Live live, const DeadRows &dr,
                           hipStream_t s)
{
    constexpr int L = 1 << LOGL;
    constexpr int ELEM = (int)sizeof(cx_t<T>);
    constexpr int NT = CB * L / fft_ept<T>();
    if constexpr (NT < 64 || NT > 1024) {
        return 1;
    } else {
        const size_t lds = (size_t)(CB * L + 128 + (n2 + 127) / 128) * ELEM;
        const int nlive = live.count(n1);
        if constexpr (sizeof(T) == 4 && LOGL >= 5 && LOGL <= 8) {
            if (CB * L / 16 == NT) {
                const size_t lds2 = (size_t)(CB * L + 128 + (n2 + 127) / 128) * ELEM;
                if (lds2 > 65536)
                    CG_HIP(hipFuncSetAttribute((const void *)k_fft_cols_step_r2<DIR, LOGL, CB>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds2));
                const int64_t ntile = (int64_t)((nlive + CB - 1) / CB) * nsub * batch;
                const int per_cu = std::max(1, std::min((int)(kLdsMax / lds2), 2048 / NT));
                const int64_t grid = std::min<int64_t>(ntile, (int64_t)256 * per_cu);
                hipLaunchKernelGGL((k_fft_cols_step_r2<DIR, LOGL, CB>), dim3((unsigned)grid),
                                   dim3(NT), lds2, s, (float2 *)data, n2, n1, se, sj, nsub, batch,
                                   twmode, (const float2 *)tw2, live, dr.rowmask, dr.mode_in,
                                   dr.mode_out, dr.spec_a, dr.spec_b);
                CG_HIP(hipGetLastError());
                return CG_OK;
            }
        }
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute((const void *)k_fft_cols_step_pf<T, DIR, LOGL, CB>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const int64_t ntile = (int64_t)((nlive + CB - 1) / CB) * nsub * batch;
        const int per_cu = std::max(1, std::min((int)(kLdsMax / lds), 2048 / NT));
        const int64_t grid = std::min<int64_t>(ntile, (int64_t)256 * per_cu);
        hipLaunchKernelGGL((k_fft_cols_step_pf<T, DIR, LOGL, CB>), dim3((unsigned)grid), dim3(NT),
                           lds, s, data, n2, n1, se, sj, nsub, batch, twmode, tw2, live,
                           dr.rowmask, dr.mode_in, dr.mode_out, dr.spec_a, dr.spec_b);
        CG_HIP(hipGetLastError());
        return CG_OK;
    }
}

// CB adjacent columns per workgroup: 512-byte row segments within 64 KB of
// LDS
template <typename T, int DIR, int LOGL>
static int cols_step_ct(cx_t<T> *data, int batch, int n2, int n1, int se, int sj, int nsub,
                        int twmode, const cx_t<T> *tw2, Live live, const DeadRows &dr, hipStream_t s)
{
    constexpr int L = 1 << LOGL;
    constexpr int ELEM = (int)sizeof(cx_t<T>);
    constexpr int CB0 = 512 / ELEM;
    constexpr int CB = CB0 * L * ELEM <= 65536 ? CB0 : 65536 / (L * ELEM);
    return cols_step_ct_cb<T, DIR, LOGL, CB>(data, batch, n2, n1, se, sj, nsub, twmode, tw2, live,
                                             dr, s);
}

template <typename T, int DIR>
static int cols_step_ct_dispatch(cx_t<T> *data, int batch, int n2, int n1, int L, int se, int sj,
                                 int nsub, int twmode, const cx_t<T> *tw2, Live live,
                                 const DeadRows &d, hipStream_t s)
{
    switch (log2_exact(L)) {
    case 4: return cols_step_ct<T, DIR, 4>(data, batch, n2, n1, se, sj, nsub, twmode, tw2, live, d, s);
    case 5: return cols_step_ct<T, DIR, 5>(data, batch, n2, n1, se, sj, nsub, twmode, tw2, live, d, s);
    case 6: return cols_step_ct<T, DIR, 6>(data, batch, n2, n1, se, sj, nsub, twmode, tw2, live, d, s);
    case 7: return cols_step_ct<T, DIR, 7>(data, batch, n2, n1, se, sj, nsub, twmode, tw2, live, d, s);
    case 8: return cols_step_ct<T, DIR, 8>(data, batch, n2, n1, se, sj, nsub, twmode, tw2, live, d, s);
    default: return 1;
    }
}

// ---- long complex128 rows (n > 8192): X = DFT_n(x) as
//   X[2m]     = DFT_{n/2}(x[i] + x[i + n/2])[m]
//   X[2m + 1] = DFT_{n/2}((x[i] - x[i + n/2]) w^i)[m],  w = e^(dir 2 pi i / n),
// i.e. one radix-2 decimation-in-frequency step in HBM, the two halves of
// every row as rows of length n/2, and the even / odd outputs interleaved
// back to natural order; the inverse pass runs the transposed steps
// (de-interleave, half-length transforms, decimation-in-time combine).

// forward: x[i], x[i + h] <- x[i] + x[i + h], (x[i] - x[i + h]) w^i;
// inverse: x[i], x[i + h] <- y0 + y1 w^i, y0 - y1 w^i (w^i = conj(tw[i]))
template <typename T, int DIR>
__global__ __launch_bounds__(256) void k_rows_radix2(cx_t<T> *data, int64_t nrows, int n,
                                                     const cx_t<T> *__restrict__ tw)
{
    const int h = n / 2;
    const int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (idx >= nrows * h) return;
    const int64_t row = idx / h;
    const int i = (int)(idx - row * h);
    cx_t<T> *x = data + row * n;
    const cx_t<T> a = x[i], b = x[i + h];
    cx_t<T> w = tw[i];
    if (DIR < 0) w.y = -w.y;
    if (DIR > 0) {
        x[i] = mkc(a.x + b.x, a.y + b.y);
        x[i + h] = cmul<T>(mkc(a.x - b.x, a.y - b.y), w);
    } else {
        const cx_t<T> t = cmul<T>(b, w);
        x[i] = mkc(a.x + t.x, a.y + t.y);
        x[i + h] = mkc(a.x - t.x, a.y - t.y);
    }
}

// per row, in place through registers: DIR > 0 interleaves the halves
// [Y0 | Y1] to X[2m] = Y0[m], X[2m + 1] = Y1[m]; DIR < 0 splits X into
// [even | odd], reading only the live frequencies (the others are zero)
template <typename T, int DIR>
__global__ __launch_bounds__(1024) void k_rows_interleave(cx_t<T> *data, int64_t nrows, int n,
                                                          Live live)
{
    constexpr int EPR = 16;  // n <= 16384
    const int h = n / 2;
    for (int64_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        cx_t<T> *x = data + row * n;
        cx_t<T> v[EPR];
#pragma unroll
        for (int e = 0; e < EPR; e++) {
            const int j = (int)threadIdx.x + 1024 * e;
            v[e] = mkc((T)0, (T)0);
            if (j < n && (DIR > 0 || live(j))) v[e] = x[j];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EPR; e++) {
            const int j = (int)threadIdx.x + 1024 * e;
            if (j >= n) continue;
            const int d = DIR > 0 ? (j < h ? 2 * j : 2 * (j - h) + 1) : ((j & 1) ? h + (j >> 1) : j >> 1);
            x[d] = v[e];
        }
        __syncthreads();
    }
}

template <typename T, int DIR>
static int fft_rows(cx_t<T> *data, int batch, int n2, int n1, const cx_t<T> *tw, Live live,
                    const uint8_t *rowmask, hipStream_t s);

// tw holds the table of n1 followed by the table of n1 / 2
template <typename T, int DIR>
static int fft_rows_long(cx_t<T> *data, int batch, int n2, int n1, const cx_t<T> *tw,
                         Live live, hipStream_t s)
{
    const int64_t nrows = (int64_t)n2 * batch;
    const int h = n1 / 2;
    const unsigned nb2 = (unsigned)ceil_div(nrows * h, 256);
    const unsigned nbi = (unsigned)std::min<int64_t>(nrows, 4096);
    if (DIR > 0) {
        hipLaunchKernelGGL((k_rows_radix2<T, DIR>), dim3(nb2), dim3(256), 0, s, data, nrows, n1, tw);
        CG_HIP(hipGetLastError());
        CG_TRY((fft_rows<T, DIR>(data, 2 * batch, n2, h, tw + n1, Live{h, h}, nullptr, s)));
        hipLaunchKernelGGL((k_rows_interleave<T, DIR>), dim3(nbi), dim3(1024), 0, s, data, nrows, n1, live);
    } else {
        hipLaunchKernelGGL((k_rows_interleave<T, DIR>), dim3(nbi), dim3(1024), 0, s, data, nrows, n1, live);
        CG_HIP(hipGetLastError());
        CG_TRY((fft_rows<T, DIR>(data, 2 * batch, n2, h, tw + n1, Live{h, h}, nullptr, s)));
        hipLaunchKernelGGL((k_rows_radix2<T, DIR>), dim3(nb2), dim3(256), 0, s, data, nrows, n1, tw);
    }
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T, int DIR>
static int fft_rows(cx_t<T> *data, int batch, int n2, int n1, const cx_t<T> *tw, Live live,
                    const uint8_t *rowmask, hipStream_t s)
{
    if (fft_row_long<T>(n1)) return fft_rows_long<T, DIR>(data, batch, n2, n1, tw, live, s);
    {
        const int r = rows_ct_dispatch<T, DIR>(data, (int64_t)n2 * batch, n1, tw, live, rowmask, n2, s);
        if (r != 1) return r;
    }
    FftShape S{};
    S.n = n1;
    S.nrad = factor_radices(n1, S.rad);
    if (S.nrad < 0) return CG_ERR_SIZE;
    int nb = std::max(1, 4096 / n1);
    while (nb > 1 && row_lds_bytes<T>(n1, nb) > 65536) nb--;
    const int E = nb * n1;
    const size_t lds = (size_t)row_lds_bytes<T>(n1, nb);
    if (E > 1024 * fft_ept<T>() || lds > (size_t)kLdsMax) return CG_ERR_SIZE;
    const int threads = std::min(std::max(round64((E + fft_ept<T>() - 1) / fft_ept<T>()), 64), 1024);
    S.nb = nb;
    S.lt = n1;
    S.le = 1;
    S.pad = 5;
    S.tws = 1;
    const int64_t nrows = (int64_t)n2 * batch;
    const int64_t ngroups = (nrows + nb - 1) / nb;
    int per_cu = std::max<int>(1, (int)(kLdsMax / lds));
    per_cu = std::min(per_cu, std::max(1, 2048 / threads));
    const int64_t grid = std::min<int64_t>(ngroups, (int64_t)256 * per_cu);
    if (lds > 65536)
        CG_HIP(hipFuncSetAttribute((const void *)k_fft_rows<T, DIR>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_fft_rows<T, DIR>), dim3((unsigned)grid), dim3(threads), lds, s, data,
                       nrows, n1, S, tw, live);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T, int DIR>
static int fft_cols(cx_t<T> *data, int batch, int n2, int n1, const cx_t<T> *tw2, Live live,
                    const uint8_t *rowmask, int spec_a, int spec_b, int band_s, int band_len,
                    hipStream_t s)
{
    const int elem = (int)sizeof(cx_t<T>);
    const int nlive = live.count(n1);
    if (nlive <= 0 || batch <= 0) return CG_OK;
    int na, nb;
    if constexpr (sizeof(T) == 4) {
        // M = 1024 (16 columns per workgroup) for bands of <= 1024 rows, else
        // M = 2048 (8 columns); the transform buffer is 128 KB either way
        const int M = band_len <= 1024 ? 1024 : 2048, CB = M == 1024 ? 16 : 8;
        const size_t lds = sizeof(float2) * ((size_t)M * CB + M + (1 << kBandTwSh) + (n2 >> kBandTwSh));
        // (byte offsets inside a plane are 32-bit in the kernel)
        if (band_len > 0 && band_len <= M && lds <= (size_t)kLdsMax && band_s >= 0 && band_s < n2 &&
            fft_spectrum_pruned<T>(n2) && (n2 & (n2 - 1)) == 0 &&
            (size_t)n2 * n1 * sizeof(float2) <= 0xffffffffu && n2 >= 2 * M && fft_col_split(n2, na, nb)) {
            int nblk = (nlive + CB - 1) / CB;
            if (M == 2048) nblk = (nblk + 15) / 16 * 16;  // whole groups of the block remap
            const dim3 grid((unsigned)nblk, (unsigned)batch);
            auto go = [&](const void *fn, auto kern) -> int {
                CG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(kern, grid, dim3(1024), lds, s, (float2 *)data, n2, n1, n2 / M,
                                   band_s, band_len, live, rowmask, ilog2(na), nb, spec_a, spec_b,
                                   (const float2 *)tw2);
                return CG_OK;
            };
            if (M == 1024) CG_TRY(go((const void *)k_fft_cols_band<DIR, 1024>, k_fft_cols_band<DIR, 1024>));
            else CG_TRY(go((const void *)k_fft_cols_band<DIR, 2048>, k_fft_cols_band<DIR, 2048>));
            CG_HIP(hipGetLastError());
            return CG_OK;
        }
    }
    if (!fft_col_split(n2, na, nb)) {
        FftShape S{};
        S.n = n2;
        S.nrad = factor_radices(n2, S.rad);
        if (S.nrad < 0) return CG_ERR_SIZE;
        // power-of-two block of adjacent columns, up to 64 KB of LDS
        int cb = 1;
        while (cb * 2 <= 64 && (int64_t)cb * 2 * n2 * elem <= 65536 && cb * 2 <= nlive) cb *= 2;
        const int E = cb * n2;
        const size_t lds = (size_t)E * elem;
        if (E > 1024 * fft_ept<T>() || lds > (size_t)kLdsMax) return CG_ERR_SIZE;
        const int threads = std::min(std::max(round64((E + fft_ept<T>() - 1) / fft_ept<T>()), 64), 1024);
        S.nb = cb;
        S.nbsh = ilog2(cb);
        S.bfast = 1;
        S.lt = 1;
        S.le = cb;
        S.tws = 1;
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute((const void *)k_fft_cols<T, DIR>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_fft_cols<T, DIR>), dim3((unsigned)((nlive + cb - 1) / cb), (unsigned)batch),
                           dim3(threads), lds, s, data, n2, n1, S, tw2, live);
        CG_HIP(hipGetLastError());
        return CG_OK;
    }
    // dead rows: forward, the natural-order step skips loading masked-out
    // rows and the split-order step skips storing dead frequencies; inverse,
    // the split-order step skips loading dead frequencies and the
    // natural-order step skips storing masked-out rows
    auto step = [&](int L, int se, int sj, int nsub, int twmode) -> int {
        {
            DeadRows d{rowmask, spec_a, spec_b, 0, 0};
            const bool natural = sj == 1;
            const bool active = natural ? rowmask != nullptr : spec_a < spec_b;
            if (active) {
                const bool in = (DIR > 0) == natural;
                d.mode_in = in ? 1 : 0;
                d.mode_out = in ? 0 : 1;
            }
            const int r = cols_step_ct_dispatch<T, DIR>(data, batch, n2, n1, L, se, sj, nsub, twmode,
                                                        tw2, live, d, s);
            if (r != 1) return r;
        }
        FftShape S{};
        S.n = L;
        S.nrad = factor_radices(L, S.rad);
        if (S.nrad < 0) return CG_ERR_SIZE;
        // 512-byte row segments, at most 64 KB of LDS
        int cb = 512 / elem;
        while (cb > 1 && (int64_t)cb * L * elem > 65536) cb /= 2;
        const int E = cb * L;
        const int threads = std::min(std::max(round64((E + fft_ept<T>() - 1) / fft_ept<T>()), 64), 1024);
        if ((int64_t)threads * fft_ept<T>() < E) return CG_ERR_SIZE;
        S.nb = cb;
        S.nbsh = ilog2(cb);
        S.bfast = 1;
        S.lt = 1;
        S.le = cb;
        S.tws = n2 / L;
        dim3 grid((unsigned)((nlive + cb - 1) / cb), (unsigned)nsub, (unsigned)batch);
        hipLaunchKernelGGL((k_fft_cols_step<T, DIR>), grid, dim3(threads), (size_t)E * elem, s,
                           data, n2, n1, S, se, sj, twmode, tw2, live);
        CG_HIP(hipGetLastError());
        return CG_OK;
    };
    int r;
    if (DIR > 0) {
        // x[a nb + b] at row a nb + b -> Y_b[ka] w^(b ka) at row ka nb + b
        if ((r = step(na, nb, 1, nb, 1))) return r;
        // Y over b -> X[ka + na kb] at row ka nb + kb
        if ((r = step(nb, 1, nb, na, 0))) return r;
    } else {
        // X[ka + na kb] at row ka nb + kb -> Z_ka[b] at row ka nb + b
        if ((r = step(nb, 1, nb, na, 0))) return r;
        // Z_ka[b] w^(b ka) over ka -> x[a nb + b] at row a nb + b
        if ((r = step(na, nb, 1, nb, 2))) return r;
    }
    return CG_OK;
}

template <typename T>
int launch_fft_pass(cx_t<T> *data, int batch, int n2, int n1, int dir, bool rows,
                    const cx_t<T> *tw, int live_a, int live_b, const uint8_t *rowmask,
                    int spec_a, int spec_b, hipStream_t s, int band_s, int band_len)
{
    if (batch <= 0) return CG_OK;
    Live live{live_a, live_b};
    if (rows) return dir > 0 ? fft_rows<T, 1>(data, batch, n2, n1, tw, live, rowmask, s)
                             : fft_rows<T, -1>(data, batch, n2, n1, tw, live, rowmask, s);
    if (n2 <= 1) return CG_OK;
    return dir > 0 ? fft_cols<T, 1>(data, batch, n2, n1, tw, live, rowmask, spec_a, spec_b, band_s,
                                    band_len, s)
                   : fft_cols<T, -1>(data, batch, n2, n1, tw, live, rowmask, spec_a, spec_b, band_s,
                                     band_len, s);
}

template <typename T>
bool fft_rows_masked(int n1)
{
    // the conditions under which fft_rows takes rows_ct (which honours the
    // row and column masks)
    const int l = log2_exact(n1);
    if (l < 9 || l > 14) return false;
    const int nt = n1 / fft_ept<T>();
    if (nt < 64 || nt > 1024) return false;
    const int sh = l / 2;
    return (size_t)(n1 + (n1 >> 5) + (1 << sh) + (n1 >> sh)) * sizeof(cx_t<T>) <= (size_t)kLdsMax;
}

template <typename T>
bool fft_spectrum_pruned(int n2)
{
    int na, nb;
    if (n2 <= 1 || !fft_col_split(n2, na, nb)) return false;
    const int ln = log2_exact(na), lb = log2_exact(nb);
    return ln >= 4 && ln <= 8 && lb >= 4 && lb <= 8;
}

template <typename T>
static int col_permute(cx_t<T> *data, cx_t<T> *tmp, int batch, int n2, int n1, int na, int nb,
                       int to_natural, hipStream_t s)
{
    const int64_t total = (int64_t)batch * n2 * n1;
    hipLaunchKernelGGL(k_col_permute<T>, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, data,
                       tmp, total, n2, n1, na, nb, to_natural);
    CG_HIP(hipGetLastError());
    CG_HIP(hipMemcpyAsync(data, tmp, sizeof(cx_t<T>) * (size_t)total, hipMemcpyDeviceToDevice, s));
    return CG_OK;
}

template <typename T>
int launch_fft2d(cx_t<T> *data, int batch, int n2, int n1, int dir, const cx_t<T> *tw1,
                 const cx_t<T> *tw2, cx_t<T> *tmp, hipStream_t s)
{
    if (batch <= 0) return CG_OK;
    int na = 0, nb = 0;
    const bool split = n2 > 1 && fft_col_split(n2, na, nb);
    if (split && !tmp) return CG_ERR_ARG;
    int r;
    if (dir > 0) {
        if ((r = launch_fft_pass<T>(data, batch, n2, n1, dir, true, tw1, n1, n1, nullptr, n2, n2, s))) return r;
        if ((r = launch_fft_pass<T>(data, batch, n2, n1, dir, false, tw2, n1, n1, nullptr, n2, n2, s))) return r;
        if (split) return col_permute<T>(data, tmp, batch, n2, n1, na, nb, 1, s);
        return CG_OK;
    }
    if (split && (r = col_permute<T>(data, tmp, batch, n2, n1, na, nb, 0, s))) return r;
    if ((r = launch_fft_pass<T>(data, batch, n2, n1, dir, false, tw2, n1, n1, nullptr, n2, n2, s))) return r;
    return launch_fft_pass<T>(data, batch, n2, n1, dir, true, tw1, n1, n1, nullptr, n2, n2, s);
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

template int launch_fft2d<float>(float2 *, int, int, int, int, const float2 *, const float2 *, float2 *, hipStream_t);
template int launch_fft2d<double>(double2 *, int, int, int, int, const double2 *, const double2 *, double2 *, hipStream_t);
template int launch_fft_pass<float>(float2 *, int, int, int, int, bool, const float2 *, int, int, const uint8_t *, int, int, hipStream_t, int, int);
template int launch_fft_pass<double>(double2 *, int, int, int, int, bool, const double2 *, int, int, const uint8_t *, int, int, hipStream_t, int, int);
template bool fft_rows_masked<float>(int);
template bool fft_rows_masked<double>(int);
template bool fft_spectrum_pruned<float>(int);
template bool fft_spectrum_pruned<double>(int);
template bool fft_supported<float>(int);
template bool fft_supported<double>(int);
template bool fft_row_long<float>(int);
template bool fft_row_long<double>(int);
template int launch_twiddles<float>(float2 *, int, hipStream_t);
template int launch_twiddles<double>(double2 *, int, hipStream_t);

}  // namespace cg
