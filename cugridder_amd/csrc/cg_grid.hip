// ES-kernel gridding (ms2dirty) and degridding (dirty2ms) on LDS tiles.
//
// Gridding computes the reference's periodic 3-D spreading sum
//   F[p][b2][b1] = sum_j c_j psi(b1-g1_j) psi(b2-g2_j) psi(p-g3_j)
// (src/fourier/conv.cu:522-627 conv_3d_outputdriven_t1, equal to the
// input-driven conv.cu:166-244), psi(d) = exp(beta sqrt(1 - c d^2)) for
// |d| < kw/2 (conv.cu:32-36, unnormalised as in the reference).
//
// MI355X design: one workgroup owns a tx*ty tile of grid cells for a chunk of
// w-planes and accumulates it in LDS (real parts, then imaginary parts, so
// the lanes' random cells spread over twice as many banks as interleaved
// pairs would) with FP64 ds_add atomics (on gfx950 the
// FP32 LDS atomic add runs ~20x slower than the FP64 one, and FP64 also makes
// the accumulation order-insensitive at FP32 output precision); it reads the points
// of the bins whose footprints reach the tile (points are binned by the tile
// of their footprint start, so at most sbx*sby bins feed a tile) and writes
// the finished tile once with coalesced stores.  No HBM atomics and no grid
// memset: every cell is written exactly once.
//
// Degridding is the adjoint interpolation (src/fourier/interp.cu:132-210):
// a workgroup stages its tile plus a (kw-1) halo of every plane in LDS and
// each thread interpolates its points from LDS; the result is multiplied by
// the per-point factor (weight and w-phase, ra_exec.cu:55-93) and scattered
// back to the caller's row order.
#include <cstdlib>
#include <type_traits>

#include "cg_internal.hpp"

namespace cg {

__device__ __forceinline__ uint32_t geo_bits(float w) { return __float_as_uint(w); }
__device__ __forceinline__ uint32_t geo_bits(double w) { return (uint32_t)__double_as_longlong(w); }

__device__ __forceinline__ double es_val(double d, double beta, double c, double hw)
{
    return fabs(d) < hw ? exp(beta * sqrt(1.0 - c * d * d)) : 0.0;
}

// FP32 kernel values divided by e^beta (the MFMA gridders scale their sums
// back once, the other FP32 kernels every value):
//   psi(d) / e^beta = exp(beta (sqrt(1 - u) - 1) + y) = 2^(-beta log2(e) u / (1 + sqrt(1 - u)) + log2(e) y),
// u = c d^2, y the learned kernel's polynomial in s = 2 u - 1 (0 for the ES
// kernel).  The rewritten exponent has no cancellation near the centre, so
// the hardware square root, reciprocal and exp2 (about ten instructions
// instead of the library's correctly rounded expf and sqrtf) keep the large
// central taps within a few ulp; only the taps near the support edge, whose
// values are ~e^-beta of the centre, see errors up to ~beta * 1e-7.
struct KernelEvalN {
    float nb, ec, hw;
    int learned;
    const float *q;  // the launch argument's polynomial coefficients
    __device__ __forceinline__ explicit KernelEvalN(const GridArgs &g, int kw)
        : nb(g.f_nb), ec(g.f_ec), hw(0.5f * (float)kw), learned(g.learned), q(g.kq)
    {
    }
    // LEARNED: -1 follows the launch's flag, 0 / 1 fix the kernel at compile
    // time (no learned-polynomial code in the ES instantiations)
    template <int LEARNED = -1>
    __device__ __forceinline__ float at(float off0, int t) const
    {
        const float d = off0 + (float)t;
        const float u = ec * d * d;
        float y = nb * u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f - u));
        if (LEARNED == 1 || (LEARNED < 0 && learned)) {
            const float s = 2.0f * u - 1.0f;
            float p = q[LEARN_MAXDEG];
#pragma unroll
            for (int k = LEARN_MAXDEG - 1; k >= 0; k--) p = p * s + q[k];
            y += 1.4426950408889634f * p;
        }
        const float v = __builtin_amdgcn_exp2f(y);
        return fabsf(d) < hw ? v : 0.0f;
    }
    __device__ __forceinline__ float operator()(float off0, int t) const { return at(off0, t); }
};

// Kernel-value source of the complex64 MFMA kernels, fixed per
// instantiation: the ES exp2 form or the per-tap polynomials (tp_on; KernelEvalN
// also takes kKerLearned for the exact learned form).
enum : int { kKerES = 0, kKerLearned = 1, kKerPoly = 2 };
// the MFMA kernels are instantiated for the ES form and the polynomials: a
// learned kernel always comes with its polynomials (grid_args), so a
// learned launch without them is refused
inline int ker_mode(const GridArgs &g) { return g.tp_on ? kKerPoly : g.learned ? -1 : kKerES; }

// FP32 kernel values psi(off0 + t) / e^beta (the normalisation of
// KernelEvalN) from per-tap polynomials in x = 2 off0 + kw - 1 (host fit,
// tap_polynomials), coefficients in LDS; zero outside the taps.  Eleven FMAs
// instead of the square root, reciprocal and exp2 of KernelEvalN, and the
// same cost for the ES and the learned kernel.
struct TapPolyN {
    const float *tab;  // kw rows of kTapStride coefficients
    int kw;
    __device__ __forceinline__ float operator()(float off0, int t) const
    {
        if (t < 0 || t >= kw) return 0.f;
        const float x = 2.f * off0 + (float)(kw - 1);
        const float4 *c = reinterpret_cast<const float4 *>(tab + t * kTapStride);
        const float4 c0 = c[0], c1 = c[1], c2 = c[2];
        float p = c2.z;
        p = p * x + c2.y;
        p = p * x + c2.x;
        p = p * x + c1.w;
        p = p * x + c1.z;
        p = p * x + c1.y;
        p = p * x + c1.x;
        p = p * x + c0.w;
        p = p * x + c0.z;
        p = p * x + c0.y;
        p = p * x + c0.x;
        return p;
    }
};

// Kernel values of a point: the ES kernel exp(beta sqrt(1 - c d^2)) at the
// tap offsets d, or the learned kernel, whose exponent adds the network's
// polynomial sum_k q_k s^k in s = 2 c d^2 - 1 (c = (2/kw)^2).  The choice is
// uniform across a launch and the coefficients are launch arguments, so
// they stay in scalar registers.  FP64 evaluates the formula directly; FP32
// is KernelEvalN times e^beta.
template <typename T>
struct KernelEval {
    T beta, ec, hw;
    int learned;
    T q[LEARN_MAXDEG + 1];
    __device__ __forceinline__ explicit KernelEval(const GridArgs &g, int kw)
        : beta((T)g.beta), ec((T)g.es_c), hw((T)(0.5 * kw)), learned(g.learned)
    {
#pragma unroll
        for (int k = 0; k <= LEARN_MAXDEG; k++) q[k] = (T)g.kq[k];
    }
    // value of tap t of a point whose first tap sits at offset off0
    __device__ __forceinline__ T operator()(T off0, int t) const
    {
        const T d = off0 + (T)t;
        if (!learned) return es_val(d, beta, ec, hw);
        if (!(fabs(d) < hw)) return (T)0;
        const T r = (T)1 - ec * d * d;
        const T s = (T)1 - (T)2 * r;
        T y = q[LEARN_MAXDEG];
#pragma unroll
        for (int k = LEARN_MAXDEG - 1; k >= 0; k--) y = y * s + q[k];
        return exp(beta * sqrt(r) + y);
    }
};
template <>
struct KernelEval<float> : KernelEvalN {
    float eb;  // e^beta
    __device__ __forceinline__ explicit KernelEval(const GridArgs &g, int kw)
        : KernelEvalN(g, kw), eb(g.f_e1)
    {
    }
    __device__ __forceinline__ float operator()(float off0, int t) const
    {
        return KernelEvalN::operator()(off0, t) * eb;
    }
};

// gridding accumulates in FP64 whatever the grid precision: 16 bytes per cell
int grid_lds_bytes(const GridArgs &g, int) { return g.pc * g.tx * g.ty * 16; }
__host__ __device__ inline int degrid_row_stride(int tx, int kw) { return (tx + kw - 1) | 1; }
int degrid_lds_bytes(const GridArgs &g, int kw, int elem)
{
    return g.pc * degrid_row_stride(g.tx, kw) * (g.ty + kw - 1) * elem;
}

// ------------------------------------------------------------ gridding

// Footprint of a point relative to a tile (or plane chunk) [0, t) of an axis
// of length n, given its start r relative to the tile origin (in [0, n)):
// first tap position s (negative when the footprint starts in an earlier
// tile) and the taps [lo, hi) that land inside the tile.  With one tile
// along the axis (t == n) the footprint wraps inside it.
struct Span {
    int s, lo, hi;
};
template <int KW>
__device__ __forceinline__ Span tile_span(int r, int t, int n)
{
    Span sp;
    if (t == n) {
        sp.s = r;
        sp.lo = 0;
        sp.hi = KW;
    } else {
        sp.s = r < t ? r : r - n;
        sp.lo = max(0, -sp.s);
        sp.hi = min(KW, t - sp.s);
    }
    return sp;
}

// most source bins feeding one tile: (1 + ceil(15/2))^2 for tiles >= 2
constexpr int kMaxBins = 81;

__global__ void k_zero_i32(int32_t *p, int64_t second);

template <typename T, int KW>
__global__ __launch_bounds__(1024) void k_grid(const geo_t<T> *__restrict__ geo,
                                               const cx_t<T> *__restrict__ val,
                                               const int32_t *__restrict__ tile_start,
                                               cx_t<T> *__restrict__ fw, GridArgs g, int is2d)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *acc = reinterpret_cast<double *>(smem_raw);
    __shared__ int s_beg[kMaxBins], s_cnt[kMaxBins], s_ox[kMaxBins], s_oy[kMaxBins];
    __shared__ int s_total;
    // work item = (tile, plane chunk): one per workgroup, or, with a tile
    // list (the tiles some FFT pass reads), the workgroups walk the list
    const int nwork = g.tile_list ? g.tile_list[0] * g.nchunk : (int)gridDim.x;
    for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int chunk = wi % g.nchunk;
    const int tile = g.tile_list ? g.tile_list[1 + wi / g.nchunk] : wi / g.nchunk;
    const int bx = tile % g.ntx, by = tile / g.ntx;
    // planes [p0, p0 + np) of the slab [zb, zb + zn) held by the buffer
    const int zn = slab_planes(g);
    const int p0 = g.zb + chunk * g.pc;
    const int np = min(g.pc, g.zb + zn - p0);
    const int cells = np * g.ty * g.tx;
    for (int k = threadIdx.x; k < 2 * cells; k += blockDim.x) acc[k] = 0.0;
    // source bins whose footprints can reach this tile: bin (bx - sx, by - sy)
    const int nbins = g.sbx * g.sby;
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int b = 0; b < nbins; b++) {
            const int sy = b / g.sbx, sx = b - sy * g.sbx;
            int sby_ = by - sy, sbx_ = bx - sx;
            if (sby_ < 0) sby_ += g.nty;
            if (sbx_ < 0) sbx_ += g.ntx;
            int oy = (sby_ - by) * g.ty, ox = (sbx_ - bx) * g.tx;
            if (oy < 0) oy += g.nf2;
            if (ox < 0) ox += g.nf1;
            const int sidx = sby_ * g.ntx + sbx_;
            s_beg[b] = tile_start[sidx];
            s_cnt[b] = tile_start[sidx + 1] - s_beg[b];
            // dense bins are added afterwards by k_grid_dense
            if (g.dense_min > 0 && s_cnt[b] >= g.dense_min) s_cnt[b] = 0;
            s_ox[b] = ox;
            s_oy[b] = oy;
            tot += s_cnt[b];
        }
        s_total = tot;
    }
    __syncthreads();

    const KernelEval<T> ker(g, KW);
    const int total = s_total;
    // all points of all source bins as one list, so every thread has work;
    // points are sorted by overlap class inside a bin, so the lanes of a
    // wavefront mostly share their tap ranges and skip the same taps
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        int b = 0, j = i;
        while (b < nbins - 1 && j >= s_cnt[b]) {
            j -= s_cnt[b];
            b++;
        }
        const int idx = s_beg[b] + j;
        const geo_t<T> G = geo[idx];
        const uint32_t pk = geo_bits(G.w);
        int rx = s_ox[b] + (int)(pk & 0xffu);
        if (rx >= g.nf1) rx -= g.nf1;
        int ry = s_oy[b] + (int)((pk >> 8) & 0xffu);
        if (ry >= g.nf2) ry -= g.nf2;
        const Span X = tile_span<KW>(rx, g.tx, g.nf1);
        const Span Y = tile_span<KW>(ry, g.ty, g.nf2);
        if (X.lo >= X.hi || Y.lo >= Y.hi) continue;
        Span Z{0, 0, 1};
        if (!is2d) {
            int rz = (int)(pk >> 16) - p0;
            if (rz < 0) rz += g.nf3;
            Z = tile_span<KW>(rz, np, g.nf3);
            if (Z.lo >= Z.hi) continue;
        }
        const cx_t<T> v = val[idx];
        T kx[KW], ky[KW];
#pragma unroll
        for (int t = 0; t < KW; t++) {
            kx[t] = ker(G.x, t);
            ky[t] = ker(G.y, t);
        }
        for (int tz = Z.lo; tz < Z.hi; tz++) {
            int zz = Z.s + tz;
            if (zz >= np) zz -= np;  // wraps only when one chunk holds every plane
            const T kz = is2d ? (T)1 : ker(G.z, tz);
            const T vzr = v.x * kz, vzi = v.y * kz;
#pragma unroll
            for (int t2 = 0; t2 < KW; t2++) {
                if (t2 < Y.lo || t2 >= Y.hi) continue;
                int yy = Y.s + t2;
                if (yy >= g.nf2) yy -= g.nf2;  // wraps only with one tile along nf2
                const T vyr = vzr * ky[t2], vyi = vzi * ky[t2];
                double *row = acc + (zz * g.ty + yy) * g.tx;
#pragma unroll
                for (int t1 = 0; t1 < KW; t1++) {
                    if (t1 < X.lo || t1 >= X.hi) continue;
                    int xx = X.s + t1;
                    if (xx >= g.nf1) xx -= g.nf1;
                    atomicAdd(row + xx, (double)(vyr * kx[t1]));
                    atomicAdd(row + cells + xx, (double)(vyi * kx[t1]));
                }
            }
        }
    }
    __syncthreads();
    const int plane_cells = g.ty * g.tx;
    for (int k = threadIdx.x; k < cells; k += blockDim.x) {
        int lz = k / plane_cells;
        int rem = k - lz * plane_cells;
        int y = rem / g.tx, x = rem - y * g.tx;
        size_t gi = ((size_t)(p0 - g.zb + lz) * g.nf2 + (size_t)(by * g.ty + y)) * g.nf1 + bx * g.tx + x;
        fw[gi] = mkc((T)acc[k], (T)acc[cells + k]);
    }
    __syncthreads();  // acc and the bin table are reused by the next item
    }
}

// ------------------------------------------------------------ 2-D on MFMA
//
// One plane (nf3 == 1), complex64: LDS FP64 atomics are what bounds the
// sparse kernel (about 2.5 random lane-ops per clock per CU on gfx950, 72 per
// point at kw 6).  Here a wave owns a 16 x 16 quadrant of a tile and
// accumulates it in registers as a matrix product over the points that
// touch it:
//   C[i][j] = sum_k psi_y,k(i) * (c_k psi_x,k(j)),
// four points per v_mfma_f32_16x16x4f32 (A[i][k] from lane i + 16 k, B[k][j]
// from lane j + 16 k, D[4 (l / 16) + v][l % 16] in lane l, register v), real
// and imaginary parts as two products.  The wave reads the points of the
// bins that can reach its quadrant as one list, 64 at a time with the next
// 64 already in flight, compacts the touching ones into an LDS queue (ballot
// + prefix count, so the order and the sums are reproducible) and runs one
// MFMA step per four queued points, carrying the rest (< 4) over.  The
// quadrant is written once with plain stores.  The four waves of a
// workgroup work on independent quadrants (no barriers).  Sums over points
// are FP32 (MFMA accumulation).

constexpr int kQ2 = 16;      // quadrant edge
constexpr int kQueue = 68;   // queued points: < 4 carried + 64 new
constexpr int kQWaves = 4;   // independent waves per workgroup

typedef float v4f __attribute__((ext_vector_type(4)));

struct Queue2d {
    int xs[kQueue], ys[kQueue];     // first tap relative to the quadrant origin
    float gx[kQueue], gy[kQueue];   // offset of the first tap from the point
    float vr[kQueue], vi[kQueue];   // strength
    int beg[4], cnt[4], ox[4], oy[4];  // source bins
};

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int KW>
__global__ __launch_bounds__(64 * kQWaves) void k_grid2d_mfma(const float4 *__restrict__ geo,
                                                             const float2 *__restrict__ val,
                                                             const int32_t *__restrict__ tile_start,
                                                             float2 *__restrict__ fw, GridArgs g)
{
    __shared__ Queue2d Qs[kQWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Queue2d &Q = Qs[wave];
    const int nqx = g.tx / kQ2, nq = nqx * (g.ty / kQ2);
    const int work = blockIdx.x * kQWaves + wave;
    const int item = work / nq, quad = work - item * nq;
    const int nitems = g.tile_list ? g.tile_list[0] : g.ntx * g.nty;
    if (item >= nitems) return;  // the waves share no barrier
    const int tile = g.tile_list ? g.tile_list[1 + item] : item;
    const int qx = quad % nqx, qy = quad / nqx;
    const int bx = tile % g.ntx, by = tile / g.ntx;
    // source bins (sx, sy) in {0, 1}^2: this tile's, the left, upper and
    // upper-left neighbours' (the tile edge is >= kw - 1).  Left bins reach
    // only the quadrants of the tile's first 16 columns, upper bins only
    // those of its first 16 rows (kw - 1 <= 16): others get count 0.
    if (lane < 4) {
        const int b = lane, sx = b & 1, sy = b >> 1;
        int sby_ = by - sy, sbx_ = bx - sx;
        if (sby_ < 0) sby_ += g.nty;
        if (sbx_ < 0) sbx_ += g.ntx;
        int oy = (sby_ - by) * g.ty, ox = (sbx_ - bx) * g.tx;
        if (oy < 0) oy += g.nf2;
        if (ox < 0) ox += g.nf1;
        const int sidx = sby_ * g.ntx + sbx_;
        const bool reach = !((sx && qx != 0) || (sy && qy != 0));
        Q.beg[b] = tile_start[sidx];
        Q.cnt[b] = reach ? tile_start[sidx + 1] - Q.beg[b] : 0;
        Q.ox[b] = ox;
        Q.oy[b] = oy;
    }
    wave_lds_sync();
    const int n0 = Q.cnt[0], n1 = n0 + Q.cnt[1], n2 = n1 + Q.cnt[2], total = n2 + Q.cnt[3];
    // point i of the list: its bin and position
    auto fetch = [&](int i, float4 &G, float2 &v, int &b) {
        b = i < n0 ? 0 : i < n1 ? 1 : i < n2 ? 2 : 3;
        if (i < total) {
            const int idx = Q.beg[b] + i - (b == 0 ? 0 : b == 1 ? n0 : b == 2 ? n1 : n2);
            G = geo[idx];
            v = val[idx];
        }
    };

    const KernelEvalN ker(g, KW);  // values / e^beta: the sums are scaled by e^(2 beta)
    const int k = lane >> 4, r = lane & 15;
    v4f cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
    // one MFMA step over queued points e0 .. e0 + 3
    auto step = [&](int e0) {
        const int e = e0 + k;
        const int dy = r - Q.ys[e], dx = r - Q.xs[e];
        const float a = (dy >= 0 && dy < KW) ? ker(Q.gy[e], dy) : 0.f;
        const float kx = (dx >= 0 && dx < KW) ? ker(Q.gx[e], dx) : 0.f;
        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Q.vr[e] * kx, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Q.vi[e] * kx, cim, 0, 0, 0);
    };
    int queued = 0;  // wave-uniform, < 4 between chunks
    float4 G = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 v = make_float2(0.f, 0.f);
    int b = 0;
    fetch(lane, G, v, b);
    for (int base = 0; base < total; base += 64) {
        // the next 64 points are loaded while these are queued and stepped
        float4 Gn = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 vn = make_float2(0.f, 0.f);
        int bn = 0;
        fetch(base + 64 + lane, Gn, vn, bn);
        bool touch = false;
        int xq = 0, yq = 0;
        if (base + lane < total) {
            const uint32_t pk = __float_as_uint(G.w);
            int rx = Q.ox[b] + (int)(pk & 0xffu);
            if (rx >= g.nf1) rx -= g.nf1;
            int ry = Q.oy[b] + (int)((pk >> 8) & 0xffu);
            if (ry >= g.nf2) ry -= g.nf2;
            // first tap relative to the tile (negative: an earlier tile),
            // then to the quadrant
            xq = (rx < g.tx ? rx : rx - g.nf1) - qx * kQ2;
            yq = (ry < g.ty ? ry : ry - g.nf2) - qy * kQ2;
            touch = xq > -KW && xq < kQ2 && yq > -KW && yq < kQ2;
        }
        const uint64_t mask = __ballot(touch);
        if (touch) {
            const int pos = queued + (int)__popcll(mask & ((1ull << lane) - 1ull));
            Q.xs[pos] = xq;
            Q.ys[pos] = yq;
            Q.gx[pos] = G.x;
            Q.gy[pos] = G.y;
            Q.vr[pos] = v.x;
            Q.vi[pos] = v.y;
        }
        wave_lds_sync();
        queued += (int)__popcll(mask);
        const int full = queued & ~3;
#pragma unroll 2
        for (int e0 = 0; e0 < full; e0 += 4) step(e0);
        const int rest = queued - full;
        if (rest && full && lane < rest) {
            // carry the last (< 4) queued points to the front (source
            // entries >= 4, destinations < 3: the ranges are disjoint)
            const int src = full + lane;
            Q.xs[lane] = Q.xs[src];
            Q.ys[lane] = Q.ys[src];
            Q.gx[lane] = Q.gx[src];
            Q.gy[lane] = Q.gy[src];
            Q.vr[lane] = Q.vr[src];
            Q.vi[lane] = Q.vi[src];
        }
        queued = rest;
        wave_lds_sync();
        G = Gn;
        v = vn;
        b = bn;
    }
    if (queued) {
        // pad the last step with zero-strength points
        if (lane >= queued && lane < 4) {
            Q.xs[lane] = 0;
            Q.ys[lane] = 0;
            Q.gx[lane] = 0.f;
            Q.gy[lane] = 0.f;
            Q.vr[lane] = 0.f;
            Q.vi[lane] = 0.f;
        }
        wave_lds_sync();
        step(0);
    }
    // lane l holds rows 4 (l / 16) + v of column l % 16
    const int gx0 = bx * g.tx + qx * kQ2 + r;
    const int gy0 = by * g.ty + qy * kQ2 + 4 * k;
    const float sc = g.f_e2;
#pragma unroll
    for (int v4 = 0; v4 < 4; v4++)
        fw[(size_t)(gy0 + v4) * g.nf1 + gx0] = make_float2(cre[v4] * sc, cim[v4] * sc);
}

bool grid2d_mfma_supported(const GridArgs &g, int kw)
{
    return g.mfma2d && g.nf3 == 1 && kw >= 2 && kw <= kQ2 && (g.tx == 16 || g.tx == 32) &&
           (g.ty == 16 || g.ty == 32) && g.ntx >= 2 && g.nty >= 2 && g.sbx == 2 && g.sby == 2 &&
           g.nchunk == 1;
}

// list[0] = number of tiles with a masked-in row and a masked-in column,
// list[1..] = those tiles (any order; each is gridded independently)
__global__ __launch_bounds__(256) void k_need_tiles(const uint8_t *__restrict__ rowmask,
                                                    const uint8_t *__restrict__ colmask, int ntx,
                                                    int64_t ntiles, int tx, int ty,
                                                    int32_t *__restrict__ list)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= ntiles) return;
    const int bx = (int)(t % ntx), by = (int)(t / ntx);
    bool need_r = false, need_c = false;
    for (int r = 0; r < ty && !need_r; r++) need_r = rowmask[by * ty + r] != 0;
    for (int c = 0; c < tx && !need_c; c++) need_c = colmask[bx * tx + c] != 0;
    if (need_r && need_c) list[1 + atomicAdd(list, 1)] = (int32_t)t;
}

// threads per gridding workgroup: enough waves per CU when the tile's LDS
// leaves room for few workgroups
static int grid_threads(size_t lds)
{
    const int per_cu = std::max<int>(1, (int)(163840 / std::max<size_t>(lds, 1)));
    if (per_cu >= 4) return 256;
    if (per_cu >= 2) return 512;
    return 1024;
}

template <typename T, int KW>
static int grid_kw(const geo_t<T> *geo, const cx_t<T> *val, const int32_t *ts, cx_t<T> *fw,
                   const GridArgs &g, hipStream_t s)
{
    if (g.sbx * g.sby > kMaxBins) return CG_ERR_SIZE;
    int is2d = g.nf3 == 1 ? 1 : 0;
    size_t lds = (size_t)grid_lds_bytes(g, (int)sizeof(cx_t<T>));
    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    unsigned nblk = (unsigned)(ntiles * g.nchunk);
    if (g.tile_list) {
        // build the list of tiles to write; workgroups then walk it
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, const_cast<int32_t *>(g.tile_list),
                           (int64_t)0);
        hipLaunchKernelGGL(k_need_tiles, dim3((unsigned)ceil_div(ntiles, 256)), dim3(256), 0, s,
                           g.rowmask, g.colmask, g.ntx, ntiles, g.tx, g.ty,
                           const_cast<int32_t *>(g.tile_list));
        const int threads = grid_threads(lds);
        const int per_cu = std::max(1, std::min((int)(163840 / std::max<size_t>(lds, 1)),
                                                2048 / threads));
        nblk = (unsigned)std::min<int64_t>(ntiles * g.nchunk, (int64_t)256 * per_cu);
    }
    if constexpr (sizeof(T) == 4 && KW <= kQ2) {
        if (grid2d_mfma_supported(g, KW)) {
            // one wave per quadrant of every tile (the hardware balances
            // clustered tiles); with a tile list, waves past it return
            const int64_t nwork = ntiles * (g.tx / kQ2) * (g.ty / kQ2);
            const dim3 grid((unsigned)ceil_div(nwork, kQWaves)), block(64 * kQWaves);
            hipLaunchKernelGGL(k_grid2d_mfma<KW>, grid, block, 0, s, (const float4 *)geo,
                               (const float2 *)val, ts, (float2 *)fw, g);
            CG_HIP(hipGetLastError());
            return CG_OK;
        }
    }
    if (lds > 65536)
        CG_HIP(hipFuncSetAttribute((const void *)k_grid<T, KW>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_grid<T, KW>), dim3(nblk), dim3(grid_threads(lds)), lds, s, geo, val, ts,
                       fw, g, is2d);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// ------------------------------------------------------------ dense bins
//
// A bin holding many points per start cell (dense uv coverage) is gridded as
// matrix products: the points j of a group of start cells share a footprint
// frame, so their contributions are
//   C[tap][p] = sum_j A[tap][j] B[j][p],  A = psi_x psi_y,  B = v_j psi_z,
// FP32 MFMA steps (k_grid_dense2 below).  The frame is added into the bin's
// footprint region in LDS (FP64), and the region once per bin into the grid
// with FP32 global atomics, on top of what the sparse kernel wrote.

// Work units of the dense kernels: every bin with at least dense_min points,
// cut into chunks of at most `chunk` consecutive sorted points, so that one
// bin holding a large share of the points (a small image with many rows) is
// spread over many workgroups.  list[0] counts the units, list[1] is the
// dense kernels' work counter, unit e is (bin, first point offset) at
// list[2 + 2 e], list[3 + 2 e] (any order).  Needs dense_list_ints ints.
__global__ __launch_bounds__(256) void k_dense_bins(const int32_t *__restrict__ tile_start,
                                                    int64_t ntiles, int dense_min, int chunk,
                                                    int32_t *__restrict__ list)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= ntiles) return;
    const int n = tile_start[t + 1] - tile_start[t];
    if (n < dense_min || n == 0) return;
    const int nu = (int)(((int64_t)n + chunk - 1) / chunk);
    const int e0 = atomicAdd(list, nu);
    for (int u = 0; u < nu; u++) {
        list[2 + 2 * (e0 + u)] = (int32_t)t;
        list[3 + 2 * (e0 + u)] = u * chunk;
    }
}

int dense_chunk(int64_t m)
{
    // about 1024 units over the rows (four per CU), at least 512 points each
    int64_t c = 512;
    while (c < m / 1024 && c < (1 << 30)) c *= 2;
    return (int)c;
}

int64_t dense_list_ints(int64_t ntiles, int64_t m)
{
    const int64_t units = ntiles + (m + dense_chunk(m) - 1) / dense_chunk(m);
    return std::max<int64_t>(2 + 2 * units, ntiles + 2);  // also the sparse tiles' list
}

__global__ void k_zero_i32(int32_t *p, int64_t second)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p[0] = 0;
        p[second] = 0;
    }
}

// Dense bins, second form.  A workgroup takes work units (k_dense_bins: a
// bin, or a chunk of a crowded one) from a device counter; its waves take
// pieces (order_pieces) of the unit's work items.  Work items are quads,
// 2 x 2 start cells (single cells when kw = 8): with kw <= 7 the footprints
// of a quad's points lie in one 8 x 8 tap frame.  A quad's points (two runs
// of the sorted order) are
// first ordered by w-window (window w holds the points whose first plane z0
// lies in [S w, S w + S), S = 9 - kw, so every footprint of the window fits
// the eight planes [S w, S w + 8)) with a stable counting sort of the wave in
// LDS (ballot ranks, so the order and the sums are reproducible).  The
// window's points are then gridded four at a time with
// v_mfma_f32_16x16x4f32:
//   C[tap][col] += A[tap][k] B[k][col],  A = kx(x - dx) ky(y - dy) of point k
//   for the 16 taps (x = tap % 8, y = 2 blk + tap / 8) of tap block blk,
//   (dx, dy) its start cell in the quad,
//   B = Re / Im(v_k) kz(plane) for the eight planes (col = 2 plane + part,
//   part 0 real, 1 imaginary),
// so the complex value rides in B.  Each lane evaluates two kernel values of
// its point per step (x or y position r % 8, and the plane of its column);
// the others come by lane shuffles.  Moving from window w to w + 1, the S
// planes that leave the frame are added to the FP64 LDS footprint region
// (x fastest) and the frame shifts by S planes, 2 S lanes down each 16-lane
// row (one DPP move per register); the last frame is added whole.
constexpr int kDense2Threads = 512;
constexpr int kDense2Max = 512;  // points of a quad sorted at once

// Stable counting sort by w-window (window = first plane z0 / S) of the n
// (<= kDense2Max) points at positions c0 .. c0 + n - 1 of a work item whose
// points are the sorted runs [a0, a0 + n0) then [a1, ...): lst[] receives the
// positions p < n in window order, wb[w] .. wb[w + 1] bounds window w.  One
// wavefront; ranks come from ballots, so the order is reproducible.
struct Runs {
    int a0, n0, a1, n;  // n = all points of the item
    __device__ __forceinline__ int row(int i) const { return i < n0 ? a0 + i : a1 + (i - n0); }
};

template <int S>
__device__ __forceinline__ void sort_by_window(const float4 *__restrict__ geo, const Runs &R, int c0,
                                               int n, unsigned short *lst, int *cnt, int *wb, int lane)
{
    const uint64_t lt = (1ull << lane) - 1ull;
    if (lane < 33) cnt[lane] = 0;
    wave_lds_sync();
    auto window_of = [&](int p, bool &ok) {
        ok = p < n;
        const int z0 = ok ? (int)(__float_as_uint(geo[R.row(c0 + p)].w) >> 16) : 0;
        return ok ? min(z0 / S, 31) : 31;
    };
    auto match = [&](int wv, bool ok) {
        uint64_t m = __ballot(ok);
#pragma unroll
        for (int bt = 0; bt < 5; bt++) {
            const bool bit = (wv >> bt) & 1;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        return m;
    };
    for (int i0 = 0; i0 < n; i0 += 64) {
        bool ok;
        const int wv = window_of(i0 + lane, ok);
        const uint64_t m = match(wv, ok);
        if (ok && (m & lt) == 0) cnt[wv] += __popcll(m);  // one leader per window
        wave_lds_sync();
    }
    if (lane < 32) {
        // exclusive prefix over the 32 window counts
        const int v = cnt[lane];
        int x = v;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const int y = __shfl_up(x, d, 32);
            if (lane >= d) x += y;
        }
        wb[lane] = x - v;
        cnt[lane] = x - v;
        if (lane == 31) wb[32] = x;
    }
    wave_lds_sync();
    for (int i0 = 0; i0 < n; i0 += 64) {
        bool ok;
        const int wv = window_of(i0 + lane, ok);
        const uint64_t m = match(wv, ok);
        const int base = ok ? cnt[wv] : 0;
        wave_lds_sync();
        if (ok) {
            lst[base + __popcll(m & lt)] = (unsigned short)(i0 + lane);
            if ((m & lt) == 0) cnt[wv] = base + __popcll(m);
        }
        wave_lds_sync();
    }
}

// first sorted position of every start cell of a tile (run[0 .. ncell]):
// the tile's keys (tile << cshift | cell, ascending) are scanned once with
// coalesced loads; position i starts the cells after its predecessor's up
// to its own, and end starts the cells after the last one up to ncell.
__device__ __forceinline__ void cell_runs(const uint32_t *__restrict__ keys, int beg, int end,
                                          uint32_t key0, int cshift, int ncell, int *run)
{
    const uint32_t kb = key0 << cshift;
    for (int i = beg + (int)threadIdx.x; i <= end; i += blockDim.x) {
        const int cprev = i > beg ? (int)min(keys[i - 1] - kb, (uint32_t)ncell) : -1;
        const int ccur = i < end ? (int)min(keys[i] - kb, (uint32_t)ncell) : ncell;
        for (int c = cprev + 1; c <= ccur; c++) run[c] = i;
    }
}

// the points of work item q of a tile: quad (QS = 2) or cell (QS = 1)
template <int QS>
__device__ __forceinline__ Runs item_runs(const int *run, int q, int tx, int &lx0, int &ly0)
{
    const int qx = tx / QS;
    lx0 = (q % qx) * QS;
    ly0 = (q / qx) * QS;
    const int c = ly0 * tx + lx0;
    Runs R;
    R.a0 = run[c];
    R.n0 = run[c + QS] - R.a0;
    R.a1 = QS == 2 ? run[c + tx] : 0;
    R.n = R.n0 + (QS == 2 ? run[c + tx + 2] - R.a1 : 0);
    return R;
}

// Work items of a tile in descending order of their point counts (ties by
// index), so the waves' dynamic queue starts on the largest: ord[0..nitem)
// (nitem <= kDenseItems); cnt is scratch.  Uses the whole workgroup; ends
// with a barrier.
constexpr int kDenseItems = 256;
template <int QS>
__device__ __forceinline__ void order_items(const int *run, int nitem, int tx, int *cnt,
                                            unsigned short *ord)
{
    for (int i = threadIdx.x; i < nitem; i += blockDim.x) {
        int lx0, ly0;
        cnt[i] = item_runs<QS>(run, i, tx, lx0, ly0).n;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nitem; i += blockDim.x) {
        const int ci = cnt[i];
        int rank = 0;
        for (int j = 0; j < nitem; j++) {
            const int cj = cnt[j];
            rank += cj > ci || (cj == ci && j < i);
        }
        ord[rank] = (unsigned short)i;
    }
    __syncthreads();
}

// The waves' queue over a work unit: pieces of at most psz consecutive
// points of one item, items in order_items' order, so a unit whose points
// crowd into a few start cells still occupies every wave.  piece[p] =
// ordered item | part << 8 (points [part psz, min(n, (part + 1) psz)) of
// it); empty items have no piece.  The pieces number at most nitem + unit
// points / psz, i.e. <= kDensePieces when psz >= unit points / 16 (parts
// per item <= 17).  Call after order_items (cnt, ord); the first wave
// builds the table (a prefix sum over the items by lane shuffles); ends
// with a barrier.
constexpr int kDensePieces = kDenseItems + 16;
__device__ __forceinline__ void order_pieces(const int *cnt, const unsigned short *ord, int nitem,
                                             int psz, unsigned short *piece, int *npiece)
{
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int base = 0;
        for (int k0 = 0; k0 < nitem; k0 += 64) {
            const int k = k0 + lane;
            const int parts = k < nitem ? (cnt[ord[k]] + psz - 1) / psz : 0;
            int x = parts;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(x, d);
                if (lane >= d) x += y;
            }
            const int first = base + x - parts;
            for (int q = 0; q < parts; q++) piece[first + q] = (unsigned short)(k | q << 8);
            base += __shfl(x, 63);
        }
        if (lane == 0) *npiece = base;
    }
    __syncthreads();
}

template <int KW, int QS, int KM>
__global__ __launch_bounds__(kDense2Threads) void k_grid_dense2(const float4 *__restrict__ geo,
                                                                const float2 *__restrict__ val,
                                                                const uint32_t *__restrict__ keys,
                                                                const int32_t *__restrict__ tile_start,
                                                                int32_t *__restrict__ list,
                                                                int chunk, int psz,
                                                                float2 *__restrict__ fw, GridArgs g)
{
    constexpr int S = 9 - KW;             // planes between windows
    constexpr int SPAN = KW + QS - 1;     // frame positions per axis (<= 8)
    constexpr int NB = (SPAN + 1) / 2;    // 16-tap blocks (x stride 8, two y per block)
    constexpr int NWAVES = kDense2Threads / 64;
    static_assert(SPAN <= 8, "frame");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // footprint region, x fastest: [part][plane][y][x] (part 0 real, 1 imaginary)
    const int RX = g.tx + KW - 1, RY = g.ty + KW - 1, NP = g.nf3;
    const int rcells = RX * RY * NP;
    double *are = reinterpret_cast<double *>(smem_raw);
    int *run = reinterpret_cast<int *>(are + 2 * rcells);  // tx*ty + 1 run starts
    __shared__ unsigned short s_lst[NWAVES][kDense2Max];
    __shared__ int s_cnt[NWAVES][33], s_wb[NWAVES][33];
    __shared__ int s_next, s_tile;
    __shared__ int s_icnt[kDenseItems];
    __shared__ unsigned short s_ord[kDenseItems];
    __shared__ unsigned short s_piece[kDensePieces];
    __shared__ int s_npiece;
    const int ncell = g.tx * g.ty;
    const int nitem = ncell / (QS * QS);
    const int nlist = list[0];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // lane (k, r): point slot k of a group; MFMA column r = frame plane
    // col = r / 2, part r % 2; A row r = x position r (r < 8) or y parity
    const int k = lane >> 4, r = lane & 15, col = r >> 1, part = r & 1;
    const bool isx = r < 8;
    const int NW = (NP + S - 1) / S;  // windows (<= 32)
    // kernel values / e^beta (sums scaled by e^(3 beta)): the exp2 form or
    // the per-tap polynomials
    __shared__ __align__(16) float s_tp[8 * kTapStride];
    if (KM == kKerPoly && threadIdx.x < 8 * kTapStride) s_tp[threadIdx.x] = g.tp[threadIdx.x];
    __syncthreads();
    const KernelEvalN kex(g, KW);
    const TapPolyN kpo{s_tp, KW};
    auto ker = [&](float off0, int t) {
        if constexpr (KM == kKerPoly)
            return kpo(off0, t);
        else
            return kex.template at<KM>(off0, t);
    };
    const double sc = g.e3;
    unsigned short *lst = s_lst[wave];
    int *cnt = s_cnt[wave], *wb = s_wb[wave];
    // lane constants: the e1 position (x of A row r, or y of r - 8) and the
    // start-cell byte of the point record it subtracts
    const int e1pos = r & 7, e1shift = isx ? 0 : 8;
    const int src_kx = isx ? lane : lane - 8;  // x value of A row r

    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&list[1], 1);
        __syncthreads();
        const int li = s_tile;
        if (li >= nlist) break;
        // the unit: points [beg, end) of bin `tile`
        const int tile = __builtin_amdgcn_readfirstlane(list[2 + 2 * li]);
        const int tb = tile_start[tile], te = tile_start[tile + 1];
        const int beg = __builtin_amdgcn_readfirstlane(tb + list[3 + 2 * li]);
        const int end = __builtin_amdgcn_readfirstlane((int)min((int64_t)te, (int64_t)beg + chunk));
        const int bx = tile % g.ntx, by = tile / g.ntx;
        for (int q = threadIdx.x; q < 2 * rcells; q += blockDim.x) are[q] = 0.0;
        cell_runs(keys, beg, end, (uint32_t)g.kr.lower(tile, ntiles), g.cshift, ncell, run);
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        order_items<QS>(run, nitem, g.tx, s_icnt, s_ord);
        order_pieces(s_icnt, s_ord, nitem, psz, s_piece, &s_npiece);
        const int npiece = s_npiece;
        for (;;) {
            int it = 0;
            if (lane == 0) it = atomicAdd(&s_next, 1);
            it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
            if (it >= npiece) break;
            const int pc = __builtin_amdgcn_readfirstlane((int)s_piece[it]);
            int lx0, ly0;
            Runs R = item_runs<QS>(run, __builtin_amdgcn_readfirstlane((int)s_ord[pc & 255]), g.tx, lx0, ly0);
            R.a0 = __builtin_amdgcn_readfirstlane(R.a0);
            R.n0 = __builtin_amdgcn_readfirstlane(R.n0);
            R.a1 = __builtin_amdgcn_readfirstlane(R.a1);
            R.n = __builtin_amdgcn_readfirstlane(R.n);
            // the item's records from a0: position i at i, or i + d1 past the
            // first run (d1 >= 0)
            const float4 *gi = geo + R.a0;
            const float2 *vi = val + R.a0;
            const unsigned d1 = (unsigned)(R.a1 - R.a0 - R.n0);
            const int l0 = isx ? lx0 : ly0;
            // the piece: item points [pb, pe)
            const int pb = (pc >> 8) * psz, pe = min(R.n, pb + psz);
            for (int c0 = pb; c0 < pe; c0 += kDense2Max) {
                const int n = min(kDense2Max, pe - c0);
                sort_by_window<S>(geo, R, c0, n, lst, cnt, wb, lane);
                // ---- windows present
                const bool has = lane < NW && wb[lane + 1] > wb[lane];
                const uint64_t hm = __ballot(has);
                if (hm == 0) continue;  // no point in range (cannot happen for valid records)
                const int wfirst = __ffsll((unsigned long long)hm) - 1;
                const int wlast = 63 - __clzll(hm);
                v4f C[NB];
#pragma unroll
                for (int q = 0; q < NB; q++) C[q] = v4f{0.f, 0.f, 0.f, 0.f};
                // add frame planes rel [0, np) of window w to the region: the
                // lane's registers (blk q, v) hold x = 4 (k % 2) + v, y = 2 q + k / 2
                auto flush = [&](int w, int np) {
                    if (col >= np) return;
                    int p = S * w + col;
                    if (p >= NP) p -= NP;
                    double *b = are + part * rcells + ((p * RY + ly0 + (k >> 1)) * RX + lx0 + 4 * (k & 1));
#pragma unroll
                    for (int q = 0; q < NB; q++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int x = 4 * (k & 1) + v, y = 2 * q + (k >> 1);
                            if (x < SPAN && y < SPAN) atomicAdd(b + 2 * q * RX + v, (double)C[q][v]);
                        }
                };
                // slot k of the next group, loaded one group ahead: groups
                // never cross a window, and the group after the last one of
                // window w starts at wb[w + 1] whichever window is next;
                // positions past the item's end read its last point
                auto fetch = [&](int pos, float4 &Gq, float2 &vq) {
                    const unsigned i = (unsigned)(c0 + lst[min(pos, n - 1)]);
                    const unsigned o = i < (unsigned)R.n0 ? i : i + d1;
                    Gq = gi[o];
                    vq = vi[o];
                };
                float4 Gn;
                float2 vn;
                fetch(k, Gn, vn);  // wb[wfirst] = 0
                for (int w = wfirst; w <= wlast; w++) {
                    if (w > wfirst) {
                        flush(w - 1, S);
                        // shift the frame down by S planes: DPP row_shl:2S
                        // reads lane r + 2S of the same 16-lane row, and reads
                        // past the row (planes col + S >= 8) give zero
#pragma unroll
                        for (int q = 0; q < NB; q++)
#pragma unroll
                            for (int v = 0; v < 4; v++)
                                C[q][v] = __int_as_float(__builtin_amdgcn_update_dpp(
                                    0, __float_as_int(C[q][v]), 0x100 + 2 * S, 0xf, 0xf, true));
                    }
                    const int e = wb[w + 1];
                    const int cw = col + S * w;  // plane of the lane's column
                    for (int j0 = wb[w]; j0 < e; j0 += 4) {
                        const bool ok = j0 + k < e;
                        const float4 G = Gn;
                        const float2 v = vn;
                        fetch((j0 + 4 < e ? j0 + 4 : e) + k, Gn, vn);
                        // tap index of the lane's e1 position (x or y) and of
                        // its plane: position minus the point's start cell
                        const uint32_t pk = __float_as_uint(G.w);
                        const int dq = (int)((pk >> e1shift) & 0xffu) - l0;
                        const float e1 = ker(isx ? G.x : G.y, e1pos - dq);
                        const float ex = __shfl(e1, src_kx);
                        const float kz = ker(G.z, cw - (int)(pk >> 16));
                        // an empty slot (past the window's end) holds another
                        // point: its kernel values are finite, and a zero B
                        // row drops it
                        const float bv = ok ? (part ? v.y : v.x) * kz : 0.f;
#pragma unroll
                        for (int q = 0; q < NB; q++) {
                            const float ey = __shfl(e1, 16 * k + 8 + 2 * q + (r >> 3));
                            C[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ex * ey, bv, C[q], 0, 0, 0);
                        }
                    }
                }
                flush(wlast, 8);
            }
        }
        __syncthreads();
        // footprint region -> grid, x fastest so a wavefront covers row runs
        for (int q = threadIdx.x; q < rcells; q += blockDim.x) {
            const int rx = q % RX;
            const int rest = q / RX;
            const int ry = rest % RY, p = rest / RY;
            int gx = bx * g.tx + rx;
            if (gx >= g.nf1) gx -= g.nf1;
            int gy = by * g.ty + ry;
            if (gy >= g.nf2) gy -= g.nf2;
            float2 *dst = fw + ((size_t)p * g.nf2 + gy) * g.nf1 + gx;
            atomicAdd(&dst->x, (float)(are[q] * sc));
            atomicAdd(&dst->y, (float)(are[rcells + q] * sc));
        }
        __syncthreads();
    }
}

// static LDS of k_grid_dense2 (sorted lists and window counters of 8 waves)
constexpr size_t kDense2Static = (size_t)(kDense2Threads / 64) * (2 * kDense2Max + 2 * 33 * 4) + 64 + 6 * kDenseItems +
                                 2 * kDensePieces + 4;

static size_t dense_lds_bytes(const GridArgs &g, int kw)
{
    const size_t rcells = (size_t)(g.tx + kw - 1) * (g.ty + kw - 1) * g.nf3;
    return 16 * rcells + 4 * (size_t)(g.tx * g.ty + 1);
}

bool grid_dense_supported(const GridArgs &g, int kw, int elem)
{
    return elem == 8 && kw >= 2 && kw <= 8 && g.nf3 >= 2 * kw && g.nf3 <= 32 && g.cshift > 0 &&
           g.tx * g.ty <= kDenseItems && g.tx + kw - 1 <= g.nf1 && g.ty + kw - 1 <= g.nf2 &&
           dense_lds_bytes(g, kw) + kDense2Static <= 163840 - 64;
}


template <int KW>
static int grid_dense2_kw(const float4 *geo, const float2 *val, const uint32_t *keys,
                          const int32_t *ts, int32_t *list, float2 *fw, const GridArgs &g,
                          hipStream_t s)
{
    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    const int chunk = dense_chunk(g.m);
    const int64_t units = ntiles + ceil_div((int64_t)g.m, (int64_t)chunk);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, list, 1);
    hipLaunchKernelGGL(k_dense_bins, dim3((unsigned)ceil_div(ntiles, 256)), dim3(256), 0, s, ts,
                       ntiles, g.dense_min, chunk, list);
    CG_HIP(hipGetLastError());
    const size_t lds = dense_lds_bytes(g, KW);
    const int per_cu = std::max<int>(1, (int)(163840 / (lds + kDense2Static)));
    const unsigned nblk = (unsigned)std::min<int64_t>(units, (int64_t)256 * per_cu);
    // pieces: about one per wave of a full unit
    const int psz = (int)ceil_div(chunk, kDense2Threads / 64);
    // quads of start cells where their footprints fit the 8 x 8 frame
    auto go = [&](const void *fn, auto kern) -> int {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(kDense2Threads), lds, s, geo, val, keys, ts, list,
                           chunk, psz, fw, g);
        CG_HIP(hipGetLastError());
        return CG_OK;
    };
    const int km = ker_mode(g);
    auto pick = [&](auto qs) -> int {
        constexpr int QS = decltype(qs)::value;
        if (km == kKerPoly) return go((const void *)k_grid_dense2<KW, QS, kKerPoly>, k_grid_dense2<KW, QS, kKerPoly>);
        if (km == kKerES) return go((const void *)k_grid_dense2<KW, QS, kKerES>, k_grid_dense2<KW, QS, kKerES>);
        return CG_ERR_ARG;
    };
    if constexpr (KW <= 7)
        if (g.tx % 2 == 0 && g.ty % 2 == 0) return pick(std::integral_constant<int, 2>{});
    return pick(std::integral_constant<int, 1>{});
}

int launch_grid_dense(const geo_t<float> *geo, const cx_t<float> *val, const uint32_t *keys,
                      const int32_t *ts, int32_t *list, cx_t<float> *fw, const GridArgs &g, int kw,
                      hipStream_t s)
{
    if (!grid_dense_supported(g, kw, 8) || g.dense_min <= 0) return CG_ERR_ARG;
    switch (kw) {
#define CG_CASE(K) \
    case K: return grid_dense2_kw<K>(geo, val, keys, ts, list, fw, g, s);
        CG_CASE(2) CG_CASE(3) CG_CASE(4) CG_CASE(5) CG_CASE(6) CG_CASE(7) CG_CASE(8)
#undef CG_CASE
    default: return CG_ERR_ARG;
    }
}

template <typename T>
int launch_grid(const geo_t<T> *geo, const cx_t<T> *val, const int32_t *ts, cx_t<T> *fw,
                const GridArgs &g, int kw, hipStream_t s)
{
    switch (kw) {
#define CG_CASE(K) \
    case K: return grid_kw<T, K>(geo, val, ts, fw, g, s);
        CG_CASE(2) CG_CASE(3) CG_CASE(4) CG_CASE(5) CG_CASE(6) CG_CASE(7) CG_CASE(8) CG_CASE(9)
        CG_CASE(10) CG_CASE(11) CG_CASE(12) CG_CASE(13) CG_CASE(14) CG_CASE(15) CG_CASE(16)
#undef CG_CASE
    default: return CG_ERR_ARG;
    }
}

// ------------------------------------------------------------ degridding

// Tiles of the degridder's list: the non-empty ones below dense_min (all
// non-empty ones when dense_min <= 0).  list[0] counts them, list[1..]
// holds them (any order), list[ntiles + 1] is the work counter.
__global__ __launch_bounds__(256) void k_sparse_tiles(const int32_t *__restrict__ tile_start,
                                                      int64_t ntiles, int dense_min,
                                                      int32_t *__restrict__ list)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= ntiles) return;
    const int n = tile_start[t + 1] - tile_start[t];
    if (n > 0 && (dense_min <= 0 || n < dense_min)) list[1 + atomicAdd(list, 1)] = (int32_t)t;
}

template <typename T, int KW, int PPT>
__device__ __forceinline__ void degrid_tile(const geo_t<T> *__restrict__ geo,
                                            const cx_t<T> *__restrict__ fac,
                                            const int32_t *__restrict__ perm,
                                            const int32_t *__restrict__ tile_start,
                                            const cx_t<T> *__restrict__ fw,
                                            cx_t<T> *__restrict__ out, const GridArgs &g, int is2d,
                                            int tile)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *lds = reinterpret_cast<cx_t<T> *>(smem_raw);
    const int beg = tile_start[tile], end = tile_start[tile + 1];
    if (beg == end) return;
    if (g.dense_min > 0 && end - beg >= g.dense_min) return;  // k_degrid_dense's
    const int bx = tile % g.ntx, by = tile / g.ntx;
    const int hx = g.tx + KW - 1, hy = g.ty + KW - 1;
    // LDS row stride: odd, so that cells a power-of-two number of rows or
    // columns apart do not share banks
    const int hxs = degrid_row_stride(g.tx, KW);
    const int x0 = bx * g.tx, y0 = by * g.ty;
    const KernelEval<T> ker(g, KW);
    const bool single_stage = g.nchunk == 1;
    const int nzt = is2d ? 1 : KW;

    auto stage = [&](int p0, int np) {
        const int pcells = hy * hx;
        const int n = np * pcells;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            int lz = k / pcells;
            int rem = k - lz * pcells;
            int y = rem / hx, x = rem - y * hx;
            int gx = x0 + x;
            if (gx >= g.nf1) gx -= g.nf1;
            int gy = y0 + y;
            if (gy >= g.nf2) gy -= g.nf2;
            lds[(lz * hy + y) * hxs + x] = fw[((size_t)(p0 - g.zb + lz) * g.nf2 + gy) * g.nf1 + gx];
        }
    };

    // the buffer holds planes [zb, zb + zn); stage() takes absolute planes
    const int zn = slab_planes(g);
    if (single_stage) {
        stage(g.zb, zn);
        __syncthreads();
    }
    for (int base = beg; base < end; base += (int)blockDim.x * PPT) {
        T ar[PPT], ai[PPT];
#pragma unroll
        for (int k = 0; k < PPT; k++) ar[k] = ai[k] = (T)0;
        for (int c = 0; c < g.nchunk; c++) {
            const int p0 = g.zb + c * g.pc;
            const int np = min(g.pc, g.zb + zn - p0);
            if (!single_stage) {
                __syncthreads();
                stage(p0, np);
                __syncthreads();
            }
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                const int i = base + (int)threadIdx.x + k * (int)blockDim.x;
                if (i >= end) continue;
                const geo_t<T> G = geo[i];
                const uint32_t pk = geo_bits(G.w);
                const int lx = min((int)(pk & 0xffu), g.tx - 1);
                const int ly = min((int)((pk >> 8) & 0xffu), g.ty - 1);
                int rz = (int)(pk >> 16) - p0;
                if (rz < 0) rz += g.nf3;
                if (is2d) rz = 0;
                else if (rz >= np && rz + KW - 1 < g.nf3) continue;
                T kx[KW], ky[KW];
#pragma unroll
                for (int t = 0; t < KW; t++) {
                    kx[t] = ker(G.x, t);
                    ky[t] = ker(G.y, t);
                }
                for (int tz = 0; tz < nzt; tz++) {
                    int zz = rz + tz;
                    if (zz >= g.nf3) zz -= g.nf3;
                    if (zz >= np) continue;
                    const T kzv = is2d ? (T)1 : ker(G.z, tz);
                    const cx_t<T> *pl = lds + ((size_t)zz * hy + ly) * hxs + lx;
                    T sr = 0, si = 0;
#pragma unroll
                    for (int t2 = 0; t2 < KW; t2++) {
                        T rr = 0, ri = 0;
#pragma unroll
                        for (int t1 = 0; t1 < KW; t1++) {
                            const cx_t<T> f = pl[t2 * hxs + t1];
                            rr += f.x * kx[t1];
                            ri += f.y * kx[t1];
                        }
                        sr += rr * ky[t2];
                        si += ri * ky[t2];
                    }
                    ar[k] += sr * kzv;
                    ai[k] += si * kzv;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int i = base + (int)threadIdx.x + k * (int)blockDim.x;
            if (i >= end) continue;
            const cx_t<T> f = fac[i];
            const int j = perm[i];
            if (j >= 0 && j < g.m) {  // a valid permutation always passes
                cx_t<T> r = mkc(ar[k] * f.x - ai[k] * f.y, ar[k] * f.y + ai[k] * f.x);
                if (g.accumulate) {
                    const cx_t<T> prev = out[j];
                    r.x += prev.x;
                    r.y += prev.y;
                }
                out[j] = r;
            }
        }
    }
}

// Dense bins of the degridder (complex64), the adjoint of k_grid_dense2.  A
// workgroup takes work units as k_grid_dense2 does, stages the unit's tile
// plus the (kw-1) halo of every plane in LDS (as k_degrid does) and its
// waves take pieces of quads (2 x 2 start cells; cells at kw = 8).  A piece
// with at least kDegrid2Item points is ordered by w-window
// (sort_by_window) and each window's points are interpolated sixteen at a
// time as one 16 x 16 product on v_mfma_f32_16x16x4f32:
//   R[j][col] = sum_tap W[j][tap] F[tap][col],  W = kx_j(x - dx_j) ky_j(y - dy_j)
// over the 8 x 8 frame taps (four per step: x = 4 (step % 2) + k, y =
// step / 2), with F the staged grid values of the window's eight planes
// (col = 2 plane + part, part 0 real, 1 imaginary; loaded once per window
// and shared by its groups), then v_j = sum_plane kz_j(plane) R[j][plane],
// a DPP reduction across the lanes of each 16-lane row.  Each lane evaluates
// six kernel values of one point per group (two each along x, y and z); the
// y and z values reach the lanes that need them through a small per-wave
// LDS table.  Pieces with fewer points give each point a lane.
constexpr int kDegrid2Threads = 1024;
constexpr int kDegrid2Item = 48;

template <int KW, int QS, int KM>
__global__ __launch_bounds__(kDegrid2Threads) void k_degrid_dense(const float4 *__restrict__ geo,
                                                                  const float2 *__restrict__ fac,
                                                                  const int32_t *__restrict__ perm,
                                                                  const uint32_t *__restrict__ keys,
                                                                  const int32_t *__restrict__ tile_start,
                                                                  int32_t *__restrict__ list,
                                                                  int chunk, int psz,
                                                                  const float2 *__restrict__ fw,
                                                                  float2 *__restrict__ out, GridArgs g)
{
    constexpr int S = 9 - KW;
    constexpr int SPAN = KW + QS - 1;
    constexpr int NCH = 2 * SPAN;  // tap steps: y = ch / 2 < SPAN
    constexpr int NWAVES = kDegrid2Threads / 64;
    static_assert(SPAN <= 8, "frame");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *lds = reinterpret_cast<float2 *>(smem_raw);
    const float *ldsf = reinterpret_cast<const float *>(smem_raw);
    __shared__ unsigned short s_lst[NWAVES][kDense2Max];
    __shared__ int s_cnt[NWAVES][33], s_wb[NWAVES][33];
    __shared__ int s_run[257];
    __shared__ int s_next, s_tile;
    __shared__ int s_icnt[kDenseItems];
    __shared__ unsigned short s_ord[kDenseItems];
    __shared__ unsigned short s_piece[kDensePieces];
    __shared__ int s_npiece;
    // per-wave exchange of a group's kernel values: ky[point][y], kz[plane][point]
    __shared__ __align__(16) float s_ky[NWAVES][16 * 8];
    __shared__ __align__(16) float s_kz[NWAVES][8 * 16];
    const int NP = g.nf3;
    const int hx = g.tx + KW - 1, hy = g.ty + KW - 1;
    const int hxs = degrid_row_stride(g.tx, KW);
    const int ncell = g.tx * g.ty;  // <= 256
    const int nitem = ncell / (QS * QS);
    const int nlist = list[0];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // lane (hi, lo): A operand point lo, tap hi of a step; B operand column
    // lo (plane lo / 2, part lo % 2); result points 4 hi + v at column lo
    const int hi = lane >> 4, lo = lane & 15;
    const int NW = (NP + S - 1) / S;
    // kernel values / e^beta (sums of products scaled by e^(3 beta)): the
    // exp2 form or the per-tap polynomials
    __shared__ __align__(16) float s_tp[8 * kTapStride];
    if (KM == kKerPoly && threadIdx.x < 8 * kTapStride) s_tp[threadIdx.x] = g.tp[threadIdx.x];
    __syncthreads();
    const KernelEvalN kex(g, KW);
    const TapPolyN kpo{s_tp, KW};
    auto ker = [&](float off0, int t) {
        if constexpr (KM == kKerPoly)
            return kpo(off0, t);
        else
            return kex.template at<KM>(off0, t);
    };
    const float sc = (float)g.e3;
    unsigned short *lst = s_lst[wave];
    int *cnt = s_cnt[wave], *wb = s_wb[wave];
    float *ky_t = s_ky[wave], *kz_t = s_kz[wave];

    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&list[1], 1);
        __syncthreads();
        const int li = s_tile;
        if (li >= nlist) break;
        // the unit: points [beg, end) of bin `tile`
        const int tile = __builtin_amdgcn_readfirstlane(list[2 + 2 * li]);
        const int tb = tile_start[tile], te = tile_start[tile + 1];
        const int beg = __builtin_amdgcn_readfirstlane(tb + list[3 + 2 * li]);
        const int end = __builtin_amdgcn_readfirstlane((int)min((int64_t)te, (int64_t)beg + chunk));
        const int bx = tile % g.ntx, by = tile / g.ntx;
        const int x0 = bx * g.tx, y0 = by * g.ty;
        // the tile and its halo of every plane
        const int pcells = hy * hx;
        for (int q = threadIdx.x; q < NP * pcells; q += blockDim.x) {
            const int lz = q / pcells;
            const int rem = q - lz * pcells;
            const int y = rem / hx, x = rem - y * hx;
            int gx = x0 + x;
            if (gx >= g.nf1) gx -= g.nf1;
            int gy = y0 + y;
            if (gy >= g.nf2) gy -= g.nf2;
            lds[(lz * hy + y) * hxs + x] = fw[((size_t)lz * g.nf2 + gy) * g.nf1 + gx];
        }
        cell_runs(keys, beg, end, (uint32_t)g.kr.lower(tile, ntiles), g.cshift, ncell, s_run);
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        order_items<QS>(s_run, nitem, g.tx, s_icnt, s_ord);
        order_pieces(s_icnt, s_ord, nitem, psz, s_piece, &s_npiece);
        const int npiece = s_npiece;
        for (;;) {
            int it = 0;
            if (lane == 0) it = atomicAdd(&s_next, 1);
            it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
            if (it >= npiece) break;
            const int pc = __builtin_amdgcn_readfirstlane((int)s_piece[it]);
            int lx0, ly0;
            Runs R = item_runs<QS>(s_run, __builtin_amdgcn_readfirstlane((int)s_ord[pc & 255]), g.tx, lx0, ly0);
            R.a0 = __builtin_amdgcn_readfirstlane(R.a0);
            R.n0 = __builtin_amdgcn_readfirstlane(R.n0);
            R.a1 = __builtin_amdgcn_readfirstlane(R.a1);
            R.n = __builtin_amdgcn_readfirstlane(R.n);
            // the piece: item points [pb, pe)
            const int pb = (pc >> 8) * psz, pe = min(R.n, pb + psz);
            if (pe - pb < kDegrid2Item) {
                // one lane per point: all KW^3 taps from LDS
                for (int i = pb + lane; i < pe; i += 64) {
                    const int row = R.row(i);
                    const float4 G = geo[row];
                    const uint32_t pk = __float_as_uint(G.w);
                    const int lx = (int)(pk & 0xffu), ly = (int)((pk >> 8) & 0xffu);
                    float kx[KW], ky[KW];
#pragma unroll
                    for (int t = 0; t < KW; t++) {
                        kx[t] = ker(G.x, t);
                        ky[t] = ker(G.y, t);
                    }
                    const int z0 = (int)(pk >> 16);
                    float ar = 0.f, ai = 0.f;
                    for (int tz = 0; tz < KW; tz++) {
                        int zz = z0 + tz;
                        if (zz >= NP) zz -= NP;
                        const float kzv = ker(G.z, tz);
                        const float2 *pl = lds + ((size_t)zz * hy + ly) * hxs + lx;
                        float sr = 0.f, si = 0.f;
#pragma unroll
                        for (int t2 = 0; t2 < KW; t2++) {
                            float rr = 0.f, ri = 0.f;
#pragma unroll
                            for (int t1 = 0; t1 < KW; t1++) {
                                const float2 f = pl[t2 * hxs + t1];
                                rr += f.x * kx[t1];
                                ri += f.y * kx[t1];
                            }
                            sr += rr * ky[t2];
                            si += ri * ky[t2];
                        }
                        ar += sr * kzv;
                        ai += si * kzv;
                    }
                    ar *= sc;
                    ai *= sc;
                    const float2 f = fac[row];
                    const int dst = perm[row];
                    if (dst >= 0 && dst < g.m) out[dst] = make_float2(ar * f.x - ai * f.y, ar * f.y + ai * f.x);
                }
                continue;
            }
            // the item's records from a0: position i at i, or i + d1 past the
            // first run (d1 >= 0)
            const float4 *gi = geo + R.a0;
            const unsigned d1 = (unsigned)(R.a1 - R.a0 - R.n0);
            auto rec = [&](int i) -> unsigned {
                return (unsigned)i < (unsigned)R.n0 ? (unsigned)i : (unsigned)i + d1;
            };
            for (int c0 = pb; c0 < pe; c0 += kDense2Max) {
                const int n = min(kDense2Max, pe - c0);
                sort_by_window<S>(geo, R, c0, n, lst, cnt, wb, lane);
                const bool has = lane < NW && wb[lane + 1] > wb[lane];
                const uint64_t hm = __ballot(has);
                if (hm == 0) continue;
                const int wfirst = __ffsll((unsigned long long)hm) - 1;
                const int wlast = 63 - __clzll(hm);
                // point lo of the next group, loaded one group ahead (the
                // group after the last one of window w starts at wb[w + 1]);
                // positions past the item's end read its last point
                auto fetch = [&](int pos) { return gi[rec(c0 + lst[min(pos, n - 1)])]; };
                float4 Gn = fetch(lo);  // wb[wfirst] = 0
                for (int w = wfirst; w <= wlast; w++) {
                    const int e = wb[w + 1];
                    if (wb[w] == e) continue;
                    // the window's grid values: step ch holds tap (x, y) =
                    // (4 (ch % 2) + hi, ch / 2) of column lo
                    int p = S * w + (lo >> 1);
                    if (p >= NP) p -= NP;
                    const float *fb = ldsf + 2 * (((size_t)p * hy + ly0) * hxs + lx0 + hi) + (lo & 1);
                    float B[NCH];
#pragma unroll
                    for (int ch = 0; ch < NCH; ch++) {
                        const int x = 4 * (ch & 1) + hi;
                        B[ch] = x < SPAN ? fb[2 * ((ch >> 1) * hxs + 4 * (ch & 1))] : 0.f;
                    }
                    const int zw = S * w;
                    for (int j0 = wb[w]; j0 < e; j0 += 16) {
                        const float4 G = Gn;
                        Gn = fetch((j0 + 16 < e ? j0 + 16 : e) + lo);
                        // the writer lanes (lo even, lo < 8) own point
                        // 4 hi + lo / 2 of the group: its record, early
                        const int pw = 4 * hi + (lo >> 1);
                        const bool wr = (lo & 9) == 0 && j0 + pw < e;
                        unsigned ow = 0;
                        float2 fw2 = make_float2(0.f, 0.f);
                        int dst = -1;
                        if (wr) {
                            ow = R.a0 + rec(c0 + lst[j0 + pw]);
                            fw2 = fac[ow];
                            dst = perm[ow];
                        }
                        const uint32_t pk = __float_as_uint(G.w);
                        const int dx = (int)(pk & 0xffu) - lx0, dy = (int)((pk >> 8) & 0xffu) - ly0;
                        const int dz = (int)(pk >> 16) - zw;
                        const float kxa = ker(G.x, hi - dx), kxb = ker(G.x, hi + 4 - dx);
                        const float2 kyv = make_float2(ker(G.y, 2 * hi - dy), ker(G.y, 2 * hi + 1 - dy));
                        const float kz0 = ker(G.z, 2 * hi - dz), kz1 = ker(G.z, 2 * hi + 1 - dz);
                        wave_lds_sync();  // the previous group's reads are done
                        *reinterpret_cast<float2 *>(ky_t + lo * 8 + 2 * hi) = kyv;
                        kz_t[(2 * hi) * 16 + lo] = kz0;
                        kz_t[(2 * hi + 1) * 16 + lo] = kz1;
                        wave_lds_sync();
                        const float4 ya = *reinterpret_cast<const float4 *>(ky_t + lo * 8);
                        const float4 yb = *reinterpret_cast<const float4 *>(ky_t + lo * 8 + 4);
                        const float4 kzv = *reinterpret_cast<const float4 *>(kz_t + (lo >> 1) * 16 + 4 * hi);
                        const float ky[8] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y, yb.z, yb.w};
                        v4f D0 = v4f{0.f, 0.f, 0.f, 0.f}, D1 = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int ch = 0; ch < NCH; ch += 2) {
                            D0 = __builtin_amdgcn_mfma_f32_16x16x4f32(kxa * ky[ch >> 1], B[ch], D0, 0, 0, 0);
                            D1 = __builtin_amdgcn_mfma_f32_16x16x4f32(kxb * ky[ch >> 1], B[ch + 1], D1, 0, 0, 0);
                        }
                        // z weights, then the sum over the eight planes of
                        // each part: lanes of equal parity in the row
                        float s[4] = {(D0[0] + D1[0]) * kzv.x, (D0[1] + D1[1]) * kzv.y,
                                      (D0[2] + D1[2]) * kzv.z, (D0[3] + D1[3]) * kzv.w};
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            s[v] += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s[v]), 0x122, 0xf, 0xf, false));
                            s[v] += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s[v]), 0x124, 0xf, 0xf, false));
                            s[v] += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s[v]), 0x128, 0xf, 0xf, false));
                        }
                        // lane lo keeps point 4 hi + (lo / 2) % 4; an even lane
                        // takes the imaginary part from lane lo + 1
                        const int vs = (lo >> 1) & 3;
                        const float mine = vs == 0 ? s[0] : vs == 1 ? s[1] : vs == 2 ? s[2] : s[3];
                        const float other = __int_as_float(
                            __builtin_amdgcn_update_dpp(0, __float_as_int(mine), 0xb1, 0xf, 0xf, false));
                        if (wr && dst >= 0 && dst < g.m) {
                            const float sr = mine * sc, si = other * sc;
                            out[dst] = make_float2(sr * fw2.x - si * fw2.y, sr * fw2.y + si * fw2.x);
                        }
                    }
                }
            }
        }
        __syncthreads();  // the staged tile and the run table are reused
    }
}

// static LDS of k_degrid_dense
constexpr size_t kDegrid2Static =
    (size_t)(kDegrid2Threads / 64) * (2 * kDense2Max + 2 * 33 * 4 + 2 * 128 * 4) + 257 * 4 + 64 +
    6 * kDenseItems + 2 * kDensePieces + 4;

template <int KW>
static int degrid_dense_kw(const float4 *geo, const float2 *fac, const int32_t *perm,
                           const uint32_t *keys, const int32_t *ts, int32_t *list, const float2 *fw,
                           float2 *out, const GridArgs &g, hipStream_t s)
{
    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    const int chunk = dense_chunk(g.m);
    const int64_t units = ntiles + ceil_div((int64_t)g.m, (int64_t)chunk);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, list, 1);
    hipLaunchKernelGGL(k_dense_bins, dim3((unsigned)ceil_div(ntiles, 256)), dim3(256), 0, s, ts,
                       ntiles, g.dense_min, chunk, list);
    CG_HIP(hipGetLastError());
    const size_t lds = (size_t)degrid_lds_bytes(g, KW, 8);
    const unsigned nblk = (unsigned)std::min<int64_t>(units, 256);
    // pieces: about one per wave of a full unit, at least 64 points (the
    // MFMA path's share; the piece table needs psz >= chunk / 16)
    const int psz = std::max<int>(64, (int)ceil_div(chunk, kDegrid2Threads / 64));
    auto go = [&](const void *fn, auto kern) -> int {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(kDegrid2Threads), lds, s, geo, fac, perm, keys, ts,
                           list, chunk, psz, fw, out, g);
        CG_HIP(hipGetLastError());
        return CG_OK;
    };
    const int km = ker_mode(g);
    auto pick = [&](auto qs) -> int {
        constexpr int QS = decltype(qs)::value;
        if (km == kKerPoly) return go((const void *)k_degrid_dense<KW, QS, kKerPoly>, k_degrid_dense<KW, QS, kKerPoly>);
        if (km == kKerES) return go((const void *)k_degrid_dense<KW, QS, kKerES>, k_degrid_dense<KW, QS, kKerES>);
        return CG_ERR_ARG;
    };
    if constexpr (KW <= 7)
        if (g.tx % 2 == 0 && g.ty % 2 == 0) return pick(std::integral_constant<int, 2>{});
    return pick(std::integral_constant<int, 1>{});
}

bool degrid_dense_supported(const GridArgs &g, int kw)
{
    return kw >= 2 && kw <= 8 && g.nf3 >= 2 * kw && g.nf3 <= 32 && g.cshift > 0 && g.nchunk == 1 &&
           g.zn == 0 && g.tx * g.ty <= 256 && g.tx + kw - 1 <= g.nf1 && g.ty + kw - 1 <= g.nf2 &&
           (size_t)degrid_lds_bytes(g, kw, 8) + kDegrid2Static <= 163840 - 64;
}

// One tile per workgroup (LISTED false: tile blockIdx.x), or a persistent
// grid taking the tiles of list (k_sparse_tiles) one at a time.  Two
// instantiations, so the one-tile form keeps its scalar registers free of
// the work counter's state.
template <typename T, int KW, int PPT, bool LISTED>
__global__ __launch_bounds__(1024) void k_degrid(const geo_t<T> *__restrict__ geo,
                                                const cx_t<T> *__restrict__ fac,
                                                const int32_t *__restrict__ perm,
                                                const int32_t *__restrict__ tile_start,
                                                const cx_t<T> *__restrict__ fw,
                                                cx_t<T> *__restrict__ out, GridArgs g, int is2d,
                                                int32_t *__restrict__ list, int64_t ntiles)
{
    if constexpr (!LISTED) {
        degrid_tile<T, KW, PPT>(geo, fac, perm, tile_start, fw, out, g, is2d, blockIdx.x);
    } else {
        __shared__ int s_tile;
        const int nlist = list[0];
        for (;;) {
            if (threadIdx.x == 0) s_tile = atomicAdd(&list[ntiles + 1], 1);
            __syncthreads();
            const int li = s_tile;
            if (li >= nlist) break;
            // uniform across the workgroup: tile-derived addresses stay scalar
            degrid_tile<T, KW, PPT>(geo, fac, perm, tile_start, fw, out, g, is2d,
                                    __builtin_amdgcn_readfirstlane(list[1 + li]));
            __syncthreads();  // the staged planes and s_tile are reused
        }
    }
}

int launch_degrid_dense(const float4 *geo, const float2 *fac, const int32_t *perm,
                        const uint32_t *keys, const int32_t *ts, int32_t *list, const float2 *fw,
                        float2 *out, const GridArgs &g, int kw, hipStream_t s)
{
    if (!degrid_dense_supported(g, kw) || g.dense_min <= 0) return CG_ERR_ARG;
    switch (kw) {
#define CG_CASE(K) \
    case K: return degrid_dense_kw<K>(geo, fac, perm, keys, ts, list, fw, out, g, s);
        CG_CASE(2) CG_CASE(3) CG_CASE(4) CG_CASE(5) CG_CASE(6) CG_CASE(7) CG_CASE(8)
#undef CG_CASE
    default: return CG_ERR_ARG;
    }
}

template <typename T, int KW>
static int degrid_kw(const geo_t<T> *geo, const cx_t<T> *fac, const int32_t *perm,
                     const int32_t *ts, const cx_t<T> *fw, cx_t<T> *out, const GridArgs &g,
                     int32_t *list, hipStream_t s)
{
    int is2d = g.nf3 == 1 ? 1 : 0;
    size_t lds = (size_t)degrid_lds_bytes(g, KW, (int)sizeof(cx_t<T>));
    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    unsigned nblk = (unsigned)ntiles;
    const int threads = grid_threads(lds);
    if (list) {
        // a persistent grid over the listed tiles: as many workgroups as fit
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, list, ntiles + 1);
        hipLaunchKernelGGL(k_sparse_tiles, dim3((unsigned)ceil_div(ntiles, 256)), dim3(256), 0, s,
                           ts, ntiles, g.dense_min, list);
        CG_HIP(hipGetLastError());
        const int per_cu = std::max(1, std::min((int)(163840 / std::max<size_t>(lds, 1)), 2048 / threads));
        nblk = (unsigned)std::min<int64_t>(ntiles, (int64_t)256 * per_cu);
    }
    auto go = [&](const void *fn, auto kern) -> int {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(threads), lds, s, geo, fac, perm, ts, fw, out, g,
                           is2d, list, ntiles);
        CG_HIP(hipGetLastError());
        return CG_OK;
    };
    if (list) return go((const void *)k_degrid<T, KW, 4, true>, k_degrid<T, KW, 4, true>);
    return go((const void *)k_degrid<T, KW, 4, false>, k_degrid<T, KW, 4, false>);
}

template <typename T, typename TV>
int launch_degrid(const geo_t<T> *geo, const cx_t<T> *fac, const int32_t *perm,
                  const int32_t *ts, const cx_t<T> *fw, TV *out, const GridArgs &g, int kw,
                  int32_t *list, hipStream_t s)
{
    switch (kw) {
#define CG_CASE(K) \
    case K: return degrid_kw<T, K>(geo, fac, perm, ts, fw, out, g, list, s);
        CG_CASE(2) CG_CASE(3) CG_CASE(4) CG_CASE(5) CG_CASE(6) CG_CASE(7) CG_CASE(8) CG_CASE(9)
        CG_CASE(10) CG_CASE(11) CG_CASE(12) CG_CASE(13) CG_CASE(14) CG_CASE(15) CG_CASE(16)
#undef CG_CASE
    default: return CG_ERR_ARG;
    }
}

template int launch_grid<float>(const float4 *, const float2 *, const int32_t *, float2 *, const GridArgs &, int, hipStream_t);
template int launch_grid<double>(const double4 *, const double2 *, const int32_t *, double2 *, const GridArgs &, int, hipStream_t);
template int launch_degrid<float, float2>(const float4 *, const float2 *, const int32_t *, const int32_t *, const float2 *, float2 *, const GridArgs &, int, int32_t *, hipStream_t);
template int launch_degrid<double, double2>(const double4 *, const double2 *, const int32_t *, const int32_t *, const double2 *, double2 *, const GridArgs &, int, int32_t *, hipStream_t);

}  // namespace cg
