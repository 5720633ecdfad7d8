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
    __device__ __forceinline__ float operator()(float off0, int t) const
    {
        const float d = off0 + (float)t;
        const float u = ec * d * d;
        float y = nb * u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f - u));
        if (learned) {
            const float s = 2.0f * u - 1.0f;
            float p = q[LEARN_MAXDEG];
#pragma unroll
            for (int k = LEARN_MAXDEG - 1; k >= 0; k--) p = p * s + q[k];
            y += 1.4426950408889634f * p;
        }
        const float v = __builtin_amdgcn_exp2f(y);
        return fabsf(d) < hw ? v : 0.0f;
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
// a product per start cell: the points j of one cell share the footprint
// origin, so their contributions to the cell's kw x kw x nf3 footprint are
//   C[tap][p] = sum_j A[tap][j] B[j][p],  A = v_j psi_x psi_y,  B = psi_z,
// a (64 taps) x (2 points) x (32 planes) FP32 MFMA step per point pair
// (v_mfma_f32_32x32x2f32, real and imaginary parts as two products; taps are
// t1 + 8 t2, planes the whole w axis).  C is then added once per cell into
// the bin's footprint region in LDS (FP64), and the region once per bin into
// the grid with FP32 global atomics, on top of what the sparse kernel wrote.

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kDenseThreads = 512;

__global__ __launch_bounds__(256) void k_dense_bins(const int32_t *__restrict__ tile_start,
                                                    int64_t ntiles, int dense_min,
                                                    int32_t *__restrict__ list)
{
    // list[0] counts the dense bins, list[1..] holds them (any order);
    // list[ntiles + 1] is the work counter of k_grid_dense
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= ntiles) return;
    if (tile_start[t + 1] - tile_start[t] >= dense_min) list[1 + atomicAdd(list, 1)] = (int32_t)t;
}

__global__ void k_zero_i32(int32_t *p, int64_t second)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p[0] = 0;
        p[second] = 0;
    }
}

// BATCH: the points of a start cell are loaded 64 at a time (one per lane)
// and handed to the lanes of each point pair by shuffles, so the MFMA loop
// waits on no global load; bins are taken from a work counter.
template <int KW, bool BATCH>
__global__ __launch_bounds__(kDenseThreads) void k_grid_dense(const float4 *__restrict__ geo,
                                                              const float2 *__restrict__ val,
                                                              const uint32_t *__restrict__ keys,
                                                              const int32_t *__restrict__ tile_start,
                                                              int32_t *__restrict__ list,
                                                              int64_t ntiles,
                                                              float2 *__restrict__ fw, GridArgs g)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int RX = g.tx + KW - 1, RY = g.ty + KW - 1, NP = g.nf3;
    const int rcells = RX * RY * NP;
    double *are = reinterpret_cast<double *>(smem_raw);
    double *aim = are + rcells;
    int *run = reinterpret_cast<int *>(aim + rcells);  // tx*ty + 1 run starts
    __shared__ int s_next, s_tile;
    const int ncell = g.tx * g.ty;
    const int nlist = list[0];
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5, i = lane & 31;
    const KernelEvalN ker(g, KW);  // values / e^beta: the sums are scaled by e^(3 beta)
    const double sc = g.e3;
    // lane roles: A[i][half] with tap i = t1 + 8 t2 (two 32-tap blocks),
    // B[half][j] with plane j = i.  Each lane evaluates one kernel value of
    // its point (i < 8: x tap i, 8 <= i < 16: y tap i - 8, 16 <= i < 24: z tap
    // i - 16) and the operands are gathered from those lanes.
    const int t1 = i & 7, t2a = i >> 3, t2b = 4 + (i >> 3);
    const int src_kx = half * 32 + t1, src_ky0 = half * 32 + 8 + t2a, src_ky1 = half * 32 + 8 + t2b;
    const int tap = i & 7;
    const int axis = i >> 3;  // 0 x, 1 y, 2 z, 3 none

    for (int li0 = blockIdx.x;; li0 += gridDim.x) {
        int li = li0;
        if (BATCH) {
            if (threadIdx.x == 0) s_tile = atomicAdd(&list[ntiles + 1], 1);
            __syncthreads();
            li = s_tile;
        }
        if (li >= nlist) break;
        const int tile = list[1 + li];
        const int bx = tile % g.ntx, by = tile / g.ntx;
        const int beg = tile_start[tile], end = tile_start[tile + 1];
        for (int k = threadIdx.x; k < 2 * rcells; k += blockDim.x) are[k] = 0.0;
        // first sorted position of every start cell of the bin
        for (int c = threadIdx.x; c <= ncell; c += blockDim.x) {
            const uint32_t key = ((uint32_t)g.kr.lower(tile, ntiles) << g.cshift) + (uint32_t)c;
            int lo = beg, hi = end;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (keys[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            run[c] = lo;
        }
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        for (;;) {
            int c = 0;
            if (lane == 0) c = atomicAdd(&s_next, 1);
            c = __shfl(c, 0);
            if (c >= ncell) break;
            const int a = run[c], b = run[c + 1];
            if (a >= b) continue;
            v16f r0 = {}, q0 = {}, r1 = {}, q1 = {};
            // one MFMA step for the point pair (G, v) of this half
            auto pair = [&](bool ok, const float4 &G, const float2 &v) {
                float a0r = 0.f, a0i = 0.f, a1r = 0.f, a1i = 0.f, bz = 0.f;
                // this lane's kernel value, then the operands from the lanes
                // holding them (every lane takes part in the shuffles)
                const float off = axis == 0 ? G.x : axis == 1 ? G.y : G.z;
                const float kv = (ok && axis < 3 && tap < KW) ? ker(off, tap) : 0.f;
                const float kx = __shfl(kv, src_kx);
                const float ky0 = __shfl(kv, src_ky0);
                const float ky1 = __shfl(kv, src_ky1);
                int d = i - (int)(__float_as_uint(G.w) >> 16);
                if (d < 0) d += NP;
                const float kz = __shfl(kv, half * 32 + 16 + (d & 7));
                if (ok) {
                    const float w0 = kx * ky0, w1 = kx * ky1;
                    a0r = v.x * w0;
                    a0i = v.y * w0;
                    a1r = v.x * w1;
                    a1i = v.y * w1;
                    if (i < NP && d < KW) bz = kz;
                }
                r0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0r, bz, r0, 0, 0, 0);
                q0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0i, bz, q0, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1r, bz, r1, 0, 0, 0);
                q1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1i, bz, q1, 0, 0, 0);
            };
            if (BATCH) {
                for (int kb = a; kb < b; kb += 64) {
                    const int n = min(64, b - kb);
                    float4 Gl = make_float4(0.f, 0.f, 0.f, 0.f);
                    float2 vl = make_float2(0.f, 0.f);
                    if (lane < n) {
                        Gl = geo[kb + lane];
                        vl = val[kb + lane];
                    }
                    for (int j0 = 0; j0 < n; j0 += 2) {
                        const int src = j0 + half;  // <= 63
                        float4 G;
                        float2 v;
                        G.x = __shfl(Gl.x, src);
                        G.y = __shfl(Gl.y, src);
                        G.z = __shfl(Gl.z, src);
                        G.w = __shfl(Gl.w, src);
                        v.x = __shfl(vl.x, src);
                        v.y = __shfl(vl.y, src);
                        pair(src < n, G, v);
                    }
                }
            } else {
                for (int k0 = a; k0 < b; k0 += 2) {
                    const int k = k0 + half;
                    const bool ok = k < b;
                    float4 G = make_float4(0.f, 0.f, 0.f, 0.f);
                    float2 v = make_float2(0.f, 0.f);
                    if (ok) {
                        G = geo[k];
                        v = val[k];
                    }
                    pair(ok, G, v);
                }
            }
            // D[tap][plane]: lane holds plane i, taps (v/4)*8 + half*4 + v%4 (+32)
            const int lx = c % g.tx, ly = c / g.tx;
            if (i < NP) {
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int tap = (v >> 2) * 8 + half * 4 + (v & 3);
                    const int u1 = tap & 7, u2 = tap >> 3;  // block 0: t2 < 4; block 1: t2 + 4
                    if (u1 < KW && u2 < KW) {
                        const int idx = ((ly + u2) * RX + lx + u1) * NP + i;
                        atomicAdd(are + idx, (double)r0[v] * sc);
                        atomicAdd(aim + idx, (double)q0[v] * sc);
                    }
                    if (u1 < KW && u2 + 4 < KW) {
                        const int idx = ((ly + u2 + 4) * RX + lx + u1) * NP + i;
                        atomicAdd(are + idx, (double)r1[v] * sc);
                        atomicAdd(aim + idx, (double)q1[v] * sc);
                    }
                }
            }
        }
        __syncthreads();
        // footprint region -> grid, x fastest so a wavefront covers row runs
        for (int k = threadIdx.x; k < rcells; k += blockDim.x) {
            const int rx = k % RX;
            const int rest = k / RX;
            const int ry = rest % RY, p = rest / RY;
            const int idx = (ry * RX + rx) * NP + p;
            int gx = bx * g.tx + rx;
            if (gx >= g.nf1) gx -= g.nf1;
            int gy = by * g.ty + ry;
            if (gy >= g.nf2) gy -= g.nf2;
            float2 *dst = fw + ((size_t)p * g.nf2 + gy) * g.nf1 + gx;
            atomicAdd(&dst->x, (float)are[idx]);
            atomicAdd(&dst->y, (float)aim[idx]);
        }
        __syncthreads();
    }
}

static size_t dense_lds_bytes(const GridArgs &g, int kw)
{
    const size_t rcells = (size_t)(g.tx + kw - 1) * (g.ty + kw - 1) * g.nf3;
    return 16 * rcells + 4 * (size_t)(g.tx * g.ty + 1);
}

bool grid_dense_supported(const GridArgs &g, int kw, int elem)
{
    return elem == 8 && kw >= 2 && kw <= 8 && g.nf3 >= 2 * kw && g.nf3 <= 32 && g.cshift > 0 &&
           g.tx + kw - 1 <= g.nf1 && g.ty + kw - 1 <= g.nf2 && dense_lds_bytes(g, kw) <= 163840 - 64;
}

template <int KW>
static int grid_dense_kw(const float4 *geo, const float2 *val, const uint32_t *keys,
                         const int32_t *ts, int32_t *list, float2 *fw, const GridArgs &g,
                         hipStream_t s)
{
    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, list, ntiles + 1);
    hipLaunchKernelGGL(k_dense_bins, dim3((unsigned)ceil_div(ntiles, 256)), dim3(256), 0, s, ts,
                       ntiles, g.dense_min, list);
    CG_HIP(hipGetLastError());
    const size_t lds = dense_lds_bytes(g, KW);
    const int per_cu = std::max<int>(1, (int)(163840 / lds));
    const unsigned nblk = (unsigned)std::min<int64_t>(ntiles, (int64_t)256 * per_cu);
    const char *e = std::getenv("CG_DENSE_BATCH");
    if (!e || std::atoi(e) != 0) {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute((const void *)k_grid_dense<KW, true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_grid_dense<KW, true>), dim3(nblk), dim3(kDenseThreads), lds, s, geo,
                           val, keys, ts, list, ntiles, fw, g);
    } else {
        if (lds > 65536)
            CG_HIP(hipFuncSetAttribute((const void *)k_grid_dense<KW, false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_grid_dense<KW, false>), dim3(nblk), dim3(kDenseThreads), lds, s, geo,
                           val, keys, ts, list, ntiles, fw, g);
    }
    CG_HIP(hipGetLastError());
    return CG_OK;
}

int launch_grid_dense(const geo_t<float> *geo, const cx_t<float> *val, const uint32_t *keys,
                      const int32_t *ts, int32_t *list, cx_t<float> *fw, const GridArgs &g, int kw,
                      hipStream_t s)
{
    if (!grid_dense_supported(g, kw, 8) || g.dense_min <= 0) return CG_ERR_ARG;
    switch (kw) {
#define CG_CASE(K) \
    case K: return grid_dense_kw<K>(geo, val, keys, ts, list, fw, g, s);
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

template <typename T, int KW, int PPT>
__global__ __launch_bounds__(1024) void k_degrid(const geo_t<T> *__restrict__ geo,
                                                const cx_t<T> *__restrict__ fac,
                                                const int32_t *__restrict__ perm,
                                                const int32_t *__restrict__ tile_start,
                                                const cx_t<T> *__restrict__ fw,
                                                cx_t<T> *__restrict__ out, GridArgs g, int is2d)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx_t<T> *lds = reinterpret_cast<cx_t<T> *>(smem_raw);
    const int tile = blockIdx.x;
    const int beg = tile_start[tile], end = tile_start[tile + 1];
    if (beg == end) return;
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

template <typename T, int KW>
static int degrid_kw(const geo_t<T> *geo, const cx_t<T> *fac, const int32_t *perm,
                     const int32_t *ts, const cx_t<T> *fw, cx_t<T> *out, const GridArgs &g,
                     hipStream_t s)
{
    int is2d = g.nf3 == 1 ? 1 : 0;
    size_t lds = (size_t)degrid_lds_bytes(g, KW, (int)sizeof(cx_t<T>));
    unsigned nblk = (unsigned)((int64_t)g.ntx * g.nty);
    if (lds > 65536)
        CG_HIP(hipFuncSetAttribute((const void *)k_degrid<T, KW, 4>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_degrid<T, KW, 4>), dim3(nblk), dim3(grid_threads(lds)), lds, s, geo, fac,
                       perm, ts, fw, out, g, is2d);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T, typename TV>
int launch_degrid(const geo_t<T> *geo, const cx_t<T> *fac, const int32_t *perm,
                  const int32_t *ts, const cx_t<T> *fw, TV *out, const GridArgs &g, int kw,
                  hipStream_t s)
{
    switch (kw) {
#define CG_CASE(K) \
    case K: return degrid_kw<T, K>(geo, fac, perm, ts, fw, out, g, s);
        CG_CASE(2) CG_CASE(3) CG_CASE(4) CG_CASE(5) CG_CASE(6) CG_CASE(7) CG_CASE(8) CG_CASE(9)
        CG_CASE(10) CG_CASE(11) CG_CASE(12) CG_CASE(13) CG_CASE(14) CG_CASE(15) CG_CASE(16)
#undef CG_CASE
    default: return CG_ERR_ARG;
    }
}

template int launch_grid<float>(const float4 *, const float2 *, const int32_t *, float2 *, const GridArgs &, int, hipStream_t);
template int launch_grid<double>(const double4 *, const double2 *, const int32_t *, double2 *, const GridArgs &, int, hipStream_t);
template int launch_degrid<float, float2>(const float4 *, const float2 *, const int32_t *, const int32_t *, const float2 *, float2 *, const GridArgs &, int, hipStream_t);
template int launch_degrid<double, double2>(const double4 *, const double2 *, const int32_t *, const int32_t *, const double2 *, double2 *, const GridArgs &, int, hipStream_t);

}  // namespace cg
