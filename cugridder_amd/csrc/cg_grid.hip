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

template <typename T>
__device__ __forceinline__ T es_val(T d, T beta, T c, T hw);
template <>
__device__ __forceinline__ float es_val<float>(float d, float beta, float c, float hw)
{
    return fabsf(d) < hw ? expf(beta * sqrtf(1.0f - c * d * d)) : 0.0f;
}
template <>
__device__ __forceinline__ double es_val<double>(double d, double beta, double c, double hw)
{
    return fabs(d) < hw ? exp(beta * sqrt(1.0 - c * d * d)) : 0.0;
}

// Kernel values of a point: the ES kernel exp(beta sqrt(1 - c d^2)) at the
// tap offsets d, or the learned kernel, whose exponent adds the network's
// polynomial sum_k q_k s^k in s = 2 c d^2 - 1 (c = (2/kw)^2).  The choice is
// uniform across a launch and the coefficients are launch arguments, so
// they stay in scalar registers.
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
        if (!learned) return es_val<T>(d, beta, ec, hw);
        if (!(fabs(d) < hw)) return (T)0;
        const T r = (T)1 - ec * d * d;
        const T s = (T)1 - (T)2 * r;
        T y = q[LEARN_MAXDEG];
#pragma unroll
        for (int k = LEARN_MAXDEG - 1; k >= 0; k--) y = y * s + q[k];
        return exp(beta * sqrt(r) + y);
    }
};

// gridding accumulates in FP64 whatever the grid precision: 16 bytes per cell
int grid_lds_bytes(const GridArgs &g, int) { return g.pc * g.tx * g.ty * 16; }
int degrid_lds_bytes(const GridArgs &g, int kw, int elem)
{
    return g.pc * (g.tx + kw - 1) * (g.ty + kw - 1) * elem;
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
    const int p0 = chunk * g.pc;
    const int np = min(g.pc, g.nf3 - p0);
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
        size_t gi = ((size_t)(p0 + lz) * g.nf2 + (size_t)(by * g.ty + y)) * g.nf1 + bx * g.tx + x;
        fw[gi] = mkc((T)acc[k], (T)acc[cells + k]);
    }
    __syncthreads();  // acc and the bin table are reused by the next item
    }
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
    const KernelEval<float> ker(g, KW);
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
            const uint32_t key = ((uint32_t)tile << g.cshift) + (uint32_t)c;
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
                        atomicAdd(are + idx, (double)r0[v]);
                        atomicAdd(aim + idx, (double)q0[v]);
                    }
                    if (u1 < KW && u2 + 4 < KW) {
                        const int idx = ((ly + u2 + 4) * RX + lx + u1) * NP + i;
                        atomicAdd(are + idx, (double)r1[v]);
                        atomicAdd(aim + idx, (double)q1[v]);
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
            lds[k] = fw[((size_t)(p0 + lz) * g.nf2 + gy) * g.nf1 + gx];
        }
    };

    if (single_stage) {
        stage(0, g.nf3);
        __syncthreads();
    }
    for (int base = beg; base < end; base += (int)blockDim.x * PPT) {
        T ar[PPT], ai[PPT];
#pragma unroll
        for (int k = 0; k < PPT; k++) ar[k] = ai[k] = (T)0;
        for (int c = 0; c < g.nchunk; c++) {
            const int p0 = c * g.pc;
            const int np = min(g.pc, g.nf3 - p0);
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
                    const cx_t<T> *pl = lds + (size_t)zz * hy * hx + (size_t)ly * hx + lx;
                    T sr = 0, si = 0;
#pragma unroll
                    for (int t2 = 0; t2 < KW; t2++) {
                        T rr = 0, ri = 0;
#pragma unroll
                        for (int t1 = 0; t1 < KW; t1++) {
                            const cx_t<T> f = pl[t2 * hx + t1];
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
            if (j >= 0 && j < g.m)  // a valid permutation always passes
                out[j] = mkc(ar[k] * f.x - ai[k] * f.y, ar[k] * f.y + ai[k] * f.x);
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
