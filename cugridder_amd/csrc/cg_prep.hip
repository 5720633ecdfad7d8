// Coordinate preparation: w range, tile keys and sorted point records.
//
// All coordinate arithmetic is FP64 (the uvw scaling to grid coordinates): the reference's
// get_effective_coordinate (src/astro/precomp.cu:48-80), rescaling_real
// (src/utils.cu:532-556), pre_stage_2 (src/fourier/cuft.cu:82-103) and
// SHIFT_RESCALE (include/conv.h:24) are folded into one pass that emits, per
// visibility, the integer footprint start on the grid plus the FP32/FP64
// offset of the first tap.  The reference's per-cell histogram/scan/gather
// sort (src/utils.cu:105-132,457-486) is replaced by tile keys + a stable
// radix sort (cg_sort.hip).
#include "cg_internal.hpp"

// Coordinates are computed with separately rounded IEEE operations (no FMA
// contraction), so tile keys equal a host FP64 evaluation bit for bit.
#pragma clang fp contract(off)

namespace cg {

CoordArgs make_coord_args(const WPlan &p, const Tiling &t)
{
    CoordArgs a{};
    a.scale_uv = 2.0 * kPi * p.f_over_c;
    a.scale_w = 2.0 * kPi * p.f_over_c * p.sign;
    a.nchan = 1;
    a.chan_scale = nullptr;
    a.pix_x = p.pix_x;
    a.pix_y = p.pix_y;
    a.i_center = p.i_center;
    a.gamma = p.gamma;
    a.o_center = p.o_center;
    a.nf1 = p.nf1;
    a.nf2 = p.nf2;
    a.nf3 = p.nf3;
    a.kw = p.conv.kw;
    a.pirange = 0;
    a.has_w = p.wstack;
    set_coord_tiling(a, t);
    return a;
}

static int bits_for(int64_t n)
{
    int b = 1;
    while (((int64_t)1 << b) < n) b++;
    return b;
}

void set_coord_tiling(CoordArgs &a, const Tiling &t)
{
    a.tx = t.tx;
    a.ty = t.ty;
    a.ntx = t.ntx;
    a.nty = t.nty;
    int cs = 0;
    while ((1 << cs) < t.tx * t.ty) cs++;
    a.cshift = bits_for(t.ntiles()) + cs <= 32 ? cs : 0;
}

int sort_key_bits(const CoordArgs &a, int64_t ntiles)
{
    return bits_for(a.cnbx ? (int64_t)a.cnby * a.cnbx : ntiles) + a.cshift;
}

// include/conv.h:24 SHIFT_RESCALE (p = pirange)
__device__ __forceinline__ double shift_rescale(double x, int N, int pirange)
{
    const double twopi = 6.28318530717958648;
    double y = x;
    if (!pirange) {
        double t = x - floor(x / twopi) * twopi;
        y = t - (t >= kPi ? twopi : 0.0);
    }
    return (y * 0.159154943091895336 + 0.5) * N;
}

struct Foot {
    int x0, y0, z0;      // wrapped footprint starts
    double d1, d2, d3;   // offset of the first tap: start - g
};

__device__ __forceinline__ int wrapi(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

__device__ __forceinline__ void footprint_1d(double g, int n, int kw, int &i0, double &d)
{
    double s = ceil(g - 0.5 * kw);
    d = s - g;
    i0 = wrapi((int)s, n);
}

// point j -> row of uvw and the coordinate scales of its channel
struct PointScale {
    int64_t row;
    double suv, sw;
};
__device__ __forceinline__ PointScale point_scale(int64_t j, const CoordArgs &a)
{
    if (a.nchan <= 1) return PointScale{j, a.scale_uv, a.scale_w};
    const int64_t r = j / a.nchan;
    const double s = a.chan_scale[j - r * a.nchan];
    return PointScale{r, s, a.scale_w < 0 ? -s : s};
}

// grid axis 1 (fast) <- v * pix_y, axis 2 <- u * pix_x (cugridder.cu:191)
__device__ __forceinline__ Foot make_foot(double u, double v, double w, const CoordArgs &a,
                                          double suv, double sw)
{
    Foot f;
    double xs = v * suv * a.pix_y;
    double ys = u * suv * a.pix_x;
    footprint_1d(shift_rescale(xs, a.nf1, 0), a.nf1, a.kw, f.x0, f.d1);
    footprint_1d(shift_rescale(ys, a.nf2, 0), a.nf2, a.kw, f.y0, f.d2);
    if (a.has_w) {
        double wp = (w * sw - a.i_center) / a.gamma;
        footprint_1d(shift_rescale(wp, a.nf3, 0), a.nf3, a.kw, f.z0, f.d3);
    } else {
        f.z0 = 0;
        f.d3 = 0.0;
    }
    return f;
}

__device__ __forceinline__ uint32_t tile_key(const Foot &f, const CoordArgs &a)
{
    const int by = f.y0 / a.ty, bx = f.x0 / a.tx;
    const uint32_t tile = a.cnbx ? (uint32_t)((by - a.cby0) * a.cnbx + (bx - a.cbx0))
                                 : (uint32_t)(by * a.ntx + bx);
    if (!a.cshift) return tile;
    return (tile << a.cshift) | (uint32_t)((f.y0 % a.ty) * a.tx + f.x0 % a.tx);
}

// ------------------------------------------------------------ w min/max

// min/max of w * scale_w, of u and of v over the rows: part[1024 q + block]
// for q = wmin, wmax, umin, umax, vmin, vmax; finals at part[6144 + q]
template <typename TU>
__global__ __launch_bounds__(256) void k_wminmax(const TU *uvw, int64_t nrow, double scale_w,
                                                 double *part)
{
    __shared__ double sv[6][256];
    double r[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x; j < nrow; j += (int64_t)gridDim.x * 256) {
        const double w = (double)uvw[3 * j + 2] * scale_w;
        const double u = (double)uvw[3 * j], v = (double)uvw[3 * j + 1];
        r[0] = fmin(r[0], w);
        r[1] = fmax(r[1], w);
        r[2] = fmin(r[2], u);
        r[3] = fmax(r[3], u);
        r[4] = fmin(r[4], v);
        r[5] = fmax(r[5], v);
    }
#pragma unroll
    for (int q = 0; q < 6; q++) sv[q][threadIdx.x] = r[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int q = 0; q < 6; q++)
                sv[q][threadIdx.x] = (q & 1) ? fmax(sv[q][threadIdx.x], sv[q][threadIdx.x + s])
                                             : fmin(sv[q][threadIdx.x], sv[q][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[1024 * threadIdx.x + blockIdx.x] = sv[threadIdx.x][0];
}

__global__ __launch_bounds__(1024) void k_minmax_final(double *part, int nb)
{
    __shared__ double sv[6][1024];
    int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 6; q++) sv[q][t] = t < nb ? part[1024 * q + t] : ((q & 1) ? -INFINITY : INFINITY);
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int q = 0; q < 6; q++)
                sv[q][t] = (q & 1) ? fmax(sv[q][t], sv[q][t + s]) : fmin(sv[q][t], sv[q][t + s]);
        }
        __syncthreads();
    }
    if (t < 6) part[6144 + t] = sv[t][0];
}

template <typename TU>
int launch_wminmax(const TU *uvw, int64_t nrow, double scale_w, double *d_part, double &wmin,
                   double &wmax, hipStream_t s, double *uv)
{
    if (nrow == 0) {
        wmin = wmax = 0.0;
        if (uv) uv[0] = uv[1] = uv[2] = uv[3] = 0.0;
        return CG_OK;
    }
    int nb = (int)std::min<int64_t>(1024, ceil_div(nrow, 256));
    hipLaunchKernelGGL(k_wminmax<TU>, dim3(nb), dim3(256), 0, s, uvw, nrow, scale_w, d_part);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(1024), 0, s, d_part, nb);
    CG_HIP(hipGetLastError());
    double h[6];
    CG_HIP(hipMemcpyAsync(h, d_part + 6144, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    wmin = h[0];
    wmax = h[1];
    if (uv) {
        for (int q = 0; q < 4; q++) uv[q] = h[2 + q];
    }
    return CG_OK;
}

// ------------------------------------------------------------ keys

template <typename TU>
__global__ __launch_bounds__(256) void k_keys(const TU *uvw, int64_t nrow, CoordArgs a,
                                              uint32_t *keys, int32_t *vals)
{
    // j runs over the points: rows x channels
    int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (j >= nrow) return;
    CoordArgs b = a;
    b.has_w = 0;  // the key only depends on the (u, v) footprint
    const PointScale ps = point_scale(j, a);
    Foot f = make_foot((double)uvw[3 * ps.row], (double)uvw[3 * ps.row + 1], 0.0, b, ps.suv, ps.sw);
    keys[j] = tile_key(f, a);
    if (vals) vals[j] = (int32_t)j;
}

template <typename TU>
int launch_keys(const TU *uvw, int64_t nrow, const CoordArgs &a, uint32_t *keys, int32_t *vals,
                hipStream_t s)
{
    if (nrow == 0) return CG_OK;
    hipLaunchKernelGGL(k_keys<TU>, dim3((unsigned)ceil_div(nrow, 256)), dim3(256), 0, s, uvw, nrow,
                       a, keys, vals);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// 2-D NUFFT: x -> axis 1 (fast), y -> axis 2, coordinates in [-pi,pi)
__device__ __forceinline__ Foot make_foot_xy(double x, double y, const CoordArgs &a)
{
    Foot f;
    footprint_1d(shift_rescale(x, a.nf1, a.pirange), a.nf1, a.kw, f.x0, f.d1);
    footprint_1d(shift_rescale(y, a.nf2, a.pirange), a.nf2, a.kw, f.y0, f.d2);
    f.z0 = 0;
    f.d3 = 0.0;
    return f;
}

template <typename TU>
__global__ __launch_bounds__(256) void k_keys_xy(const TU *x, const TU *y, int64_t m, CoordArgs a,
                                                 uint32_t *keys, int32_t *vals)
{
    int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (j >= m) return;
    Foot f = make_foot_xy((double)x[j], (double)y[j], a);
    keys[j] = tile_key(f, a);
    if (vals) vals[j] = (int32_t)j;
}

template <typename TU>
int launch_keys_xy(const TU *x, const TU *y, int64_t m, const CoordArgs &a, uint32_t *keys,
                   int32_t *vals, hipStream_t s)
{
    if (m == 0) return CG_OK;
    hipLaunchKernelGGL(k_keys_xy<TU>, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, x, y, m, a,
                       keys, vals);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// ------------------------------------------------------------ tile bounds

// start[t] = first sorted position whose tile comes at or after grid tile t
// (with a compact key map: at or after the first rectangle tile >= t)
__global__ __launch_bounds__(256) void k_tile_bounds(const uint32_t *keys, int64_t m,
                                                     int64_t ntiles, int shift, KeyRect kr,
                                                     int32_t *start)
{
    int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t > ntiles) return;
    const int64_t c = kr.lower(t, ntiles);
    int64_t lo = 0, hi = m;  // first index with keys[i] >> shift >= c
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)(keys[mid] >> shift) < c) lo = mid + 1;
        else hi = mid;
    }
    start[t] = (int32_t)lo;
}

int launch_tile_bounds(const uint32_t *sorted_keys, int64_t m, int64_t ntiles, int shift,
                       const KeyRect &kr, int32_t *tile_start, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)ceil_div(ntiles + 1, 256)), dim3(256), 0, s,
                       sorted_keys, m, ntiles, shift, kr, tile_start);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// ------------------------------------------------------------ row mask

// mask[y] = 1 for every grid row y (along nf2) that the footprint of a point
// of a non-empty tile can reach: rows [y0, y0 + ty + kw - 1) of the tile,
// periodic.  All other rows of the grid are zero (ms2dirty) or never read
// (dirty2ms), so the FFT passes may skip them.
__global__ __launch_bounds__(256) void k_fill_u8(uint8_t *p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// mask[y] (y < nf2) marks the grid rows and mask[nf2 + x] (x < nf1) the
// grid columns that the footprints of the points of non-empty tiles reach
__global__ __launch_bounds__(256) void k_row_mask(const int32_t *__restrict__ tile_start,
                                                  int64_t ntiles, int ntx, int tx, int ty, int kw,
                                                  int nf1, int nf2, uint8_t *__restrict__ mask)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= ntiles || tile_start[t + 1] == tile_start[t]) return;
    const int y0 = (int)(t / ntx) * ty;
    const int spany = min(ty + kw - 1, nf2);
    for (int r = 0; r < spany; r++) {
        int y = y0 + r;
        if (y >= nf2) y -= nf2;
        mask[y] = 1;
    }
    const int x0 = (int)(t % ntx) * tx;
    const int spanx = min(tx + kw - 1, nf1);
    for (int r = 0; r < spanx; r++) {
        int x = x0 + r;
        if (x >= nf1) x -= nf1;
        mask[nf2 + x] = 1;
    }
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s)
{
    if (n <= 0) return CG_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

int launch_row_mask(const int32_t *tile_start, const Tiling &t, int kw, int nf1, int nf2,
                    uint8_t *mask, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)ceil_div(nf1 + nf2, 256)), dim3(256), 0, s, mask,
                       nf1 + nf2);
    const int64_t ntiles = t.ntiles();
    hipLaunchKernelGGL(k_row_mask, dim3((unsigned)ceil_div(ntiles, 256)), dim3(256), 0, s, tile_start,
                       ntiles, t.ntx, t.tx, t.ty, kw, nf1, nf2, mask);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// ------------------------------------------------------------ records

template <typename T> __device__ __forceinline__ geo_t<T> make_geo(double d1, double d2, double d3, uint32_t pk);
template <> __device__ __forceinline__ float4 make_geo<float>(double d1, double d2, double d3, uint32_t pk)
{
    return make_float4((float)d1, (float)d2, (float)d3, __uint_as_float(pk));
}
template <> __device__ __forceinline__ double4 make_geo<double>(double d1, double d2, double d3, uint32_t pk)
{
    return make_double4(d1, d2, d3, __longlong_as_double((long long)pk));
}

__device__ __forceinline__ double2 ld_c(const float2 *p, int64_t j)
{
    float2 v = p[j];
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_c(const double2 *p, int64_t j) { return p[j]; }

template <typename T, typename TU, typename TV, typename TW>
__global__ __launch_bounds__(256) void k_records(const TU *uvw, const TV *vis, const TW *wgt,
                                                 const int32_t *perm, int64_t nrow, CoordArgs a,
                                                 int adjoint, geo_t<T> *geo, cx_t<T> *val)
{
    int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= nrow) return;
    // point j = row * nchan + channel; vis and wgt are [rows][channels]
    int64_t j = perm[i];
    if (j < 0 || j >= nrow) j = 0;  // a valid permutation never takes this
    const PointScale ps = point_scale(j, a);
    const int64_t r = ps.row;
    double u = (double)uvw[3 * r], v = (double)uvw[3 * r + 1], w = (double)uvw[3 * r + 2];
    Foot f = make_foot(u, v, w, a, ps.suv, ps.sw);
    uint32_t pk = (uint32_t)(f.x0 % a.tx) | ((uint32_t)(f.y0 % a.ty) << 8) | ((uint32_t)f.z0 << 16);
    store_nt(geo + i, make_geo<T>(f.d1, f.d2, f.d3, pk));
    double wt = wgt ? (double)wgt[j] : 1.0;
    double ws = w * ps.sw;
    double2 c;
    if (!adjoint) {
        // ms2dirty: vis * wgt (ra_exec.cu:103-111) * e^{i o_c w} (cuft.cu:46-80)
        double2 x = ld_c(vis, j);
        double s, co;
        sincos(a.o_center * ws, &s, &co);
        c = make_double2(wt * (x.x * co - x.y * s), wt * (x.x * s + x.y * co));
    } else {
        // dirty2ms: factor wgt * e^{-i (w' gamma + i_c) o_c}  (ra_exec.cu:55-93)
        double wp = (ws - a.i_center) / a.gamma;
        double s, co;
        sincos(-(wp * a.gamma + a.i_center) * a.o_center, &s, &co);
        c = make_double2(wt * co, wt * s);
    }
    store_nt(val + i, mkc((T)c.x, (T)c.y));
}

// The same records for complex64 in two passes: every point's record is
// computed in the caller's order (coalesced reads of uvw, vis, wgt) into a
// 32-byte slot {geo, val, pad}, then one aligned 32-byte slot is gathered
// per sorted position (instead of three scattered reads).  Bit-identical to
// k_records.
template <typename TU, typename TV, typename TW>
__device__ __forceinline__ void prerecord(const TU *uvw, const TV *vis, const TW *wgt, int64_t j,
                                          const CoordArgs &a, int adjoint, float4 *__restrict__ rec)
{
    const PointScale ps = point_scale(j, a);
    const int64_t r = ps.row;
    double u = (double)uvw[3 * r], v = (double)uvw[3 * r + 1], w = (double)uvw[3 * r + 2];
    Foot f = make_foot(u, v, w, a, ps.suv, ps.sw);
    uint32_t pk = (uint32_t)(f.x0 % a.tx) | ((uint32_t)(f.y0 % a.ty) << 8) | ((uint32_t)f.z0 << 16);
    double wt = wgt ? (double)wgt[j] : 1.0;
    double ws = w * ps.sw;
    double2 c;
    if (!adjoint) {
        double2 x = ld_c(vis, j);
        double sn, co;
        sincos(a.o_center * ws, &sn, &co);
        c = make_double2(wt * (x.x * co - x.y * sn), wt * (x.x * sn + x.y * co));
    } else {
        double wp = (ws - a.i_center) / a.gamma;
        double sn, co;
        sincos(-(wp * a.gamma + a.i_center) * a.o_center, &sn, &co);
        c = make_double2(wt * co, wt * sn);
    }
    rec[2 * j] = make_geo<float>(f.d1, f.d2, f.d3, pk);
    rec[2 * j + 1] = make_float4((float)c.x, (float)c.y, 0.f, 0.f);
}

template <typename TU, typename TV, typename TW>
__global__ __launch_bounds__(256) void k_prerecords(const TU *uvw, const TV *vis, const TW *wgt,
                                                    int64_t nrow, CoordArgs a, int adjoint,
                                                    float4 *__restrict__ rec)
{
    const int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (j >= nrow) return;
    prerecord(uvw, vis, wgt, j, a, adjoint, rec);
}

// the key pass (k_keys) and k_prerecords in one pass over uvw
template <typename TU, typename TV, typename TW>
__global__ __launch_bounds__(256) void k_keys_prerecords(const TU *uvw, const TV *vis,
                                                         const TW *wgt, int64_t nrow, CoordArgs a,
                                                         int adjoint, uint32_t *__restrict__ keys,
                                                         int32_t *__restrict__ vals,
                                                         float4 *__restrict__ rec)
{
    const int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (j >= nrow) return;
    CoordArgs b = a;
    b.has_w = 0;
    const PointScale ps = point_scale(j, a);
    const Foot f = make_foot((double)uvw[3 * ps.row], (double)uvw[3 * ps.row + 1], 0.0, b, ps.suv, ps.sw);
    keys[j] = tile_key(f, a);
    if (vals) vals[j] = (int32_t)j;
    prerecord(uvw, vis, wgt, j, a, adjoint, rec);
}

template <typename TU, typename TV, typename TW>
int launch_keys_prerecords(const TU *uvw, const TV *vis, const TW *wgt, int64_t nrow,
                           const CoordArgs &a, int adjoint, uint32_t *keys, int32_t *vals,
                           float4 *rec, hipStream_t s)
{
    if (nrow == 0) return CG_OK;
    hipLaunchKernelGGL((k_keys_prerecords<TU, TV, TW>), dim3((unsigned)ceil_div(nrow, 256)), dim3(256),
                       0, s, uvw, vis, wgt, nrow, a, adjoint, keys, vals, rec);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

int launch_gather_records(const float4 *rec, const int32_t *perm, int64_t nrow, float4 *geo,
                          float2 *val, hipStream_t s);

__global__ __launch_bounds__(256) void k_gather_records(const float4 *__restrict__ rec,
                                                        const int32_t *__restrict__ perm,
                                                        int64_t nrow, float4 *__restrict__ geo,
                                                        float2 *__restrict__ val)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= nrow) return;
    int64_t j = perm[i];
    if (j < 0 || j >= nrow) j = 0;  // a valid permutation never takes this
    store_nt(geo + i, rec[2 * j]);
    const float4 t = rec[2 * j + 1];
    store_nt(val + i, make_float2(t.x, t.y));
}

int launch_gather_records(const float4 *rec, const int32_t *perm, int64_t nrow, float4 *geo,
                          float2 *val, hipStream_t s)
{
    if (nrow == 0) return CG_OK;
    hipLaunchKernelGGL(k_gather_records, dim3((unsigned)ceil_div(nrow, 256)), dim3(256), 0, s, rec,
                       perm, nrow, geo, val);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T, typename TU, typename TV, typename TW>
int launch_records(const TU *uvw, const TV *vis, const TW *wgt, const int32_t *perm, int64_t nrow,
                   const CoordArgs &a, int adjoint, geo_t<T> *geo, cx_t<T> *val, float4 *rec,
                   hipStream_t s)
{
    if (nrow == 0) return CG_OK;
    if constexpr (sizeof(T) == 4) {
        if (rec) {
            hipLaunchKernelGGL((k_prerecords<TU, TV, TW>), dim3((unsigned)ceil_div(nrow, 256)),
                               dim3(256), 0, s, uvw, vis, wgt, nrow, a, adjoint, rec);
            hipLaunchKernelGGL(k_gather_records, dim3((unsigned)ceil_div(nrow, 256)), dim3(256), 0, s,
                               (const float4 *)rec, perm, nrow, geo, val);
            CG_HIP(hipGetLastError());
            return CG_OK;
        }
    }
    hipLaunchKernelGGL((k_records<T, TU, TV, TW>), dim3((unsigned)ceil_div(nrow, 256)), dim3(256), 0,
                       s, uvw, vis, wgt, perm, nrow, a, adjoint, geo, val);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T, typename TU, typename TV>
__global__ __launch_bounds__(256) void k_records_xy(const TU *x, const TU *y, const TV *c,
                                                    const int32_t *perm, int64_t m, CoordArgs a,
                                                    geo_t<T> *geo, cx_t<T> *val)
{
    int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= m) return;
    int64_t j = perm[i];
    if (j < 0 || j >= m) j = 0;  // a valid permutation never takes this
    Foot f = make_foot_xy((double)x[j], (double)y[j], a);
    uint32_t pk = (uint32_t)(f.x0 % a.tx) | ((uint32_t)(f.y0 % a.ty) << 8);
    store_nt(geo + i, make_geo<T>(f.d1, f.d2, 0.0, pk));
    if (c) {
        double2 v = ld_c(c, j);
        store_nt(val + i, mkc((T)v.x, (T)v.y));
    } else {
        store_nt(val + i, mkc((T)1, (T)0));
    }
}

// 2-D complex64 records in two coalescing passes: the key pass also writes
// every point's record in the caller's order (a 16-byte geo slot, plus a
// 16-byte strength slot when c is given), the record pass gathers the slots
// by the sorted permutation.  Bit-identical to k_keys_xy + k_records_xy.
__global__ __launch_bounds__(256) void k_keys_prerecords_xy(const float *__restrict__ x,
                                                            const float *__restrict__ y,
                                                            const float2 *__restrict__ c,
                                                            int64_t m, CoordArgs a,
                                                            uint32_t *__restrict__ keys,
                                                            int32_t *__restrict__ vals,
                                                            float4 *__restrict__ rec)
{
    const int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (j >= m) return;
    const Foot f = make_foot_xy((double)x[j], (double)y[j], a);
    keys[j] = tile_key(f, a);
    if (vals) vals[j] = (int32_t)j;
    const uint32_t pk = (uint32_t)(f.x0 % a.tx) | ((uint32_t)(f.y0 % a.ty) << 8);
    const float4 G = make_geo<float>(f.d1, f.d2, 0.0, pk);
    if (c) {
        const float2 v = c[j];
        rec[2 * j] = G;
        rec[2 * j + 1] = make_float4(v.x, v.y, 0.f, 0.f);
    } else {
        rec[j] = G;
    }
}

__global__ __launch_bounds__(256) void k_gather_records_xy(const float4 *__restrict__ rec,
                                                           const int32_t *__restrict__ perm,
                                                           int64_t m, int has_c,
                                                           float4 *__restrict__ geo,
                                                           float2 *__restrict__ val)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= m) return;
    int64_t j = perm[i];
    if (j < 0 || j >= m) j = 0;  // a valid permutation never takes this
    if (has_c) {
        store_nt(geo + i, rec[2 * j]);
        const float4 t = rec[2 * j + 1];
        store_nt(val + i, make_float2(t.x, t.y));
    } else {
        store_nt(geo + i, rec[j]);
        store_nt(val + i, make_float2(1.f, 0.f));
    }
}

int launch_keys_prerecords_xy(const float *x, const float *y, const float2 *c, int64_t m,
                              const CoordArgs &a, uint32_t *keys, int32_t *vals, float4 *rec,
                              hipStream_t s)
{
    if (m == 0) return CG_OK;
    hipLaunchKernelGGL(k_keys_prerecords_xy, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, x, y,
                       c, m, a, keys, vals, rec);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

int launch_gather_records_xy(const float4 *rec, const int32_t *perm, int64_t m, int has_c,
                             float4 *geo, float2 *val, hipStream_t s)
{
    if (m == 0) return CG_OK;
    hipLaunchKernelGGL(k_gather_records_xy, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, rec,
                       perm, m, has_c, geo, val);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

template <typename T, typename TU, typename TV>
int launch_records_xy(const TU *x, const TU *y, const TV *c, const int32_t *perm, int64_t m,
                      const CoordArgs &a, geo_t<T> *geo, cx_t<T> *val, hipStream_t s)
{
    if (m == 0) return CG_OK;
    hipLaunchKernelGGL((k_records_xy<T, TU, TV>), dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s,
                       x, y, c, perm, m, a, geo, val);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// explicit instantiations
template int launch_wminmax<float>(const float *, int64_t, double, double *, double &, double &, hipStream_t, double *);
template int launch_wminmax<double>(const double *, int64_t, double, double *, double &, double &, hipStream_t, double *);
template int launch_keys<float>(const float *, int64_t, const CoordArgs &, uint32_t *, int32_t *, hipStream_t);
template int launch_keys<double>(const double *, int64_t, const CoordArgs &, uint32_t *, int32_t *, hipStream_t);
template int launch_keys_xy<float>(const float *, const float *, int64_t, const CoordArgs &, uint32_t *, int32_t *, hipStream_t);
template int launch_keys_xy<double>(const double *, const double *, int64_t, const CoordArgs &, uint32_t *, int32_t *, hipStream_t);
template int launch_keys_prerecords<float, float2, float>(const float *, const float2 *, const float *, int64_t, const CoordArgs &, int, uint32_t *, int32_t *, float4 *, hipStream_t);
template int launch_keys_prerecords<double, float2, float>(const double *, const float2 *, const float *, int64_t, const CoordArgs &, int, uint32_t *, int32_t *, float4 *, hipStream_t);
template int launch_records<float, float, float2, float>(const float *, const float2 *, const float *, const int32_t *, int64_t, const CoordArgs &, int, float4 *, float2 *, float4 *, hipStream_t);
template int launch_records<float, double, float2, float>(const double *, const float2 *, const float *, const int32_t *, int64_t, const CoordArgs &, int, float4 *, float2 *, float4 *, hipStream_t);
template int launch_records<double, double, double2, double>(const double *, const double2 *, const double *, const int32_t *, int64_t, const CoordArgs &, int, double4 *, double2 *, float4 *, hipStream_t);
template int launch_records_xy<float, float, float2>(const float *, const float *, const float2 *, const int32_t *, int64_t, const CoordArgs &, float4 *, float2 *, hipStream_t);
template int launch_records_xy<double, double, double2>(const double *, const double *, const double2 *, const int32_t *, int64_t, const CoordArgs &, double4 *, double2 *, hipStream_t);

}  // namespace cg
