// Deterministic stable LSD radix sort of (tile key, row index) pairs and an
// exclusive int scan.
//
// Replaces the reference's per-cell counting sort (histogram_3d_sparse with
// an atomicAdd in-bin rank, src/utils.cu:105-132; thrust exclusive_scan,
// src/common_utils.cu:156-170; mapping_based_gather_3d, src/utils.cu:457-486).
// That sort's in-bin order depends on atomic timing; here every pass ranks
// items inside a 64-wide wavefront with ballots (a match on each digit bit)
// and across wavefronts with an LDS prefix, so the permutation equals a host
// std::stable_sort by key, bit for bit.
#include "cg_internal.hpp"

namespace cg {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 4096 items / block

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// ------------------------------------------------------------ scan

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// exclusive scan of per-thread values across the block; returns this
// thread's exclusive prefix and writes the block total to *total
__device__ __forceinline__ int block_excl_scan(int v, int *s_wave, int *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int incl = wave_incl_scan(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        int w = lane < nw ? s_wave[lane] : 0;
        int wi = wave_incl_scan(w);
        if (lane < nw) s_wave[lane] = wi - w;
        if (lane == nw - 1) s_wave[64] = wi;
    }
    __syncthreads();
    int r = s_wave[wave] + incl - v;
    *total = s_wave[64];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t *d, int64_t n,
                                                              int32_t *bsum)
{
    __shared__ int s_wave[65];
    int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanItems;
    int v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) v += d[base + k];
    int tot;
    block_excl_scan(v, s_wave, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// exclusive scan of one tile per block, adding offs[block] (offs may be null)
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(int32_t *d, int64_t n,
                                                            const int32_t *offs)
{
    __shared__ int s_wave[65];
    int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanItems;
    int x[kScanItems];
    int v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        x[k] = base + k < n ? d[base + k] : 0;
        v += x[k];
    }
    int tot;
    int pre = block_excl_scan(v, s_wave, &tot);
    if (offs) pre += offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) d[base + k] = pre;
        pre += x[k];
    }
}

int64_t scan_tmp_size(int64_t n)
{
    int64_t s = 0;
    while (n > kScanTile) {
        n = ceil_div(n, kScanTile);
        s += n;
    }
    return s + 1;
}

int launch_exclusive_scan(int32_t *data, int64_t n, int32_t *tmp, hipStream_t s)
{
    if (n <= 0) return CG_OK;
    if (n <= kScanTile) {
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanThreads), 0, s, data, n,
                           (const int32_t *)nullptr);
        CG_HIP(hipGetLastError());
        return CG_OK;
    }
    int64_t nb = ceil_div(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, data, n, tmp);
    CG_HIP(hipGetLastError());
    int r = launch_exclusive_scan(tmp, nb, tmp + nb, s);
    if (r) return r;
    hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)nb), dim3(kScanThreads), 0, s, data, n,
                       (const int32_t *)tmp);
    CG_HIP(hipGetLastError());
    return CG_OK;
}

// ------------------------------------------------------------ radix sort

int64_t radix_hist_size(int64_t m) { return (int64_t)kRadix * std::max<int64_t>(1, ceil_div(m, kSortTile)); }

// The tile a workgroup of the radix passes takes: workgroups are dealt to
// the 8 XCDs round-robin, so workgroup b runs on XCD b % 8; XCD x takes the
// contiguous tiles [x q + min(x, r), ...) (nblocks = 8 q + r), so that the
// histogram entries of adjacent tiles, and the adjacent runs of a digit that
// adjacent tiles write, meet in one L2.  A bijection of [0, nblocks).
__device__ __forceinline__ int radix_tile(int nblocks)
{
    const int b = blockIdx.x, x = b & 7, q = nblocks >> 3, r = nblocks & 7;
    return x * q + min(x, r) + (b >> 3);
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t *keys, int64_t m,
                                                             int shift, int32_t *hist, int nblocks)
{
    __shared__ int cnt[kRadix];
    cnt[threadIdx.x] = 0;
    const int tile = radix_tile(nblocks);
    int64_t base = tile * (int64_t)kSortTile;
    // all loads first: one memory latency per block, not one per round
    uint32_t kk[kSortRounds];
#pragma unroll
    for (int k = 0; k < kSortRounds; k++) {
        const int64_t i = base + k * kSortThreads + threadIdx.x;
        kk[k] = i < m ? keys[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortRounds; k++) {
        const int64_t i = base + k * kSortThreads + threadIdx.x;
        if (i < m) atomicAdd(&cnt[(kk[k] >> shift) & (kRadix - 1)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + tile] = cnt[threadIdx.x];
}

// One pass: the block ranks its 4096 items by digit in LDS and writes them
// out so that the items of one digit leave as one contiguous run at the
// digit's global offset for this block (coalesced stores instead of one
// scattered store per item).  Wavefront w owns the contiguous quarter
// [1024 w, 1024 w + 1024) of the tile, 64 items per round: it ranks each
// item among its own items of the same digit with ballots (a match on each
// digit bit) and a running count per digit in its LDS row, with no
// workgroup barrier between rounds.  One prefix over the four rows per
// digit then places the quarters in order, so the order of a digit's items
// is the input order (stable).  offs is the scanned histogram (digit-major,
// block-minor).
__device__ __forceinline__ void wave_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t *kin, const int32_t *vin, uint32_t *kout, int32_t *vout, int64_t m, int shift,
    const int32_t *offs, int nblocks)
{
    constexpr int NW = kSortThreads / 64;
    constexpr int QW = kSortTile / NW;  // items of a wavefront's quarter
    constexpr int R = QW / 64;          // rounds
    static_assert(kSortThreads == kRadix, "one thread per digit");
    __shared__ int s_gbase[kRadix];
    __shared__ int s_lstart[kRadix];
    __shared__ int s_wcnt[NW][kRadix];
    __shared__ int s_wave[65];
    __shared__ uint32_t s_key[kSortTile];
    __shared__ int32_t s_val[kSortTile];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int tile = radix_tile(nblocks);
    s_gbase[tid] = offs[(int64_t)tid * nblocks + tile];
#pragma unroll
    for (int w = 0; w < NW; w++) s_wcnt[w][tid] = 0;
    const uint64_t lt = __lanemask_lt();
    const int64_t tbase = tile * (int64_t)kSortTile;
    const int64_t qbase = tbase + (int64_t)wave * QW;
    // all loads first: one memory latency per block, not one per round
    uint32_t kk[R];
    int32_t vv[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t i = qbase + 64 * r + lane;
        kk[r] = i < m ? kin[i] : 0u;
        vv[r] = i < m ? (vin ? vin[i] : (int32_t)i) : 0;  // no vin: the identity
    }
    __syncthreads();  // counters zeroed
    int *cnt = s_wcnt[wave];
    int rk[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool valid = qbase + 64 * r + lane < m;
        const int d = (int)((kk[r] >> shift) & (kRadix - 1));
        uint64_t mask = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            const int bit = (d >> b) & 1;
            const uint64_t bb = __ballot(bit);
            mask &= bit ? bb : ~bb;
        }
        const int before = __popcll(mask & lt);
        const int prev = cnt[d];
        rk[r] = prev + before;
        wave_fence();  // every lane has read its digit's count
        if (valid && before == 0) cnt[d] = prev + __popcll(mask);
        wave_fence();
    }
    __syncthreads();
    {
        // quarters before this one, per digit, and the block's digit starts
        int acc = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const int t = s_wcnt[w][tid];
            s_wcnt[w][tid] = acc;
            acc += t;
        }
        int tot;
        s_lstart[tid] = block_excl_scan(acc, s_wave, &tot);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (qbase + 64 * r + lane < m) {
            const int d = (int)((kk[r] >> shift) & (kRadix - 1));
            const int lp = s_lstart[d] + cnt[d] + rk[r];
            if (lp >= 0 && lp < kSortTile) {  // always true for a consistent histogram
                s_key[lp] = kk[r];
                s_val[lp] = vv[r];
            }
        }
    }
    __syncthreads();
    const int nvalid = (int)std::min<int64_t>(kSortTile, m - tbase);
    for (int p = tid; p < nvalid; p += kSortThreads) {
        const uint32_t key = s_key[p];
        const int d = (int)((key >> shift) & (kRadix - 1));
        const int64_t pos = (int64_t)s_gbase[d] + (p - s_lstart[d]);
        if (pos >= 0 && pos < m) {  // always true for a consistent histogram
            kout[pos] = key;
            vout[pos] = s_val[p];
        }
    }
}

int radix_passes(int key_bits) { return (key_bits + kRadixBits - 1) / kRadixBits; }

__global__ __launch_bounds__(256) void k_iota(int32_t *v, int64_t m)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i < m) v[i] = (int32_t)i;
}

int launch_radix_sort(uint32_t *keys, int32_t *vals, uint32_t *tmp_keys, int32_t *tmp_vals,
                      int64_t m, int key_bits, int32_t *hist, int32_t *scan_tmp, hipStream_t s,
                      int copy_back, int identity_vals)
{
    const int passes = radix_passes(key_bits);
    if (identity_vals && (m <= 1 || passes == 0)) {
        // no pass reads the values: write the identity where the result is
        if (m > 0)
            hipLaunchKernelGGL(k_iota, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, vals, m);
        CG_HIP(hipGetLastError());
    }
    if (m <= 1) {
        // the one element (if any) goes where the passes would have left it
        if (!copy_back && (passes & 1) && m == 1) {
            CG_HIP(hipMemcpyAsync(tmp_keys, keys, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            CG_HIP(hipMemcpyAsync(tmp_vals, vals, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        }
        return CG_OK;
    }
    int nblocks = (int)ceil_div(m, kSortTile);
    uint32_t *ka = keys, *kb = tmp_keys;
    int32_t *va = vals, *vb = tmp_vals;
    for (int p = 0; p < passes; p++) {
        int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, s, ka, m, shift,
                           hist, nblocks);
        CG_HIP(hipGetLastError());
        int r = launch_exclusive_scan(hist, (int64_t)kRadix * nblocks, scan_tmp, s);
        if (r) return r;
        // with identity_vals the first pass generates the values
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, s, ka,
                           p == 0 && identity_vals ? nullptr : va, kb, vb, m, shift, hist, nblocks);
        CG_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (copy_back && ka != keys) {
        CG_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, s));
        CG_HIP(hipMemcpyAsync(vals, va, sizeof(int32_t) * m, hipMemcpyDeviceToDevice, s));
    }
    return CG_OK;
}

}  // namespace cg
