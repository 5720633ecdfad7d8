// Orchestration of ms2dirty / dirty2ms on one MI355X and the C-ABI.
//
// The eight reference entry points (include/cugridder.h:20-46, defined at
// src/astro/cugridder.cu:540-571) keep their names, argument order and
// meaning: host buffers in, host buffers out, every device allocation made
// and released inside the call, int ier returned (0 ok, 1 eps clamped).
// Unlike the reference, failures return a non-zero code instead of exit().
//
// Device-pointer entry points (cg_*_dev) run the same pipeline on inputs
// already resident in HBM with a reusable workspace held by a context.
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "cg_internal.hpp"
#include "../../include/curafft.h"

namespace cg {

void report_hip_error(hipError_t e, const char *what, const char *file, int line)
{
    if (std::getenv("CG_QUIET")) return;
    std::fprintf(stderr, "curafft: HIP error '%s' (%d) at %s:%d in %s\n", hipGetErrorString(e),
                 (int)e, file, line, what);
}

// A fixed device region that a context's buffers are carved from instead of
// hipMalloc (the allocation-free entry point); reset at the start of a call.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, off = 0;
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    Arena *arena = nullptr;  // non-null: carved from the arena, never freed
};

enum Stage { ST_PLAN = 0, ST_SORT, ST_RECORDS, ST_GRID, ST_FFT, ST_FFT2, ST_IMAGE, ST_COUNT };

struct Ctx {
    hipStream_t stream = nullptr;
    int lds_budget = 155648;      // degridding tile + halo staging
    int grid_lds_budget = 65536;  // gridding accumulators
    int tile_grid = 8;
    int tile_grid2d = 32;  // 2-D (one plane) gridding tiles
    int mfma2d = 1;        // 2-D complex64 gridding as per-quadrant MFMA products
    int tile_degrid = 16;
    int tile_degrid2d = 32;  // 2-D NUFFT type 2 degridding tiles (two radix passes at 8192^2)
    int dense_min = 64;      // points per bin that send it to the MFMA gridding path (0: off)
    int degrid_dense_min = 2048;  // points per degridding tile for the MFMA degridder (0: off)
    int tap_poly = 0;        // complex64 MFMA kernels: kernel values from per-tap polynomials
    int row_mask = 1;        // let the FFT skip grid rows no point reaches
    int prerec = 1;          // complex64 records in two coalescing passes
    int band_fft = 1;        // band-limited column passes when the rows points reach allow
    int compact_keys = 1;    // sort keys number only the tiles the u / v ranges allow
    // grid buffer limit in MiB (0: what the device has free); a W-stacking
    // grid larger than the limit is processed in slabs of w-planes
    int grid_limit_mb = 0;
    int last_slab = 0, last_nslab = 1;  // planes per slab and slabs of the last call
    int stop_after = 0;      // test hook: 1 = after gridding/prologue, 2 = after FFT
    // geometry of the current call beyond (fov, nx, ny): a separate pixel
    // size along ny and plain 2-D gridding; set only for the duration of a
    // call by the entry points that take them
    PlanGeom geom;
    int64_t last_m = 0;
    int last_cshift = 0;
    // workspace (grow-only)
    DevBuf keys, keys2, vals, vals2, hist, scan, tstart, geo, val, grid, zq, phi3, phi1, phi2,
        part, tw1, tw2, fftmp, dlist, rmask, chans, tlist;
    int tw1_n = 0, tw2_n = 0, tw_bytes = 0;
    // stage timing (HIP events on the context stream)
    bool timing = false;
    // stage marks ev[0..ST_COUNT]; ev[ST_COUNT + 1, + 2] bracket the dense
    // gridding kernel inside the ms2dirty grid stage
    hipEvent_t ev[ST_COUNT + 3] = {};
    bool ev_ok = false, dense_timed = false;
    double stage_ms[2][ST_COUNT] = {};
    double dense_ms = 0, degrid_dense_ms = 0;  // the MFMA kernels' parts of the grid stages
    int64_t stage_calls[2] = {0, 0};
    WPlan last_plan;
    Tiling last_tiling;
    // learned kernels by support width: used (single / mixed precision) when
    // certified at an eps no larger than the call's and narrower than ES's
    struct Learned {
        double eps = 0;  // 0: none
        // the oversampling factor the kernel was trained and certified for:
        // training (learn_beta, the training samples) and certification run
        // at sigma = 2, so calls at another sigma keep the ES kernel
        double sigma = 2.0;
        LearnedKernel K;
    } learned[LEARN_MAXKW + 1];
    int use_learned = 1;
    // set by use_arena(): every buffer is carved from this region
    Arena arena;
    bool arena_mode = false;

    template <typename F>
    void each_buf(F f)
    {
        DevBuf *bufs[] = {&keys, &keys2, &vals, &vals2, &hist, &scan, &tstart, &geo, &val,
                          &grid, &zq, &phi3, &phi1, &phi2, &part, &tw1, &tw2, &fftmp, &dlist,
                          &rmask, &chans, &tlist};
        for (DevBuf *b : bufs) f(*b);
    }
    // switches this (fresh) context to carving its buffers from [base,
    // base + cap); nothing is allocated or freed from then on
    void use_arena(void *base, size_t cap)
    {
        arena.base = (char *)base;
        arena.cap = cap;
        arena.off = 0;
        arena_mode = true;
        each_buf([this](DevBuf &b) { b.arena = &arena; });
    }
    // forgets every carved buffer (their contents included, so cached
    // twiddle tables are rebuilt): the next call carves from offset 0
    void reset_arena()
    {
        arena.off = 0;
        each_buf([](DevBuf &b) {
            b.p = nullptr;
            b.bytes = 0;
        });
        tw1_n = tw2_n = tw_bytes = 0;
    }

    ~Ctx()
    {
        each_buf([](DevBuf &b) {
            if (b.p && !b.arena) (void)hipFree(b.p);
        });
        if (ev_ok)
            for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

constexpr size_t kArenaAlign = 256;
inline size_t arena_round(size_t bytes) { return (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

static int ensure(DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (b.bytes >= bytes) return CG_OK;
    if (b.arena) {
        // a buffer that grows within a call takes a new region; the old one
        // stays unused until the arena is reset
        Arena &A = *b.arena;
        const size_t need = arena_round(bytes);
        if (A.off > A.cap || need > A.cap - A.off) return CG_ERR_SIZE;
        b.p = A.base + A.off;
        b.bytes = need;
        A.off += need;
        return CG_OK;
    }
    if (b.p) {
        CG_HIP(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return CG_ERR_ALLOC;
    }
    b.bytes = bytes;
    return CG_OK;
}

template <typename X>
static X *as(DevBuf &b)
{
    return reinterpret_cast<X *>(b.p);
}


static void env_overrides(Ctx &C)
{
    if (const char *e = std::getenv("CG_LDS_BUDGET")) C.lds_budget = std::atoi(e);
    if (const char *e = std::getenv("CG_GRID_LDS")) C.grid_lds_budget = std::atoi(e);
    if (const char *e = std::getenv("CG_DENSE_MIN")) C.dense_min = std::atoi(e);
    if (const char *e = std::getenv("CG_TILE_GRID")) C.tile_grid = std::atoi(e);
    if (const char *e = std::getenv("CG_TILE_GRID2D")) C.tile_grid2d = std::atoi(e);
    if (const char *e = std::getenv("CG_PREREC")) C.prerec = std::atoi(e);
    if (const char *e = std::getenv("CG_MFMA2D")) C.mfma2d = std::atoi(e);
    if (const char *e = std::getenv("CG_BAND_FFT")) C.band_fft = std::atoi(e);
    if (const char *e = std::getenv("CG_COMPACT_KEYS")) C.compact_keys = std::atoi(e);
    if (const char *e = std::getenv("CG_TILE_DEGRID")) C.tile_degrid = std::atoi(e);
    if (const char *e = std::getenv("CG_TILE_DEGRID2D")) C.tile_degrid2d = std::atoi(e);
    if (const char *e = std::getenv("CG_GRID_LIMIT_MB")) C.grid_limit_mb = std::atoi(e);
    if (const char *e = std::getenv("CG_TAP_POLY")) C.tap_poly = std::atoi(e);
}

struct StageClock {
    Ctx &C;
    int dir;
    explicit StageClock(Ctx &c, int d) : C(c), dir(d) {}
    int mark(int i)
    {
        if (!C.timing) return CG_OK;
        if (!C.ev_ok) {
            for (auto &e : C.ev) CG_HIP(hipEventCreate(&e));
            C.ev_ok = true;
        }
        CG_HIP(hipEventRecord(C.ev[i], C.stream));
        return CG_OK;
    }
    int finish()
    {
        if (!C.timing) return CG_OK;
        CG_HIP(hipEventSynchronize(C.ev[ST_COUNT]));
        for (int i = 0; i < ST_COUNT; i++) {
            float ms = 0.f;
            CG_HIP(hipEventElapsedTime(&ms, C.ev[i], C.ev[i + 1]));
            C.stage_ms[dir][i] += ms;
        }
        if (C.dense_timed) {
            float ms = 0.f;
            CG_HIP(hipEventElapsedTime(&ms, C.ev[ST_COUNT + 1], C.ev[ST_COUNT + 2]));
            C.dense_ms += ms;
            C.dense_timed = false;
        }
        C.stage_calls[dir]++;
        return CG_OK;
    }
};

// rows (mask[0, nf2)) and columns (mask[nf2, nf2 + nf1)) of the grid that
// points of non-empty tiles reach (null: all); needs the tile starts of the
// last sort
static int row_mask(Ctx &C, const Tiling &t, int kw, int nf1, int nf2, const uint8_t *&mask)
{
    mask = nullptr;
    if (!C.row_mask) return CG_OK;
    CG_TRY(ensure(C.rmask, (size_t)nf1 + nf2));
    CG_TRY(launch_row_mask(as<int32_t>(C.tstart), t, kw, nf1, nf2, as<uint8_t>(C.rmask), C.stream));
    mask = as<uint8_t>(C.rmask);
    return CG_OK;
}

// lets the gridder leave the tiles the FFT will not read unwritten, when
// every pass of this FFT honours the masks (test hook stop_after = 1 keeps
// the whole grid written)
template <typename T>
static int grid_masks(Ctx &C, GridArgs &g, const uint8_t *mask, int nf1, int nf2)
{
    if (mask && C.stop_after != 1 && fft_rows_masked<T>(nf1) && fft_spectrum_pruned<T>(nf2)) {
        CG_TRY(ensure(C.tlist, sizeof(int32_t) * ((size_t)g.ntx * g.nty + 1)));
        g.rowmask = mask;
        g.colmask = mask + nf2;
        g.tile_list = as<int32_t>(C.tlist);
    }
    return CG_OK;
}

// scratch for the two-pass records (complex64 ms2dirty, whose records read
// uvw, vis and wgt; dirty2ms reads less and keeps the one-pass gather): the
// grid buffer, which is not in use until after the records are built
template <typename T>
static int records_scratch(Ctx &C, size_t grid_bytes, int64_t M, int dir, float4 *&rec)
{
    rec = nullptr;
    if (sizeof(T) != 4 || M <= 0 || dir != 0 || !C.prerec) return CG_OK;
    CG_TRY(ensure(C.grid, std::max(grid_bytes, (size_t)32 * M)));
    rec = as<float4>(C.grid);
    return CG_OK;
}

// twiddle tables (cached per length and precision)
template <typename T>
static int twiddles(Ctx &C, int n1, int n2)
{
    const int eb = (int)sizeof(cx_t<T>);
    if (C.tw_bytes != eb || C.tw1_n != n1) {
        // long rows also need the half-length table, stored after the first
        const bool lng = fft_row_long<T>(n1);
        CG_TRY(ensure(C.tw1, (size_t)(lng ? n1 + n1 / 2 : n1) * eb));
        CG_TRY(launch_twiddles<T>(as<cx_t<T>>(C.tw1), n1, C.stream));
        if (lng) CG_TRY(launch_twiddles<T>(as<cx_t<T>>(C.tw1) + n1, n1 / 2, C.stream));
        C.tw1_n = n1;
    }
    if (C.tw_bytes != eb || C.tw2_n != n2) {
        CG_TRY(ensure(C.tw2, (size_t)n2 * eb));
        CG_TRY(launch_twiddles<T>(as<cx_t<T>>(C.tw2), n2, C.stream));
        C.tw2_n = n2;
    }
    C.tw_bytes = eb;
    return CG_OK;
}

// sort rows by the tile of their footprint start; leaves the permutation in
// C.vals and the tile offsets in C.tstart.  With rec (complex64 ms2dirty),
// the key pass also writes the caller-order record slots of vis and wgt.
template <typename TU>
static int sort_rows(Ctx &C, const TU *uvw, int64_t M, const CoordArgs &a, const Tiling &t,
                     const float2 *vis = nullptr, const float *wgt = nullptr, float4 *rec = nullptr)
{
    const int64_t ntiles = t.ntiles();
    CG_TRY(ensure(C.keys, sizeof(uint32_t) * M));
    CG_TRY(ensure(C.keys2, sizeof(uint32_t) * M));
    CG_TRY(ensure(C.vals, sizeof(int32_t) * M));
    CG_TRY(ensure(C.vals2, sizeof(int32_t) * M));
    const int64_t hs = radix_hist_size(M);
    CG_TRY(ensure(C.hist, sizeof(int32_t) * hs));
    CG_TRY(ensure(C.scan, sizeof(int32_t) * std::max(scan_tmp_size(hs), scan_tmp_size(ntiles + 1))));
    CG_TRY(ensure(C.tstart, sizeof(int32_t) * (ntiles + 1)));
    // an odd number of radix passes starts in the second buffers, so the
    // sorted keys and permutation end in C.keys / C.vals without a copy
    const int bits = sort_key_bits(a, ntiles);
    const bool odd = radix_passes(bits) & 1;
    uint32_t *k0 = as<uint32_t>(odd ? C.keys2 : C.keys), *k1 = as<uint32_t>(odd ? C.keys : C.keys2);
    int32_t *v0 = as<int32_t>(odd ? C.vals2 : C.vals), *v1 = as<int32_t>(odd ? C.vals : C.vals2);
    // the permutation starts as the identity, generated by the first pass
    if (rec)
        CG_TRY((launch_keys_prerecords<TU, float2, float>(uvw, vis, wgt, M, a, 0, k0, nullptr, rec,
                                                           C.stream)));
    else
        CG_TRY(launch_keys<TU>(uvw, M, a, k0, nullptr, C.stream));
    CG_TRY(launch_radix_sort(k0, v0, k1, v1, M, bits, as<int32_t>(C.hist), as<int32_t>(C.scan), C.stream, 0,
                             1));
    CG_TRY(launch_tile_bounds(as<uint32_t>(C.keys), M, ntiles, a.cshift, key_rect(a), as<int32_t>(C.tstart),
                              C.stream));
    return CG_OK;
}

// compact tile numbering of the sort key when both bands are known, do not
// wrap, and the rectangle of their tiles needs fewer key bits than the grid
static void set_key_rect(CoordArgs &a, const WPlan &P, const Tiling &t, bool on)
{
    a.cby0 = a.cnby = a.cbx0 = a.cnbx = 0;
    if (!on || P.band_len <= 0 || P.cband_len <= 0 || P.band_s + P.band_len > P.nf2 ||
        P.cband_s + P.cband_len > P.nf1)
        return;
    const int by0 = P.band_s / t.ty, by1 = (P.band_s + P.band_len - 1) / t.ty;
    const int bx0 = P.cband_s / t.tx, bx1 = (P.cband_s + P.cband_len - 1) / t.tx;
    const int64_t n = (int64_t)(by1 - by0 + 1) * (bx1 - bx0 + 1);
    if (n * 2 > t.ntiles()) return;  // saves less than a key bit
    a.cby0 = by0;
    a.cnby = by1 - by0 + 1;
    a.cbx0 = bx0;
    a.cnbx = bx1 - bx0 + 1;
}

static GridArgs grid_args(Ctx &C, const WPlan &P, const Tiling &t, int64_t m, int cshift)
{
    GridArgs g{};
    if (P.conv.tab_nq > 0) {
        const LearnedKernel &K = C.learned[P.conv.kw].K;
        g.learned = 1;
        for (int k = 0; k <= LEARN_MAXDEG; k++) g.kq[k] = K.coef[k];
    }
    g.cshift = cshift;
    g.m = (int)m;
    g.nf1 = P.nf1;
    g.nf2 = P.nf2;
    g.nf3 = P.nf3;
    g.tx = t.tx;
    g.ty = t.ty;
    g.ntx = t.ntx;
    g.nty = t.nty;
    g.sbx = t.sbx;
    g.sby = t.sby;
    g.pc = t.pc;
    g.nchunk = t.nchunk;
    g.beta = P.conv.beta;
    g.es_c = P.conv.es_c;
    g.f_nb = (float)(-g.beta * 1.4426950408889634);
    g.f_ec = (float)g.es_c;
    g.f_e1 = (float)std::exp(g.beta);
    g.f_e2 = (float)std::exp(2.0 * g.beta);
    g.e3 = std::exp(3.0 * g.beta);
    // per-tap polynomials of the same normalised values (complex64 MFMA
    // kernels): on request, and always for a learned kernel, whose exponent
    // polynomial would otherwise cost 16 more FMAs per tap
    g.tp_on = 0;
    if ((C.tap_poly || g.learned) && P.conv.kw <= 8) {
        tap_polynomials(P.conv.kw, P.conv.beta, P.conv.es_c, g.learned ? g.kq : nullptr,
                        LEARN_MAXDEG, 10, g.tp);
        g.tp_on = 1;
    }
    return g;
}

// Plan for nrow rows observed at nchan frequencies (host array freq): the
// w-range of the points is the raw w-range of the rows scaled per channel
// (exact: an FP64 product is monotonic in each factor).
template <typename T, typename TU>
static int plan_call(Ctx &C, WPlan &P, int64_t nrow, int nchan, const double *freq, int nx,
                     int ny, double fov, const TU *uvw, double eps, double sigma, int sign,
                     bool single)
{
    if (nrow < 0 || nchan < 1 || !freq) return CG_ERR_ARG;
    if (nrow * (int64_t)nchan > (int64_t)INT32_MAX - 1) return CG_ERR_SIZE;
    for (int c = 0; c < nchan; c++)
        if (!(freq[c] > 0)) return CG_ERR_ARG;
    CG_TRY(ensure(C.part, sizeof(double) * 6152));
    double wlo = 0, whi = 0, wmin = 0, wmax = 0, uv[4] = {0, 0, 0, 0};
    CG_TRY(launch_wminmax<TU>(uvw, nrow, 1.0, as<double>(C.part), wlo, whi, C.stream, uv));
    if (nrow > 0) {
        wmin = INFINITY;
        wmax = -INFINITY;
        for (int c = 0; c < nchan; c++) {
            const double s = 2.0 * kPi * (freq[c] / kC) * sign;
            const double a = wlo * s, b = whi * s;
            wmin = std::min(wmin, std::min(a, b));
            wmax = std::max(wmax, std::max(a, b));
        }
    }
    int ier = make_wplan(P, nrow * nchan, nx, ny, fov, freq[0], eps, sigma, sign, single, wmin, wmax,
                         nullptr, &C.geom);
    if (ier >= 2) return ier;
    // the narrowest learned kernel certified for eps and no wider than the
    // ES kernel (at equal support it is the more accurate one)
    if (sizeof(T) == 4 && C.use_learned) {
        for (int kw = 2; kw <= P.conv.kw; kw++) {
            const Ctx::Learned &L = C.learned[kw];
            if (!(L.eps > 0) || L.eps > eps || sigma != L.sigma) continue;
            // the gridders evaluate exp(beta sqrt(1 - x^2) + y): the node
            // values take the same normalisation (e^beta times psi)
            ConvOpts o = P.conv;
            o.kw = kw;
            o.halfwidth = kw / 2.0;
            o.beta = L.K.beta;
            o.es_c = 4.0 / ((double)kw * kw);
            o.tab_nq = L.K.nq;
            for (int q = 0; q < L.K.nq; q++) o.tab_psi[q] = L.K.qpsi[q] * std::exp(L.K.beta);
            ier = make_wplan(P, nrow * nchan, nx, ny, fov, freq[0], eps, sigma, sign, single, wmin,
                             wmax, &o, &C.geom);
            if (ier >= 2) return ier;
            break;
        }
    }
    if (!fft_supported<T>(P.nf1) || !fft_supported<T>(P.nf2)) return CG_ERR_SIZE;
    // grid rows and columns the points can reach: from the u and v ranges
    // over all channels
    if (nrow > 0) {
        double smin = INFINITY, smax = 0;
        for (int c = 0; c < nchan; c++) {
            const double sc = 2.0 * kPi * (freq[c] / kC);
            smin = std::min(smin, sc);
            smax = std::max(smax, sc);
        }
        auto lo = [&](double x) { return x < 0 ? x * smax : x * smin; };
        auto hi = [&](double x) { return x > 0 ? x * smax : x * smin; };
        plan_bands(P, lo(uv[0]), hi(uv[1]), lo(uv[2]), hi(uv[3]));
    }
    return ier;
}

// per-channel coordinate scales 2 pi f_c / c on the device (nchan > 1)
static int channel_scales(Ctx &C, int nchan, const double *freq, CoordArgs &a)
{
    a.nchan = nchan;
    a.chan_scale = nullptr;
    if (nchan <= 1) return CG_OK;
    std::vector<double> h(nchan);
    for (int c = 0; c < nchan; c++) h[c] = 2.0 * kPi * (freq[c] / kC);
    CG_TRY(ensure(C.chans, sizeof(double) * nchan));
    CG_HIP(hipMemcpyAsync(C.chans.p, h.data(), sizeof(double) * nchan, hipMemcpyHostToDevice, C.stream));
    CG_HIP(hipStreamSynchronize(C.stream));
    a.chan_scale = as<double>(C.chans);
    return CG_OK;
}

// ------------------------------------------------------------ ms2dirty

// Tuning of one call: tile edges, LDS budgets and the dense threshold
// (defaults from the context options, or per item in a batch).
struct Tune {
    int tile_grid, tile_degrid, grid_lds, lds, dense_min;
};
static Tune ctx_tune(const Ctx &C)
{
    return Tune{C.tile_grid, C.tile_degrid, C.grid_lds_budget, C.lds_budget, C.dense_min};
}

// Everything a call decides before its kernels run.  Planning reads the
// w-range back to the host (it synchronises); executing a plan only
// launches kernels and copies on the context stream, so a sequence of
// executions can be captured into a graph once the workspace has grown.
struct CallPlan {
    WPlan P;
    Tiling t;
    CoordArgs a;
    int64_t M = 0;
    int dense_min = 0;  // > 0: dense bins go through the MFMA gridder
    int slab = 0;       // > 0: w-planes per slab of the grid buffer (out-of-core flow)
    int ier = 0;
};

// Planes per slab when the W-stacking grid does not fit the grid limit (the
// context option, else 90% of what the device has free, counting the grid
// buffer the context already holds, less room for the call's other
// buffers): 0 when it fits whole.  The slabs are processed one after the
// other through one buffer, as the reference's mem_limit pipeline does
// (src/astro/ra_exec.cu:292-389, 520-613).
static int plan_slab(Ctx &C, const WPlan &P, int eb, int64_t M, int &slab)
{
    slab = 0;
    if (!P.wstack || P.nf3 <= 1) return CG_OK;
    const size_t plane = (size_t)P.nf1 * P.nf2 * eb;
    const size_t whole = plane * P.nf3;
    size_t limit;
    if (C.grid_limit_mb > 0) {
        limit = (size_t)C.grid_limit_mb << 20;
    } else {
        size_t fr = 0, tot = 0;
        CG_HIP(hipMemGetInfo(&fr, &tot));
        const size_t avail = fr + C.grid.bytes;
        const size_t other = (size_t)M * 96 + ((size_t)256 << 20);
        limit = avail > other ? (size_t)((avail - other) * 0.9) : 0;
    }
    if (whole <= limit) return CG_OK;
    if (plane > limit) return CG_ERR_ALLOC;  // not even one plane fits
    const int per = (int)std::min<size_t>(limit / plane, (size_t)P.nf3);
    const int nslab = (P.nf3 + per - 1) / per;
    slab = (P.nf3 + nslab - 1) / nslab;  // balanced slabs
    return CG_OK;
}

// bytes of the grid buffer: the whole grid, or one slab of planes
static size_t slab_bytes(const WPlan &P, int slab, int eb)
{
    return (size_t)P.nf1 * P.nf2 * (slab ? slab : P.nf3) * eb;
}

// vis and wgt are [nrow][nchan]; every (row, channel) is one point
template <typename T, typename TU>
static int plan_ms2dirty(Ctx &C, CallPlan &cp, int64_t nrow, int nchan, const double *freq,
                         int nx, int ny, double fov, const TU *uvw, double eps, double sigma,
                         int sign, bool single, const Tune &tn)
{
    WPlan &P = cp.P;
    cp.ier = plan_call<T, TU>(C, P, nrow, nchan, freq, nx, ny, fov, uvw, eps, sigma, sign, single);
    if (cp.ier >= 2) return cp.ier;
    const int eb = (int)sizeof(cx_t<T>);
    cp.M = nrow * nchan;
    CG_TRY(plan_slab(C, P, eb, cp.M, cp.slab));
    // gridding accumulates in FP64 (16 bytes per cell) whatever the precision
    cp.t = choose_tiling(P.nf1, P.nf2, cp.slab ? cp.slab : P.nf3, P.conv.kw, 16, tn.grid_lds,
                         tn.tile_grid, 0);
    cp.a = make_coord_args(P, cp.t);
    set_key_rect(cp.a, P, cp.t, C.compact_keys != 0);
    CG_TRY(channel_scales(C, nchan, freq, cp.a));
    // bins with many points per start cell go through the MFMA path, added
    // on top of what the sparse kernel writes; it needs the points of a bin
    // ordered by start cell, which the sparse kernel alone is better without
    // (same-cell neighbours in a wavefront collide on the same LDS words).
    // It holds every plane of its footprint region, so not with slabs.
    const bool dense = sizeof(T) == 4 && tn.dense_min > 0 && cp.M > 0 && !cp.slab &&
                       grid_dense_supported(grid_args(C, P, cp.t, cp.M, cp.a.cshift), P.conv.kw, eb);
    if (!dense) cp.a.cshift = 0;
    cp.dense_min = dense ? tn.dense_min : 0;
    return cp.ier;
}

template <typename T, typename TU, typename TW>
static int exec_ms2dirty(Ctx &C, const CallPlan &cp, const TU *uvw, const cx_t<T> *vis,
                         const TW *wgt, cx_t<T> *dirty, StageClock &clk)
{
    const WPlan &P = cp.P;
    const Tiling &t = cp.t;
    const CoordArgs &a = cp.a;
    const int64_t M = cp.M;
    const int eb = (int)sizeof(cx_t<T>);
    C.last_plan = P;
    C.last_tiling = t;
    C.last_m = M;
    C.last_cshift = a.cshift;
    CG_TRY(clk.mark(ST_SORT));
    // complex64: the key pass writes the caller-order records too, and the
    // record stage only gathers them
    float4 *rec = nullptr;
    if (M > 0) {
        if constexpr (sizeof(T) == 4 && std::is_same<TW, float>::value) {
            CG_TRY(records_scratch<T>(C, slab_bytes(P, cp.slab, eb), M, 0, rec));
            if (rec) CG_TRY(sort_rows<TU>(C, uvw, M, a, t, (const float2 *)vis, wgt, rec));
        }
        if (!rec) CG_TRY(sort_rows<TU>(C, uvw, M, a, t));
    } else {
        CG_TRY(ensure(C.tstart, sizeof(int32_t) * (t.ntiles() + 1)));
        CG_TRY(launch_fill_i32(as<int32_t>(C.tstart), t.ntiles() + 1, 0, C.stream));
    }
    CG_TRY(clk.mark(ST_RECORDS));
    CG_TRY(ensure(C.geo, sizeof(geo_t<T>) * M));
    CG_TRY(ensure(C.val, sizeof(cx_t<T>) * M));
    if (rec)
        CG_TRY(launch_gather_records(rec, as<int32_t>(C.vals), M, (float4 *)C.geo.p, (float2 *)C.val.p,
                                     C.stream));
    else
        CG_TRY((launch_records<T, TU, cx_t<T>, TW>(uvw, vis, wgt, as<int32_t>(C.vals), M, a, 0,
                                                    as<geo_t<T>>(C.geo), as<cx_t<T>>(C.val), nullptr,
                                                    C.stream)));
    CG_TRY(clk.mark(ST_GRID));
    CG_TRY(ensure(C.grid, slab_bytes(P, cp.slab, eb)));
    const uint8_t *rm = nullptr;
    CG_TRY(row_mask(C, t, P.conv.kw, P.nf1, P.nf2, rm));
    GridArgs g = grid_args(C, P, t, M, a.cshift);
    g.kr = key_rect(a);
    g.dense_min = cp.dense_min;
    CG_TRY(grid_masks<T>(C, g, rm, P.nf1, P.nf2));
    // correction factors and the quarter tables, for the image stage
    WArgs w = make_wargs(P);
    CG_TRY(ensure(C.phi1, sizeof(double) * (P.nf1 / 2 + 1)));
    CG_TRY(ensure(C.phi2, sizeof(double) * (P.nf2 / 2 + 1)));
    const int64_t nq = (int64_t)(P.N1 / 2 + 1) * (P.N2 / 2 + 1);
    if (P.wstack) {
        CG_TRY(ensure(C.zq, sizeof(double) * nq));
        CG_TRY(ensure(C.phi3, sizeof(double) * nq));
    }
    CG_TRY(twiddles<T>(C, P.nf1, P.nf2));
    // only the columns holding image pixels are transformed along n2, only
    // the rows points reach hold non-zero data, and only the frequencies of
    // image rows are kept along n2
    const int la = P.N1 - P.N1 / 2, lb = P.nf1 - P.N1 / 2;
    const int sa = P.N2 - P.N2 / 2, sb = P.nf2 - P.N2 / 2;
    // slabs of planes [zb, zb + zn) one after the other through the grid
    // buffer (one slab holding every plane without the out-of-core flow):
    // grid, transform, and add the slab's part of the w-DFT to the image
    const int per = cp.slab ? cp.slab : P.nf3;
    const int nslab = (P.nf3 + per - 1) / per;
    C.last_slab = cp.slab;
    C.last_nslab = nslab;
    for (int zb = 0; zb < P.nf3; zb += per) {
        const int zn = std::min(per, P.nf3 - zb);
        const bool last = zb + zn >= P.nf3;
        if (cp.slab) {
            g.zb = zb;
            g.zn = zn;
            g.nchunk = (zn + t.pc - 1) / t.pc;
        }
        CG_TRY(launch_grid<T>(as<geo_t<T>>(C.geo), as<cx_t<T>>(C.val), as<int32_t>(C.tstart),
                              as<cx_t<T>>(C.grid), g, P.conv.kw, C.stream));
        if (cp.dense_min > 0) {
            CG_TRY(ensure(C.dlist, sizeof(int32_t) * dense_list_ints(t.ntiles(), g.m)));
            CG_TRY(clk.mark(ST_COUNT + 1));
            CG_TRY(launch_grid_dense((const geo_t<float> *)C.geo.p, (const cx_t<float> *)C.val.p,
                                     as<uint32_t>(C.keys), as<int32_t>(C.tstart), as<int32_t>(C.dlist),
                                     (cx_t<float> *)C.grid.p, g, P.conv.kw, C.stream));
            CG_TRY(clk.mark(ST_COUNT + 2));
            C.dense_timed = C.timing;
        }
        if (C.stop_after == 1) return P.ier;
        // with slabs, the whole loop is reported as the grid stage
        if (!cp.slab) CG_TRY(clk.mark(ST_FFT));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), zn, P.nf2, P.nf1, +1, true,
                                  as<cx_t<T>>(C.tw1), la, lb, rm, sa, sb, C.stream));
        if (!cp.slab) CG_TRY(clk.mark(ST_FFT2));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), zn, P.nf2, P.nf1, +1, false,
                                  as<cx_t<T>>(C.tw2), la, lb, rm, sa, sb, C.stream, P.band_s,
                                  C.band_fft ? P.band_len : 0));
        if (C.stop_after == 2) return P.ier;
        if (!cp.slab) CG_TRY(clk.mark(ST_IMAGE));
        if (zb == 0) {
            CG_TRY(launch_fseries_grid(w.qd, P.nf1, as<double>(C.phi1), C.stream));
            CG_TRY(launch_fseries_grid(w.qd, P.nf2, as<double>(C.phi2), C.stream));
            if (P.wstack) CG_TRY(launch_quarter_table(w, as<double>(C.zq), as<double>(C.phi3), C.stream));
        }
        if (!P.wstack) {
            // plain 2-D: the u/v deconvolution only (deconv_2d type 1)
            CG_TRY((launch_deconv2d<T, cx_t<T>>(as<cx_t<T>>(C.grid), dirty, as<double>(C.phi1),
                                                as<double>(C.phi2), P.N1, P.N2, P.nf1, P.nf2, 1,
                                                C.stream)));
        } else {
            w.zb = zb;
            w.zn = cp.slab ? zn : 0;
            w.accumulate = zb > 0;
            CG_TRY((launch_ms2dirty_epilogue<T, cx_t<T>>(as<cx_t<T>>(C.grid), as<double>(C.zq),
                                                          as<double>(C.phi3), as<double>(C.phi1),
                                                          as<double>(C.phi2), dirty, w, C.stream)));
        }
        if (last && cp.slab) {
            CG_TRY(clk.mark(ST_FFT));
            CG_TRY(clk.mark(ST_FFT2));
            CG_TRY(clk.mark(ST_IMAGE));
        }
    }
    CG_TRY(clk.mark(ST_COUNT));
    return P.ier;
}

template <typename T, typename TU, typename TW>
int ms2dirty_dev(Ctx &C, int64_t nrow, int nchan, const double *freq, int nx, int ny, double fov,
                 const TU *uvw, const cx_t<T> *vis, const TW *wgt, cx_t<T> *dirty, double eps,
                 double sigma, int sign, bool single)
{
    StageClock clk(C, 0);
    CG_TRY(clk.mark(ST_PLAN));
    CallPlan cp;
    const int ier = plan_ms2dirty<T, TU>(C, cp, nrow, nchan, freq, nx, ny, fov, uvw, eps, sigma,
                                         sign, single, ctx_tune(C));
    if (ier >= 2) return ier;
    CG_TRY((exec_ms2dirty<T, TU, TW>(C, cp, uvw, vis, wgt, dirty, clk)));
    if (C.stop_after == 0) CG_TRY(clk.finish());
    return cp.ier;
}

// ------------------------------------------------------------ dirty2ms

template <typename T, typename TU>
static int plan_dirty2ms(Ctx &C, CallPlan &cp, int64_t nrow, int nchan, const double *freq,
                         int nx, int ny, double fov, const TU *uvw, double eps, double sigma,
                         int sign, bool single, const Tune &tn)
{
    WPlan &P = cp.P;
    cp.ier = plan_call<T, TU>(C, P, nrow, nchan, freq, nx, ny, fov, uvw, eps, sigma, sign, single);
    if (cp.ier >= 2) return cp.ier;
    const int eb = (int)sizeof(cx_t<T>);
    const int kw = P.conv.kw;
    cp.M = nrow * nchan;
    CG_TRY(plan_slab(C, P, eb, cp.M, cp.slab));
    cp.t = choose_tiling(P.nf1, P.nf2, cp.slab ? cp.slab : P.nf3, kw, eb, tn.lds, tn.tile_degrid,
                         kw - 1);
    cp.a = make_coord_args(P, cp.t);
    set_key_rect(cp.a, P, cp.t, C.compact_keys != 0);
    CG_TRY(channel_scales(C, nchan, freq, cp.a));
    cp.dense_min = 0;
    return cp.ier;
}

template <typename T, typename TU, typename TW>
static int exec_dirty2ms(Ctx &C, const CallPlan &cp, const TU *uvw, cx_t<T> *vis, const TW *wgt,
                         const cx_t<T> *dirty, StageClock &clk)
{
    const WPlan &P = cp.P;
    const Tiling &t = cp.t;
    const CoordArgs &a = cp.a;
    const int64_t M = cp.M;
    const int eb = (int)sizeof(cx_t<T>);
    const int kw = P.conv.kw;
    C.last_plan = P;
    C.last_tiling = t;
    C.last_m = M;
    C.last_cshift = a.cshift;
    CG_TRY(clk.mark(ST_SORT));
    if (M > 0) CG_TRY(sort_rows<TU>(C, uvw, M, a, t));
    CG_TRY(clk.mark(ST_RECORDS));
    CG_TRY(ensure(C.geo, sizeof(geo_t<T>) * M));
    CG_TRY(ensure(C.val, sizeof(cx_t<T>) * M));
    float4 *rec = nullptr;
    CG_TRY(records_scratch<T>(C, slab_bytes(P, cp.slab, eb), M, 1, rec));
    CG_TRY((launch_records<T, TU, cx_t<T>, TW>(uvw, (const cx_t<T> *)nullptr, wgt,
                                                as<int32_t>(C.vals), M, a, 1, as<geo_t<T>>(C.geo),
                                                as<cx_t<T>>(C.val), rec, C.stream)));
    CG_TRY(ensure(C.phi1, sizeof(double) * (P.nf1 / 2 + 1)));
    CG_TRY(ensure(C.phi2, sizeof(double) * (P.nf2 / 2 + 1)));
    CG_TRY(ensure(C.grid, slab_bytes(P, cp.slab, eb)));
    WArgs w = make_wargs(P);
    const int64_t nq = (int64_t)(P.N1 / 2 + 1) * (P.N2 / 2 + 1);
    if (P.wstack) {
        CG_TRY(ensure(C.zq, sizeof(double) * nq));
        CG_TRY(ensure(C.phi3, sizeof(double) * nq));
    }
    CG_TRY(twiddles<T>(C, P.nf1, P.nf2));
    // the degridder reads only the rows its tiles' points reach
    const uint8_t *rm = nullptr;
    if (M > 0) CG_TRY(row_mask(C, t, kw, P.nf1, P.nf2, rm));
    GridArgs g = grid_args(C, P, t, M, a.cshift);
    g.kr = key_rect(a);
    // columns first (only those holding image pixels; frequencies outside
    // the image rows are zero), then the rows the degridder reads
    const int la = P.N1 - P.N1 / 2, lb = P.nf1 - P.N1 / 2;
    const int sa = P.N2 - P.N2 / 2, sb = P.nf2 - P.N2 / 2;
    // slabs of planes [zb, zb + zn) one after the other through the grid
    // buffer (one slab holding every plane without the out-of-core flow):
    // w-IDFT onto the slab, transform, and add the slab's taps to vis
    const int per = cp.slab ? cp.slab : P.nf3;
    C.last_slab = cp.slab;
    C.last_nslab = (P.nf3 + per - 1) / per;
    CG_TRY(clk.mark(ST_IMAGE));
    CG_TRY(launch_fseries_grid(w.qd, P.nf1, as<double>(C.phi1), C.stream));
    CG_TRY(launch_fseries_grid(w.qd, P.nf2, as<double>(C.phi2), C.stream));
    if (P.wstack) CG_TRY(launch_quarter_table(w, as<double>(C.zq), as<double>(C.phi3), C.stream));
    if (cp.slab) {
        // the whole loop is reported as the grid stage
        CG_TRY(clk.mark(ST_FFT2));
        CG_TRY(clk.mark(ST_FFT));
        CG_TRY(clk.mark(ST_GRID));
    }
    for (int zb = 0; zb < P.nf3; zb += per) {
        const int zn = std::min(per, P.nf3 - zb);
        if (!P.wstack) {
            // plain 2-D: the u/v deconvolution only (deconv_2d type 2)
            CG_TRY((launch_deconv2d<T, cx_t<T>>(as<cx_t<T>>(C.grid), const_cast<cx_t<T> *>(dirty),
                                                as<double>(C.phi1), as<double>(C.phi2), P.N1, P.N2,
                                                P.nf1, P.nf2, 2, C.stream)));
        } else {
            w.zb = zb;
            w.zn = cp.slab ? zn : 0;
            CG_TRY((launch_dirty2ms_prologue<T, cx_t<T>>(dirty, as<double>(C.zq), as<double>(C.phi3),
                                                          as<double>(C.phi1), as<double>(C.phi2),
                                                          as<cx_t<T>>(C.grid), w, C.stream)));
        }
        if (C.stop_after == 1) return P.ier;
        if (!cp.slab) CG_TRY(clk.mark(ST_FFT2));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), zn, P.nf2, P.nf1, -1, false,
                                  as<cx_t<T>>(C.tw2), la, lb, rm, sa, sb, C.stream, P.band_s,
                                  C.band_fft ? P.band_len : 0));
        if (!cp.slab) CG_TRY(clk.mark(ST_FFT));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), zn, P.nf2, P.nf1, -1, true,
                                  as<cx_t<T>>(C.tw1), la, lb, rm, sa, sb, C.stream));
        if (C.stop_after == 2) return P.ier;
        if (!cp.slab) CG_TRY(clk.mark(ST_GRID));
        if (cp.slab) {
            g.zb = zb;
            g.zn = zn;
            g.nchunk = (zn + t.pc - 1) / t.pc;
            g.accumulate = zb > 0;
        }
        // complex64: tiles with many points go to the MFMA degridder
        g.dense_min = 0;
        if (sizeof(T) == 4 && M > 0 && C.degrid_dense_min > 0 && degrid_dense_supported(g, kw))
            g.dense_min = C.degrid_dense_min;
        CG_TRY(ensure(C.dlist, sizeof(int32_t) * dense_list_ints(t.ntiles(), g.m)));
        if (M > 0)
            CG_TRY((launch_degrid<T, cx_t<T>>(as<geo_t<T>>(C.geo), as<cx_t<T>>(C.val),
                                              as<int32_t>(C.vals), as<int32_t>(C.tstart),
                                              as<cx_t<T>>(C.grid), vis, g, kw,
                                              as<int32_t>(C.dlist), C.stream)));
        if (g.dense_min > 0) {
            CG_TRY(clk.mark(ST_COUNT + 1));
            C.dense_timed = C.timing && !cp.slab;
            CG_TRY(launch_degrid_dense((const float4 *)C.geo.p, (const float2 *)C.val.p,
                                       as<int32_t>(C.vals), as<uint32_t>(C.keys),
                                       as<int32_t>(C.tstart), as<int32_t>(C.dlist),
                                       (const float2 *)C.grid.p, (float2 *)vis, g, kw, C.stream));
            CG_TRY(clk.mark(ST_COUNT + 2));
        }
    }
    CG_TRY(clk.mark(ST_COUNT));
    return P.ier;
}

// dirty2ms runs its stages in the order image, column FFT, row FFT, grid;
// each is still reported under its own slot
static int finish_dirty2ms_clock(Ctx &C)
{
    if (!C.timing) return CG_OK;
    CG_HIP(hipEventSynchronize(C.ev[ST_COUNT]));
    float ms[ST_COUNT] = {};
    CG_HIP(hipEventElapsedTime(&ms[ST_PLAN], C.ev[ST_PLAN], C.ev[ST_SORT]));
    CG_HIP(hipEventElapsedTime(&ms[ST_SORT], C.ev[ST_SORT], C.ev[ST_RECORDS]));
    CG_HIP(hipEventElapsedTime(&ms[ST_RECORDS], C.ev[ST_RECORDS], C.ev[ST_IMAGE]));
    CG_HIP(hipEventElapsedTime(&ms[ST_IMAGE], C.ev[ST_IMAGE], C.ev[ST_FFT2]));
    CG_HIP(hipEventElapsedTime(&ms[ST_FFT2], C.ev[ST_FFT2], C.ev[ST_FFT]));
    CG_HIP(hipEventElapsedTime(&ms[ST_FFT], C.ev[ST_FFT], C.ev[ST_GRID]));
    CG_HIP(hipEventElapsedTime(&ms[ST_GRID], C.ev[ST_GRID], C.ev[ST_COUNT]));
    for (int i = 0; i < ST_COUNT; i++) C.stage_ms[1][i] += ms[i];
    if (C.dense_timed) {
        float dm = 0.f;
        CG_HIP(hipEventElapsedTime(&dm, C.ev[ST_COUNT + 1], C.ev[ST_COUNT + 2]));
        C.degrid_dense_ms += dm;
        C.dense_timed = false;
    }
    C.stage_calls[1]++;
    return CG_OK;
}

template <typename T, typename TU, typename TW>
int dirty2ms_dev(Ctx &C, int64_t nrow, int nchan, const double *freq, int nx, int ny, double fov,
                 const TU *uvw, cx_t<T> *vis, const TW *wgt, const cx_t<T> *dirty, double eps,
                 double sigma, int sign, bool single)
{
    StageClock clk(C, 1);
    CG_TRY(clk.mark(ST_PLAN));
    CallPlan cp;
    const int ier = plan_dirty2ms<T, TU>(C, cp, nrow, nchan, freq, nx, ny, fov, uvw, eps, sigma,
                                         sign, single, ctx_tune(C));
    if (ier >= 2) return ier;
    CG_TRY((exec_dirty2ms<T, TU, TW>(C, cp, uvw, vis, wgt, dirty, clk)));
    if (C.stop_after == 0) CG_TRY(finish_dirty2ms_clock(C));
    return cp.ier;
}

// ------------------------------------------------------------ batches
//
// Many independent channels (each its own uvw, vis, image size and
// frequency, complex64 with FP64 uvw) run as one launch: every item is
// planned once (optionally autotuned), executed once eagerly so that the
// batch's private workspace reaches its final size, then the executions of
// all items are captured into one HIP graph that cg_batch_run replays.

// installs a tabulated kernel in the context; eps <= 0 removes the kernel
// of that support
static int set_learned(Ctx &C, const LearnedKernel &K, double eps)
{
    Ctx::Learned &L = C.learned[K.kw];
    if (!(eps > 0)) {
        L.eps = 0;
        return CG_OK;
    }
    L.K = K;
    L.eps = eps;
    return CG_OK;
}

// Items run on `lanes` independent branches, each with its own workspace
// (the graph holds its pointers) and stream, so small channels overlap;
// items are spread over the lanes by estimated cost.
struct Batch {
    Ctx C;                // planning and option template
    std::vector<std::unique_ptr<Ctx>> lane_ctx;
    std::vector<hipStream_t> lane_stream;
    std::vector<hipEvent_t> lane_done;
    hipEvent_t fork = nullptr;
    std::vector<int> lane;  // lane of each item
    int dir = 0;            // 0 ms2dirty, 1 dirty2ms
    std::vector<cg_batch_item> items;
    std::vector<CallPlan> plans;
    std::vector<Tune> tunes;
    std::vector<float> tuned_ms;
    double eps = 0, sigma = 2, tune_ms = 0;
    int sign = -1;
    hipStream_t user = nullptr, cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        for (hipStream_t st : lane_stream) (void)hipStreamDestroy(st);
        for (hipEvent_t e : lane_done) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
    }
};

static int batch_plan(Batch &B, size_t i, const Tune &tn, CallPlan &cp)
{
    const cg_batch_item &it = B.items[i];
    const double f = it.freq;
    if (B.dir == 0)
        return plan_ms2dirty<float, double>(B.C, cp, it.nrow, 1, &f, it.nxdirty, it.nydirty, it.fov,
                                            (const double *)it.d_uvw, B.eps, B.sigma, B.sign, true, tn);
    return plan_dirty2ms<float, double>(B.C, cp, it.nrow, 1, &f, it.nxdirty, it.nydirty, it.fov,
                                        (const double *)it.d_uvw, B.eps, B.sigma, B.sign, true, tn);
}

// item i on its lane's workspace and current stream
static int batch_exec(Batch &B, size_t i)
{
    const cg_batch_item &it = B.items[i];
    Ctx &L = *B.lane_ctx[B.lane[i]];
    StageClock clk(L, B.dir);
    if (B.dir == 0)
        return exec_ms2dirty<float, double, float>(L, B.plans[i], (const double *)it.d_uvw,
                                                   (const float2 *)it.d_vis, (const float *)it.d_wgt,
                                                   (float2 *)it.d_dirty, clk);
    return exec_dirty2ms<float, double, float>(L, B.plans[i], (const double *)it.d_uvw,
                                               (float2 *)it.d_vis, (const float *)it.d_wgt,
                                               (const float2 *)it.d_dirty, clk);
}

// every item, lanes forked from and joined back into stream s (captured
// into the graph as parallel branches when s is capturing)
static int batch_run_lanes(Batch &B, hipStream_t s)
{
    const size_t nl = B.lane_ctx.size();
    CG_HIP(hipEventRecord(B.fork, s));
    for (size_t k = 0; k < nl; k++) {
        CG_HIP(hipStreamWaitEvent(B.lane_stream[k], B.fork, 0));
        B.lane_ctx[k]->stream = B.lane_stream[k];
    }
    int r = CG_OK;
    for (size_t i = 0; i < B.items.size() && !r; i++) r = batch_exec(B, i);
    for (size_t k = 0; k < nl; k++) {
        if (hipEventRecord(B.lane_done[k], B.lane_stream[k]) != hipSuccess ||
            hipStreamWaitEvent(s, B.lane_done[k], 0) != hipSuccess)
            r = r ? r : CG_ERR_HIP;
    }
    return r;
}

// lanes by longest-first greedy assignment of a cost estimate (rows plus
// grid cells), each lane a copy of the template context's options
static int batch_lanes(Batch &B, int nlanes)
{
    const size_t n = B.items.size();
    nlanes = std::max(1, std::min<int>(nlanes, (int)n));
    for (int k = 0; k < nlanes; k++) {
        std::unique_ptr<Ctx> L(new Ctx());
        L->lds_budget = B.C.lds_budget;
        L->grid_lds_budget = B.C.grid_lds_budget;
        L->tile_grid = B.C.tile_grid;
        L->tile_degrid = B.C.tile_degrid;
        L->dense_min = B.C.dense_min;
        L->row_mask = B.C.row_mask;
        L->use_learned = B.C.use_learned;
        for (int kw = 0; kw <= LEARN_MAXKW; kw++) L->learned[kw] = B.C.learned[kw];
        L->stream = B.user;
        B.lane_ctx.push_back(std::move(L));
        hipStream_t st = nullptr;
        CG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        B.lane_stream.push_back(st);
        hipEvent_t e = nullptr;
        CG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B.lane_done.push_back(e);
    }
    CG_HIP(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
    std::vector<size_t> order(n);
    std::vector<double> cost(n), load(nlanes, 0.0);
    for (size_t i = 0; i < n; i++) {
        order[i] = i;
        cost[i] = (double)B.items[i].nrow + 0.25 * (double)B.plans[i].P.grid_cells();
    }
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    B.lane.assign(n, 0);
    for (size_t i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        B.lane[i] = k;
        load[k] += cost[i];
    }
    return CG_OK;
}

// time one execution of item i under tuning tn (after one untimed run)
static int batch_time(Batch &B, size_t i, const Tune &tn, float &ms)
{
    CallPlan cp;
    int ier = batch_plan(B, i, tn, cp);
    if (ier >= 2) return ier;
    std::swap(B.plans[i], cp);
    hipEvent_t e0, e1;
    CG_HIP(hipEventCreate(&e0));
    CG_HIP(hipEventCreate(&e1));
    const hipStream_t st = B.lane_ctx[B.lane[i]]->stream;
    int r = batch_exec(B, i);
    if (!r && hipEventRecord(e0, st) != hipSuccess) r = CG_ERR_HIP;
    if (!r) r = batch_exec(B, i);
    if (!r && hipEventRecord(e1, st) != hipSuccess) r = CG_ERR_HIP;
    if (!r && hipEventSynchronize(e1) != hipSuccess) r = CG_ERR_HIP;
    if (!r && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) r = CG_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::swap(B.plans[i], cp);
    return r;
}

// candidate tunings per direction; results are shared by items of the same
// image size and power-of-two visibility count
static int batch_autotune(Batch &B)
{
    std::vector<Tune> cands;
    const Tune d = ctx_tune(B.C);
    if (B.dir == 0) {
        for (int dm : {32, 64, 256}) cands.push_back(Tune{8, d.tile_degrid, d.grid_lds, d.lds, dm});
        cands.push_back(Tune{16, d.tile_degrid, d.grid_lds, d.lds, d.dense_min});
    } else {
        for (int td : {16, 32}) cands.push_back(Tune{d.tile_grid, td, d.grid_lds, d.lds, d.dense_min});
    }
    std::vector<std::pair<long long, size_t>> seen;  // shape key -> item tuned
    for (size_t i = 0; i < B.items.size(); i++) {
        const cg_batch_item &it = B.items[i];
        int lg = 0;
        while ((1LL << (lg + 1)) <= (long long)std::max(1, it.nrow)) lg++;
        const long long key = ((long long)it.nxdirty << 40) | ((long long)it.nydirty << 20) | lg;
        bool done = false;
        for (auto &kv : seen)
            if (kv.first == key) {
                B.tunes[i] = B.tunes[kv.second];
                B.tuned_ms[i] = B.tuned_ms[kv.second];
                done = true;
                break;
            }
        if (done) continue;
        float best = 1e30f;
        for (const Tune &tn : cands) {
            float ms = 0;
            CG_TRY(batch_time(B, i, tn, ms));
            if (ms < best) {
                best = ms;
                B.tunes[i] = tn;
            }
        }
        B.tuned_ms[i] = best;
        seen.push_back({key, i});
    }
    return CG_OK;
}

static int batch_build(Batch &B, int flags, int nlanes)
{
    const size_t n = B.items.size();
    B.plans.assign(n, CallPlan());
    B.tunes.assign(n, ctx_tune(B.C));
    B.tuned_ms.assign(n, 0.f);
    for (size_t i = 0; i < n; i++) CG_TRY(batch_plan(B, i, B.tunes[i], B.plans[i]));
    CG_TRY(batch_lanes(B, nlanes));
    if (flags & CG_BATCH_AUTOTUNE) CG_TRY(batch_autotune(B));
    int ier = 0;
    for (size_t i = 0; i < n; i++) {
        const int r = batch_plan(B, i, B.tunes[i], B.plans[i]);
        if (r >= 2) return r;
        ier = std::max(ier, r);
    }
    // one eager pass: every lane's workspace reaches its final size
    for (size_t i = 0; i < n; i++) CG_TRY(batch_exec(B, i));
    CG_HIP(hipStreamSynchronize(B.user));
    if (flags & CG_BATCH_GRAPH) {
        // capture on a private stream (the context's may be the null stream)
        CG_HIP(hipStreamCreateWithFlags(&B.cap, hipStreamNonBlocking));
        CG_HIP(hipStreamBeginCapture(B.cap, hipStreamCaptureModeThreadLocal));
        const int r = batch_run_lanes(B, B.cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(B.cap, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        CG_HIP(e);
        B.graph = g;
        CG_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    }
    return ier;
}

// ------------------------------------------------------------ 2-D NUFFT

template <typename T, typename TU>
int nufft2d_dev(Ctx &C, int type, int64_t M, const TU *x, const TU *y, cx_t<T> *c, int N1, int N2,
                cx_t<T> *fk, double eps, double sigma, bool single)
{
    if (M < 0 || M > (int64_t)INT32_MAX - 1 || N1 <= 0 || N2 <= 0) return CG_ERR_ARG;
    StageClock clk(C, type == 1 ? 0 : 1);
    CG_TRY(clk.mark(ST_PLAN));
    ConvOpts conv;
    int ier = setup_conv_opts(eps, sigma, single, conv);
    if (ier >= 2) return ier;
    WPlan P;
    P.N1 = N1;
    P.N2 = N2;
    P.nf1 = num_cells(N1, sigma, conv.kw);
    P.nf2 = num_cells(N2, sigma, conv.kw);
    P.nf3 = 1;
    P.conv = conv;
    P.ier = ier;
    if (!fft_supported<T>(P.nf1) || !fft_supported<T>(P.nf2)) return CG_ERR_SIZE;
    const int eb = (int)sizeof(cx_t<T>);
    const int kw = conv.kw;
    Tiling t = type == 1 ? choose_tiling(P.nf1, P.nf2, 1, kw, 16, C.grid_lds_budget, C.tile_grid2d, 0)
                         : choose_tiling(P.nf1, P.nf2, 1, kw, eb, C.lds_budget, C.tile_degrid2d, kw - 1);
    C.last_plan = P;
    C.last_tiling = t;
    CoordArgs a{};
    a.nf1 = P.nf1;
    a.nf2 = P.nf2;
    a.nf3 = 1;
    a.kw = kw;
    set_coord_tiling(a, t);
    // tile keys only: the 2-D gridders do not use the start-cell order, and
    // the degridder loses less to it (1.27 vs 1.26 ms at 2^25 clustered
    // points) than a fourth radix pass costs (0.17 ms)
    a.cshift = 0;
    C.last_cshift = a.cshift;
    a.pirange = 1;
    a.has_w = 0;
    const int64_t ntiles = t.ntiles();
    CG_TRY(clk.mark(ST_SORT));
    CG_TRY(ensure(C.tstart, sizeof(int32_t) * (ntiles + 1)));
    // complex64 with FP32 coordinates: the key pass also writes the caller-order
    // records into the grid buffer, which is idle until gridding / deconvolution
    float4 *rec = nullptr;
    if constexpr (sizeof(T) == 4 && std::is_same<TU, float>::value) {
        if (M > 0 && C.prerec) {
            CG_TRY(ensure(C.grid, std::max((size_t)P.nf1 * P.nf2 * eb,
                                           (size_t)(type == 1 ? 32 : 16) * M)));
            rec = as<float4>(C.grid);
        }
    }
    if (M > 0) {
        CG_TRY(ensure(C.keys, sizeof(uint32_t) * M));
        CG_TRY(ensure(C.keys2, sizeof(uint32_t) * M));
        CG_TRY(ensure(C.vals, sizeof(int32_t) * M));
        CG_TRY(ensure(C.vals2, sizeof(int32_t) * M));
        const int64_t hs = radix_hist_size(M);
        CG_TRY(ensure(C.hist, sizeof(int32_t) * hs));
        CG_TRY(ensure(C.scan, sizeof(int32_t) * scan_tmp_size(hs)));
        // odd pass counts start in the second buffers (no copy at the end)
        const int bits = sort_key_bits(a, ntiles);
        const bool odd = radix_passes(bits) & 1;
        uint32_t *k0 = as<uint32_t>(odd ? C.keys2 : C.keys), *k1 = as<uint32_t>(odd ? C.keys : C.keys2);
        int32_t *v0 = as<int32_t>(odd ? C.vals2 : C.vals), *v1 = as<int32_t>(odd ? C.vals : C.vals2);
        if constexpr (sizeof(T) == 4 && std::is_same<TU, float>::value) {
            if (rec)
                CG_TRY(launch_keys_prerecords_xy(x, y, type == 1 ? (const float2 *)c : nullptr, M, a,
                                                 k0, nullptr, rec, C.stream));
        }
        if (!rec) CG_TRY(launch_keys_xy<TU>(x, y, M, a, k0, nullptr, C.stream));
        CG_TRY(launch_radix_sort(k0, v0, k1, v1, M, bits, as<int32_t>(C.hist), as<int32_t>(C.scan),
                                 C.stream, 0, 1));
        CG_TRY(launch_tile_bounds(as<uint32_t>(C.keys), M, ntiles, a.cshift, key_rect(a),
                                  as<int32_t>(C.tstart), C.stream));
    } else {
        CG_TRY(launch_fill_i32(as<int32_t>(C.tstart), ntiles + 1, 0, C.stream));
    }
    CG_TRY(clk.mark(ST_RECORDS));
    CG_TRY(ensure(C.geo, sizeof(geo_t<T>) * M));
    CG_TRY(ensure(C.val, sizeof(cx_t<T>) * M));
    if (rec)
        CG_TRY(launch_gather_records_xy(rec, as<int32_t>(C.vals), M, type == 1 ? 1 : 0,
                                        (float4 *)C.geo.p, (float2 *)C.val.p, C.stream));
    else
        CG_TRY((launch_records_xy<T, TU, cx_t<T>>(x, y, type == 1 ? c : (const cx_t<T> *)nullptr,
                                                  as<int32_t>(C.vals), M, a, as<geo_t<T>>(C.geo),
                                                  as<cx_t<T>>(C.val), C.stream)));
    CG_TRY(ensure(C.grid, (size_t)P.nf1 * P.nf2 * eb));
    CG_TRY(ensure(C.phi1, sizeof(double) * (P.nf1 / 2 + 1)));
    CG_TRY(ensure(C.phi2, sizeof(double) * (P.nf2 / 2 + 1)));
    Quad qd = make_quad(conv);
    CG_TRY(launch_fseries_grid(qd, P.nf1, as<double>(C.phi1), C.stream));
    CG_TRY(launch_fseries_grid(qd, P.nf2, as<double>(C.phi2), C.stream));
    CG_TRY(twiddles<T>(C, P.nf1, P.nf2));
    GridArgs g = grid_args(C, P, t, M, a.cshift);
    g.kr = key_rect(a);
    g.mfma2d = C.mfma2d;
    // fast-axis columns that hold modes; grid rows that points reach; mode
    // frequencies along nf2
    const int la = N1 - N1 / 2, lb = P.nf1 - N1 / 2;
    const int sa = N2 - N2 / 2, sb = P.nf2 - N2 / 2;
    const uint8_t *rm = nullptr;
    CG_TRY(row_mask(C, t, kw, P.nf1, P.nf2, rm));
    if (type == 1) {
        CG_TRY(grid_masks<T>(C, g, rm, P.nf1, P.nf2));
        CG_TRY(clk.mark(ST_GRID));
        CG_TRY(launch_grid<T>(as<geo_t<T>>(C.geo), as<cx_t<T>>(C.val), as<int32_t>(C.tstart),
                              as<cx_t<T>>(C.grid), g, kw, C.stream));
        CG_TRY(clk.mark(ST_FFT));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), 1, P.nf2, P.nf1, +1, true,
                                  as<cx_t<T>>(C.tw1), la, lb, rm, sa, sb, C.stream));
        CG_TRY(clk.mark(ST_FFT2));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), 1, P.nf2, P.nf1, +1, false,
                                  as<cx_t<T>>(C.tw2), la, lb, rm, sa, sb, C.stream));
        CG_TRY(clk.mark(ST_IMAGE));
        CG_TRY((launch_deconv2d<T, cx_t<T>>(as<cx_t<T>>(C.grid), fk, as<double>(C.phi1),
                                            as<double>(C.phi2), N1, N2, P.nf1, P.nf2, 1, C.stream)));
    } else {
        CG_TRY(clk.mark(ST_GRID));
        CG_TRY((launch_deconv2d<T, cx_t<T>>(as<cx_t<T>>(C.grid), fk, as<double>(C.phi1),
                                            as<double>(C.phi2), N1, N2, P.nf1, P.nf2, 2, C.stream)));
        CG_TRY(clk.mark(ST_FFT));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), 1, P.nf2, P.nf1, -1, false,
                                  as<cx_t<T>>(C.tw2), la, lb, rm, sa, sb, C.stream));
        CG_TRY(clk.mark(ST_FFT2));
        CG_TRY(launch_fft_pass<T>(as<cx_t<T>>(C.grid), 1, P.nf2, P.nf1, -1, true,
                                  as<cx_t<T>>(C.tw1), la, lb, rm, sa, sb, C.stream));
        CG_TRY(clk.mark(ST_IMAGE));
        if (M > 0)
            CG_TRY((launch_degrid<T, cx_t<T>>(as<geo_t<T>>(C.geo), as<cx_t<T>>(C.val),
                                              as<int32_t>(C.vals), as<int32_t>(C.tstart),
                                              as<cx_t<T>>(C.grid), c, g, kw, nullptr, C.stream)));
    }
    CG_TRY(clk.mark(ST_COUNT));
    CG_TRY(clk.finish());
    return ier;
}

// ------------------------------------------------------------ host wrappers

struct HostBuf {
    void *p = nullptr;
    ~HostBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes)
    {
        if (bytes == 0) bytes = 256;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return CG_ERR_ALLOC;
        }
        return CG_OK;
    }
};

// host-pointer ms2dirty / dirty2ms: H2D, run, D2H, free (reference semantics);
// vis and wgt are [nrow][nchan] for nchan frequencies freq[nchan]
template <typename T, typename TU, typename TW>
static int host_ms2dirty(int nrow, int nchan, const double *freq, int nx, int ny, double fov,
                         const TU *uvw, const void *vis, const TW *wgt, void *dirty, double eps,
                         double sigma, int sign, bool single, const PlanGeom &geom = PlanGeom())
{
    if (nrow < 0 || nchan < 1 || !freq || nx <= 0 || ny <= 0 || !uvw || !dirty || (nrow > 0 && !vis))
        return CG_ERR_ARG;
    try {
        Ctx C;
        env_overrides(C);
        C.geom = geom;
        const size_t R = (size_t)nrow, M = R * (size_t)nchan;
        HostBuf du, dv, dw, dd;
        CG_TRY(du.alloc(sizeof(TU) * 3 * R));
        CG_TRY(dv.alloc(sizeof(cx_t<T>) * M));
        if (wgt) CG_TRY(dw.alloc(sizeof(TW) * M));
        CG_TRY(dd.alloc(sizeof(cx_t<T>) * (size_t)nx * ny));
        if (M) {
            CG_HIP(hipMemcpy(du.p, uvw, sizeof(TU) * 3 * R, hipMemcpyHostToDevice));
            CG_HIP(hipMemcpy(dv.p, vis, sizeof(cx_t<T>) * M, hipMemcpyHostToDevice));
            if (wgt) CG_HIP(hipMemcpy(dw.p, wgt, sizeof(TW) * M, hipMemcpyHostToDevice));
        }
        int ier = ms2dirty_dev<T, TU, TW>(C, nrow, nchan, freq, nx, ny, fov, (const TU *)du.p,
                                          (const cx_t<T> *)dv.p, wgt ? (const TW *)dw.p : nullptr,
                                          (cx_t<T> *)dd.p, eps, sigma, sign, single);
        if (ier >= 2) return ier;
        CG_HIP(hipMemcpy(dirty, dd.p, sizeof(cx_t<T>) * (size_t)nx * ny, hipMemcpyDeviceToHost));
        return ier;
    } catch (const std::bad_alloc &) {
        return CG_ERR_ALLOC;
    }
}

template <typename T, typename TU, typename TW>
static int host_dirty2ms(int nrow, int nchan, const double *freq, int nx, int ny, double fov,
                         const TU *uvw, void *vis, const TW *wgt, const void *dirty, double eps,
                         double sigma, int sign, bool single, const PlanGeom &geom = PlanGeom())
{
    if (nrow < 0 || nchan < 1 || !freq || nx <= 0 || ny <= 0 || !uvw || !dirty || (nrow > 0 && !vis))
        return CG_ERR_ARG;
    try {
        Ctx C;
        env_overrides(C);
        C.geom = geom;
        const size_t R = (size_t)nrow, M = R * (size_t)nchan;
        HostBuf du, dv, dw, dd;
        CG_TRY(du.alloc(sizeof(TU) * 3 * R));
        CG_TRY(dv.alloc(sizeof(cx_t<T>) * M));
        if (wgt) CG_TRY(dw.alloc(sizeof(TW) * M));
        CG_TRY(dd.alloc(sizeof(cx_t<T>) * (size_t)nx * ny));
        if (M) {
            CG_HIP(hipMemcpy(du.p, uvw, sizeof(TU) * 3 * R, hipMemcpyHostToDevice));
            if (wgt) CG_HIP(hipMemcpy(dw.p, wgt, sizeof(TW) * M, hipMemcpyHostToDevice));
        }
        CG_HIP(hipMemcpy(dd.p, dirty, sizeof(cx_t<T>) * (size_t)nx * ny, hipMemcpyHostToDevice));
        int ier = dirty2ms_dev<T, TU, TW>(C, nrow, nchan, freq, nx, ny, fov, (const TU *)du.p,
                                          (cx_t<T> *)dv.p, wgt ? (const TW *)dw.p : nullptr,
                                          (const cx_t<T> *)dd.p, eps, sigma, sign, single);
        if (ier >= 2) return ier;
        if (M) CG_HIP(hipMemcpy(vis, dv.p, sizeof(cx_t<T>) * M, hipMemcpyDeviceToHost));
        return ier;
    } catch (const std::bad_alloc &) {
        return CG_ERR_ALLOC;
    }
}

template <typename T, typename TU>
static int host_nufft2d(int type, int M, const TU *x, const TU *y, void *c, int N1, int N2,
                        void *fk, double eps, double sigma, bool single)
{
    if (M < 0 || N1 <= 0 || N2 <= 0 || !fk || (M > 0 && (!x || !y || !c))) return CG_ERR_ARG;
    try {
        Ctx C;
        env_overrides(C);
        HostBuf dx, dy, dc, dk;
        CG_TRY(dx.alloc(sizeof(TU) * (size_t)M));
        CG_TRY(dy.alloc(sizeof(TU) * (size_t)M));
        CG_TRY(dc.alloc(sizeof(cx_t<T>) * (size_t)M));
        CG_TRY(dk.alloc(sizeof(cx_t<T>) * (size_t)N1 * N2));
        if (M) {
            CG_HIP(hipMemcpy(dx.p, x, sizeof(TU) * (size_t)M, hipMemcpyHostToDevice));
            CG_HIP(hipMemcpy(dy.p, y, sizeof(TU) * (size_t)M, hipMemcpyHostToDevice));
            if (type == 1) CG_HIP(hipMemcpy(dc.p, c, sizeof(cx_t<T>) * (size_t)M, hipMemcpyHostToDevice));
        }
        if (type == 2)
            CG_HIP(hipMemcpy(dk.p, fk, sizeof(cx_t<T>) * (size_t)N1 * N2, hipMemcpyHostToDevice));
        int ier = nufft2d_dev<T, TU>(C, type, M, (const TU *)dx.p, (const TU *)dy.p,
                                     (cx_t<T> *)dc.p, N1, N2, (cx_t<T> *)dk.p, eps, sigma, single);
        if (ier >= 2) return ier;
        if (type == 1)
            CG_HIP(hipMemcpy(fk, dk.p, sizeof(cx_t<T>) * (size_t)N1 * N2, hipMemcpyDeviceToHost));
        else if (M)
            CG_HIP(hipMemcpy(c, dc.p, sizeof(cx_t<T>) * (size_t)M, hipMemcpyDeviceToHost));
        return ier;
    } catch (const std::bad_alloc &) {
        return CG_ERR_ALLOC;
    }
}

// ---- allocation-free 2-D NUFFT
//
// The same pipeline as cg_nufft2d_dev, with every workspace buffer carved
// from `scratch`, a fixed device array of this library (one copy per GPU, held
// until the process ends), so the call makes no hipMalloc / hipFree and no
// other allocating call: usable under stream capture or where allocation is
// not allowed.  The host side keeps one context per device whose buffers point
// into that device's copy of the array.  Calls that overlap on one GPU (other
// threads, other streams) race on the array; calls are serialised on the host
// per device, not on the GPU.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoAllocDevice {
    std::mutex m;
    void *ptr = nullptr;  // device address of this GPU's copy of `scratch`
    Ctx *ctx = nullptr;
};
constexpr int kMaxDevices = 64;
static NoAllocDevice g_noalloc[kMaxDevices];

// An upper bound of the bytes nufft2d_dev carves for one call: every buffer
// it may ensure, each rounded to the arena alignment.
static size_t nufft2d_workspace(int type, int prec, int64_t M, int N1, int N2, double eps,
                                double sigma, int grid_lds, int lds, int tile2d, int tile_degrid)
{
    ConvOpts conv;
    if (setup_conv_opts(eps, sigma, prec == 0, conv) >= 2) return 0;
    const int nf1 = num_cells(N1, sigma, conv.kw), nf2 = num_cells(N2, sigma, conv.kw);
    const size_t eb = prec == 0 ? 8 : 16, gb = prec == 0 ? 16 : 32;
    const int kw = conv.kw;
    const Tiling t = type == 1 ? choose_tiling(nf1, nf2, 1, kw, 16, grid_lds, tile2d, 0)
                               : choose_tiling(nf1, nf2, 1, kw, (int)eb, lds, tile_degrid, kw - 1);
    const int64_t ntiles = t.ntiles();
    const int64_t hs = radix_hist_size(std::max<int64_t>(M, 1));
    size_t b = 0;
    b += arena_round(4 * (size_t)(ntiles + 1));                                       // tstart
    b += arena_round(std::max((size_t)nf1 * nf2 * eb, (size_t)32 * M));              // grid / records
    b += 4 * arena_round(4 * (size_t)std::max<int64_t>(M, 64));                       // keys, vals
    b += arena_round(4 * (size_t)hs) + arena_round(4 * (size_t)scan_tmp_size(hs));   // hist, scan
    b += arena_round(gb * (size_t)std::max<int64_t>(M, 16)) + arena_round(eb * (size_t)std::max<int64_t>(M, 32));
    b += arena_round(8 * (size_t)(nf1 / 2 + 1)) + arena_round(8 * (size_t)(nf2 / 2 + 1));  // phi
    b += arena_round(eb * (size_t)nf1) + arena_round(eb * (size_t)nf2);               // twiddles
    b += arena_round((size_t)nf1 + nf2) + arena_round(4 * (size_t)(ntiles + 1));     // masks, tile list
    return b;
}

}  // namespace cg

using namespace cg;

// ====================================================================== C-ABI

extern "C" {

// reference entry points (include/cugridder.h:20-46), double precision
int ms2dirty_2(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw,
               double *vis, double *wgt, double *dirty, double epsilon, double sigma, int sign)
{
    const double fq = freq;
    return host_ms2dirty<double, double, double>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, wgt,
                                                 dirty, epsilon, sigma, sign, false);
}
int ms2dirty_1(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw,
               double *vis, double *dirty, double epsilon, double sigma, int sign)
{
    const double fq = freq;
    return host_ms2dirty<double, double, double>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis,
                                                 nullptr, dirty, epsilon, sigma, sign, false);
}
int dirty2ms_2(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw,
               double *vis, double *wgt, double *dirty, double epsilon, double sigma, int sign)
{
    const double fq = freq;
    return host_dirty2ms<double, double, double>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, wgt,
                                                 dirty, epsilon, sigma, sign, false);
}
int dirty2ms_1(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw,
               double *vis, double *dirty, double epsilon, double sigma, int sign)
{
    const double fq = freq;
    return host_dirty2ms<double, double, double>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis,
                                                 nullptr, dirty, epsilon, sigma, sign, false);
}
// single precision (float ABI, FP32 uvw)
int ms2dirtyf_2(int nrow, int nxdirty, int nydirty, float fov, float freq, float *uvw, float *vis,
                float *wgt, float *dirty, float epsilon, float sigma, int sign)
{
    const double fq = freq;
    return host_ms2dirty<float, float, float>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, wgt,
                                              dirty, epsilon, sigma, sign, true);
}
int ms2dirtyf_1(int nrow, int nxdirty, int nydirty, float fov, float freq, float *uvw, float *vis,
                float *dirty, float epsilon, float sigma, int sign)
{
    const double fq = freq;
    return host_ms2dirty<float, float, float>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, nullptr,
                                              dirty, epsilon, sigma, sign, true);
}
int dirty2msf_2(int nrow, int nxdirty, int nydirty, float fov, float freq, float *uvw, float *vis,
                float *wgt, float *dirty, float epsilon, float sigma, int sign)
{
    const double fq = freq;
    return host_dirty2ms<float, float, float>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, wgt,
                                              dirty, epsilon, sigma, sign, true);
}
int dirty2msf_1(int nrow, int nxdirty, int nydirty, float fov, float freq, float *uvw, float *vis,
                float *dirty, float epsilon, float sigma, int sign)
{
    const double fq = freq;
    return host_dirty2ms<float, float, float>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, nullptr,
                                              dirty, epsilon, sigma, sign, true);
}

// mixed precision: FP64 uvw, complex64 vis/dirty, FP32 wgt (may be NULL)
int ms2dirty_mixed(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *uvw,
                   const float *vis, const float *wgt, float *dirty, double epsilon, double sigma,
                   int sign)
{
    const double fq = freq;
    return host_ms2dirty<float, double, float>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, wgt,
                                               dirty, epsilon, sigma, sign, true);
}
int dirty2ms_mixed(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *uvw,
                   float *vis, const float *wgt, const float *dirty, double epsilon, double sigma,
                   int sign)
{
    const double fq = freq;
    return host_dirty2ms<float, double, float>(nrow, 1, &fq, nxdirty, nydirty, fov, uvw, vis, wgt,
                                               dirty, epsilon, sigma, sign, true);
}

// several channels: freq[nchan] (host), vis and wgt [nrow][nchan]; the
// dirty image sums all channels (no counterpart in the reference C-ABI,
// whose binding passes freq[0] only, python/curagridder/cursl.py:103)
int ms2dirty_mc_2(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                  const double *uvw, const double *vis, const double *wgt, double *dirty,
                  double epsilon, double sigma, int sign)
{
    return host_ms2dirty<double, double, double>(nrow, nchan, freq, nxdirty, nydirty, fov, uvw, vis,
                                                 wgt, dirty, epsilon, sigma, sign, false);
}
int dirty2ms_mc_2(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                  const double *uvw, double *vis, const double *wgt, const double *dirty,
                  double epsilon, double sigma, int sign)
{
    return host_dirty2ms<double, double, double>(nrow, nchan, freq, nxdirty, nydirty, fov, uvw, vis,
                                                 wgt, dirty, epsilon, sigma, sign, false);
}
int ms2dirty_mc_mixed(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                      const double *uvw, const float *vis, const float *wgt, float *dirty,
                      double epsilon, double sigma, int sign)
{
    return host_ms2dirty<float, double, float>(nrow, nchan, freq, nxdirty, nydirty, fov, uvw, vis,
                                               wgt, dirty, epsilon, sigma, sign, true);
}
int dirty2ms_mc_mixed(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                      const double *uvw, float *vis, const float *wgt, const float *dirty,
                      double epsilon, double sigma, int sign)
{
    return host_dirty2ms<float, double, float>(nrow, nchan, freq, nxdirty, nydirty, fov, uvw, vis,
                                               wgt, dirty, epsilon, sigma, sign, true);
}

// 2-D NUFFT type 1 / 2 (points in [-pi,pi)), host pointers
int nufft2d1(int M, const double *x, const double *y, const double *c, int N1, int N2, double *fk,
             double epsilon, double sigma)
{
    return host_nufft2d<double, double>(1, M, x, y, (void *)c, N1, N2, fk, epsilon, sigma, false);
}
int nufft2d2(int M, const double *x, const double *y, double *c, int N1, int N2, const double *fk,
             double epsilon, double sigma)
{
    return host_nufft2d<double, double>(2, M, x, y, c, N1, N2, (void *)fk, epsilon, sigma, false);
}
int nufft2d1f(int M, const float *x, const float *y, const float *c, int N1, int N2, float *fk,
              float epsilon, float sigma)
{
    return host_nufft2d<float, float>(1, M, x, y, (void *)c, N1, N2, fk, epsilon, sigma, true);
}
int nufft2d2f(int M, const float *x, const float *y, float *c, int N1, int N2, const float *fk,
              float epsilon, float sigma)
{
    return host_nufft2d<float, float>(2, M, x, y, c, N1, N2, (void *)fk, epsilon, sigma, true);
}

// ---- contexts and device-pointer entry points

int cg_ctx_create(void **ctx, void *stream)
{
    if (!ctx) return CG_ERR_ARG;
    try {
        Ctx *C = new Ctx();
        C->stream = (hipStream_t)stream;
        env_overrides(*C);
        *ctx = C;
        return CG_OK;
    } catch (const std::bad_alloc &) {
        return CG_ERR_ALLOC;
    }
}

int cg_ctx_destroy(void *ctx)
{
    delete (Ctx *)ctx;
    return CG_OK;
}

int cg_ctx_set_option(void *ctx, const char *name, int value)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !name) return CG_ERR_ARG;
    if (!std::strcmp(name, "timing")) C->timing = value != 0;
    else if (!std::strcmp(name, "lds_budget")) C->lds_budget = value;
    else if (!std::strcmp(name, "grid_lds_budget")) C->grid_lds_budget = value;
    else if (!std::strcmp(name, "tile_grid")) C->tile_grid = value;
    else if (!std::strcmp(name, "tile_grid2d")) C->tile_grid2d = value;
    else if (!std::strcmp(name, "prerec")) C->prerec = value;
    else if (!std::strcmp(name, "mfma2d")) C->mfma2d = value;
    else if (!std::strcmp(name, "band_fft")) C->band_fft = value;
    else if (!std::strcmp(name, "compact_keys")) C->compact_keys = value;
    else if (!std::strcmp(name, "tile_degrid")) C->tile_degrid = value;
    else if (!std::strcmp(name, "tile_degrid2d")) C->tile_degrid2d = value;
    else if (!std::strcmp(name, "stop_after")) C->stop_after = value;
    else if (!std::strcmp(name, "dense_min")) C->dense_min = value;
    else if (!std::strcmp(name, "degrid_dense_min")) C->degrid_dense_min = value;
    else if (!std::strcmp(name, "tap_poly")) C->tap_poly = value;
    else if (!std::strcmp(name, "row_mask")) C->row_mask = value;
    else if (!std::strcmp(name, "use_learned")) C->use_learned = value;
    else if (!std::strcmp(name, "grid_limit_mb")) C->grid_limit_mb = value;
    else return CG_ERR_ARG;
    return CG_OK;
}

// stage times summed over calls: out[dir*7 + stage] (ms), out[14] the dense
// gridding kernel's part of the ms2dirty grid stage; returns calls
int cg_ctx_stage_times(void *ctx, double *out, long long *calls)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !out) return CG_ERR_ARG;
    for (int d = 0; d < 2; d++)
        for (int i = 0; i < ST_COUNT; i++) out[d * ST_COUNT + i] = C->stage_ms[d][i];
    out[2 * ST_COUNT] = C->dense_ms;
    out[2 * ST_COUNT + 1] = C->degrid_dense_ms;
    if (calls) {
        calls[0] = C->stage_calls[0];
        calls[1] = C->stage_calls[1];
    }
    return CG_OK;
}

int cg_ctx_reset_times(void *ctx)
{
    Ctx *C = (Ctx *)ctx;
    if (!C) return CG_ERR_ARG;
    std::memset(C->stage_ms, 0, sizeof(C->stage_ms));
    C->dense_ms = C->degrid_dense_ms = 0;
    C->stage_calls[0] = C->stage_calls[1] = 0;
    return CG_OK;
}

// last plan: out[0..15] = nf1 nf2 nf3 kw beta es_c pix_x pix_y i_center
// i_half_width o_center o_half_width gamma h tx ty
int cg_ctx_last_plan(void *ctx, double *out)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !out) return CG_ERR_ARG;
    const WPlan &P = C->last_plan;
    double v[16] = {(double)P.nf1, (double)P.nf2, (double)P.nf3, (double)P.conv.kw, P.conv.beta,
                    P.conv.es_c, P.pix_x, P.pix_y, P.i_center, P.i_half_width, P.o_center,
                    P.o_half_width, P.gamma, P.h, (double)C->last_tiling.tx,
                    (double)C->last_tiling.ty};
    std::memcpy(out, v, sizeof(v));
    return CG_OK;
}

// out[0] = w-planes per slab of the last call's grid buffer (0: the whole
// grid at once), out[1] = slabs
int cg_ctx_last_slab(void *ctx, int *out)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !out) return CG_ERR_ARG;
    out[0] = C->last_slab;
    out[1] = C->last_nslab;
    return CG_OK;
}

int cg_ctx_last_band(void *ctx, int *out)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !out) return CG_ERR_ARG;
    out[0] = C->last_plan.band_s;
    out[1] = C->band_fft ? C->last_plan.band_len : 0;
    return CG_OK;
}

// prec: 0 = complex64 vis/image with FP32 gridding and FFT, 1 = complex128;
// uvw_prec / wgt_prec: 0 float, 1 double.  d_wgt may be NULL.  freq is a
// host array of nchan frequencies; d_vis and d_wgt are [nrow][nchan].
int cg_ms2dirty_mc_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty, int nydirty,
                       double fov, const double *freq, const void *d_uvw, int uvw_prec,
                       const void *d_vis, const void *d_wgt, int wgt_prec, void *d_dirty,
                       double epsilon, double sigma, int sign)
{
    Ctx *C = (Ctx *)ctx;
    if (!C) return CG_ERR_ARG;
    bool single = prec == 0;
    if (prec == 0 && uvw_prec == 1 && wgt_prec == 0)
        return ms2dirty_dev<float, double, float>(*C, nrow, nchan, freq, nxdirty, nydirty, fov,
                                                  (const double *)d_uvw, (const float2 *)d_vis,
                                                  (const float *)d_wgt, (float2 *)d_dirty, epsilon,
                                                  sigma, sign, single);
    if (prec == 0 && uvw_prec == 0 && wgt_prec == 0)
        return ms2dirty_dev<float, float, float>(*C, nrow, nchan, freq, nxdirty, nydirty, fov,
                                                 (const float *)d_uvw, (const float2 *)d_vis,
                                                 (const float *)d_wgt, (float2 *)d_dirty, epsilon,
                                                 sigma, sign, single);
    if (prec == 1 && uvw_prec == 1 && wgt_prec == 1)
        return ms2dirty_dev<double, double, double>(*C, nrow, nchan, freq, nxdirty, nydirty, fov,
                                                    (const double *)d_uvw, (const double2 *)d_vis,
                                                    (const double *)d_wgt, (double2 *)d_dirty,
                                                    epsilon, sigma, sign, single);
    return CG_ERR_ARG;
}

int cg_dirty2ms_mc_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty, int nydirty,
                       double fov, const double *freq, const void *d_uvw, int uvw_prec, void *d_vis,
                       const void *d_wgt, int wgt_prec, const void *d_dirty, double epsilon,
                       double sigma, int sign)
{
    Ctx *C = (Ctx *)ctx;
    if (!C) return CG_ERR_ARG;
    bool single = prec == 0;
    if (prec == 0 && uvw_prec == 1 && wgt_prec == 0)
        return dirty2ms_dev<float, double, float>(*C, nrow, nchan, freq, nxdirty, nydirty, fov,
                                                  (const double *)d_uvw, (float2 *)d_vis,
                                                  (const float *)d_wgt, (const float2 *)d_dirty,
                                                  epsilon, sigma, sign, single);
    if (prec == 0 && uvw_prec == 0 && wgt_prec == 0)
        return dirty2ms_dev<float, float, float>(*C, nrow, nchan, freq, nxdirty, nydirty, fov,
                                                 (const float *)d_uvw, (float2 *)d_vis,
                                                 (const float *)d_wgt, (const float2 *)d_dirty,
                                                 epsilon, sigma, sign, single);
    if (prec == 1 && uvw_prec == 1 && wgt_prec == 1)
        return dirty2ms_dev<double, double, double>(*C, nrow, nchan, freq, nxdirty, nydirty, fov,
                                                    (const double *)d_uvw, (double2 *)d_vis,
                                                    (const double *)d_wgt, (const double2 *)d_dirty,
                                                    epsilon, sigma, sign, single);
    return CG_ERR_ARG;
}

int cg_ms2dirty_dev(void *ctx, int prec, int nrow, int nxdirty, int nydirty, double fov,
                    double freq, const void *d_uvw, int uvw_prec, const void *d_vis,
                    const void *d_wgt, int wgt_prec, void *d_dirty, double epsilon, double sigma,
                    int sign)
{
    return cg_ms2dirty_mc_dev(ctx, prec, nrow, 1, nxdirty, nydirty, fov, &freq, d_uvw, uvw_prec,
                              d_vis, d_wgt, wgt_prec, d_dirty, epsilon, sigma, sign);
}

int cg_dirty2ms_dev(void *ctx, int prec, int nrow, int nxdirty, int nydirty, double fov,
                    double freq, const void *d_uvw, int uvw_prec, void *d_vis, const void *d_wgt,
                    int wgt_prec, const void *d_dirty, double epsilon, double sigma, int sign)
{
    return cg_dirty2ms_mc_dev(ctx, prec, nrow, 1, nxdirty, nydirty, fov, &freq, d_uvw, uvw_prec,
                              d_vis, d_wgt, wgt_prec, d_dirty, epsilon, sigma, sign);
}

// NIFTY geometry: fov from pix_x as the reference's binding derives it
// (python/curagridder/cursl.py:99), pix_y honoured, plain 2-D on request
static int px_geom(int nx, double pix_x, double pix_y, int do_wstacking, double &fov, PlanGeom &g)
{
    if (!(pix_x > 0) || !(pix_y > 0) || nx <= 0) return CG_ERR_ARG;
    fov = pix_x * nx * 180.0 / kPi;
    g.pix_y = pix_y;
    g.wstack = do_wstacking ? 1 : 0;
    return CG_OK;
}

struct GeomScope {
    Ctx &C;
    GeomScope(Ctx &c, const PlanGeom &g) : C(c) { C.geom = g; }
    ~GeomScope() { C.geom = PlanGeom(); }
};

int cg_ms2dirty_px_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty, int nydirty,
                       double pix_x, double pix_y, const double *freq, const void *d_uvw,
                       int uvw_prec, const void *d_vis, const void *d_wgt, int wgt_prec,
                       void *d_dirty, double epsilon, double sigma, int sign, int do_wstacking)
{
    Ctx *C = (Ctx *)ctx;
    if (!C) return CG_ERR_ARG;
    double fov = 0;
    PlanGeom g;
    CG_TRY(px_geom(nxdirty, pix_x, pix_y, do_wstacking, fov, g));
    GeomScope scope(*C, g);
    return cg_ms2dirty_mc_dev(ctx, prec, nrow, nchan, nxdirty, nydirty, fov, freq, d_uvw,
                              uvw_prec, d_vis, d_wgt, wgt_prec, d_dirty, epsilon, sigma, sign);
}

int cg_dirty2ms_px_dev(void *ctx, int prec, int nrow, int nchan, int nxdirty, int nydirty,
                       double pix_x, double pix_y, const double *freq, const void *d_uvw,
                       int uvw_prec, void *d_vis, const void *d_wgt, int wgt_prec,
                       const void *d_dirty, double epsilon, double sigma, int sign,
                       int do_wstacking)
{
    Ctx *C = (Ctx *)ctx;
    if (!C) return CG_ERR_ARG;
    double fov = 0;
    PlanGeom g;
    CG_TRY(px_geom(nxdirty, pix_x, pix_y, do_wstacking, fov, g));
    GeomScope scope(*C, g);
    return cg_dirty2ms_mc_dev(ctx, prec, nrow, nchan, nxdirty, nydirty, fov, freq, d_uvw,
                              uvw_prec, d_vis, d_wgt, wgt_prec, d_dirty, epsilon, sigma, sign);
}

int ms2dirty_px(int prec, int nrow, int nchan, int nxdirty, int nydirty, double pix_x,
                double pix_y, const double *freq, const void *uvw, const void *vis,
                const void *wgt, void *dirty, double epsilon, double sigma, int sign,
                int do_wstacking)
{
    double fov = 0;
    PlanGeom g;
    CG_TRY(px_geom(nxdirty, pix_x, pix_y, do_wstacking, fov, g));
    if (prec == 0)
        return host_ms2dirty<float, double, float>(nrow, nchan, freq, nxdirty, nydirty, fov,
                                                   (const double *)uvw, vis, (const float *)wgt,
                                                   dirty, epsilon, sigma, sign, true, g);
    if (prec == 1)
        return host_ms2dirty<double, double, double>(nrow, nchan, freq, nxdirty, nydirty, fov,
                                                     (const double *)uvw, vis, (const double *)wgt,
                                                     dirty, epsilon, sigma, sign, false, g);
    if (prec == 2)
        return host_ms2dirty<float, float, float>(nrow, nchan, freq, nxdirty, nydirty, fov,
                                                  (const float *)uvw, vis, (const float *)wgt,
                                                  dirty, epsilon, sigma, sign, true, g);
    return CG_ERR_ARG;
}

int dirty2ms_px(int prec, int nrow, int nchan, int nxdirty, int nydirty, double pix_x,
                double pix_y, const double *freq, const void *uvw, void *vis, const void *wgt,
                const void *dirty, double epsilon, double sigma, int sign, int do_wstacking)
{
    double fov = 0;
    PlanGeom g;
    CG_TRY(px_geom(nxdirty, pix_x, pix_y, do_wstacking, fov, g));
    if (prec == 0)
        return host_dirty2ms<float, double, float>(nrow, nchan, freq, nxdirty, nydirty, fov,
                                                   (const double *)uvw, vis, (const float *)wgt,
                                                   dirty, epsilon, sigma, sign, true, g);
    if (prec == 1)
        return host_dirty2ms<double, double, double>(nrow, nchan, freq, nxdirty, nydirty, fov,
                                                     (const double *)uvw, vis, (const double *)wgt,
                                                     dirty, epsilon, sigma, sign, false, g);
    if (prec == 2)
        return host_dirty2ms<float, float, float>(nrow, nchan, freq, nxdirty, nydirty, fov,
                                                  (const float *)uvw, vis, (const float *)wgt,
                                                  dirty, epsilon, sigma, sign, true, g);
    return CG_ERR_ARG;
}

int cg_batch_create(void **batch, void *ctx, int direction, int n, const cg_batch_item *items,
                    double epsilon, double sigma, int sign, int flags)
{
    Ctx *C = (Ctx *)ctx;
    if (!batch || !C || (direction != 0 && direction != 1) || n <= 0 || !items) return CG_ERR_ARG;
    *batch = nullptr;
    for (int i = 0; i < n; i++) {
        const cg_batch_item &it = items[i];
        if (it.nrow < 0 || it.nxdirty <= 0 || it.nydirty <= 0 || !it.d_dirty ||
            (it.nrow > 0 && (!it.d_uvw || !it.d_vis)) || !(it.freq > 0) || !(it.fov > 0))
            return CG_ERR_ARG;
    }
    try {
        Batch *B = new Batch();
        B->C.lds_budget = C->lds_budget;
        B->C.grid_lds_budget = C->grid_lds_budget;
        B->C.tile_grid = C->tile_grid;
        B->C.tile_degrid = C->tile_degrid;
        B->C.dense_min = C->dense_min;
        B->C.row_mask = C->row_mask;
        B->C.stream = C->stream;
        B->C.use_learned = C->use_learned;
        for (int kw = 0; kw <= LEARN_MAXKW; kw++)
            if (C->learned[kw].eps > 0) {
                const int r = set_learned(B->C, C->learned[kw].K, C->learned[kw].eps);
                if (r) {
                    delete B;
                    return r;
                }
            }
        B->user = C->stream;
        B->dir = direction;
        B->items.assign(items, items + n);
        B->eps = epsilon;
        B->sigma = sigma;
        B->sign = sign;
        int lanes = 4;
        if (const char *e = std::getenv("CG_BATCH_LANES")) lanes = std::max(1, std::atoi(e));
        const int ier = batch_build(*B, flags, lanes);
        if (ier >= 2) {
            delete B;
            return ier;
        }
        *batch = B;
        return ier;
    } catch (const std::bad_alloc &) {
        return CG_ERR_ALLOC;
    }
}

int cg_batch_run(void *batch)
{
    Batch *B = (Batch *)batch;
    if (!B) return CG_ERR_ARG;
    if (B->exec) {
        CG_HIP(hipGraphLaunch(B->exec, B->user));
        return CG_OK;
    }
    return batch_run_lanes(*B, B->user);
}

int cg_batch_info(void *batch, double *out)
{
    Batch *B = (Batch *)batch;
    if (!B || !out) return CG_ERR_ARG;
    for (size_t i = 0; i < B->items.size(); i++) {
        const Tune &t = B->tunes[i];
        out[5 * i + 0] = t.tile_grid;
        out[5 * i + 1] = t.tile_degrid;
        out[5 * i + 2] = B->plans[i].dense_min;
        out[5 * i + 3] = B->tuned_ms[i];
        out[5 * i + 4] = B->plans[i].P.nf3;
    }
    return B->exec ? 1 : 0;
}

int cg_batch_destroy(void *batch)
{
    delete (Batch *)batch;
    return CG_OK;
}

int cg_nufft2d_dev(void *ctx, int type, int prec, int M, const void *d_x, const void *d_y,
                   void *d_c, int N1, int N2, void *d_fk, double epsilon, double sigma)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || (type != 1 && type != 2)) return CG_ERR_ARG;
    if (prec == 0)
        return nufft2d_dev<float, float>(*C, type, M, (const float *)d_x, (const float *)d_y,
                                         (float2 *)d_c, N1, N2, (float2 *)d_fk, epsilon, sigma,
                                         true);
    if (prec == 1)
        return nufft2d_dev<double, double>(*C, type, M, (const double *)d_x, (const double *)d_y,
                                           (double2 *)d_c, N1, N2, (double2 *)d_fk, epsilon, sigma,
                                           false);
    return CG_ERR_ARG;
}

// cg_nufft2d_dev on the fixed workspace `scratch` (see above)
int cg_nufft2d_noalloc(void *stream, int type, int prec, int M, const void *d_x,
                       const void *d_y, void *d_c, int N1, int N2, void *d_fk, double epsilon,
                       double sigma)
{
    if ((type != 1 && type != 2) || (prec != 0 && prec != 1) || M < 0 || N1 <= 0 || N2 <= 0)
        return CG_ERR_ARG;
    int dev = 0;
    CG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return CG_ERR_ARG;
    NoAllocDevice &D = g_noalloc[dev];
    std::lock_guard<std::mutex> lock(D.m);
    if (!D.ptr) {
        // once per device, with the device current: the address of this
        // GPU's copy of the array (the host-side name is not a device address)
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            report_hip_error(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
            (void)hipGetLastError();
            return CG_ERR_HIP;
        }
        D.ptr = p;
    }
    if (!D.ctx) {
        Ctx *C = new (std::nothrow) Ctx();
        if (!C) return CG_ERR_ALLOC;
        env_overrides(*C);
        C->use_arena(D.ptr, kScratchBytes);
        D.ctx = C;
    }
    Ctx &C = *D.ctx;
    const size_t need = nufft2d_workspace(type, prec, M, N1, N2, epsilon, sigma, C.grid_lds_budget,
                                          C.lds_budget, C.tile_grid2d, C.tile_degrid2d);
    if (need == 0) return CG_ERR_ARG;
    if (need > kScratchBytes) {
        if (!std::getenv("CG_QUIET"))
            std::fprintf(stderr,
                         "curafft: cg_nufft2d_noalloc needs %zu bytes of workspace, more than its "
                         "fixed %zu (use cg_nufft2d_dev)\n",
                         need, kScratchBytes);
        return CG_ERR_SIZE;
    }
    C.stream = (hipStream_t)stream;
    C.reset_arena();
    return cg_nufft2d_dev(&C, type, prec, M, d_x, d_y, d_c, N1, N2, d_fk, epsilon, sigma);
}

// ---- host-only planner query (no device calls): out[16] = nf1 nf2 nf3 kw
// beta es_c pix_x pix_y i_center i_half_width o_center o_half_width gamma h
// f_over_c ier, for a given range of the scaled w (w * 2 pi f/c * sign)
int cg_plan_query(int nrow, int nxdirty, int nydirty, double fov, double freq, double wmin,
                  double wmax, double epsilon, double sigma, int sign, int single, double *out)
{
    if (!out) return CG_ERR_ARG;
    WPlan P;
    int ier = make_wplan(P, nrow, nxdirty, nydirty, fov, freq, epsilon, sigma, sign, single != 0,
                         wmin, wmax);
    double v[16] = {(double)P.nf1, (double)P.nf2, (double)P.nf3, (double)P.conv.kw, P.conv.beta,
                    P.conv.es_c, P.pix_x, P.pix_y, P.i_center, P.i_half_width, P.o_center,
                    P.o_half_width, P.gamma, P.h, P.f_over_c, (double)ier};
    std::memcpy(out, v, sizeof(v));
    return ier;
}

// ---- test hooks: device pointers into the context workspace (valid until
// the next call on the context)
int cg_ctx_grid(void *ctx, void **ptr, long long *cells)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !ptr || !cells) return CG_ERR_ARG;
    *ptr = C->grid.p;
    *cells = (long long)C->last_plan.nf1 * C->last_plan.nf2 * C->last_plan.nf3;
    return CG_OK;
}

// sorted keys (uint32), permutation (int32) and tile starts (int32) of the
// last ms2dirty/dirty2ms call; n[0] = rows, n[1] = tiles
int cg_ctx_sorted(void *ctx, void **keys, void **perm, void **tstart, long long *n)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !keys || !perm || !tstart || !n) return CG_ERR_ARG;
    *keys = C->keys.p;
    *perm = C->vals.p;
    *tstart = C->tstart.p;
    n[0] = C->last_m;
    n[1] = C->last_tiling.ntiles();
    n[2] = C->last_cshift;
    return CG_OK;
}

// in-place 2-D FFT of batch planes [batch][n2][n1] on the context stream
int cg_fft2d_dev(void *ctx, int prec, void *d_data, int batch, int n2, int n1, int dir)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || !d_data || batch < 0 || n1 <= 0 || n2 <= 0 || (dir != 1 && dir != -1))
        return CG_ERR_ARG;
    if (prec == 0) {
        if (!fft_supported<float>(n1) || !fft_supported<float>(n2)) return CG_ERR_SIZE;
        CG_TRY(twiddles<float>(*C, n1, n2));
        CG_TRY(ensure(C->fftmp, sizeof(float2) * (size_t)batch * n2 * n1));
        return launch_fft2d<float>((float2 *)d_data, batch, n2, n1, dir, as<float2>(C->tw1),
                                   as<float2>(C->tw2), as<float2>(C->fftmp), C->stream);
    }
    if (prec == 1) {
        if (!fft_supported<double>(n1) || !fft_supported<double>(n2)) return CG_ERR_SIZE;
        CG_TRY(twiddles<double>(*C, n1, n2));
        CG_TRY(ensure(C->fftmp, sizeof(double2) * (size_t)batch * n2 * n1));
        return launch_fft2d<double>((double2 *)d_data, batch, n2, n1, dir, as<double2>(C->tw1),
                                    as<double2>(C->tw2), as<double2>(C->fftmp), C->stream);
    }
    return CG_ERR_ARG;
}

// stable sort of (key, value) pairs by the low key_bits bits, in place
int cg_sort_pairs_dev(void *ctx, void *d_keys, void *d_vals, int m, int key_bits)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || m < 0 || key_bits < 1 || key_bits > 32) return CG_ERR_ARG;
    if (m == 0) return CG_OK;
    CG_TRY(ensure(C->keys2, sizeof(uint32_t) * m));
    CG_TRY(ensure(C->vals2, sizeof(int32_t) * m));
    const int64_t hs = radix_hist_size(m);
    CG_TRY(ensure(C->hist, sizeof(int32_t) * hs));
    CG_TRY(ensure(C->scan, sizeof(int32_t) * scan_tmp_size(hs)));
    return launch_radix_sort((uint32_t *)d_keys, (int32_t *)d_vals, as<uint32_t>(C->keys2),
                             as<int32_t>(C->vals2), m, key_bits, as<int32_t>(C->hist),
                             as<int32_t>(C->scan), C->stream);
}

// device-to-device copy on a stream (test hook)
int cg_memcpy_dev(void *dst, const void *src, void *stream, size_t bytes)
{
    if (!bytes) return CG_OK;
    CG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return CG_OK;
}

// the per-tap polynomial coefficients the complex64 kernels use (host only)
int cg_tap_poly(int kw, double beta, double c, const float *kq, int ndeg, int deg, float *out)
{
    if (kw < 1 || kw > 16 || deg < 0 || deg >= kTapStride || ndeg < 0 || ndeg > LEARN_MAXDEG || !out)
        return CG_ERR_ARG;
    tap_polynomials(kw, beta, c, kq, ndeg, deg, out);
    return CG_OK;
}

int cg_fft_col_split(int n, int *na, int *nb)
{
    if (n < 1 || !na || !nb) return CG_ERR_ARG;
    if (n > 1 && fft_col_split(n, *na, *nb)) return 1;
    *na = *nb = 0;
    return 0;
}

int cg_version(void) { return 102; }

// ---- learned kernel
int cg_learn_nparams(void) { return LEARN_NPARAMS; }

int cg_learn_grad(int kw, const float *params, const double *u, const double *theta, int nsamp,
                  float *grad, double *loss)
{
    return learn_grad(kw, params, u, theta, nsamp, grad, loss, nullptr);
}

int cg_learn_train(int kw, int mode, const double *a, const double *b, int n, int batch,
                   int steps, double lr, float *params, double *losses)
{
    return learn_train(kw, mode, a, b, n, batch, steps, lr, params, losses, nullptr);
}

int cg_learn_eval(const float *params, const float *s, int n, float *y)
{
    return learn_eval(params, s, n, y, nullptr);
}

int cg_learn_tabulate(int kw, int deg, const float *params, float *coef, double *qpsi)
{
    if (!coef || !qpsi) return CG_ERR_ARG;
    LearnedKernel K;
    CG_TRY(learn_kernel_build(kw, deg, params, K, nullptr));
    std::memcpy(coef, K.coef, sizeof(float) * (deg + 1));
    std::memcpy(qpsi, K.qpsi, sizeof(double) * K.nq);
    return CG_OK;
}

int cg_ctx_set_kernel(void *ctx, int kw, int deg, const float *params, double eps)
{
    Ctx *C = (Ctx *)ctx;
    if (!C || kw < 2 || kw > LEARN_MAXKW) return CG_ERR_ARG;
    if (!(eps > 0)) {
        C->learned[kw].eps = 0;
        return CG_OK;
    }
    if (!params) return CG_ERR_ARG;
    LearnedKernel K;
    CG_TRY(learn_kernel_build(kw, deg, params, K, C->stream));
    return set_learned(*C, K, eps);
}

}  // extern "C"
