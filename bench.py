"""Throughput of improved W-stacking ms2dirty + dirty2ms on one MI355X.

One step = ms2dirty of the workload's visibilities followed by dirty2ms of a
fixed image, on inputs already resident in HBM, through libcurafft.so's
device-pointer entry points (the same kernels as the reference-compatible
host entry points).  Prints one JSON line (see --help for the workloads).

The default workload is the full W-stacking configuration (8192^2 image,
fov 3 deg, 1.4 GHz, 2^27 visibilities from 256-antenna Earth-rotation tracks,
eps 1e-6).  `--workload plain` is the one-channel 1024^2 case (131072 rows,
eps 1e-5) and `--workload nufft2d` the 2-D type-1 + type-2 NUFFT (2^25
points, 4096^2 modes, eps 1e-5).  `--workload batch` runs 64 channels of
mixed sizes as two graph-captured batches, and `--workload learned` the full
W-stacking configuration at eps 1e-5 with a learned support-6 kernel that is
trained on the GPU and certified against the ES support-7 kernel before the
timed steps (the ES kernel is used when certification fails).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C_LIGHT = 299792458.0
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
# points per bin from which the complex64 gridder / degridder use their MFMA
# kernels (the library's defaults, set explicitly on the timed context)
DENSE_MIN = 64
DEGRID_DENSE_MIN = 2048
METRIC = "Mvis/s ms2dirty+dirty2ms at eps 1e-5, plain vs batched/learned; HBM GB/s"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- workloads

def make_plain(torch, dev):
    """131072 rows, u,v ~ N(0, 1500 m), w ~ U(-300, 300) m, freq 1e8, fov 2
    deg, 1024^2, vis complex64 N(0,1), wgt U(0.5, 1.5) float32, eps 1e-5."""
    rng = np.random.default_rng(1)
    M = 131072
    uvw = np.empty((M, 3))
    uvw[:, 0] = rng.normal(0, 1500, M)
    uvw[:, 1] = rng.normal(0, 1500, M)
    uvw[:, 2] = rng.uniform(-300, 300, M)
    vis = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    wgt = rng.uniform(0.5, 1.5, M).astype(np.float32)
    img = rng.standard_normal((1024, 1024)).astype(np.complex64)
    t = lambda a: torch.tensor(a, device=dev)  # noqa: E731
    return dict(name="plain one-channel 131072 rows 1024^2 fov 2deg 1e8 Hz eps 1e-5",
                uvw=t(uvw), vis=t(vis), wgt=t(wgt), img=t(img), nx=1024, ny=1024, fov=2.0,
                freq=1e8, eps=1e-5, M=M)


def make_wstack_full(torch, dev, nrow=1 << 27, npix=8192, seed=3):
    """256 antennas uniform in a 4 km disk (latitude 45 deg), Earth-rotation
    tracks over 8 h at 10 s steps toward dec 45 deg, both baseline
    orientations (65280 x 2880 = 188 M samples), 2^27 drawn without
    replacement; positions scaled so max |w| = 5000 wavelengths at 1.4 GHz
    (the raw 8 km baselines reach ~37000).  vis complex64 N(0,1), weights 1."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    f64 = torch.float64
    nant = 256
    r = 4000.0 * torch.sqrt(torch.rand(nant, generator=g, device=dev, dtype=f64))
    th = 2 * np.pi * torch.rand(nant, generator=g, device=dev, dtype=f64)
    E, N = r * torch.cos(th), r * torch.sin(th)
    lat, dec = np.deg2rad(45.0), np.deg2rad(45.0)
    X, Y, Z = -np.sin(lat) * N, E, np.cos(lat) * N
    ii, jj = torch.triu_indices(nant, nant, 1, device=dev)
    bx = torch.cat([X[jj] - X[ii], X[ii] - X[jj]])
    by = torch.cat([Y[jj] - Y[ii], Y[ii] - Y[jj]])
    bz = torch.cat([Z[jj] - Z[ii], Z[ii] - Z[jj]])
    H = torch.linspace(-np.pi / 3, np.pi / 3, 2880, device=dev, dtype=f64)
    sH, cH = torch.sin(H)[None, :], torch.cos(H)[None, :]
    bx, by, bz = bx[:, None], by[:, None], bz[:, None]
    u = (sH * bx + cH * by).reshape(-1)
    v = (-np.sin(dec) * cH * bx + np.sin(dec) * sH * by + np.cos(dec) * bz).reshape(-1)
    w = (np.cos(dec) * cH * bx - np.cos(dec) * sH * by + np.sin(dec) * bz).reshape(-1)
    del bx, by, bz
    sel = torch.randperm(u.numel(), generator=g, device=dev)[:nrow]
    uvw = torch.stack([u[sel], v[sel], w[sel]], dim=1).contiguous()
    del u, v, w, sel
    freq = 1.4e9
    lam = C_LIGHT / freq
    uvw *= 5000.0 * lam / uvw[:, 2].abs().max()
    vis = torch.complex(torch.randn(nrow, generator=g, device=dev),
                        torch.randn(nrow, generator=g, device=dev))
    img = torch.complex(torch.randn(npix, npix, generator=g, device=dev),
                        torch.zeros(npix, npix, device=dev))
    return dict(name="full W-stacking %d^2 fov 3deg 1.4 GHz %d rows eps 1e-6" % (npix, nrow),
                uvw=uvw, vis=vis, wgt=None, img=img, nx=npix, ny=npix, fov=3.0, freq=freq,
                eps=1e-6, M=nrow)


def make_batch(torch, dev, nchan=64, seed=4):
    """64 independent channels: visibility counts floor(2^U), U ~ U(18, 23);
    uvw of the full W-stacking tracks (metres) observed at frequencies
    1.0 .. 1.6 GHz; images drawn from {512, 1024, 2048, 4096}^2 at the full
    configuration's pixel size (3 deg / 8192); complex64, eps 1e-5."""
    base = make_wstack_full(torch, dev, nrow=1 << 23, npix=64, seed=seed)
    rng = np.random.default_rng(seed)
    counts = np.floor(2.0 ** rng.uniform(18, 23, nchan)).astype(np.int64)
    freqs = np.linspace(1.0e9, 1.6e9, nchan)
    sizes = rng.choice([512, 1024, 2048, 4096], nchan)
    pix_deg = 3.0 / 8192
    items = []
    for c in range(nchan):
        m = int(counts[c])
        start = int(rng.integers(0, (1 << 23) - m + 1))
        n = int(sizes[c])
        g = torch.Generator(device=dev)
        g.manual_seed(seed * 1000 + c)
        vis = torch.complex(torch.randn(m, generator=g, device=dev),
                            torch.randn(m, generator=g, device=dev))
        img = torch.complex(torch.randn(n, n, generator=g, device=dev), torch.zeros(n, n, device=dev))
        items.append(dict(uvw=base["uvw"][start:start + m], vis=vis, wgt=None, img=img,
                          dirty=torch.empty(n, n, dtype=torch.complex64, device=dev),
                          out=torch.empty(m, dtype=torch.complex64, device=dev),
                          fov=pix_deg * n, freq=float(freqs[c])))
    del base
    return dict(name="batch of %d channels, 2^18..2^23 rows, images 512^2..4096^2, 1.0-1.6 GHz, "
                     "eps 1e-5" % nchan, items=items, eps=1e-5, M=int(counts.sum()))


def make_nufft2d(torch, dev, M=1 << 25, N=4096, seed=2):
    """2^25 points in [-pi,pi)^2: 80% from 16 Gaussians (means uniform,
    std 0.3, wrapped) + 20% uniform; strengths complex64 N(0,1); 4096^2."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    nb = int(0.8 * M)
    means = (torch.rand(16, 2, generator=g, device=dev) * 2 - 1) * np.pi
    comp = torch.randint(0, 16, (nb,), generator=g, device=dev)
    pts = means[comp] + 0.3 * torch.randn(nb, 2, generator=g, device=dev)
    unif = (torch.rand(M - nb, 2, generator=g, device=dev) * 2 - 1) * np.pi
    pts = torch.cat([pts, unif])
    pts = torch.remainder(pts + np.pi, 2 * np.pi) - np.pi
    c = torch.complex(torch.randn(M, generator=g, device=dev), torch.randn(M, generator=g, device=dev))
    fk = torch.complex(torch.randn(N, N, generator=g, device=dev), torch.randn(N, N, generator=g, device=dev))
    return dict(name="2-D NUFFT type 1 + type 2, 2^25 clustered points, 4096^2 modes, eps 1e-5",
                x=pts[:, 0].contiguous(), y=pts[:, 1].contiguous(), c=c, fk=fk, N=N, eps=1e-5, M=M)


# ---------------------------------------------------------------- roofline

def setup_learned(torch, ctx, W, args, kw=6):
    """Trains the support-kw kernel on the GPU (2^20 seeded visibilities of a
    256^2 patch), certifies it at the workload's eps against the ES certifier
    (support 7) and installs it in ctx when accepted; also times one step of
    the ES support-7 path (eps 1e-6, the certifier's accuracy) for comparison.
    Untimed set-up: returns what the JSON line reports about it."""
    from cugridder_amd import learned
    W["eps"] = 1e-5
    W["name"] = W["name"].replace("eps 1e-6", "eps 1e-5, learned kernel")
    t0 = time.perf_counter()
    params, losses = learned.train(kw)
    t_train = time.perf_counter() - t0
    log("trained kw %d in %.1f s, final loss %.3e" % (kw, t_train, float(np.mean(losses[-100:]))))
    cert = learned.certify(kw, params, W["eps"])
    log("certification: %s" % json.dumps(cert))
    if cert["accepted"]:
        ctx.set_kernel(kw, params, W["eps"])
    out_d = torch.empty(W["nx"], W["ny"], dtype=torch.complex64, device=W["vis"].device)
    out_v = torch.empty(W["M"], dtype=torch.complex64, device=W["vis"].device)
    from cugridder_amd import Context
    es = Context()

    def es_step(eps):
        es.ms2dirty(W["uvw"], W["vis"], W["wgt"], W["nx"], W["ny"], W["fov"], W["freq"], eps,
                    out=out_d)
        es.dirty2ms(W["uvw"], W["img"], W["wgt"], W["fov"], W["freq"], eps, out=out_v)

    def es_time(eps, reps=3):
        es_step(eps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            es_step(eps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps, es.last_plan()["kw"]

    # the certifier's accuracy (eps 1e-6, support 7) and the ES kernel at the
    # workload's own eps (1e-5: support 6, the learned kernel's support)
    es_ms, es_kw = es_time(1e-6)
    es5_ms, es5_kw = es_time(W["eps"])
    es.close()
    log("ES support %d at eps 1e-6: %.1f ms per step; ES support %d at eps %g: %.1f ms per step"
        % (es_kw, es_ms, es5_kw, W["eps"], es5_ms))
    return {"kw": kw, "accepted": cert["accepted"], "certified_rel_l2": [cert["ms2dirty"],
            cert["dirty2ms"]], "certifier": "ES support 7, eps 1e-6", "train_s": round(t_train, 2),
            "es_certifier_kw": es_kw, "es_certifier_ms_per_step": round(es_ms, 2),
            "es_certifier_Mvis_s": round(W["M"] / (es_ms * 1e-3) / 1e6, 3),
            "es_same_eps_kw": es5_kw, "es_same_eps_ms_per_step": round(es5_ms, 2),
            "es_same_eps_Mvis_s": round(W["M"] / (es5_ms * 1e-3) / 1e6, 3)}


def grid_stats(ctx, plan, kw, dense_min):
    """Per-call geometry the algorithmic byte counts need, from the last
    call's tile starts: grid rows / columns that points reach (the FFT's
    masks), gridding tiles that are written, points in dense bins."""
    ts = ctx.tile_starts().cpu().numpy().astype(np.int64)
    nf1, nf2 = plan["nf1"], plan["nf2"]
    tx, ty = plan["tile_x"], plan["tile_y"]
    ntx, nty = nf1 // tx, nf2 // ty
    cnt = np.diff(ts)[: ntx * nty].reshape(nty, ntx)
    nz = cnt > 0
    rows = np.zeros(nf2, bool)
    cols = np.zeros(nf1, bool)
    for by in np.nonzero(nz.any(1))[0]:
        rows[(by * ty + np.arange(ty + kw - 1)) % nf2] = True
    for bx in np.nonzero(nz.any(0))[0]:
        cols[(bx * tx + np.arange(tx + kw - 1)) % nf1] = True
    row_t = rows.reshape(nty, ty).any(1)
    col_t = cols.reshape(ntx, tx).any(1)
    is_dense = (cnt >= dense_min) if (dense_min > 0 and ctx.key_shift > 0) else np.zeros_like(nz)
    band_s, band_len = ctx.last_band()
    return {"rows": int(rows.sum()), "cols": int(cols.sum()),
            "tiles_written": int(row_t.sum() * col_t.sum()), "tile": tx * ty,
            "nonempty_tiles": int(nz.sum()), "dense_points": int(cnt[is_dense].sum()),
            "dense_tiles": int(is_dense.sum()), "band": band_len}


def band_kernel(nf2, band):
    """True when the column passes of a complex64 call with this row band run
    as the one-pass band kernel (M = 1024 or 2048 rows, nf2 >= 2 M)."""
    if band <= 0 or nf2 & (nf2 - 1):
        return False
    m = 1024 if band <= 1024 else 2048
    return band <= m and nf2 >= 2 * m


def kernel_table(W, plan, fw, bw, stage_ms):
    """(kernel, direction, launches per call, ms per call, algorithmic bytes
    per launch) of every timed kernel of one ms2dirty and one dirty2ms call.

    Algorithmic bytes follow the per-unit figures of the HBM model: 36 B per
    visibility (FP64 uvw 24 + complex64 vis 8 + FP32 weight 4) for the stage
    that consumes the visibilities (gridding / degridding), 8 B per grid cell
    written or read once, 8 B per cell per FFT pass read and written, and
    8 B per live pixel per plane for the w-(I)DFT plus 8 B per image pixel.
    Cells are counted only where the masks keep them (the grid rows Rm and
    columns Cm points reach, the N1 live columns, the N2 live frequencies of
    nf2, the tiles the gridder writes); sort, record and w-range passes are
    not credited.  P planes, complex64 (8 B) cells."""
    M = W["M"]
    nf1, nf2, P = plan["nf1"], plan["nf2"], plan["nf3"]
    N1, N2 = W["ny"], W["nx"]
    Md = fw["dense_points"]
    Rm, Cm = fw["rows"], fw["cols"]
    sm, sb = stage_ms["ms2dirty"], stage_ms["dirty2ms"]
    col_kernel = "k_fft_cols_step_r2" if (nf2 > 128 and nf2 & (nf2 - 1) == 0) else "k_fft_cols_step"
    cells_w = P * fw["tiles_written"] * fw["tile"]
    rows = []
    # the sparse kernel writes every kept cell once; the dense kernel adds
    # the dense bins' points onto those cells
    rows.append(("k_grid", "ms2dirty", 1, sm["grid"] - sm["grid_dense"], 36 * (M - Md) + 8 * cells_w))
    if Md:
        rows.append(("k_grid_dense", "ms2dirty", 1, sm["grid_dense"], 36 * Md))
    rows.append(("k_fft_rows_ct", "ms2dirty", 1, sm["fft_rows"], 8 * P * Rm * (Cm + N1)))
    if band_kernel(nf2, fw["band"]):
        # one pass: the masked rows of the band in, the live spectrum out
        rows.append(("k_fft_cols_band", "ms2dirty", 1, sm["fft_cols"], 8 * P * Rm * N1 + 8 * P * N2 * N1))
    else:
        rows.append((col_kernel, "ms2dirty", 2, sm["fft_cols"],
                     (8 * P * (Rm * N1 + nf2 * N1) + 8 * P * (nf2 * N1 + N2 * N1)) // 2))
    rows.append(("k_ms2dirty_epi", "ms2dirty", 1, sm["image"], 8 * P * N1 * N2 + 8 * N1 * N2))
    Rb, Cb = bw["rows"], bw["cols"]
    rows.append(("k_dirty2ms_pro", "dirty2ms", 1, sb["image"], 8 * N1 * N2 + 8 * P * N2 * N1))
    if band_kernel(nf2, bw["band"]):
        # one pass: the live spectrum in, the band's rows out
        rows.append(("k_fft_cols_band", "dirty2ms", 1, sb["fft_cols"], 8 * P * N2 * N1 + 8 * P * bw["band"] * N1))
    else:
        rows.append((col_kernel, "dirty2ms", 2, sb["fft_cols"],
                     (8 * P * (N2 * N1 + nf2 * N1) + 8 * P * (nf2 * N1 + Rb * N1)) // 2))
    rows.append(("k_fft_rows_ct", "dirty2ms", 1, sb["fft_rows"], 8 * P * Rb * (N1 + Cb)))
    # the degridders read each point once and each cell of their tiles once
    Mb, Tb = bw["dense_points"], bw["dense_tiles"]
    rows.append(("k_degrid", "dirty2ms", 1, sb["grid"] - sb["grid_dense"],
                 36 * (M - Mb) + 8 * P * max(bw["tiles_written"] - Tb, 0) * bw["tile"]))
    if Mb:
        rows.append(("k_degrid_dense", "dirty2ms", 1, sb["grid_dense"], 36 * Mb + 8 * P * Tb * bw["tile"]))
    return rows


def hbm_model(W, kernels, ms_per_step):
    """Whole-step algorithmic bytes (the kernel table summed over launches),
    the fraction of HBM peak the step reaches on them, and the Mvis/s a step
    moving exactly those bytes at peak would reach (the mask-aware ceiling)."""
    total = sum(r[2] * r[4] for r in kernels)
    return {"step_algorithmic_GB": round(total / 1e9, 3),
            "step_frac_of_hbm_peak": round(total / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "ceiling_Mvis_s_at_peak": round(W["M"] / (total / (HBM_PEAK_GBS * 1e9)) / 1e6, 1)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(W, workload, runs=5):
    """FP64 OpenMP restatement of the reference algorithm (the oracle) on the
    host cores: the median of `runs` timed ms2dirty + dirty2ms pairs after
    one untimed pair, on a bounded sample of the workload."""
    from oracle import oracle as O
    threads = O.num_threads()

    def timed(uvw, freq, vis, wgt, img, nx, ny, fov, eps):
        ts = []
        for k in range(runs + 1):
            t0 = time.perf_counter()
            O.ms2dirty(uvw, freq, vis, wgt, nx, ny, fov, eps)
            O.dirty2ms(uvw, freq, img, wgt, fov, eps)
            if k:
                ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), ts

    if workload == "plain":
        uvw = W["uvw"].cpu().numpy()
        vis = W["vis"].cpu().numpy().astype(np.complex128)
        wgt = W["wgt"].cpu().numpy().astype(np.float64)
        img = W["img"].cpu().numpy().astype(np.complex128)
        sample = "the whole workload (131072 rows, 1024^2, eps 1e-5)"
        nx, ny, fov, eps = W["nx"], W["ny"], W["fov"], W["eps"]
    elif workload in ("wstack", "learned"):
        # 2^20 rows of the same tracks, 1024^2 pixels of the same size
        # (fov 3/8 deg), same eps: the full 16384^2 x 22-plane FP64 grid does
        # not fit a host
        M = 1 << 20
        uvw = W["uvw"][:M].cpu().numpy()
        vis = W["vis"][:M].cpu().numpy().astype(np.complex128)
        wgt = None
        nx = ny = 1024
        fov, eps = W["fov"] / 8.0, W["eps"]
        img = W["img"][:nx, :ny].cpu().numpy().astype(np.complex128)
        sample = ("2^20 rows of the same tracks, 1024^2 pixels of the same size (fov 0.375 deg), "
                  "eps %g" % eps)
    else:
        return None
    t, ts = timed(uvw, W["freq"], vis, wgt, img, nx, ny, fov, eps)
    mvis = uvw.shape[0] / t / 1e6
    log("cpu baseline: median %.2f s per ms2dirty + dirty2ms over %d runs (%s) on %d threads"
        % (t, runs, ", ".join("%.2f" % x for x in ts), threads))
    return {"value": round(mvis, 4), "unit": "Mvis/s", "cores": threads, "kind": "port",
            "sample": sample + "; FP64 C/OpenMP restatement of the reference algorithm; median of "
                      "%d runs after 1 warm-up" % runs,
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "runs_s": [round(x, 3) for x in ts]}


def dft_accuracy(torch, W, d_img, v_out, npix_sample=64, nrow_sample=64, seed=0):
    """Relative L2 error of the last ms2dirty image on sampled pixels and of
    the last dirty2ms visibilities on sampled rows against an FP64 direct
    sum evaluated on the GPU (untimed).  dirty(l, m) = sum_j vis_j wgt_j
    exp(2 pi i f/c (u l + v m + sign w (n - 1))) / n with sign -1, and its
    adjoint."""
    dev = d_img.device
    f64, c128 = torch.float64, torch.complex128
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    nx, ny = W["nx"], W["ny"]
    foc = 2 * np.pi * W["freq"] / C_LIGHT
    pix_x = W["fov"] * np.pi / 180 / nx
    pix_y = W["fov"] * np.pi / 180 / ny
    uvw = W["uvw"].to(f64)
    M = uvw.shape[0]
    wgt = None if W["wgt"] is None else W["wgt"].to(f64)

    def lmn(flat):
        r, c = flat // ny, flat % ny
        l_ = (r - nx // 2).to(f64) * pix_x
        m_ = (c - ny // 2).to(f64) * pix_y
        r2 = l_ * l_ + m_ * m_
        return l_, m_, -r2 / (torch.sqrt(1 - r2) + 1)

    pix = torch.randperm(nx * ny, generator=g)[:npix_sample].to(dev)
    l_, m_, nm1 = lmn(pix)
    acc = torch.zeros(npix_sample, dtype=c128, device=dev)
    step = 1 << 22
    for j0 in range(0, M, step):
        u = uvw[j0:j0 + step]
        ph = foc * (u[:, 0:1] * l_[None] + u[:, 1:2] * m_[None] - u[:, 2:3] * nm1[None])
        cv = W["vis"][j0:j0 + step].to(c128)
        if wgt is not None:
            cv = cv * wgt[j0:j0 + step]
        acc += (cv[:, None] * torch.polar(torch.ones_like(ph), ph)).sum(0)
    ref = acc / (nm1 + 1)
    got = d_img.reshape(-1)[pix].to(c128)
    e_img = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
    rows = torch.randperm(M, generator=g)[:nrow_sample].to(dev)
    ur = uvw[rows]
    img = W["img"].reshape(-1)
    accv = torch.zeros(nrow_sample, dtype=c128, device=dev)
    pstep = 1 << 22
    for p0 in range(0, nx * ny, pstep):
        flat = torch.arange(p0, min(p0 + pstep, nx * ny), device=dev)
        l_, m_, nm1 = lmn(flat)
        ph = -foc * (ur[:, 0:1] * l_[None] + ur[:, 1:2] * m_[None] - ur[:, 2:3] * nm1[None])
        val = img[p0:p0 + pstep].to(c128) / (nm1 + 1)
        accv += (val[None] * torch.polar(torch.ones_like(ph), ph)).sum(1)
    if wgt is not None:
        accv = accv * wgt[rows]
    gotv = v_out[rows].to(c128)
    e_vis = float(torch.linalg.norm(gotv - accv) / torch.linalg.norm(accv))
    return {"ms2dirty_rel_l2_vs_dft": e_img, "dirty2ms_rel_l2_vs_dft": e_vis,
            "pixels_sampled": npix_sample, "rows_sampled": nrow_sample,
            "reference": "FP64 direct sum on the GPU (torch), sampled pixels / rows"}


def copy_bandwidth(torch, dev, nbytes=1 << 31, reps=3):
    """Device-to-device copy rate on this GPU (GB/s, bytes read plus
    written): the practical HBM ceiling the roofline is also compared with."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def traffic_from_profiles(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(kernel)
    except Exception:
        return None


# ---------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("wstack", "plain", "nufft2d", "batch", "learned"),
                    default="wstack")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the direct-sum accuracy check of the last step's outputs")
    ap.add_argument("--nrow", type=int, default=1 << 27, help="rows of the wstack workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device-to-device copy before the steps, so that a "
                         "counter run can scale FETCH_SIZE/WRITE_SIZE by a copy of known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this build has one MI355X and the "
              "multi-GPU path is a design note (w-plane slab sharding), not code; run --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from cugridder_amd import Context
    dev = "cuda:0"
    log("generating workload %s" % args.workload)
    if args.workload == "plain":
        W = make_plain(torch, dev)
    elif args.workload == "batch":
        W = make_batch(torch, dev)
    elif args.workload in ("wstack", "learned"):
        W = make_wstack_full(torch, dev, nrow=args.nrow)
    else:
        W = make_nufft2d(torch, dev)
    torch.cuda.synchronize()
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy probe: 1 GiB read + 1 GiB written")
    copy_gbs = copy_bandwidth(torch, dev)
    log("device copy: %.0f GB/s (read + write)" % copy_gbs)
    ctx = Context(timing=True, dense_min=DENSE_MIN, degrid_dense_min=DEGRID_DENSE_MIN)

    learned_info = None
    if args.workload == "learned":
        learned_info = setup_learned(torch, ctx, W, args)

    single_ms = None
    if args.workload == "batch":
        # the same channels one call at a time (context defaults, no graph)
        def single_step():
            for it in W["items"]:
                ctx.ms2dirty(it["uvw"], it["vis"], it["wgt"], *it["dirty"].shape, it["fov"],
                             it["freq"], W["eps"], out=it["dirty"])
                ctx.dirty2ms(it["uvw"], it["img"], it["wgt"], it["fov"], it["freq"], W["eps"],
                             out=it["out"])
        single_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        single_step()
        torch.cuda.synchronize()
        single_ms = (time.perf_counter() - t0) * 1e3
        log("one call per channel: %.1f ms per step" % single_ms)
        tb = time.perf_counter()
        fw = [dict(it, dirty=it["dirty"]) for it in W["items"]]
        bw = [dict(it, dirty=it["img"], vis=it["out"]) for it in W["items"]]
        b1 = ctx.batch("ms2dirty", fw, W["eps"])
        b2 = ctx.batch("dirty2ms", bw, W["eps"])
        torch.cuda.synchronize()
        log("batch build (autotune + graph capture): %.1f s" % (time.perf_counter() - tb))

        def step():
            b1.run()
            b2.run()
    elif args.workload == "nufft2d":
        out_fk = torch.empty(W["N"], W["N"], dtype=torch.complex64, device=dev)
        out_c = torch.empty(W["M"], dtype=torch.complex64, device=dev)

        def step():
            ctx.nufft2d1(W["x"], W["y"], W["c"], W["N"], W["N"], W["eps"], out=out_fk)
            ctx.nufft2d2(W["x"], W["y"], W["fk"], W["eps"], out=out_c)
    else:
        out_d = torch.empty(W["nx"], W["ny"], dtype=torch.complex64, device=dev)
        out_v = torch.empty(W["M"], dtype=torch.complex64, device=dev)

        def step():
            ctx.ms2dirty(W["uvw"], W["vis"], W["wgt"], W["nx"], W["ny"], W["fov"], W["freq"],
                         W["eps"], out=out_d)
            ctx.dirty2ms(W["uvw"], W["img"], W["wgt"], W["fov"], W["freq"], W["eps"], out=out_v)

    for i in range(args.warmup):
        step()
        log("warmup step %d done" % i)
    torch.cuda.synchronize()
    ctx.reset_times()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        step()
        log("timed step %d done" % i)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    value = W["M"] / (ms_per_step * 1e-3) / 1e6
    st = ctx.stage_times()
    plan = ctx.last_plan()
    calls = max(st["calls"][0], 1)
    stage_ms = {d: {k: v / calls for k, v in st[d].items()} for d in ("ms2dirty", "dirty2ms")}
    log("stage ms per call: %s" % json.dumps({d: {k: round(v, 3) for k, v in s.items()}
                                              for d, s in stage_ms.items()}))

    roof = None
    kernels = None
    if args.workload not in ("batch", "nufft2d"):
        # per-call geometry of each direction (untimed calls after the steps)
        ctx.ms2dirty(W["uvw"], W["vis"], W["wgt"], W["nx"], W["ny"], W["fov"], W["freq"], W["eps"],
                     out=out_d)
        fw_stats = grid_stats(ctx, ctx.last_plan(), plan["kw"], DENSE_MIN)
        ctx.dirty2ms(W["uvw"], W["img"], W["wgt"], W["fov"], W["freq"], W["eps"], out=out_v)
        bw_stats = grid_stats(ctx, ctx.last_plan(), plan["kw"], DEGRID_DENSE_MIN)
        torch.cuda.synchronize()
        kernels = kernel_table(W, plan, fw_stats, bw_stats, stage_ms)
        # the dominant kernel: most time per step (a kernel name counts the
        # launches of both directions only when they are the same template)
        name, d, launches, t_ms, work = max(kernels, key=lambda r: r[3])
        t_launch = t_ms / launches
        achieved = work / (t_launch * 1e-3) / 1e9
        traffic = traffic_from_profiles(args.workload, "%s:%s" % (name, d))
        roof = {"kernel": "%s (%s)" % (name, d), "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": work, "avg_ms": round(t_launch, 4),
                "launches_per_call": launches}
        if traffic:
            roof["traffic_GBs"] = round(traffic / (t_launch * 1e-3) / 1e9, 1)
        if name in ("k_grid_dense", "k_grid", "k_degrid", "k_degrid_dense"):
            # useful arithmetic: kw^3 taps per point, a complex value times a
            # real kernel product (4 flop) per tap
            pts = {"k_grid_dense": fw_stats["dense_points"],
                   "k_grid": W["M"] - fw_stats["dense_points"],
                   "k_degrid_dense": bw_stats["dense_points"],
                   "k_degrid": W["M"] - bw_stats["dense_points"]}[name]
            kw = plan["kw"]
            useful = 4.0 * kw ** 3 * pts
            roof["useful_TFLOP_s"] = round(useful / (t_launch * 1e-3) / 1e12, 2)
            roof["useful_frac_of_fp32_mfma_peak"] = round(useful / (t_launch * 1e-3) / 1e12 /
                                                          FP32_MFMA_PEAK_TFLOPS, 4)
        roof.update(hbm_model(W, kernels, ms_per_step))
        step_bytes = traffic_from_profiles(args.workload, "step_total")
        if step_bytes:
            # every kernel of one step by the counters (profiles/), over the
            # measured step time
            roof["step_counter_GB"] = round(step_bytes / 1e9, 2)
            roof["step_counter_GBs"] = round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1)
        log("kernels (name, direction, launches, ms per call, algorithmic bytes per launch): %s"
            % json.dumps(kernels))
    elif args.workload == "nufft2d":
        # type-1 gridding: 16 B geometry + 8 B strength read per point, one
        # complex64 write per cell of the (2N)^2 grid
        t_grid = stage_ms["ms2dirty"]["grid"]
        nbytes = 24 * W["M"] + 8 * (2 * W["N"]) ** 2
        achieved = nbytes / (t_grid * 1e-3) / 1e9
        roof = {"kernel": "k_grid2d_mfma (type 1)", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "algorithmic_bytes": nbytes, "avg_ms": round(t_grid, 4)}

    if roof is not None:
        roof["copy_GBs"] = round(copy_gbs, 1)
        if roof.get("unit") == "GB/s":
            roof["frac_of_copy"] = round(roof["achieved"] / copy_gbs, 4)

    accuracy = None
    if args.workload in ("wstack", "plain", "learned") and not args.no_accuracy:
        # the timed steps' outputs against an FP64 direct sum (untimed)
        try:
            accuracy = dft_accuracy(torch, W, out_d, out_v)
            log("accuracy: %s" % json.dumps(accuracy))
        except Exception as e:
            log("accuracy check failed: %r" % (e,))

    cpu = None
    if not args.no_cpu:
        try:
            cpu = run_cpu_baseline(W, args.workload)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %r" % (e,))

    unit = "Mpts/s" if args.workload == "nufft2d" else "Mvis/s"
    cfg = {"workload": W["name"], "nrow": W["M"], "eps": W["eps"],
           "precision": "complex64 gridding/FFT, FP64 coordinates, phases and corrections"}
    if args.workload == "batch":
        info, is_graph = b1.info()
        cfg.update({"channels": len(W["items"]), "graph": is_graph,
                    "one_call_per_channel_ms": round(single_ms, 2),
                    "one_call_per_channel_Mvis_s": round(W["M"] / (single_ms * 1e-3) / 1e6, 3),
                    "tuned_tile_grid": sorted(set(int(x) for x in info[:, 0])),
                    "tuned_dense_min": sorted(set(int(x) for x in info[:, 2]))})
    elif args.workload != "nufft2d":
        cfg.update({"image": "%dx%d" % (W["nx"], W["ny"]), "fov_deg": W["fov"], "freq_hz": W["freq"],
                    "nf": [plan["nf1"], plan["nf2"], plan["nf3"]], "kw": plan["kw"]})
    if learned_info is not None:
        cfg["learned"] = learned_info
    out = {"metric": METRIC, "value": round(value, 3), "unit": unit, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
           "data": "synthetic", "config": cfg, "stage_ms": {d: {k: round(v, 3) for k, v in s.items()}
                                                            for d, s in stage_ms.items()},
           "roofline": roof, "cpu_baseline": cpu, "accuracy": accuracy, "wall_s": round(wall, 2)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
