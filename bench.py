#!/usr/bin/env python3
"""Throughput of the coherent F-statistic search core on one MI355X.

Default workload (the network configuration): two detectors (H1-, L1-like; white noise sigma 1
and 1.3, one injection h0 = 0.1), N = 86400 samples at dt = 2 s, nfft = 2^17, fftpad = 2
(F-statistic FFT length L = 2^18), 256 sky points x 512 spindowns = 131072 templates, threshold
F = 20.

A step is one pgw_job_batch call over `--sky-per-step` sky points (all their spindowns); the
default 16 steps x 16 sky points cover the 256 sky points once.  Inputs are resident in HBM
before the timed region.  The headline value is the FP32 form (FP64 phases and spline
coefficients, FP32 series and FFTs); the FP64 path is timed the same way beside it, and the FP32
candidates and F values are compared with the FP64 ones.

--workload selects the other configurations, each printed as its own JSON line:
  plain            1 detector, N=86400, nfft=L=2^17, 4 sky x 32 spindowns
  precision        2 detectors, N=259200, nfft=2^18, fftpad 2 (L=2^19), 1024 sky x 256 spindowns
  batching         3 detectors, 64 segments of N in [60000, 86400] with 5-30 % gaps (nfft 2^16 or
                   2^17, fftpad 1), 32 sky x 512 spindowns each (1,048,576 templates), one handle
  differentiation  2 detectors, N=86400: Fstatnet and its gradient at 1024 candidate points

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import functools
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
import pgwload  # noqa: E402

pgwload.load()
from polgraw_allsky_amd import abi, pgw, synth  # noqa: E402

METRIC = "F-stat templates/s on 1 MI355X (FP32-batched vs FP64 plain); max rel ΔF vs FP64"
HBM_PEAK_GBS = 8000.0
# the template kernels of every KTIME_EVERY-th launch of the timed steps are bracketed by HIP
# events (each event record costs the following kernels a little; sampling keeps that < 1 %)
KTIME_EVERY = 4


# template-search workloads: make_segment arguments and the config text
TEMPLATE_WORKLOADS = {
    "network": (dict(N=86400, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, h0=0.1, nsky=(16, 16), nspin=512),
                "network: 2 detectors (H1, L1), N=86400, nfft=2^17, fftpad=2, L=2^18, 256 sky x 512 spindowns, trl=20"),
    "plain": (dict(N=86400, dets=("H1",), sigmas=(1.0,), fftpad=1, h0=0.3, nsky=(2, 2), nspin=32),
              "plain: 1 detector (H1), N=86400, nfft=L=2^17, fftpad=1, 4 sky x 32 spindowns, trl=20"),
    "precision": (dict(N=259200, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, h0=0.06, nsky=(32, 32), nspin=256),
                  "precision: 2 detectors (H1, L1), N=259200, nfft=2^18, fftpad=2, L=2^19, 1024 sky x 256 spindowns, "
                  "trl=20"),
}


def workload(args, name="network"):
    kw, _ = TEMPLATE_WORKLOADS[name]
    return synth.make_segment(seed=args.seed, **kw)


def run_path(h, sky, rng, opts, steps, warmup, per_step):
    """Runs warmup + timed steps; returns (templates, seconds, gpu_ms, records, kernel times)."""
    nsteps_total = warmup + steps
    chunks = [sky[(i * per_step) % len(sky):(i * per_step) % len(sky) + per_step] for i in range(nsteps_total)]
    for c in chunks[:warmup]:
        h.job_batch(c, rng, opts)
    recs, fnum, gpu_ms = [], 0, 0.0
    h.kernel_times(KTIME_EVERY)  # HIP events around the template kernels of the timed steps
    t0 = time.perf_counter()
    for c in chunks[warmup:]:
        r, f = h.job_batch(c, rng, opts)
        recs.append(r)
        fnum += f
        gpu_ms += h.last_gpu_ms
    dt = time.perf_counter() - t0
    kt = h.kernel_times(0)
    return fnum, dt, gpu_ms, (np.concatenate(recs) if recs else np.zeros((0, 5))), kt


def cand_compare(r32, r64):
    """Candidate-set difference (keys = frequency, spindown, sky fields) and F deviation."""
    k32 = {tuple(r[:4]): r[4] for r in r32}
    k64 = {tuple(r[:4]): r[4] for r in r64}
    common = set(k32) & set(k64)
    F = lambda snr: snr * snr / 2 + 2  # noqa: E731
    rel = max((abs(F(k32[k]) - F(k64[k])) / F(k64[k]) for k in common), default=0.0)
    return dict(n_fp64=len(k64), n_fp32=len(k32), only_fp64=len(set(k64) - set(k32)),
                only_fp32=len(set(k32) - set(k64)), max_rel_dF_candidates=rel)


def fstat_compare(h32, h64, seg, opts, ntempl=4):
    """max |dF| and max relative dF on bins with F >= 1 over a few full F arrays."""
    pm, mm, nn = seg.injection["pm"], int(seg.injection["grid"][3]), int(seg.injection["grid"][2])
    ss0 = int(seg.injection["grid"][1])
    mabs, mrel = 0.0, 0.0
    for ss in range(ss0 - ntempl // 2, ss0 - ntempl // 2 + ntempl):
        a = h32.fstat(pm, mm, nn, ss, opts)
        b = h64.fstat(pm, mm, nn, ss, opts)
        d = np.abs(a - b)
        mabs = max(mabs, float(d.max()))
        sel = b >= 1.0
        mrel = max(mrel, float((d[sel] / b[sel]).max()))
    return mabs, mrel


@functools.lru_cache(maxsize=None)
def measured_traffic(kernel, prec, workload, L):
    """HBM bytes per template of a kernel at F-statistic FFT length L on a workload, from the
    committed rocprofv3 counter summaries (profiles/*.md lines 'bench-traffic: {json}'), or None."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if line.startswith("bench-traffic:"):
                d = json.loads(line.split(":", 1)[1])
                if (d.get("kernel"), d.get("precision"), d.get("workload"), d.get("L")) == (kernel, prec, workload, L):
                    return d["bytes_per_template"]
    return None


def measured_valu(kernel, prec):
    """VALU busy fraction of a kernel from the committed rocprofv3 counter summary
    (profiles/*.md lines 'bench-valu: {json}'), or None."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if line.startswith("bench-valu:"):
                d = json.loads(line.split(":", 1)[1])
                if d.get("kernel") == kernel and d.get("precision") == prec:
                    return d["valu_busy"]
    return None


def measured_flops(kernel):
    """FP64 FLOPs per (point, sample) of a kernel from the committed rocprofv3 counter summary
    (profiles/*.md lines 'bench-flops: {json}'), or None."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if line.startswith("bench-flops:"):
                d = json.loads(line.split(":", 1)[1])
                if d.get("kernel") == kernel:
                    return d["flops_per_sample"]
    return None


def alg_bytes(seg, p):
    """Algorithmic bytes per template of the column and row passes: the column pass reads the
    nifo resampled series a, b (p-byte reals) and the FP64 shifts and writes two L-point spectra;
    the row pass reads the two spectra and writes F once.  p = bytes per real of the series and
    spectra."""
    s = seg.sett
    nifo, N, L = len(seg.dets), s["N"], s["nfftf"]
    M = s["nmax"] - s["nmin"]
    b_col = nifo * N * (4 * p + 8) + 4 * L * p     # read xDatma/b + shft, write two L-point spectra
    b_row = 4 * L * p + M * p                      # read the two spectra, write F once
    return b_col, b_row


def roofline(kt, segw, prec_bytes, workload, tuned=None):
    """The dominant template kernel against the HBM roofline.  kt: the library's HIP-event kernel
    timing over the timed steps (pgw_kernel_times); segw: [(segment, templates)] of those steps
    for the algorithmic bytes."""
    tot = sum(w for _, w in segw)
    b_col = sum(w * alg_bytes(sg, prec_bytes)[0] for sg, w in segw) / tot
    b_row = sum(w * alg_bytes(sg, prec_bytes)[1] for sg, w in segw) / tot
    n, nl = kt["templates"], kt["launches"]
    k = "col" if kt["col_ms"] >= kt["row_ms"] else "row"
    # row pass kernel in use: 8 elements per thread when tuned so, or by default for FP64
    row8 = tuned["row_na"] == 8 if tuned else prec_bytes == 8
    kname = "k_spin_col2" if k == "col" else ("k_spin_row8" if row8 else "k_spin_row2")
    achieved = (b_col if k == "col" else b_row) * n / (kt[k + "_ms"] * 1e-3) / 1e9
    stage_ms = kt["col_ms"] + kt["row_ms"] + kt["thr_ms"]
    prec = "fp32" if prec_bytes == 4 else "fp64"
    trs = [(measured_traffic(kname, prec, workload, sg.sett["nfftf"]), w) for sg, w in segw]
    tr = None if any(t is None for t, _ in trs) else sum(t * w for t, w in trs) / tot
    ntl = n / nl
    # the kernels also issue close to the VALU limit: counter-measured fraction of VALU issue
    # cycles (SQ_INSTS_VALU x 4 / SIMD cycles) from the committed profile
    return dict(bound="hbm", kernel=kname, achieved=round(achieved, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=(int(tr * ntl) if tr else None), traffic_unit="bytes per launch",
                # counter-measured HBM bytes over this run's kernel duration
                traffic_GBps=(round(tr * n / (kt[k + "_ms"] * 1e-3) / 1e9, 1) if tr else None),
                valu_busy=measured_valu(kname, prec),
                timing=(f"HIP events around every {KTIME_EVERY}th template launch of the timed steps; "
                        "each launch's threshold runs inside the next launch's column pass"),
                launches=nl, templates_per_launch=round(ntl, 2),
                ms_per_launch={x: round(kt[x + "_ms"] / nl, 4) for x in ("col", "row", "thr")},
                bytes_per_template=dict(col=round(b_col), row=round(b_row)),
                template_stage_GBps=round((b_col + b_row) * n / (stage_ms * 1e-3) / 1e9, 1))


def threads_env():
    """The OpenMP settings of the CPU baseline's run (thread count and pinning; unset = the
    runtime's defaults, threads unpinned)."""
    return {k: os.environ.get(k, "unset") for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import orc
    return orc


def cpu_baseline(seg, rng, opts, sky=None, budget_s=10.0):
    """The FP64 oracle on sky points of the workload (all spindowns of each, the list repeated if
    it is short), OpenMP threads, until about budget_s seconds of wall time have been spent."""
    orc = oracle()
    threads = orc.default_threads()
    o = orc.Search(seg, nthreads=threads)
    pts = sky or [(seg.injection["pm"], int(seg.injection["grid"][3]), int(seg.injection["grid"][2]))]
    fn, nsky, dt = 0, 0, 0.0
    while dt < budget_s and nsky < 100 * len(pts):  # a small workload is repeated
        pm, mm, nn = pts[nsky % len(pts)]
        t0 = time.perf_counter()
        _, f = o.job_core(pm, mm, nn, rng, opts)
        dt += time.perf_counter() - t0
        fn += f
        nsky += 1
    return dict(value=round(fn / dt, 2), unit="templates/s", cores=threads, kind="port",
                sample=f"{nsky} sky-point searches x all spindowns ({fn} templates) of the same workload (FP64 oracle, "
                       f"own Stockham FFT, OpenMP over spindowns; {dt:.1f} s wall)", cpu_model=cpu_model(),
                threads_env=threads_env(), host_cpus=len(os.sched_getaffinity(0)))


def tune_handle(args, h, segi, seg, opts):
    """The library's autotuner on the injected sky point (setup, before any timed step): kernel
    variants and templates per launch for this handle.  Returns the choice (None with --no-tune)."""
    if args.no_tune:
        return None
    pm, mm, nn = seg.injection["pm"], int(seg.injection["grid"][3]), int(seg.injection["grid"][2])
    choice = h.autotune(segi, pm, mm, nn, opts)
    print(f"autotune: {choice}", file=sys.stderr, flush=True)
    return choice


def base_line(args, value, t, cfg, dtype="fp32", unit="templates/s", metric=METRIC):
    return {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * t / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": cfg}


def bench_templates(args, name):
    """network / plain / precision: FP32 form and FP64 path over the same sky points."""
    seg = workload(args, name)
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=20.0)
    sky = synth.sky_list(seg)
    nspin = rng.spndr[1] - rng.spndr[0] + 1
    per_step = args.sky_per_step or min(len(sky), {"network": 16, "plain": 4, "precision": 64}[name])
    if args.steps is None:
        # plain: 4 sky points are one step of 0.3 ms; 256 repetitions time ~70 ms, so one-time
        # host-side costs of the runtime (a stall of a few ms after ~1000 queued operations was
        # seen) do not dominate the timed region
        args.steps = max(1, len(sky) // per_step) if name != "plain" else 256
    s = seg.sett
    print(f"workload {name}: N={s['N']} nfft={s['nfft']} fftpad={s['fftpad']} L={s['nfftf']} "
          f"sky={len(sky)} spindowns={nspin}", file=sys.stderr, flush=True)
    h32 = pgw.Handle(seg, abi.PREC_FP32)
    tune = {"fp32": tune_handle(args, h32, 0, seg, opts)}
    n32, t32, g32, r32, kt32 = run_path(h32, sky, rng, opts, args.steps, args.warmup, per_step)
    v32 = n32 / t32
    print(f"fp32: {n32} templates in {t32:.3f} s -> {v32:.0f} templates/s (gpu {g32:.1f} ms)",
          file=sys.stderr, flush=True)
    out_fp64, acc = None, None
    if not args.no_fp64:
        h64 = pgw.Handle(seg, abi.PREC_FP64)
        tune["fp64"] = tune_handle(args, h64, 0, seg, opts)
        n64, t64, g64, r64, kt64 = run_path(h64, sky, rng, opts, args.steps, args.warmup, per_step)
        v64 = n64 / t64
        print(f"fp64: {n64} templates in {t64:.3f} s -> {v64:.0f} templates/s (gpu {g64:.1f} ms)",
              file=sys.stderr, flush=True)
        out_fp64 = dict(value=round(v64, 1), ms_per_step=round(1e3 * t64 / args.steps, 3),
                        ms_per_sky_point=round(1e3 * t64 / (args.steps * per_step), 4),
                        gpu_ms_per_step=round(g64 / args.steps, 3), templates=n64,
                        roofline=roofline(kt64, [(seg, 1)], 8, name, tune.get("fp64")))
        acc = cand_compare(r32, r64)
        mabs, mrel = fstat_compare(h32, h64, seg, opts)
        acc.update(max_abs_dF=mabs, max_rel_dF_F_ge_1=mrel)
        h64.close()
    rl = roofline(kt32, [(seg, 1)], 4, name, tune.get("fp32"))
    cpu = None if args.no_cpu else cpu_baseline(seg, rng, opts, sky)
    if cpu:
        print(f"cpu: {cpu}", file=sys.stderr, flush=True)
    cfg = {"workload": TEMPLATE_WORKLOADS[name][1], "N": s["N"], "nfft": s["nfft"], "fftpad": s["fftpad"],
           "L": s["nfftf"], "detectors": len(seg.dets), "sky_points": len(sky), "spindowns": nspin,
           "sky_per_step": per_step, "tuning": tune}
    line = base_line(args, v32, t32, cfg)
    line.update({"templates": n32, "gpu_ms_per_step": round(g32 / args.steps, 3),
                 "ms_per_sky_point": round(1e3 * t32 / (args.steps * per_step), 4), "fp64": out_fp64,
                 "speedup_fp32_vs_fp64": (round(v32 / out_fp64["value"], 3) if out_fp64 else None),
                 "accuracy_fp32_vs_fp64": acc, "roofline": rl, "cpu_baseline": cpu})
    return line


def batching_segments(args, nseg=64):
    rs = np.random.default_rng(args.seed)
    segs = []
    for k in range(nseg):
        N = int(rs.integers(60000, 86401))
        segs.append(synth.make_segment(N=N, dets=("H1", "L1", "V1"), sigmas=(1.0, 1.3, 1.1), fftpad=1,
                                       seed=args.seed + 1 + k, h0=0.12, gap_frac=float(rs.uniform(0.05, 0.3)),
                                       nsky=(4, 8), nspin=512))
    return segs


def bench_batching(args):
    """64 uneven segments of 3 detectors in one handle; work items (segment, sky point) ordered by
    segment, 128 per step; FP32 and FP64 over the same items."""
    segs = batching_segments(args)
    work = [(k, pm, mm, nn) for k, sg in enumerate(segs) for pm, mm, nn in synth.sky_list(sg)]
    per_step = args.sky_per_step or 128
    if args.steps is None:
        args.steps = max(1, len(work) // per_step)
    n16 = sum(1 for sg in segs if sg.sett["nfft"] == 1 << 16)
    print(f"workload batching: {len(segs)} segments ({n16} with nfft 2^16), {len(work)} work items",
          file=sys.stderr, flush=True)
    opts = abi.make_opts(trl=20.0)

    tune = {}

    # the autotuner runs on a segment of the longer FFT length (most of the templates)
    ktune = max(range(len(segs)), key=lambda k: (segs[k].sett["nfftf"], -k))

    def run(prec):
        h = pgw.Handle(segs, prec)
        tune["fp32" if prec == abi.PREC_FP32 else "fp64"] = tune_handle(args, h, ktune, segs[ktune], opts)
        chunks = [work[(i * per_step) % len(work):(i * per_step) % len(work) + per_step]
                  for i in range(args.warmup + args.steps)]
        # each segment's spindowns come from its own range; the call takes one range, so the
        # work of a step shares the range of its first segment's grid centre
        for c in chunks[:args.warmup]:
            h.job_batch_multi(c, abi.make_range(segs[c[0][0]]), opts)
        recs, fn, gms = [], 0, 0.0
        h.kernel_times(KTIME_EVERY)
        t0 = time.perf_counter()
        for c in chunks[args.warmup:]:
            r, _, f = h.job_batch_multi(c, abi.make_range(segs[c[0][0]]), opts)
            recs.append(r)
            fn += f
            gms += h.last_gpu_ms
        dt = time.perf_counter() - t0
        kt = h.kernel_times(0)
        # algorithmic bytes: each timed work item's segment, weighted by its templates (equal)
        segw = [(segs[k], 1) for c in chunks[args.warmup:] for k, _, _, _ in c]
        return h, fn, dt, gms, np.concatenate(recs), kt, segw

    h32, n32, t32, g32, r32, kt32, sw32 = run(abi.PREC_FP32)
    v32 = n32 / t32
    print(f"fp32: {n32} templates in {t32:.3f} s -> {v32:.0f} templates/s (gpu {g32:.1f} ms)",
          file=sys.stderr, flush=True)
    out_fp64, acc = None, None
    if not args.no_fp64:
        h64, n64, t64, g64, r64, kt64, sw64 = run(abi.PREC_FP64)
        v64 = n64 / t64
        print(f"fp64: {n64} templates in {t64:.3f} s -> {v64:.0f} templates/s", file=sys.stderr, flush=True)
        out_fp64 = dict(value=round(v64, 1), ms_per_step=round(1e3 * t64 / args.steps, 3), templates=n64,
                        roofline=roofline(kt64, sw64, 8, "batching", tune.get("fp64")))
        acc = cand_compare(r32, r64)
        h64.close()
    rl = roofline(kt32, sw32, 4, "batching", tune.get("fp32"))
    cpu = None if args.no_cpu else cpu_baseline(segs[0], abi.make_range(segs[0]), opts, synth.sky_list(segs[0]))
    cfg = {"workload": "batching: 3 detectors (H1, L1, V1), 64 segments of N in [60000, 86400] with 5-30 % gaps, "
                       "nfft 2^16 or 2^17, fftpad=1, 32 sky x 512 spindowns per segment, trl=20",
           "segments": len(segs), "segments_nfft_2^16": n16, "work_items": len(work), "items_per_step": per_step,
           "tuning": tune}
    line = base_line(args, v32, t32, cfg)
    line.update({"templates": n32, "fp64": out_fp64,
                 "speedup_fp32_vs_fp64": (round(v32 / out_fp64["value"], 3) if out_fp64 else None),
                 "accuracy_fp32_vs_fp64": acc, "roofline": rl, "cpu_baseline": cpu})
    return line


def bench_differentiation(args, npts=1024):
    """Fstatnet with its gradient at npts points around an injection; a step is one batched call.
    Also times a 20-step quasi-Newton ascent of all points."""
    from polgraw_allsky_amd import followup
    seg = synth.make_segment(N=86400, dets=("H1", "L1"), sigmas=(1.0, 1.3), fftpad=2, seed=args.seed, h0=0.1,
                             nsky=(1, 1), nspin=1)
    inj = seg.injection
    truth = np.array([inj["f"], inj["fdot"], inj["delta"], inj["alpha"]])
    sc = followup.phase_scales(seg.DetSSB, seg.sett["N"], seg.sett["oms"], truth)
    pts = truth + np.random.default_rng(args.seed).uniform(-4, 4, (npts, 4)) * sc
    if args.steps is None:
        args.steps = 20
    h = pgw.Handle(seg, abi.PREC_FP64)
    for _ in range(args.warmup):
        h.fstatnet(pts)
    t0 = time.perf_counter()
    gms = 0.0
    for _ in range(args.steps):
        out, g = h.fstatnet(pts)
        gms += h.last_gpu_ms
    dt = time.perf_counter() - t0
    v = npts * args.steps / dt
    print(f"fstatnet+gradient: {npts} points x {args.steps} in {dt:.3f} s -> {v:.0f} points/s "
          f"(gpu {gms:.1f} ms)", file=sys.stderr, flush=True)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        h.fstatnet(pts, grad=False)
    v_nograd = npts * args.steps / (time.perf_counter() - t2)
    print(f"fstatnet values only: {v_nograd:.0f} points/s", file=sys.stderr, flush=True)
    G = followup.phase_metric(seg.DetSSB, seg.sett["N"], seg.sett["oms"], truth)
    t1 = time.perf_counter()
    x, F, nev = followup.ascend(followup.FStatNet(h), pts, G, steps=20)
    t_asc = time.perf_counter() - t1
    # the followup's Nelder-Mead from the same points, all simplices batched per device call
    t3 = time.perf_counter()
    xs, outs, nev_s, conv_s = followup.amoeba(followup.FStatNet(h), pts, 0.5 * sc, tol=1e-9)
    t_nm = time.perf_counter() - t3
    print(f"simplex: {npts} points, {int(nev_s.sum())} evaluations in {t_nm:.3f} s", file=sys.stderr, flush=True)
    # the followup's mesh search (invertedMADS), 80 mesh points of every candidate per device call,
    # from the first 128 points (its ~470 rounds per candidate make it the slowest of the three)
    nod = seg.sett["N"] * seg.sett["dt"] / 86400.0
    nm = min(128, npts)
    t4 = time.perf_counter()
    xm, outm, rounds_m = followup.mads(followup.FStatNet(h), pts[:nm], 2.0 * sc, nod)
    t_mads = time.perf_counter() - t4
    print(f"mads: {nm} points, {int(rounds_m.sum()) * 80} evaluations in {t_mads:.3f} s", file=sys.stderr,
          flush=True)
    cpu = None
    if not args.no_cpu:
        orc = oracle()
        threads = orc.default_threads()
        o = orc.Search(seg, nthreads=threads)
        # all points with gradient, repeated until ~10 s of wall time
        ns, tcd, reps = 0, 0.0, 0
        while tcd < 10.0 and reps < 200:
            tc = time.perf_counter()
            out_o, g_o = o.fstatnet(pts)
            tcd += time.perf_counter() - tc
            ns += len(pts)
            reps += 1
        cpu = dict(value=round(ns / tcd, 2), unit="points/s", cores=threads, kind="port",
                   sample=f"{reps} x the {len(pts)} points with gradient (FP64 oracle Fstatnet, OpenMP over points; "
                          f"{tcd:.1f} s wall)", cpu_model=cpu_model(), threads_env=threads_env(),
                   host_cpus=len(os.sched_getaffinity(0)),
                   max_rel_dF_vs_device=float(np.max(np.abs(out_o[:, 5] - out[:, 5]) / np.abs(out_o[:, 5]))))
    nifo, N = len(seg.dets), seg.sett["N"]
    fps = measured_flops("k_fstatnet_part")
    rl = None
    if fps:
        # FP64 vector roofline: counter-measured FLOPs per (point, sample) x samples per call over
        # the device time of a call (HIP events around the two kernels and the small copies)
        ach = fps * npts * nifo * N / (gms / args.steps * 1e-3) / 1e12
        rl = dict(bound="fp64-valu", kernel="k_fstatnet_part", achieved=round(ach, 2), peak=78.6, unit="TFLOP/s",
                  frac=round(ach / 78.6, 4), traffic=None, flops_per_sample=fps)
    cfg = {"workload": "differentiation: 2 detectors (H1, L1), N=86400, Fstatnet + dF/d(omega, omegadot, delta, "
                       "alpha) at 1024 candidate points", "N": N, "detectors": nifo, "points": npts}
    line = base_line(args, v, dt, cfg, dtype="fp64", unit="points/s",
                     metric="Fstatnet evaluations with gradient per second on 1 MI355X")
    line.update({"gpu_ms_per_step": round(gms / args.steps, 3), "samples_per_s": round(v * nifo * N, 1),
                 "values_only_points_per_s": round(v_nograd, 1),
                 "ascent": {"steps": 20, "evaluations_per_point": nev, "seconds": round(t_asc, 3),
                            "F_start_median": float(np.median(-out[:, 5])), "F_end_median": float(np.median(F))},
                 "simplex": {"tol": 1e-9, "evaluations_per_point_mean": round(float(nev_s.mean()), 1),
                             "converged": int(conv_s.sum()), "seconds": round(t_nm, 3),
                             "evaluations_per_s": round(float(nev_s.sum()) / t_nm, 1),
                             "F_end_median": float(np.median(-outs[:, 5]))},
                 "mads": {"points": nm, "p2": "2 phase scales", "rounds_mean": round(float(rounds_m.mean()), 1),
                          "seconds": round(t_mads, 3),
                          "evaluations_per_s": round(float(rounds_m.sum()) * 80 / t_mads, 1),
                          "F_end_median": float(np.median(-outm[:, 5]))},
                 "roofline": rl, "cpu_baseline": cpu})
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sky-per-step", type=int, default=None)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workload", default="network",
                    choices=["network", "plain", "precision", "batching", "differentiation"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (kernel variants and templates per launch) before timing")
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a plain 256 MiB device copy (scale for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sky points would shard across GPUs, one process and handle each, with no collectives)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        ms = pgw.copy_bench(256 << 20, reps=3)
        print(f"copy 256 MiB: {ms:.3f} ms ({2 * 256 * 1.048576 / ms:.0f} GB/s)", file=sys.stderr, flush=True)
    if args.workload in TEMPLATE_WORKLOADS:
        line = bench_templates(args, args.workload)
    elif args.workload == "batching":
        line = bench_batching(args)
    else:
        line = bench_differentiation(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
