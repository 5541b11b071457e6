#!/usr/bin/env python3
"""Throughput of the coherent F-statistic search core on one MI355X.

Default workload (the batching configuration, the largest): three detectors (H1-, L1-, V1-like;
white noise sigma 1, 1.3, 1.1, one injection h0 = 0.12 per segment), 64 segments of uneven valid
length N in [60000, 86400] with zeroed gaps over 5-30 % of the samples (nfft 2^16 or 2^17,
fftpad 1), 32 sky points x 512 spindowns per segment = 2048 work items x 512 = 1,048,576 templates
in one handle, threshold F = 20.

A step is one pgw_job_batch_multi call over `--sky-per-step` work items (default 256 items =
131072 templates); 8 steps cover the workload once, and the default 24 steps cover it three times.
Inputs are resident in HBM before the timed region.  The headline value is the FP32 form (FP64
phases and spline coefficients, FP32 series and FFTs); the FP64 path is timed the same way over
the same work items, and the FP32 candidates and F values are compared with the FP64 ones.

After the headline, unless --no-secondary, the same line carries `other_workloads`: FP32 and
FP64 templates/s, roofline and accuracy of the other configurations (fewer steps, no CPU leg),
and the drop-in path (pgw_job_core called once per sky point, as a reference search() would):
  plain            1 detector, N=86400, nfft=L=2^17, 4 sky x 32 spindowns
  network          2 detectors, N=86400, nfft=2^17, fftpad 2 (L=2^18), 256 sky x 512 spindowns
  precision        2 detectors, N=259200, nfft=2^18, fftpad 2 (L=2^19), 1024 sky x 256 spindowns
  differentiation  2 detectors, N=86400: Fstatnet and its gradient at 1024 points near 1024 injections
--workload selects one of them as the headline instead.

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
TEMPLATE_WORKLOADS = synth.WORKLOADS


def workload(args, name="network"):
    return synth.workload_segment(name, seed=args.seed)


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


def measured_valu(kernel, prec, workload=None):
    """VALU busy fraction of a kernel from the committed rocprofv3 counter summaries
    (profiles/*.md lines 'bench-valu: {json}'): the line of this workload if there is one, else
    one without a workload, else None."""
    generic = None
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if line.startswith("bench-valu:"):
                d = json.loads(line.split(":", 1)[1])
                if d.get("kernel") == kernel and d.get("precision") == prec:
                    if d.get("workload") == workload:
                        return d["valu_busy"]
                    if "workload" not in d and generic is None:
                        generic = d["valu_busy"]
    return generic


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
    """Algorithmic bytes of the template stage per template, split by kernel as the path's
    per-template figure B_t = nifo N (4p+8) + 8 L p + M p divides: the column pass reads the nifo
    resampled series a, b (p-byte reals) and the FP64 _tmp1 and writes two L-point spectra; the
    row pass reads the two spectra and writes F once.  p = bytes per real of the series and
    spectra.  Returns (series part, spectra part of the column pass, row pass)."""
    s = seg.sett
    nifo, N, L = len(seg.dets), s["N"], s["nfftf"]
    M = s["nmax"] - s["nmin"]
    return nifo * N * (4 * p + 8), 4 * L * p, 4 * L * p + M * p


def annotate_stream(obj, hbm):
    """Adds to every template-kernel roofline in obj the streaming bandwidth measured in this run
    (pgw_hbm_probe) and the kernel's fraction of it: the column pass mostly writes, the row pass
    mostly reads."""
    if isinstance(obj, dict):
        if obj.get("bound") == "hbm" and "achieved" in obj and "hbm_measured_GBps" in obj:
            against = "write" if obj["kernel"] == "k_spin_col2" else "read"
            obj["hbm_measured_GBps"] = hbm
            obj["frac_of_measured_stream"] = {"against": against, "GBps": hbm[against],
                                              "frac": round(obj["achieved"] / hbm[against], 4)}
        for v in obj.values():
            annotate_stream(v, hbm)
    elif isinstance(obj, list):
        for v in obj:
            annotate_stream(v, hbm)


def roofline(kt, segw, prec_bytes, workload, tuned=None, spins=512):
    """The dominant template kernel against the HBM roofline.  kt: the library's HIP-event kernel
    timing over the timed steps (pgw_kernel_times); segw: [(segment, templates)] of those steps
    for the algorithmic bytes.

    Bytes per launch: the row pass moves 4 L p + M p per template (no reuse is possible).  The
    column pass's templates of a launch are consecutive spindowns of one sky point and share its
    series, so one launch reads them once: nifo N (4p+8) per launch + 4 L p per template (the sky
    stage's B_s is amortised over a sky point's spindowns in the same way); a launch longer than
    a sky point's `spins` spindowns reads the series of each of its sky points.  Counted per
    template instead, the column pass would appear to run above the HBM peak, from L2 hits."""
    tot = sum(w for _, w in segw)
    ab = [(alg_bytes(sg, prec_bytes), w) for sg, w in segw]
    b_ser = sum(w * a[0] for a, w in ab) / tot
    b_spec = sum(w * a[1] for a, w in ab) / tot
    b_row = sum(w * a[2] for a, w in ab) / tot
    n, nl = kt["templates"], kt["launches"]
    k = "col" if kt["col_ms"] >= kt["row_ms"] else "row"
    # row pass kernel in use: 8 elements per thread when tuned so, or by default for FP64
    row8 = tuned["row_na"] == 8 if tuned else prec_bytes == 8
    kname = "k_spin_col2" if k == "col" else ("k_spin_row8" if row8 else "k_spin_row2")
    bytes_col = max(nl, n / spins) * b_ser + n * b_spec   # over the timed launches
    bytes_row = n * b_row
    achieved = (bytes_col if k == "col" else bytes_row) / (kt[k + "_ms"] * 1e-3) / 1e9
    stage_ms = kt["col_ms"] + kt["row_ms"] + kt["thr_ms"]
    prec = "fp32" if prec_bytes == 4 else "fp64"
    trs = [(measured_traffic(kname, prec, workload, sg.sett["nfftf"]), w) for sg, w in segw]
    tr = None if any(t is None for t, _ in trs) else sum(t * w for t, w in trs) / tot
    ntl = n / nl
    other = "row" if k == "col" else "col"
    ach_other = (bytes_row if k == "col" else bytes_col) / (kt[other + "_ms"] * 1e-3) / 1e9
    return dict(bound="hbm", kernel=kname, achieved=round(achieved, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                hbm_measured_GBps=None, frac_of_measured_stream=None,  # set by annotate_stream
                traffic=(int(tr * ntl) if tr else None), traffic_unit="bytes per launch",
                # counter-measured HBM bytes over this run's kernel duration
                traffic_GBps=(round(tr * n / (kt[k + "_ms"] * 1e-3) / 1e9, 1) if tr else None),
                # counter-measured fraction of VALU issue cycles (SQ_INSTS_VALU x 4 / SIMD cycles)
                valu_busy=measured_valu(kname, prec, workload),
                share_of_template_time=round(kt[k + "_ms"] / stage_ms, 4),
                other_pass={"kernel": other, "achieved_GBps": round(ach_other, 1),
                            "frac": round(ach_other / HBM_PEAK_GBS, 4),
                            "share_of_template_time": round(kt[other + "_ms"] / stage_ms, 4)},
                timing=(f"HIP events around every {KTIME_EVERY}th template launch of the timed steps; "
                        "each launch's threshold runs inside the next launch's column pass"),
                launches=nl, templates_per_launch=round(ntl, 2),
                ms_per_launch={x: round(kt[x + "_ms"] / nl, 4) for x in ("col", "row", "thr")},
                bytes_per_launch=dict(col=round(bytes_col / nl), row=round(bytes_row / nl)),
                bytes_per_template_Bt=dict(col=round(b_ser + b_spec), row=round(b_row)),
                template_stage_GBps_Bt=round((b_ser + b_spec + b_row) * n / (stage_ms * 1e-3) / 1e9, 1))


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


def tune_handle(args, h, segi, seg, opts, rng, nsky):
    """The library's autotuner on the injected sky point (setup, before any timed step), over calls
    of nsky copies of it with the templates the workload's spindown range forms there: kernel
    variants and templates per launch for this handle.  Returns the choice (None with --no-tune)."""
    if args.no_tune:
        return None
    pm, mm, nn = seg.injection["pm"], int(seg.injection["grid"][3]), int(seg.injection["grid"][2])
    choice = h.autotune(segi, pm, mm, nn, opts, rng, min(nsky, 32))
    choice["variants"] = h.tune_string()
    print(f"autotune: {choice}", file=sys.stderr, flush=True)
    return choice


def base_line(args, value, t, cfg, dtype="fp32", unit="templates/s", metric=METRIC, steps=None, warmup=None):
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    return {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(1e3 * t / steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": cfg}


def bench_templates(args, name, steps=None, warmup=None, per_step=None, cpu=True):
    """network / plain / precision: FP32 form and FP64 path over the same sky points."""
    seg = workload(args, name)
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=20.0)
    sky = synth.sky_list(seg)
    nspin = rng.spndr[1] - rng.spndr[0] + 1
    per_step = per_step or min(len(sky), {"network": 16, "plain": 4, "precision": 64}[name])
    warmup = args.warmup if warmup is None else warmup
    if steps is None:
        # plain: 4 sky points are one step of ~0.2 ms; 1024 repetitions time ~0.2 s, after at least
        # 32 untimed ones: the device's clocks ramp up over the first ~100 steps, and one-time
        # host-side costs of the runtime (a stall of a few ms after ~1000 queued operations was
        # seen) do not dominate the timed region
        steps = max(1, len(sky) // per_step) if name != "plain" else 1024
        if name == "plain":
            warmup = max(warmup, 32)
    s = seg.sett
    print(f"workload {name}: N={s['N']} nfft={s['nfft']} fftpad={s['fftpad']} L={s['nfftf']} "
          f"sky={len(sky)} spindowns={nspin}", file=sys.stderr, flush=True)
    h32 = pgw.Handle(seg, abi.PREC_FP32)
    tune = {"fp32": tune_handle(args, h32, 0, seg, opts, rng, per_step)}
    n32, t32, g32, r32, kt32 = run_path(h32, sky, rng, opts, steps, warmup, per_step)
    v32 = n32 / t32
    print(f"fp32: {n32} templates in {t32:.3f} s -> {v32:.0f} templates/s (gpu {g32:.1f} ms)",
          file=sys.stderr, flush=True)
    out_fp64, acc = None, None
    if not args.no_fp64:
        h64 = pgw.Handle(seg, abi.PREC_FP64)
        tune["fp64"] = tune_handle(args, h64, 0, seg, opts, rng, per_step)
        n64, t64, g64, r64, kt64 = run_path(h64, sky, rng, opts, steps, warmup, per_step)
        v64 = n64 / t64
        print(f"fp64: {n64} templates in {t64:.3f} s -> {v64:.0f} templates/s (gpu {g64:.1f} ms)",
              file=sys.stderr, flush=True)
        out_fp64 = dict(value=round(v64, 1), ms_per_step=round(1e3 * t64 / steps, 3),
                        ms_per_sky_point=round(1e3 * t64 / (steps * per_step), 4),
                        gpu_ms_per_step=round(g64 / steps, 3), templates=n64,
                        roofline=roofline(kt64, [(seg, 1)], 8, name, tune.get("fp64"), spins=nspin))
        acc = cand_compare(r32, r64)
        mabs, mrel = fstat_compare(h32, h64, seg, opts)
        acc.update(max_abs_dF=mabs, max_rel_dF_F_ge_1=mrel)
        h64.close()
    rl = roofline(kt32, [(seg, 1)], 4, name, tune.get("fp32"), spins=nspin)
    cpu = None if (args.no_cpu or not cpu) else cpu_baseline(seg, rng, opts, sky)
    if cpu:
        print(f"cpu: {cpu}", file=sys.stderr, flush=True)
    cfg = {"workload": TEMPLATE_WORKLOADS[name][1], "N": s["N"], "nfft": s["nfft"], "fftpad": s["fftpad"],
           "L": s["nfftf"], "detectors": len(seg.dets), "sky_points": len(sky), "spindowns": nspin,
           "sky_per_step": per_step, "tuning": tune}
    line = base_line(args, v32, t32, cfg, steps=steps, warmup=warmup)
    line.update({"templates": n32, "gpu_ms_per_step": round(g32 / steps, 3),
                 "ms_per_sky_point": round(1e3 * t32 / (steps * per_step), 4), "fp64": out_fp64,
                 "speedup_fp32_vs_fp64": (round(v32 / out_fp64["value"], 3) if out_fp64 else None),
                 "accuracy_fp32_vs_fp64": acc, "roofline": rl, "cpu_baseline": cpu})
    return line


def batching_segments(args, nseg=64):
    return synth.batching_segments(seed=args.seed, nseg=nseg)


def bench_batching(args):
    """64 uneven segments of 3 detectors in one handle; work items (segment, sky point) ordered by
    segment, 256 per step; FP32 and FP64 over the same items."""
    segs = batching_segments(args)
    work = [(k, pm, mm, nn) for k, sg in enumerate(segs) for pm, mm, nn in synth.sky_list(sg)]
    per_step = args.sky_per_step or 256
    if args.steps is None:
        # three passes over the 2048 work items: >= 2 s of timed work per precision
        args.steps = 3 * max(1, len(work) // per_step)
    n16 = sum(1 for sg in segs if sg.sett["nfft"] == 1 << 16)
    print(f"workload batching: {len(segs)} segments ({n16} with nfft 2^16), {len(work)} work items",
          file=sys.stderr, flush=True)
    opts = abi.make_opts(trl=20.0)

    tune = {}

    # the autotuner runs on a segment of the longer FFT length (most of the templates)
    ktune = max(range(len(segs)), key=lambda k: (segs[k].sett["nfftf"], -k))

    def run(prec):
        h = pgw.Handle(segs, prec)
        tune["fp32" if prec == abi.PREC_FP32 else "fp64"] = tune_handle(args, h, ktune, segs[ktune], opts,
                                                                       abi.make_range(segs[ktune]), per_step)
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
        # full F arrays of a few templates around the first segment's injection
        mabs, mrel = fstat_compare(h32, h64, segs[0], opts)
        acc.update(max_abs_dF=mabs, max_rel_dF_F_ge_1=mrel)
        h64.close()
    h32.close()
    rl = roofline(kt32, sw32, 4, "batching", tune.get("fp32"))
    cpu = None if args.no_cpu else cpu_baseline(segs[0], abi.make_range(segs[0]), opts, synth.sky_list(segs[0]))
    cfg = {"workload": synth.BATCHING_TEXT,
           "segments": len(segs), "segments_nfft_2^16": n16, "work_items": len(work), "items_per_step": per_step,
           "tuning": tune}
    line = base_line(args, v32, t32, cfg)
    line.update({"templates": n32, "fp64": out_fp64,
                 "speedup_fp32_vs_fp64": (round(v32 / out_fp64["value"], 3) if out_fp64 else None),
                 "accuracy_fp32_vs_fp64": acc, "roofline": rl, "cpu_baseline": cpu})
    return line


def differentiation_setup(seed, nseg=16, per_seg=64, mismatch=0.25):
    return synth.differentiation_workload(seed=seed, nseg=nseg, per_seg=per_seg, mismatch=mismatch)


def param_errors(x, truth, cells):
    """|x - truth| per parameter in cells (right ascension wrapped to (-pi, pi])."""
    d = x - truth
    d[:, 3] = (d[:, 3] + np.pi) % (2 * np.pi) - np.pi
    return np.abs(d) / cells


def err_summary(e):
    return {"median_cells": [round(float(v), 4) for v in np.median(e, axis=0)],
            "p90_cells": [round(float(v), 4) for v in np.percentile(e, 90, axis=0)]}


def bench_differentiation(args, steps=None, warmup=None, cpu=True, mads_points=64):
    """Fstatnet with its gradient at the candidate points of 1024 injections; a step is one
    evaluation of all points (one device call per segment).  Then the gradient against central
    differences (1e-3 of a grid cell per parameter) and the FP64 oracle, and a 20-step gradient
    ascent against the followup's simplex (its tolerance 1e-7 and first steps of twice the grid
    cell, followup/src/followup.c:1192,1555-1567,1727) on recovered F and parameter error."""
    from polgraw_allsky_amd import followup
    steps = (args.steps or 20) if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t_setup = time.perf_counter()
    segs, truth, pts, cells, seg_of, h0 = differentiation_setup(args.seed)
    npts = len(pts)
    print(f"differentiation: {len(segs)} segments x {npts // len(segs)} injections ({time.perf_counter() - t_setup:.1f} s "
          f"setup)", file=sys.stderr, flush=True)
    h = pgw.Handle(segs, abi.PREC_FP64)
    fs = followup.MultiFStatNet(h, seg_of)
    for _ in range(warmup):
        fs(pts)
    t0 = time.perf_counter()
    for _ in range(steps):
        out, g = fs(pts)
    dt = time.perf_counter() - t0
    v = npts * steps / dt
    t2 = time.perf_counter()
    for _ in range(steps):
        fs(pts, grad=False)
    v_nograd = npts * steps / (time.perf_counter() - t2)
    print(f"fstatnet+gradient: {npts} points x {steps} in {dt:.3f} s -> {v:.0f} points/s; values only "
          f"{v_nograd:.0f} points/s", file=sys.stderr, flush=True)
    F0 = -out[:, 5]
    # gradient vs central differences, step 1e-3 of the grid cell along each parameter
    fd = np.zeros_like(g)
    for k in range(4):
        e = np.zeros(4)
        e[k] = 1.0
        hk = 1e-3 * cells[:, k]
        fp = -fs(pts + hk[:, None] * e, grad=False)[0][:, 5]
        fm = -fs(pts - hk[:, None] * e, grad=False)[0][:, 5]
        fd[:, k] = (fp - fm) / (2 * hk)
    # per point and parameter: the gradient difference over one cell, relative to F
    fd_err = np.abs(g - fd) * cells / F0[:, None]
    # 20-step ascent from the candidate points
    G = np.array([followup.phase_metric(segs[seg_of[i]].DetSSB, 86400, segs[seg_of[i]].sett["oms"], pts[i])
                  for i in range(npts)])
    t1 = time.perf_counter()
    xa, Fa, nev = followup.ascend(fs, pts, G, steps=20)
    t_asc = time.perf_counter() - t1
    # the followup's simplex from the same points
    t3 = time.perf_counter()
    xs, outs, nev_s, conv_s = followup.amoeba(fs, pts, 2.0 * cells, tol=1e-7)
    t_nm = time.perf_counter() - t3
    Fs = -outs[:, 5]
    print(f"ascent {t_asc:.2f} s, simplex {t_nm:.2f} s ({int(nev_s.sum())} evaluations)", file=sys.stderr, flush=True)
    # the followup's mesh search from the first points (its hundreds of rounds make it slowest)
    nm = min(mads_points, npts)
    nod = 86400 * 2.0 / 86400.0
    t4 = time.perf_counter()
    xm, outm, rounds_m = followup.mads(followup.MultiFStatNet(h, seg_of[:nm]), pts[:nm], 2.0 * cells[:nm], nod)
    t_mads = time.perf_counter() - t4
    e0, ea, es, em = (param_errors(x, t, c) for x, t, c in ((pts, truth, cells), (xa, truth, cells),
                                                             (xs, truth, cells), (xm, truth[:nm], cells[:nm])))
    rel = (Fa - Fs) / Fs
    cpu_line, orc_grad = None, None
    if cpu and not args.no_cpu:
        orc = oracle()
        threads = orc.default_threads()
        # FP64 reference gradient of the oracle at the points of the first segment
        sel = np.nonzero(seg_of == 0)[0]
        o = orc.Search(segs[0], nthreads=threads)
        tc = time.perf_counter()
        out_o, g_o = o.fstatnet(pts[sel])
        tcd = time.perf_counter() - tc
        ns, reps = len(sel), 1
        while tcd < 10.0 and reps < 200:
            tc = time.perf_counter()
            o.fstatnet(pts[sel])
            tcd += time.perf_counter() - tc
            ns += len(sel)
            reps += 1
        orc_grad = dict(points=len(sel),
                        max_rel_dF=float(np.max(np.abs(out_o[:, 5] - out[sel, 5]) / np.abs(out_o[:, 5]))),
                        max_grad_diff_per_cell_over_F=float(np.max(np.abs(g_o - g[sel]) * cells[sel] / F0[sel, None])))
        cpu_line = dict(value=round(ns / tcd, 2), unit="points/s", cores=threads, kind="port",
                        sample=f"{reps} x the {len(sel)} points of one segment with gradient (FP64 oracle Fstatnet, "
                               f"OpenMP over points; {tcd:.1f} s wall)", cpu_model=cpu_model(),
                        threads_env=threads_env(), host_cpus=len(os.sched_getaffinity(0)))
    nifo, N = 2, 86400
    fps = measured_flops("k_fstatnet_part")
    rl = None
    if fps:
        # FP64 vector roofline: counter-measured FLOPs per (point, sample) x samples per step over
        # the wall time of a step (device calls of all segments and their small copies)
        ach = fps * npts * nifo * N / (dt / steps) / 1e12
        rl = dict(bound="fp64-valu", kernel="k_fstatnet_part", achieved=round(ach, 2), peak=78.6, unit="TFLOP/s",
                  frac=round(ach / 78.6, 4), traffic=None, flops_per_sample=fps)
    cfg = {"workload": "differentiation: 2 detectors (H1, L1), N=86400, fftpad 2, 1024 injections (16 noise "
                       "realisations x 64, h0 in [0.05, 0.3] sigma), Fstatnet + dF/d(omega, omegadot, delta, alpha) "
                       "at one candidate point within phase-metric mismatch 0.25 of each",
           "cell": "1 / sqrt(G_kk) of the phase metric G at the injection: one radian rms of phase per parameter",
           "N": N, "detectors": nifo, "points": npts}
    line = base_line(args, v, dt, cfg, dtype="fp64", unit="points/s", steps=steps, warmup=warmup,
                     metric="Fstatnet evaluations with gradient per second on 1 MI355X")
    line.update({
        "values_only_points_per_s": round(v_nograd, 1), "samples_per_s": round(v * nifo * N, 1),
        "gradient_check": {
            "fd_step": "1e-3 cell per parameter, central differences",
            "max_abs_grad_minus_fd_per_cell_over_F": float(fd_err.max()),
            "median_abs_grad_minus_fd_per_cell_over_F": float(np.median(fd_err)),
            "vs_fp64_oracle": orc_grad},
        "ascent": {"steps": 20, "evaluations_per_point": nev, "seconds": round(t_asc, 3),
                   "F_median": float(np.median(Fa)), "param_error": err_summary(ea)},
        "simplex": {"tol": 1e-7, "first_steps": "two cells", "evaluations_per_point_mean":
                    round(float(nev_s.mean()), 1), "converged": int(conv_s.sum()), "seconds": round(t_nm, 3),
                    "F_median": float(np.median(Fs)), "param_error": err_summary(es)},
        "ascent_vs_simplex": {"F_rel_diff_median": float(np.median(rel)), "F_rel_diff_min": float(rel.min()),
                              "ascent_F_ge_simplex_minus_1e-6": int(np.sum(rel >= -1e-6)),
                              "ascent_F_ge_simplex_minus_1e-3": int(np.sum(rel >= -1e-3))},
        "start": {"F_median": float(np.median(F0)), "param_error": err_summary(e0)},
        "mads": {"points": nm, "rounds_mean": round(float(rounds_m.mean()), 1), "seconds": round(t_mads, 3),
                 "F_median": float(np.median(-outm[:, 5])), "param_error": err_summary(em)},
        "truth_h0_range": [float(h0.min()), float(h0.max())],
        "roofline": rl, "cpu_baseline": cpu_line})
    h.close()
    return line


def bench_dropin(args, nsky=64):
    """The drop-in path as a reference search() would drive it: pgw_job_core once per sky point
    (search/network/src-cpu/jobcore.c:107-120), each call returning its records, on the network
    workload; beside it pgw_job_batch over the same sky points.  FP64 and FP32 handles."""
    seg = workload(args, "network")
    rng = abi.make_range(seg)
    opts = abi.make_opts(trl=20.0)
    sky = synth.sky_list(seg)[:nsky]
    out = {}
    for name, prec in (("fp64", abi.PREC_FP64), ("fp32", abi.PREC_FP32)):
        h = pgw.Handle(seg, prec)
        tune_handle(args, h, 0, seg, opts, rng, 1)
        for pm, mm, nn in sky[:2]:
            h.job_core(pm, mm, nn, rng, opts)
        t0 = time.perf_counter()
        fn, rec1 = 0, []
        for pm, mm, nn in sky:
            r, f = h.job_core(pm, mm, nn, rng, opts)
            fn += f
            rec1.append(r)
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        fb = 0
        recb = []
        for i in range(0, len(sky), 16):
            r, f = h.job_batch(sky[i:i + 16], rng, opts)
            fb += f
            recb.append(r)
        tb = time.perf_counter() - t0
        same = bool(np.array_equal(np.concatenate(rec1), np.concatenate(recb)))
        out[name] = {"job_core_templates_per_s": round(fn / t1, 1), "job_core_ms_per_sky_point": round(1e3 * t1 / nsky, 4),
                     "job_batch_templates_per_s": round(fb / tb, 1), "records_equal": same}
        print(f"drop-in {name}: job_core {fn / t1:.0f} templates/s, job_batch {fb / tb:.0f}", file=sys.stderr,
              flush=True)
        h.close()
    out["workload"] = f"network, first {nsky} sky points x 512 spindowns"
    return out


def secondary(args):
    """Summaries of the other configurations (no CPU leg; one pass each) and the drop-in path."""
    res = {}
    for name in ("plain", "network", "precision"):
        ln = bench_templates(args, name, steps=None, warmup=1, cpu=False)
        f64 = ln["fp64"] or {}
        res[name] = {"config": ln["config"]["workload"], "fp32_templates_per_s": ln["value"],
                     "fp64_templates_per_s": f64.get("value"), "speedup_fp32_vs_fp64": ln["speedup_fp32_vs_fp64"],
                     "ms_per_sky_point_fp32": ln["ms_per_sky_point"], "ms_per_sky_point_fp64": f64.get("ms_per_sky_point"),
                     "templates_timed": ln["templates"], "accuracy_fp32_vs_fp64": ln["accuracy_fp32_vs_fp64"],
                     "roofline_fp32": ln["roofline"], "roofline_fp64": f64.get("roofline"),
                     "tuning": ln["config"]["tuning"]}
    res["dropin"] = bench_dropin(args)
    d = bench_differentiation(args, steps=5, warmup=1, cpu=False, mads_points=32)
    res["differentiation"] = {k: d[k] for k in ("value", "unit", "values_only_points_per_s", "gradient_check",
                                                "ascent", "simplex", "ascent_vs_simplex", "start", "mads", "roofline")}
    res["differentiation"]["config"] = d["config"]["workload"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sky-per-step", type=int, default=None)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workload", default="batching",
                    choices=["network", "plain", "precision", "batching", "differentiation"])
    ap.add_argument("--no-secondary", action="store_true",
                    help="only the headline workload (no other_workloads summary)")
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
        line = bench_templates(args, args.workload, steps=args.steps, per_step=args.sky_per_step)
    elif args.workload == "batching":
        line = bench_batching(args)
    else:
        line = bench_differentiation(args)
    if args.workload == "batching" and not args.no_secondary:
        line["other_workloads"] = secondary(args)
    # streaming copy / read / write kernels over 1 GiB: the bandwidth this device reaches (the
    # roofline's peak stays the published 8 TB/s).  Run after the timed work: the small plain
    # workload measured up to 20 % slower in a process that had run them first.
    hbm = pgw.hbm_probe(1 << 30, reps=5)
    print(f"hbm probe (GB/s): {hbm}", file=sys.stderr, flush=True)
    annotate_stream(line, hbm)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
