#!/usr/bin/env python3
"""Called sites per second of the SOAPsnp calling path on one MI355X.

Workload: the 10 Mb plain-GPU configuration (seeded reference, 30x 100 bp
pairs, 3% two-hit reads, 1% known-SNP prior file, refine mode on), generated
by gsnp_amd's seeded generator.  A step is one calling pass over the whole
chromosome: site matrix -> per-site sort -> FP64 likelihood -> posterior ->
.cns rows formatted in device memory, with the reads already resident in HBM.
The timed region excludes parsing, calibration and the host copy of the text.

Prints one JSON line (metric, value, roofline of the dominant kernel from HIP
events, and the CPU oracle timed on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
TRAFFIC_PROFILE = ROOT / "profiles" / "calling_path_10Mb_hbm_traffic.md"
# kernels of each stage as named in the traffic profile
STAGE_KERNELS = {"fused": ("k_bucket", "k_fused<"), "site_build": ("k_read_ranges", "k_site_build"),
                 "format": ("k_rowlen", "k_gtext", "k_scan_blocks", "k_format")}


ISSUE_PROFILE = ROOT / "profiles" / "calling_kernels_10Mb_issue.md"


def measured_issue() -> dict | None:
    """Share of SIMD cycles the fused kernels spend issuing VALU on the 10 Mb
    workload, from the committed SQ counter summary (its last column)."""
    if not ISSUE_PROFILE.exists():
        return None
    out = {}
    for line in ISSUE_PROFILE.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) == 7 and cols[0].startswith("`k_fused<") and cols[6].endswith("%"):
            out[cols[0].strip("`")] = float(cols[6].rstrip(" %")) / 100.0
    return out or None


def measured_traffic(stage: str):
    """HBM bytes per pass of a stage on the 10 Mb workload (config 2, scale 1),
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (measured MB
    read + written, copy-calibrated); None when the summary is absent."""
    if stage not in STAGE_KERNELS or not TRAFFIC_PROFILE.exists():
        return None
    total = 0.0
    found = False
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) != 5 or not cols[0].startswith("`"):
            continue
        name = cols[0].strip("`")
        if any(name == k or (k.endswith("<") and name.startswith(k)) for k in STAGE_KERNELS[stage]):
            total += (float(cols[3]) + float(cols[4])) * 1e6
            found = True
    return int(total) if found else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--mode", choices=["batched", "plain"], default="batched")
    ap.add_argument("--per", type=int, default=0, help="windows per batched span (0: tuned default)")
    ap.add_argument("--compare", action="store_true", help="also time the other mode")
    ap.add_argument("--cpu-sites", type=int, default=2 * 262144, help="sites in the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default fused launch shape")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg (2 Mb miscalibrated workload)")
    ap.add_argument("--no-columns", action="store_true", help="skip the -C leg (column chunks instead of text)")
    ap.add_argument("--adjoint-evals", type=int, default=10)
    ap.add_argument("--fd-entries", type=int, default=256)
    ap.add_argument("--refit-steps", type=int, default=50)
    ap.add_argument("--workdir", default=None)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "the calling path runs on one GPU; --gpus must be 1 (sharding across GPUs is "
                                   "designed, not built)"}))
        return 2

    import gsnp_amd

    if not gsnp_amd.LIB_PATH.exists():
        import __graft_entry__

        __graft_entry__.build()

    work = Path(args.workdir) if args.workdir else Path(tempfile.mkdtemp(prefix="gsnp_bench_"))
    work.mkdir(parents=True, exist_ok=True)
    pre = work / f"c{args.config}"
    t0 = time.time()
    n_reads, rl = gsnp_amd.synth(args.config, args.seed, args.scale, pre)
    fa, so = f"{pre}.fa", f"{pre}.soap"
    snp = f"{pre}.snp" if Path(f"{pre}.snp").exists() else None
    flags = dict(read_length=rl, refine_mode=1 if snp else 0)
    print(f"[bench] generated config {args.config}: {n_reads} reads in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    t0 = time.time()
    job = gsnp_amd.Job(fa, so, dbsnp=snp, **flags)
    name, chr_len, chr_reads = job.info(0)
    job.select(0)
    print(f"[bench] parsed + calibrated + uploaded {name} ({chr_len} bp, {chr_reads} reads) in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    per = args.per or int(os.environ.get("GSNP_BATCH_WINDOWS", "0")) or (chr_len // gsnp_amd.WIN + 1)
    tuned = None
    if args.mode == "batched" and not args.no_tune:
        t0 = time.time()
        tuned = job.autotune(0, min(chr_len, per * gsnp_amd.WIN))
        print(f"[bench] autotuned fused launch in {time.time() - t0:.1f}s: {tuned}", file=sys.stderr, flush=True)

    def timed(mode: str):
        """Seconds per pass with the stage timers off (the batched span and its
        formatting replay as captured graphs), then the same number of passes
        with the stage timers on (eager launches) for the per-stage times."""
        batched = per if mode == "batched" else 0
        for _ in range(args.warmup):
            job.call(batched=batched)
        t_start = time.perf_counter()
        nbytes = 0
        for _ in range(args.steps):
            nbytes = job.call(batched=batched)
        t_end = time.perf_counter()
        job.records(reset=True)
        job.profile(enable=True)
        for _ in range(args.steps):
            job.call(batched=batched)
        prof = job.profile(enable=False)
        recs = job.records(reset=True) // max(1, args.steps)
        secs = (t_end - t_start) / args.steps
        return secs, prof, recs, nbytes

    secs, prof, recs, text_bytes = timed(args.mode)
    sites_per_s = chr_len / secs
    other = None
    if args.compare:
        o_mode = "plain" if args.mode == "batched" else "batched"
        o_secs, _, _, _ = timed(o_mode)
        other = {"mode": o_mode, "value": chr_len / o_secs, "ms_per_step": o_secs * 1e3}

    # dominant stage and its algorithmic bytes per launch
    steps = args.steps
    stage_ms = {k: v[0] / steps for k, v in prof.items() if v[1]}
    launches = {k: v[1] / steps for k, v in prof.items() if v[1]}
    dom = max(stage_ms, key=stage_ms.get)
    sites = chr_len
    packed = chr_reads * rl  # one packed byte per read base
    algo = {
        # reads (packed bases + 8 B position/metadata) in, records + 60 B of counters per site out
        "site_build": packed + 8 * chr_reads + 4 * recs + 60 * sites,
        # records read and written back, record counts read
        "sort": 8 * recs + 4 * sites,
        # records and count in, 10 FP64 log-likelihoods out (4n + 84 per site)
        "likelihood": 4 * recs + 84 * sites,
        # counters, reference code and likelihoods in; calls out
        "posterior": 121 * sites + 15 * sites,
        "rank": 0,
        # call columns and counters in, .cns text out
        "format": 72 * sites + text_bytes,
        "recycle": 0,
        "fused": 4 * recs + 80 * sites,
    }
    a_bytes = algo.get(dom, 0) / max(1e-9, launches[dom])
    t_launch = stage_ms[dom] / 1e3 / max(1e-9, launches[dom])
    achieved = a_bytes / t_launch / 1e9
    # traffic: measured HBM bytes per pass of the stage, from the committed
    # counter runs over this same workload (default launch shape); counters
    # cannot be read inside the timed run
    traffic = measured_traffic(dom) if (args.config == 2 and args.scale == 1.0 and args.mode == "batched") else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": TRAFFIC_PROFILE.relative_to(ROOT).as_posix() if traffic else None,
                "bytes_per_launch": int(a_bytes), "ms_per_launch": round(t_launch * 1e3, 4)}
    # what bounds the dominant stage instead: VALU issue (committed counters)
    issue = measured_issue() if (dom == "fused" and args.config == 2 and args.scale == 1.0) else None
    compute_bound = {"bound": "valu_issue", "valu_issue_frac": issue,
                     "source": ISSUE_PROFILE.relative_to(ROOT).as_posix()} if issue else None
    # the calling kernels together at B_site = 4n + 80 per called site
    call_ms = sum(stage_ms.get(k, 0) for k in ("sort", "likelihood", "posterior", "rank"))
    call_gbs = (4 * recs + 80 * sites) / (call_ms / 1e3) / 1e9 if call_ms else None

    # -C leg: the same pass with each window's column chunks encoded on the
    # device and copied to host memory (no file writes) instead of the text
    columns = None
    if not args.no_columns:
        batched = per if args.mode == "batched" else 0
        job.columns(batched=batched)
        job.profile(enable=True)
        t0 = time.perf_counter()
        c_bytes = 0
        for _ in range(args.steps):
            c_bytes = job.columns(batched=batched)
        c_secs = (time.perf_counter() - t0) / args.steps
        c_prof = job.profile(enable=False)
        enc_ms = c_prof["columns"][0] / args.steps
        columns = {"metric": "called sites/sec with -C column chunks", "value": round(chr_len / c_secs, 1),
                   "unit": "sites/s", "ms_per_step": round(c_secs * 1e3, 3), "encoder_ms_per_step": round(enc_ms, 4),
                   "chunk_bytes": c_bytes,
                   "note": "encoder = the device kernels of the chunks; the step also copies them to host memory"}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib

        oracle_lib.ensure_built()
        out = work / "cpu.cns"
        cmd = [str(oracle_lib.ORACLE_CLI), "-i", so, "-d", fa, "-o", str(out), "-L", str(rl), "-X", str(args.cpu_sites),
               "-R"]
        if snp:
            cmd += ["-s", snp, "-2"]
        res = subprocess.run(cmd, capture_output=True, text=True, check=True)
        st = json.loads(res.stdout.strip().splitlines()[-1])
        cpu = {"value": round(st["sites"] / st["seconds_call"], 1), "unit": "sites/s", "cores": 1, "kind": "port",
               "sample": f"first {st['sites']} sites ({st['windows']} windows) of the same workload, "
                         f"single-threaded CPU oracle, call stage only"}

    adjoint = None if args.no_adjoint else adjoint_leg(args, work)

    line = {
        "metric": "called sites/sec at bit-exact .cns",
        "value": round(sites_per_s, 1),
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, config %d)" % args.config,
        "config": {"workload": "10 Mb reference, 30x 100 bp pairs, 3% 2-hit, 1% known SNPs, refine on"
                   if args.config == 2 else f"config {args.config} x{args.scale}",
                   "sites": chr_len, "reads": chr_reads, "records": recs, "mode": args.mode,
                   "windows_per_span": per if args.mode == "batched" else 1, "text_bytes": text_bytes,
                   "autotuned": tuned},
        "roofline": roofline,
        "compute_bound": compute_bound,
        "call_kernels_gbs": round(call_gbs, 2) if call_gbs else None,
        "stage_ms_per_step": {k: round(v, 4) for k, v in stage_ms.items()},
        "cpu_baseline": cpu,
    }
    if other:
        line["other_mode"] = other
    if columns:
        line["columns"] = columns
    if adjoint:
        line["adjoint"] = adjoint
    print(json.dumps(line), flush=True)
    job.close()
    return 0


def adjoint_leg(args, work: Path) -> dict:
    """Gradient evaluations per second of the adjoint on the 2 Mb miscalibrated
    workload (30x, 100 bp), its central-difference check on random table
    entries, and the refit judged by the plain calling path."""
    import math

    import numpy as np

    import gsnp_amd
    from gsnp_amd.adjoint import Adjoint, calls_changed, central_differences, refit, write_matrix

    pre = work / "c4"
    t0 = time.time()
    _, rl = gsnp_amd.synth(4, args.seed, 1.0, pre)
    fa, so = f"{pre}.fa", f"{pre}.soap"
    p0 = gsnp_amd.matrix_gen(fa, so, read_length=rl)
    job = gsnp_amd.Job(fa, so, read_length=rl)
    try:
        adj = Adjoint(job)
        lh = math.log(job._p.het_novel_r)
        print(f"[bench] adjoint workload ready ({adj.n_sites} sites, {adj.n_records} records) in "
              f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        for _ in range(3):
            adj.eval(p0, lh)
        job.profile(enable=True)
        t_start = time.perf_counter()
        for _ in range(args.adjoint_evals):
            ll, grad, dh = adj.eval(p0, lh)
        t_eval = (time.perf_counter() - t_start) / args.adjoint_evals
        prof = job.profile(enable=False)
        k_ms = prof["adjoint"][0] / max(1, prof["adjoint"][1])
        # central differences (h = 1e-6) on random entries whose gradient is at
        # least 1e-3 of the largest (smaller ones sit below the differences'
        # rounding floor, about 1e-9 absolute)
        flat = grad.reshape(-1)
        nz = np.flatnonzero(np.abs(flat) >= 1e-3 * np.abs(flat).max())
        rng = np.random.default_rng(args.seed)
        pick = rng.choice(nz, size=min(args.fd_entries, nz.size), replace=False)
        fd, fd_het = central_differences(adj, p0, lh, pick, h=1e-6)
        rel = np.abs(fd - flat[pick]) / np.maximum(np.abs(flat[pick]), 1e-300)
        print(f"[bench] adjoint: {1 / t_eval:.1f} evals/s, finite differences on {len(pick)} entries done",
              file=sys.stderr, flush=True)
        fit = None
        if args.refit_steps > 0:
            p1, trace = refit(adj, p0, lh, steps=args.refit_steps)
            m0, m1 = work / "c4.p0.matrix", work / "c4.p1.matrix"
            write_matrix(m0, p0, rl)
            write_matrix(m1, p1, rl)
            c0, c1 = work / "c4.plain0.cns", work / "c4.plain1.cns"
            gsnp_amd.run(fa, so, c0, matrix_in=m0, read_length=rl)
            gsnp_amd.run(fa, so, c1, matrix_in=m1, read_length=rl)
            ng, nq = calls_changed(c0, c1)
            fit = {"steps": args.refit_steps, "objective_start": trace[0], "objective_end": trace[-1],
                   "genotype_calls_changed_vs_plain": ng, "consensus_quality_changed_vs_plain": nq}
        recs, sites = adj.n_records, adj.n_sites
    finally:
        job.close()
    # algorithmic bytes per evaluation: record keys read twice, four keyed FP64
    # contributions written and read back plus a rank, 80 B per site, 2 MB gradient
    algo = 2 * 4 * recs + 48 * recs + 80 * sites + 64 * 256 * 16 * 8
    return {"metric": "adjoint gradient evaluations/sec", "value": round(1 / t_eval, 2), "unit": "evals/s",
            "ms_per_eval": round(t_eval * 1e3, 3), "kernel_ms_per_eval": round(k_ms, 3),
            "workload": "2 Mb reference, 30x 100 bp, miscalibrated qualities (delta 3 above q30, cycle offset to 2)",
            "sites": sites, "records": recs, "loglik": ll, "dloglik_dlog_het": dh,
            "hbm_gbs_kernels": round(algo / (k_ms / 1e3) / 1e9, 2) if k_ms else None,
            "fd_entries": int(len(pick)), "fd_h": 1e-6, "fd_max_rel_err": float(rel.max()),
            "fd_het_rel_err": abs(fd_het - dh) / abs(dh) if dh else None, "refit": fit}


if __name__ == "__main__":
    sys.exit(main())
