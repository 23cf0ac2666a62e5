#!/usr/bin/env python3
"""Called sites per second of the SOAPsnp calling path on one MI355X.

Workload (default): the uneven-depth configuration -- 24 seeded contigs with
lengths log-uniform in 1-20 Mb (184 Mb at seed 2024), 100 bp pairs, depth per
10 kb block ~ Gamma(4, 7.5), 1% repeat-like blocks at depth 200-255 with 1-3
hits per read, 2% uncovered blocks -- generated in memory by gsnp_amd's seeded
generator and calibrated on the device.  Every contig's reads stay resident in
HBM.

A step is one calling pass over every contig: site matrix -> depth classes ->
fused sort + FP64 likelihood + posterior -> .cns rows formatted on the device
and copied to pinned host memory, each span's copy running beside the next
span's kernels.  The timed region starts with the reads in HBM and ends with
every row in host memory.  Beside it, in the same run: the per-window (plain)
form with one synchronous copy per window, the device-only time (rows left in
HBM), the -C chunks instead of rows, per-stage HIP-event times with the
roofline of the dominant stage, the CPU oracle on a sample of the same
workload (one thread, and window-parallel on the host cores), and the adjoint
leg on the 2 Mb miscalibrated workload.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
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
WORKLOADS = {
    1: "100 kb reference, 20x 100 bp reads",
    2: "10 Mb reference, 30x 100 bp pairs (insert 500+-50), 3% 2-hit reads, 1% known SNPs, refine mode",
    3: "24 contigs log-uniform 1-20 Mb, 100 bp pairs, depth per 10 kb ~ Gamma(4, 7.5), 1% repeat-like blocks "
       "at depth 200-255 with 1-3 hits, 2% uncovered blocks",
    4: "2 Mb reference, 30x 100 bp reads, miscalibrated qualities (delta 3 above q30, cycle offset to 2)",
    5: "248 Mb reference, 30x 150 bp pairs (insert 400+-40), cycle-decaying qualities, rank-sum test",
}
TRAFFIC_PROFILE = {3: ROOT / "profiles" / "fused_stage_c3_hbm_traffic.md"}


def host_share() -> int:
    """Host threads this run may use: the box's CPU share (OMP_NUM_THREADS)
    when set, else the machine's."""
    n = os.cpu_count() or 1
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    return max(1, min(n, omp) if omp > 0 else n)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def device_sync() -> None:
    """hipDeviceSynchronize through the HIP runtime libgsnp.so loaded."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize: {rc}")


def measured_traffic(config: int, scale: float):
    """HBM bytes per pass of the fused stage on this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated);
    counters cannot be read inside the timed run."""
    prof = TRAFFIC_PROFILE.get(config)
    if prof is None or scale != 1.0 or not prof.exists():
        return None, None
    for line in prof.read_text().splitlines():
        if line.startswith("| fused stage total"):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            return int(float(cols[1]) * 1e6), prof.relative_to(ROOT).as_posix()
    return None, None


def timed_passes(job, steps: int, warmup: int, per: int, mode: int):
    for _ in range(warmup):
        job.run_pass(batched=per, emit_text=mode)
    device_sync()
    t0 = time.perf_counter()
    nb = ns = 0
    for _ in range(steps):
        nb, ns = job.run_pass(batched=per, emit_text=mode)
    device_sync()
    return (time.perf_counter() - t0) / steps, nb, ns


def cpu_leg(args, work: Path, lens: list[int], sites_total: int) -> dict:
    """The CPU oracle on a sample of the same workload: a few whole contigs
    written as files; one thread over its first windows (about 10 s), and the
    window-parallel run over the whole sample on the host cores."""
    import gsnp_amd

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib

    oracle_lib.ensure_built()
    order = sorted(range(len(lens)), key=lambda i: lens[i])
    pick, tot = [], 0
    for i in order:  # the smallest contigs until the sample holds args.cpu_mb Mb
        if tot >= args.cpu_mb * 1e6:
            break
        pick.append(i)
        tot += lens[i]
    pick.sort()
    pre = work / "cpu_sample"
    t0 = time.time()
    scale = args.scale
    if tot > 2 * args.cpu_mb * 1e6:
        # one contig far larger than the sample (the 248 Mb workload): the same
        # generator at a smaller scale, which keeps its depth, read and quality
        # model, gives a contig of about cpu_mb
        scale = args.scale * args.cpu_mb * 1e6 / tot
        pick = [0]
        tot = gsnp_amd.synth_contigs(args.config, args.seed, scale)[0]
    gsnp_amd.synth_subset(args.config, args.seed, scale, pick, pre)
    print(f"[bench] cpu sample: contigs {pick} at scale {scale:.4g} ({tot / 1e6:.1f} Mb) written in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    base = [str(oracle_lib.ORACLE_CLI), "-i", f"{pre}.soap", "-d", f"{pre}.fa", "-L", str(args.read_len), "-R"]
    if args.config == 5:
        base += ["-u"]
    snp = Path(f"{pre}.snp")
    if snp.exists():
        base += ["-s", str(snp), "-2"]
    res = subprocess.run(base + ["-o", str(work / "cpu1.cns"), "-X", str(args.cpu_sites)], capture_output=True,
                         text=True, check=True)
    one = json.loads(res.stdout.strip().splitlines()[-1])
    threads = host_share()
    res = subprocess.run(base + ["-o", str(work / "cpuN.cns"), "-J", str(threads)], capture_output=True, text=True,
                         check=True)
    many = json.loads(res.stdout.strip().splitlines()[-1])
    (work / "cpu1.cns").unlink(missing_ok=True)
    cli = cli_leg(args, pre, work / "cpuN.cns", many["sites"])
    (work / "cpuN.cns").unlink(missing_ok=True)
    names = ", ".join(f"chr{i + 1}" for i in pick)
    what = (f"contigs {names}" if scale == args.scale else
            f"one contig of the same generator at scale {scale:.4g}")
    return {"value": round(many["sites"] / many["seconds_call"], 1), "unit": "sites/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{what} ({many['sites']} sites, {many['windows']} windows) of the same workload as "
                      f"files; the CPU oracle calling windows on {threads} threads, .cns text written; call stage "
                      f"only (parsing excluded)",
            "single_thread": {"value": round(one["sites"] / one["seconds_call"], 1), "unit": "sites/s", "cores": 1,
                              "sample": f"first {one['sites']} sites ({one['windows']} windows) of the same sample"},
            "gpu_cli": cli}


def cli_leg(args, pre: Path, want: Path, sites: int) -> dict:
    """The drop-in command line end to end on the CPU sample's files: parse
    the FASTA and SOAP files, calibrate, upload, call in batched spans and
    write the .cns text; its output is compared with the CPU oracle's."""
    import filecmp

    import gsnp_amd

    exe = gsnp_amd.LIB_PATH.parent / "gsnp"
    out = pre.parent / "cli.cns"
    cmd = [str(exe), "-i", f"{pre}.soap", "-d", f"{pre}.fa", "-o", str(out), "-L", str(args.read_len),
           "-B", str(args.per if args.per > 0 else 1 << 20)]
    if args.config == 5:
        cmd += ["-u"]
    snp = Path(f"{pre}.snp")
    if snp.exists():
        cmd += ["-s", str(snp), "-2"]
    t0 = time.perf_counter()
    subprocess.run(cmd, check=True, capture_output=True)
    wall = time.perf_counter() - t0
    same = filecmp.cmp(out, want, shallow=False)
    out.unlink(missing_ok=True)
    print(f"[bench] gsnp command line on the CPU sample: {wall:.2f}s end to end, output equals the oracle's: {same}",
          file=sys.stderr, flush=True)
    return {"wall_s": round(wall, 3), "sites": sites, "sites_per_s": round(sites / wall, 1), "matches_oracle": same,
            "note": "the gsnp binary end to end (parse, calibrate, upload, call in batched spans, write .cns) on the "
                    "CPU sample's files, output compared with the oracle's"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--per", type=int, default=32, help="windows per batched span (-1: one span per contig)")
    ap.add_argument("--plain-steps", type=int, default=1, help="timed passes of the per-window form")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-mb", type=float, default=12.0, help="Mb of the workload in the CPU sample")
    ap.add_argument("--cpu-sites", type=int, default=16 * 262144, help="sites of the single-thread CPU run")
    ap.add_argument("--tune", action="store_true", help="autotune the fused launch shape on the largest contig "
                                                         "first (default: the library's default shape)")
    ap.add_argument("--no-tune", action="store_true", help=argparse.SUPPRESS)  # the default; kept for old commands
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg (2 Mb miscalibrated workload)")
    ap.add_argument("--no-columns", action="store_true", help="skip the -C leg (column chunks instead of rows)")
    ap.add_argument("--no-stages", action="store_true", help="skip the per-stage timing passes")
    ap.add_argument("--adjoint-evals", type=int, default=10)
    ap.add_argument("--fd-entries", type=int, default=256)
    ap.add_argument("--refit-steps", type=int, default=50)
    ap.add_argument("--workdir", default=None)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "gsnp_amd runs on one GPU: --gpus must be 1 (sharding across GPUs is designed "
                                   "in the documentation, not built)"}))
        return 2

    import gsnp_amd

    if not gsnp_amd.LIB_PATH.exists():
        import __graft_entry__

        __graft_entry__.build()

    work = Path(args.workdir) if args.workdir else Path(tempfile.mkdtemp(prefix="gsnp_bench_"))
    work.mkdir(parents=True, exist_ok=True)
    flags = {"rank_sum_mode": 1} if args.config == 5 else {}
    if args.config == 2:
        flags["refine_mode"] = 1
    t0 = time.time()
    job = gsnp_amd.Job(synth=(args.config, args.seed, args.scale), **flags)
    args.read_len = job.read_len
    contigs = [job.info(i) for i in range(len(job))]
    sites = sum(c[1] for c in contigs)
    reads = sum(c[2] for c in contigs)
    lens = gsnp_amd.synth_contigs(args.config, args.seed, args.scale)
    print(f"[bench] config {args.config} x{args.scale}: {len(contigs)} contigs, {sites} sites, {reads} reads "
          f"generated, calibrated and uploaded in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    per = args.per
    tuned = None
    if args.tune:
        t0 = time.time()
        big = max(range(len(contigs)), key=lambda i: contigs[i][1])
        job.select(big)
        span = contigs[big][1] if per < 0 else min(contigs[big][1], per * gsnp_amd.WIN)
        tuned = job.autotune(0, span)
        print(f"[bench] autotuned fused launch on {contigs[big][0]} in {time.time() - t0:.1f}s: {tuned}",
              file=sys.stderr, flush=True)

    # the headline: rows in pinned host memory
    job.counters(reset=True)
    secs, text_bytes, ns = timed_passes(job, args.steps, args.warmup, per, 2)
    assert ns == sites
    value = sites / secs
    cnt = job.counters(reset=True)
    recs = cnt["records"] // max(1, args.steps + args.warmup)
    fixups = {"host_fixups_per_pass": cnt["host_fixups"] // max(1, args.steps + args.warmup),
              "device_normal_per_pass": cnt["device_normal"] // max(1, args.steps + args.warmup)}
    print(f"[bench] batched, rows to host: {secs * 1e3:.1f} ms/pass, {value:.4g} sites/s", file=sys.stderr,
          flush=True)
    d_secs, _, _ = timed_passes(job, args.steps, 1, per, 1)
    print(f"[bench] batched, rows left in HBM: {d_secs * 1e3:.1f} ms/pass", file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        p_secs, p_bytes, _ = timed_passes(job, args.plain_steps, 0, 0, 2)
        assert p_bytes == text_bytes, (p_bytes, text_bytes)
        plain = {"mode": "plain (per-window path, one synchronous copy of each window's rows)",
                 "value": round(sites / p_secs, 1), "ms_per_step": round(p_secs * 1e3, 3), "steps": args.plain_steps,
                 "batched_over_plain": round(p_secs / secs, 3)}
        print(f"[bench] plain: {p_secs * 1e3:.1f} ms/pass", file=sys.stderr, flush=True)

    # per-stage HIP-event times (eager launches, rows left in HBM)
    stage_ms, launches = {}, {}
    if not args.no_stages:
        job.profile(enable=True)
        for _ in range(args.steps):
            job.run_pass(batched=per, emit_text=1)
        prof = job.profile(enable=False)
        stage_ms = {k: v[0] / args.steps for k, v in prof.items() if v[1]}
        launches = {k: v[1] / args.steps for k, v in prof.items() if v[1]}
    # roofline of the dominant stage: the fused sort + likelihood + posterior
    # (4 B per record read once, 64 B of per-site counters and reference in,
    # 16 B of calls out: 4n + 80 per site)
    fused_bytes = 4 * recs + 80 * sites
    roofline = None
    if stage_ms.get("fused"):
        t_f = stage_ms["fused"] / 1e3
        achieved = fused_bytes / t_f / 1e9
        traffic, src = measured_traffic(args.config, args.scale)
        t_call_gbs = fused_bytes / secs / 1e9
        roofline = {"bound": "hbm", "kernel": "k_fused<S>: sort + likelihood + posterior, one launch per depth class",
                    "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": src,
                    "bytes_per_pass": int(fused_bytes), "ms_per_pass": round(t_f * 1e3, 4),
                    "spans_per_pass": launches.get("fused"),
                    "note": "one k_fused launch per depth class and span; achieved is over their HIP-event time",
                    "t_call": {"achieved": round(t_call_gbs, 2), "frac": round(t_call_gbs / HBM_PEAK_GBS, 4),
                               "note": "the same 4n + 80 bytes per site over the whole timed pass, rows to host"}}

    # the window build (reads -> records), as its own figure: 2 B per
    # observed base read (base and quality), 4 B per record written, 12 B of
    # metadata per read
    site_build = None
    if stage_ms.get("site_build"):
        sb_bytes = 6 * recs + 12 * reads
        site_build = {"ms_per_pass": round(stage_ms["site_build"], 4), "bytes_per_pass": int(sb_bytes),
                      "achieved": round(sb_bytes / (stage_ms["site_build"] / 1e3) / 1e9, 2), "unit": "GB/s",
                      "note": "2 B per base and 4 B per record over the reads' 12 B of metadata each; k_read_ranges + "
                              "k_site_build"}

    columns = None
    if not args.no_columns:
        c_secs, c_bytes, _ = timed_passes(job, args.steps, 1, per, 3)
        columns = {"metric": "called sites/sec with -C column chunks in host memory", "value": round(sites / c_secs, 1),
                   "unit": "sites/s", "ms_per_step": round(c_secs * 1e3, 3), "chunk_bytes": c_bytes}
        print(f"[bench] -C: {c_secs * 1e3:.1f} ms/pass", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_leg(args, work, lens, sites)
    job.close()
    adjoint = None if args.no_adjoint else adjoint_leg(args, work)

    line = {
        "metric": "called sites/sec at bit-exact .cns",
        "value": round(value, 1),
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded generator, config {args.config}, seed {args.seed}, scale {args.scale}), in memory",
        "config": {"workload": WORKLOADS.get(args.config, f"config {args.config}") +
                   ("" if args.scale == 1.0 else f", scaled x{args.scale}"),
                   "contigs": len(contigs), "sites": sites, "reads": reads, "records": recs,
                   "mode": "batched", "windows_per_span": per, "text_bytes": text_bytes,
                   "timed_region": "reads resident in HBM -> .cns rows in pinned host memory", "flags": flags,
                   "autotuned": tuned},
        "device_only": {"value": round(sites / d_secs, 1), "ms_per_step": round(d_secs * 1e3, 3),
                        "note": "the same pass with the rows left in HBM"},
        "other_mode": plain,
        "roofline": roofline,
        "stage_ms_per_step": {k: round(v, 4) for k, v in stage_ms.items()},
        "site_build": site_build,
        "rank_sum": fixups,
        "cpu_baseline": cpu,
    }
    if columns:
        line["columns"] = columns
    if adjoint:
        line["adjoint"] = adjoint
    print(json.dumps(line), flush=True)
    return 0


def adjoint_leg(args, work: Path) -> dict:
    """Gradient evaluations per second of the adjoint on the 2 Mb miscalibrated
    workload (30x, 100 bp), its central-difference check on random table
    entries, and the refit judged by the plain calling path and against the
    miscalibration the generator used."""
    import math

    import numpy as np

    import gsnp_amd
    from gsnp_amd.adjoint import Adjoint, calls_changed, central_differences, refit, truth_distance, write_matrix

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
                   "genotype_calls_changed_vs_plain": ng, "consensus_quality_changed_vs_plain": nq,
                   "distance_to_generator_truth": {"counting_table": truth_distance(p0, rl),
                                                   "refit_table": truth_distance(p1, rl)},
                   "truth_metric": "RMS over q 10-40 and every cycle of log10(row error rate) - log10 of the "
                                   "generator's 10^(-(q - delta)/10)"}
        recs, sites = adj.n_records, adj.n_sites
    finally:
        job.close()
    if fit is not None:
        # the same refit where the calling model matches the generator (no
        # dependency adjustment: -p 1 -g 1, independent errors)
        job = gsnp_amd.Job(fa, so, read_length=rl, pcr_dependency=1.0, global_dependency=1.0)
        try:
            adj = Adjoint(job)
            p1i, trace_i = refit(adj, p0, math.log(job._p.het_novel_r), steps=args.refit_steps)
            fit["independent_errors_model"] = {
                "objective_start": trace_i[0], "objective_end": trace_i[-1],
                "distance_to_generator_truth": {"counting_table": truth_distance(p0, rl),
                                                "refit_table": truth_distance(p1i, rl)}}
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
