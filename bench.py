#!/usr/bin/env python3
"""Benchmark of the pileup path on one MI355X.

A step is one pass of the batched form over the whole workload: every region
of the configuration piled (window kernel), its indel evidence sorted and
called, Pass 1 finalised — one pl_batch_run with the reads and contigs
already resident in HBM.  The default workload is the many-contig batching
configuration (3,000 log-normal contigs, ~60 Mb, 2x150 bp at 60x).

Prints ONE JSON line.  value = aligned bases piled per second (M/=/X read
bases whose locus lies in the region, counted once per region) over the K
timed steps.  Beside it: the plain per-region launch on the same workload,
the window kernel's HBM roofline fraction (algorithmic bytes / kernel time,
HIP events on the library's stream), and the CPU oracle timed on a bounded
sample on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "aligned bases piled/sec on 1 MI355X (batched/served vs plain); p99 latency"
WORKLOADS = {
    "c1": "one 100 kb contig, 2x150 bp at 40x, 200 SNVs + 20 indels",
    "c2": "one 5 Mb contig at 65% GC with 10 MAPQ-0 repeat copies, 2x150 bp at 100x",
    "c3": "3,000 contigs log-normal (median 8 kb, ~60 Mb), 2x150 bp at 60x, one persistent launch",
    "c4": "50 x 4 Mb contigs resident, 2x150 bp at 50x",
    "c5": "20 x 25 Mb contigs (60 chunk regions), 2x250 bp at 80x, diploid, hotspots",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_sources_sha256():
    """Hash of the window kernel's sources (a traffic profile is valid for these only)."""
    import hashlib
    h = hashlib.sha256()
    for f in ("pl_common.h", "pl_core.h", "pl_kernels.h", "pl_kernels.hip"):
        with open(os.path.join(ROOT, "pilon_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic(path, workload, window, block):
    """HBM bytes per k_window launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE profile, when it was taken on these sources and this launch."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return None, None
    sha = re.search(r"kernel_sources_sha256: (\w+)", text)
    cfg = re.search(r"launch: (\w+) window=(\d+) block=(\d+)", text)
    tb = re.search(r"traffic_bytes_per_launch: (\d+)", text)
    if not (sha and cfg and tb) or sha.group(1) != kernel_sources_sha256():
        return None, None
    if (cfg.group(1), int(cfg.group(2)), int(cfg.group(3))) != (workload, window, block):
        return None, None
    return int(tb.group(1)), os.path.relpath(path, ROOT)


def measured_copy_bandwidth():
    """The device's copy bandwidth, measured in a child process that loads
    only torch (a process that has loaded the engine's HIP runtime does not
    also get torch's): GB/s read plus written, or None."""
    import subprocess
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--measure-copy"], capture_output=True,
                             text=True, timeout=300)
        return float(out.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError):
        return None


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """A device-to-device copy of nbytes, timed with events over reps copies
    after one untimed; bytes read plus bytes written per second, in GB/s."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms / 1000.0) / 1e9


def regions_for(name, data, rids):
    """The regions Pilon would process: each contig split in chunks of at most
    10 Mb by the library's pl_contig_regions (GenomeFile.contigRegions,
    GenomeFile.scala:67-74)."""
    from pilon_amd.engine import contig_regions
    out = []
    for ci in range(data.n):
        for s, e in contig_regions(len(data.draft(ci))):
            out.append((ci, s, e, rids[ci]))
    return out


def read_bytes(soa):
    """Per-read algorithmic bytes: 16 B header + 4 B per CIGAR op + 2-bit bases + 1 B quals."""
    import ctypes as C
    n = soa.n_reads
    co = np.ctypeslib.as_array(C.cast(soa.cigar_off, C.POINTER(C.c_int64)), shape=(n + 1,))
    so = np.ctypeslib.as_array(C.cast(soa.seq_off, C.POINTER(C.c_int64)), shape=(n + 1,))
    L = np.diff(so)
    return int(16 * n + 4 * int(co[-1]) + int(((L + 3) // 4).sum()) + int(L.sum()))


class _ContigReads:
    """The synthetic reads of one contig, as the oracle takes a read set."""

    def __init__(self, data, ci):
        self.data, self.ci = data, ci

    def soa(self):
        return self.data.soa(self.ci)


def cpu_baseline(cfg, data, regions, seconds):
    """The oracle over the workload's regions in order, one region per task on
    a pool of host threads (regions are independent, as in Pilon's region
    loop), until `seconds` of wall time have passed; the regions started by
    then are finished and counted."""
    from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
    from oracle import oracle as O
    # a one-GPU box gives its process 16 CPUs, whatever the affinity mask
    # shows; the pool takes all of them (host_cpus reports the mask)
    host_cpus = len(os.sched_getaffinity(0))
    nth = max(1, min(16, host_cpus))
    O.lib()
    drafts = {}

    def one(k):
        ci, s, e, _ = regions[k]
        t = time.perf_counter()
        status, st = O.run_region_stats(cfg, drafts[ci], s, e, [_ContigReads(data, ci)], name=data.names[ci])
        assert status == 0
        return st["aligned_bases"], time.perf_counter() - t

    # contig bytes are fetched before the clock starts (a few per thread ahead)
    for ci, _, _, _ in regions[:4 * nth]:
        drafts[ci] = data.draft(ci)
    done, busy, nreg, nxt = 0, 0.0, 0, 0
    with ThreadPoolExecutor(nth) as pool:
        t0 = time.perf_counter()
        live = set()
        while nxt < len(regions) and len(live) < 2 * nth:
            live.add(pool.submit(one, nxt)); nxt += 1
        while live:
            fin, live = wait(live, return_when=FIRST_COMPLETED)
            for f in fin:
                b, dt = f.result()
                done += b; busy += dt; nreg += 1
            while nxt < len(regions) and len(live) < 2 * nth and time.perf_counter() - t0 < seconds:
                ci = regions[nxt][0]
                if ci not in drafts:
                    drafts[ci] = data.draft(ci)
                live.add(pool.submit(one, nxt)); nxt += 1
        wall = time.perf_counter() - t0
    out = {"value": done / wall, "unit": "bases/s", "cores": nth, "host_cpus": host_cpus, "kind": "port",
           "sample": f"oracle (C restatement of Pilon's path) on the first {nreg} of {len(regions)} regions, "
                     f"one region per task on {nth} threads: {done} aligned bases in {wall:.1f}s wall "
                     f"({done / busy / 1e6:.1f} M bases/s per thread)"}
    out["rows"] = cpu_rows(cfg, data, regions, nth, drafts)
    return out


def cpu_rows(cfg, data, regions, nth, drafts):
    """The oracle on the widened rows beside the device figures: VCF records
    and the 13 tracks with Pilon.bed of the first 200 regions (one region per
    task on nth threads), and the assembler's K-mer graph of the reads the
    K-mer leg uses (one thread)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    sample = regions[:200]
    for ci, _, _, _ in sample:
        if ci not in drafts:
            drafts[ci] = data.draft(ci)

    def one(k, flags):
        ci, s, e, _ = sample[k]
        r = O.run_region(cfg, drafts[ci], s, e, [_ContigReads(data, ci)], name=data.names[ci], vcf=flags)
        assert r["status"] == 0
        return r["vcf"].count("\n") if flags & O.VCF_ON else e - s + 1

    rows = {}
    for key, flags, unit in (("vcf", O.VCF_ON, "records/s"), ("tracks_bed", O.TRACKS | O.BED, "loci/s")):
        with ThreadPoolExecutor(nth) as pool:
            t0 = time.perf_counter()
            n = sum(pool.map(lambda k: one(k, flags), range(len(sample))))
            wall = time.perf_counter() - t0
        rows[key] = {"value": n / wall, "unit": unit, "cores": nth,
                     "sample": f"{len(sample)} regions, {n} {unit.split('/')[0]} in {wall:.2f}s wall (pileup included)"}
    ci = max(range(data.n), key=lambda i: len(data.draft(i)))
    b = data.batch(ci)
    b = b.slice(0, min(99_999, b.n))
    occ = int(2 * np.clip(np.diff(b.seq_off) - 47, 0, None).sum())
    t0 = time.perf_counter()
    st, g = O.kmer_graph(cfg, b, K=47, min_depth=5)
    wall = time.perf_counter() - t0
    assert st == 0
    rows["kmer_graph"] = {"value": occ / wall, "unit": "occurrences/s", "cores": 1,
                          "sample": f"{b.n} reads, {occ} K-mer occurrences, {len(g)} graph K-mers in {wall:.2f}s"}
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--window", type=int, default=0, help="positions per window; 0 = autotuned")
    ap.add_argument("--persistent", type=int, default=1)
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup; 0 = autotuned")
    ap.add_argument("--graph", type=int, default=1, help="time the steps as hipGraph replays")
    ap.add_argument("--repeat", type=int, default=1,
                    help="with --graph, queue the timed replays back to back (pl_batch_repeat) instead of one "
                         "library call per step")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-region comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU oracle sample")
    ap.add_argument("--serve", type=int, default=1, help="also run the serving workload (c4 request stream)")
    ap.add_argument("--serve-scale", type=float, default=1.0, help="genome scale of the serving workload")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=20000.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-lanes", type=int, default=4)
    ap.add_argument("--vcf", type=int, default=1, help="also format the batch's VCF records on the device")
    ap.add_argument("--tracks", type=int, default=1, help="also compute the batch's 13 per-locus tracks on the device")
    ap.add_argument("--kmer", type=int, default=1,
                    help="also build the assembler's K-mer graph of 99,999 reads of the largest contig on the device")
    ap.add_argument("--bam", type=int, default=1,
                    help="also run the BAM producer: the workload written as a BAM, decoded, uploaded and piled")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB on the device three times (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--measure-copy", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--traffic-profile", default=os.path.join(ROOT, "profiles", "k_window_c3_traffic.md"))
    args = ap.parse_args()
    if args.measure_copy:
        print(copy_bandwidth(), flush=True)
        return
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied 3 x {4 << 28} bytes", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the path would shard "
              "by region with no communication (regions are independent)", file=sys.stderr)
        sys.exit(2)

    from pilon_amd.abi import default_config
    from pilon_amd.engine import Context
    from pilon_amd.synth import SynthData, preset

    cfg = default_config()
    t0 = time.time()
    params = preset(args.config, scale=args.scale)
    data = SynthData(params)
    log(f"bench: generated {args.config} x{args.scale}: {data.n} contigs in {time.time() - t0:.1f}s")
    ctx = Context(cfg, device=0)
    rids = []
    alg_read_bytes = 0
    n_reads = 0
    t0 = time.time()
    soas = []
    for ci in range(data.n):
        ctx.ref_upload(ci, data.names[ci], data.draft(ci))
        soa = data.soa(ci)
        soas.append(soa)
        alg_read_bytes += read_bytes(soa)
        n_reads += soa.n_reads
    rids = ctx.readset_upload_many([(ci, soas[ci], "frags") for ci in range(data.n)])
    del soas
    log(f"bench: uploaded {n_reads} reads in {time.time() - t0:.1f}s")
    regions = regions_for(args.config, data, rids)
    descs = ctx.batch_descs(regions)
    G = sum(e - s + 1 for _, s, e, _ in regions)

    # one checked pass for the unit (aligned bases) and the sizes; with
    # window/block 0 it also runs the autotuner's search (cached by shape)
    _, _, stats = ctx.batch_run(regions, window=args.window, persistent=args.persistent, copy_out=False,
                                want_stats=True, descs=descs, block=args.block)
    aligned = sum(s["aligned_bases"] for s in stats)
    info = ctx.batch_info()
    W, B = info["window"], info["block"]
    log(f"bench: {len(regions)} regions, G={G}, aligned={aligned}, records={info['n_records']}, "
        f"windows={info['n_windows']}, window {W} block {B}{' (autotuned)' if not (args.window and args.block) else ''}")

    def step(graph):
        ctx.batch_run(regions, window=W, persistent=args.persistent, copy_out=False, want_stats=False,
                      descs=descs, block=B, use_graph=bool(graph))

    for _ in range(max(args.warmup, 1 if args.graph else 0)):
        step(args.graph)
    pipe = []
    # with graphs the K steps are K replays of the captured pipeline queued
    # back to back (pl_batch_repeat: every pass's counters are checked after
    # they drain); otherwise each step returns after the library's stream has
    # drained.  The device synchronisations around the loop make the bracket
    # explicit.
    ctx.synchronize()
    t0 = time.perf_counter()
    if args.graph and args.repeat:
        ctx.batch_repeat(args.steps)
        pipe.append(ctx.batch_info()["pipeline_ms"])
    else:
        for _ in range(args.steps):
            step(args.graph)
            pipe.append(ctx.batch_info()["pipeline_ms"])
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = aligned / (ms_per_step / 1000.0)
    pipe_ms = float(np.mean(pipe))
    # the window kernel's own time: HIP events around it on the library's
    # stream, in eager launches of the same batch (a graph replay has none)
    win = []
    for _ in range(max(3, args.steps // 2)):
        step(False)
        win.append(ctx.batch_info()["window_ms"])
    win_ms = float(np.mean(win))
    b_alg = alg_read_bytes + G * 80.25 + 32 * info["n_records"]
    achieved = b_alg / (win_ms / 1000.0) / 1e9
    traffic, traffic_src = measured_traffic(args.traffic_profile, args.config, W, B) if args.scale == 1.0 else (None, None)
    mode = ("graph replays queued back to back" if args.repeat else "graph") if args.graph else "eager"
    log(f"bench: {ms_per_step:.3f} ms/step wall ({mode}), device pipeline {pipe_ms:.3f} ms, "
        f"window kernel {win_ms:.3f} ms, {value / 1e9:.2f} G bases/s, window kernel {achieved:.0f} GB/s algorithmic")

    extra = {}
    if not args.no_plain:
        t0 = time.perf_counter()
        dev = 0.0
        for r in regions:
            ctx.batch_run([r], window=512, persistent=0, copy_out=False, want_stats=False, block=256)
            dev += ctx.batch_info()["pipeline_ms"]
        plain_s = time.perf_counter() - t0
        extra["plain_per_region"] = {"wall_ms": plain_s * 1000.0, "device_ms": dev,
                                     "bases_per_s": aligned / plain_s,
                                     "batched_vs_plain": (aligned / (ms_per_step / 1000.0)) / (aligned / plain_s)}
        log(f"bench: plain per-region launches: {plain_s * 1000:.1f} ms wall ({dev:.1f} ms device) -> "
            f"batched is {extra['plain_per_region']['batched_vs_plain']:.2f}x")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(cfg, data, regions, args.cpu_seconds)
        log(f"bench: CPU oracle {cpu['value'] / 1e6:.1f} M bases/s on {cpu['cores']} threads ({cpu['sample']})")
        for k, v in cpu["rows"].items():
            log(f"bench: CPU oracle {k}: {v['value'] / 1e6:.2f} M {v['unit']} on {v['cores']} threads ({v['sample']})")

    vcf = None
    if args.vcf:
        # the last batch on the context is the timed one (all regions)
        ctx.batch_run(regions, window=W, block=B, persistent=1, copy_out=False, want_stats=False)
        nbytes, lines, vms = ctx.batch_vcf(decode=False)
        wall = ctx.last_vcf_wall_ms
        # the same text handed to a consumer through the two pinned staging
        # buffers, 32 MB at a time (the consumer here only counts the bytes)
        got = [0]

        def consume(_addr, n):
            got[0] += n
        t0 = time.perf_counter()
        sbytes, sms = ctx.batch_vcf_stream(consume, raw=True)
        swall = (time.perf_counter() - t0) * 1000.0
        assert sbytes == nbytes == got[0]
        vcf = {"records": lines, "bytes": nbytes, "device_ms": vms, "wall_ms": wall,
               "stream_wall_ms": swall, "stream_GB_per_s": nbytes / (swall / 1000.0) / 1e9,
               "records_per_s": lines / (vms / 1000.0), "text_GB_per_s": nbytes / (vms / 1000.0) / 1e9,
               "note": "Vcf.writeRecord lines of every locus (GenomeRegion.writeVcf) formatted on the device; "
                       "device_ms covers the mark, indel, length, scan and write passes; wall_ms is pl_batch_vcf "
                       "with the text copied into one host buffer it allocates; stream_wall_ms is "
                       "pl_batch_vcf_stream handing the same text to a consumer in 32 MB pieces"}
        log(f"bench: VCF: {lines} records, {nbytes / 1e9:.2f} GB in {vms:.1f} ms on the device "
            f"({vcf['records_per_s'] / 1e9:.2f} G records/s), {wall:.0f} ms into one host buffer, "
            f"{swall:.0f} ms streamed ({vcf['stream_GB_per_s']:.1f} GB/s)")

    tracks = None
    if args.tracks:
        # the last batch on the context is still the timed one; the first
        # call also builds each contig's GC buffer values (once per contig)
        t0 = time.perf_counter()
        ctx.batch_tracks(copy=False)
        first_ms = (time.perf_counter() - t0) * 1000.0
        tms = [ctx.batch_tracks(copy=False)[1] for _ in range(3)]
        t0 = time.perf_counter()
        wig, wms = ctx.batch_wig(3)
        wall = time.perf_counter() - t0
        tracks = {"loci": G, "device_ms": float(np.mean(tms)), "first_call_wall_ms": first_ms,
                  "loci_per_s": G / (float(np.mean(tms)) / 1000.0), "wig_coverage_bytes": len(wig),
                  "wig_device_ms": wms, "wig_wall_ms": wall * 1000.0,
                  "note": "Tracks.standardTracks values (13 per locus, copy number and GC included) on the device; "
                          "device_ms covers the fragment-coverage, prefix-sum and value passes; the wig figures "
                          "are the Coverage track's text (values again, then length, scan and write passes)"}
        t0 = time.perf_counter()
        bed, bms = ctx.batch_bed()
        bwall = time.perf_counter() - t0
        tracks.update({"bed_lines": bed.count("\n") - 1, "bed_device_ms": bms, "bed_wall_ms": bwall * 1000.0})
        tracks["note"] += ("; Pilon.bed: the issue bits of every locus on the device with the values, their runs "
                           "summarised on the host's threads (bed_wall_ms includes the mask copy and the summary)")
        del wig, bed
        log(f"bench: tracks: {G} loci x 13 in {tracks['device_ms']:.2f} ms on the device "
            f"({tracks['loci_per_s'] / 1e9:.2f} G loci/s; first call {first_ms:.1f} ms wall with the GC buffers), "
            f"Coverage.wig {tracks['wig_coverage_bytes'] / 1e6:.0f} MB in {wms:.1f} ms, "
            f"Pilon.bed {tracks['bed_lines']} lines in {bwall * 1000:.1f} ms wall")

    kmer = None
    if args.kmer:
        ci = max(range(data.n), key=lambda i: len(data.draft(i)))
        b = data.batch(ci)
        b = b.slice(0, min(99_999, b.n))
        lens = np.diff(b.seq_off)
        occ = int(2 * np.clip(lens - 47, 0, None).sum())
        ctx.kmer_graph(b, K=47)
        kms = []
        for _ in range(3):
            g, ms = ctx.kmer_graph(b, K=47)
            kms.append(ms)
        kmer = {"reads": int(b.n), "K": 47, "occurrences": occ, "graph_kmers": len(g),
                "device_ms": float(np.mean(kms)), "occurrences_per_s": occ / (float(np.mean(kms)) / 1000.0),
                "note": "Assembler.addReads + buildGraph (K-mer pileups of local reassembly) on the device; "
                        "device_ms covers emission, the radix sort by K-mer hash, the sums and the calls"}
        del g
        log(f"bench: K-mer graph: {b.n} reads, {occ} K-mer occurrences -> {kmer['graph_kmers']} graph K-mers "
            f"in {kmer['device_ms']:.2f} ms on the device ({kmer['occurrences_per_s'] / 1e9:.2f} G occurrences/s)")

    producer = None
    if args.bam:
        producer = bam_producer(data, ctx, args, W, B, aligned)
        log(f"bench: BAM producer: {producer['records']} records; device loader {producer['device_loader_s']:.2f}s "
            f"(file to pileup {producer['end_to_end_bases_per_s'] / 1e9:.2f} G bases/s); host decode "
            f"{producer['host_decode_s']:.2f}s + upload {producer['host_upload_s']:.2f}s "
            f"({producer['host_end_to_end_bases_per_s'] / 1e9:.2f} G bases/s)")

    copy_gbs = measured_copy_bandwidth()
    if copy_gbs:
        log(f"bench: device copy of 2 GiB: {copy_gbs:.0f} GB/s read + written "
            f"(k_window at {achieved / copy_gbs:.1%} of it)")

    serving = None
    if args.serve:
        ctx.close()
        del data
        serving = serve_workload(cfg, args)

    line = {
        "metric": METRIC, "value": value, "unit": "bases/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{args.config}: {WORKLOADS[args.config]}", "scale": args.scale,
                   "regions": len(regions), "positions": G, "reads": n_reads, "aligned_bases": aligned,
                   "window": W, "block": B, "autotuned": not (args.window and args.block),
                   "persistent": args.persistent, "graph": bool(args.graph), "repeat": bool(args.graph and args.repeat)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_window", "alg_bytes_per_launch": b_alg, "kernel_ms": win_ms,
                     "measured_copy_GB_per_s": copy_gbs,
                     "frac_of_measured_copy": achieved / copy_gbs if copy_gbs else None},
        "device_pipeline_ms": pipe_ms,
        "cpu_baseline": cpu,
        "p99_latency_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
        "bam_producer": producer,
        "vcf": vcf,
        "tracks": tracks,
        "kmer_graph": kmer,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)
    if not args.serve:
        ctx.close()


def bam_producer(data, ctx, args, W, B, aligned):
    """The BAM -> read set producer on this workload: its reads written as a
    coordinate-sorted BAM (BGZF, deflate level 1, one reference per contig),
    then read two ways and piled in one batched pass each: the device
    loader (inflate on the host's threads into pinned staging, records
    decoded on the device: pl_loader_*), and the host decoder (libpilonbam)
    with one upload pass (pl_readset_upload_many).  End to end = load +
    pile, from a file in the page cache."""
    import shutil
    import tempfile
    from pilon_amd.bam import BamFile
    # a one-GPU box gives its process 16 CPUs, whatever the affinity mask
    # shows; the pool takes all of them (host_cpus reports the mask)
    host_cpus = len(os.sched_getaffinity(0))
    nth = max(1, min(16, host_cpus))
    d = tempfile.mkdtemp(prefix="pilon_bam_")
    path = os.path.join(d, f"{args.config}.bam")
    try:
        t0 = time.perf_counter()
        stream = data.write_bam(path, level=1)
        t_write = time.perf_counter() - t0
        fsize = os.path.getsize(path)
        t0 = time.perf_counter()
        bam = BamFile(path, threads=nth)
        t_dec = time.perf_counter() - t0
        st = bam.stats
        t0 = time.perf_counter()
        rids = ctx.readset_upload_many([(ci, bam.reads(ci), "frags") for ci in range(data.n)])   # returns synchronised
        t_up = time.perf_counter() - t0
        regions = regions_for(args.config, data, rids)
        t0 = time.perf_counter()
        _, _, stats = ctx.batch_run(regions, window=W, block=B, persistent=1, copy_out=False, want_stats=True)
        t_pile = time.perf_counter() - t0
        got = sum(x["aligned_bases"] for x in stats)
        assert got == aligned, ("aligned bases through the BAM differ", got, aligned)
        bam.close()
        del bam
        # the same file through the device loader: host inflate into pinned
        # staging, record decode on the device
        t0 = time.perf_counter()
        dids = ctx.readsets_from_bam(path, list(range(data.n)), threads=nth)                   # returns synchronised
        t_load = time.perf_counter() - t0
        lst = ctx.last_bam_stats
        dregions = regions_for(args.config, data, dids)
        t0 = time.perf_counter()
        _, _, dstats = ctx.batch_run(dregions, window=W, block=B, persistent=1, copy_out=False, want_stats=True)
        t_dpile = time.perf_counter() - t0
        got = sum(x["aligned_bases"] for x in dstats)
        assert got == aligned, ("aligned bases through the device loader differ", got, aligned)
        assert dstats == stats, "the device loader's regions differ from the host decoder's"
    finally:
        shutil.rmtree(d, ignore_errors=True)
    e2e_host = t_dec + t_up + t_pile
    e2e = t_load + t_dpile
    return {"file_bytes": fsize, "bam_bytes": stream, "records": st["records"], "threads": nth,
            "write_s": t_write,
            "device_loader_s": t_load, "loader_inflate_s": lst["inflate_s"], "loader_delimit_s": lst["decode_s"],
            "loader_phases_s": ctx.last_load_s,
            "pile_s": t_dpile, "end_to_end_bases_per_s": aligned / e2e,
            "host_decode_s": t_dec, "host_inflate_s": st["inflate_s"], "host_parse_s": st["decode_s"],
            "host_upload_s": t_up, "host_pile_s": t_pile, "host_end_to_end_bases_per_s": aligned / e2e_host,
            "host_decode_records_per_s": st["records"] / t_dec,
            "note": "aligned bases per second of the whole path from a BAM in the page cache: the device loader "
                    "(host inflate into pinned staging, records decoded into read sets on the device) and one "
                    "batched pileup (not a graph replay); host_* is the same file decoded on the host "
                    "(libpilonbam) and uploaded in one pass"}


def serve_workload(cfg, args):
    """The serving form on the c4 configuration: the genome and its reads
    resident, a stream of polish requests (log-uniform 2 kb - 1 Mb at uniform
    offsets) submitted at Poisson arrival times, admitted by estimated device
    memory and run with continuous batching.  Latency = submit to result
    ready (regions' stats on the host); throughput = aligned bases of all
    requests over the time from the first submission to the last result."""
    from pilon_amd.engine import Context, Server
    from pilon_amd.synth import SynthData, poisson_arrivals, preset, serve_requests
    t0 = time.time()
    data = SynthData(preset("c4", scale=args.serve_scale))
    ctx = Context(cfg, device=0)
    for ci in range(data.n):
        ctx.ref_upload(ci, data.names[ci], data.draft(ci))
    rids = ctx.readset_upload_many([(ci, data.soa(ci), "frags") for ci in range(data.n)])
    G_res = sum(len(data.draft(ci)) for ci in range(data.n))
    log(f"serve: c4 x{args.serve_scale}: {data.n} contigs, {G_res} bp resident, ready in {time.time() - t0:.1f}s")
    reqs = serve_requests(data, args.serve_requests, 2_000, 1_000_000, seed=1004)
    arrive = poisson_arrivals(len(reqs), args.serve_rate, seed=1004)
    srv = Server(ctx, lanes=args.serve_lanes, max_batch=64)
    # warm the lanes (first batches allocate their buffers)
    for ci, s, e in reqs[:2 * args.serve_lanes]:
        srv.submit(ci, s, e, rids[ci], tag=-1)
    srv.drain()
    srv.poll(max_n=4096)

    def stream(times):
        """submit every request at its arrival time (all at once if None)"""
        done = []
        t_start = time.perf_counter()
        for i, (ci, s, e) in enumerate(reqs):
            if times is not None:
                while time.perf_counter() - t_start < times[i]:
                    pass
            srv.submit(ci, s, e, rids[ci], tag=i)
            if i % 64 == 0:
                done += srv.poll(max_n=4096)
        while len(done) < len(reqs):
            done += srv.poll(max_n=4096, wait_ms=100)
        t_end = time.perf_counter()
        # PL_E_EDGE (4): regions on which Pilon itself throws (an indel
        # normalised before the region start); the request ran, its result is that error
        bad = [r for r in done if r["status"] not in (0, 4)]
        assert not bad, bad[:3]
        lat = np.array([r["latency_ms"] for r in done])
        bases = sum(r["stats"]["aligned_bases"] for r in done)
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "bases_per_s": bases / (t_end - t_start), "makespan_s": t_end - t_start,
                "mean_batch": float(np.mean([r["batch_size"] for r in done])),
                "edge_status": sum(1 for r in done if r["status"] == 4)}

    cap = stream(None)
    log(f"serve: capacity (all {len(reqs)} requests at once): {cap['bases_per_s'] / 1e9:.1f} G bases/s, "
        f"mean batch {cap['mean_batch']:.1f}")
    at = stream(arrive)
    srv.close()
    ctx.close()
    out = {"workload": "c4: 50 x 4 Mb contigs resident, 2x150 bp at 50x; 10,000 requests log-uniform 2 kb-1 Mb, "
                       "Poisson arrivals",
           "scale": args.serve_scale, "requests": len(reqs), "rate_per_s": args.serve_rate, "lanes": args.serve_lanes,
           "p50_ms": at["p50_ms"], "p99_ms": at["p99_ms"], "mean_batch": at["mean_batch"],
           "bases_per_s": at["bases_per_s"], "makespan_s": at["makespan_s"], "edge_status": at["edge_status"],
           "capacity_bases_per_s": cap["bases_per_s"], "capacity_mean_batch": cap["mean_batch"],
           "positions": int(sum(e - s + 1 for _, s, e in reqs))}
    log(f"serve: {len(reqs)} requests at {args.serve_rate:.0f}/s: p50 {out['p50_ms']:.2f} ms, p99 {out['p99_ms']:.2f} ms, "
        f"{out['bases_per_s'] / 1e9:.1f} G bases/s, mean batch {out['mean_batch']:.1f}")
    return out


if __name__ == "__main__":
    main()
