"""Throughput of the calling path (pileup -> realignment -> calls) on one MI355X.

Workload: a 5 Mb synthetic contig with 2% tandem-repeat tracts, 60x of 150 bp paired-end reads,
SNVs at 1/1000 bp, 1-30 bp indels at 1/8000 bp, 200 deletions of 50-300 bp with the reads across
their breakpoints soft-clipped; one region covering the whole contig, --fisher.  The reads are made
by rabbitvar_amd/_build/rvsynth (seeded) and uploaded to HBM before timing; one step is one full
region call (device pileup, host realignment of the candidates, device scoring, text).

Prints one JSON line (see the keys below).  `cpu_baseline` times the CPU restatement of the same
path (oracle/_build/rvcpu) on this host's cores: the whole contig as 10 kb tiles (the reference's
--auto_resize pieces) on every available core, as the reference's OpenMP region loop runs a BED
list; `single_thread` is the same restatement on one core over one 2 Mb region.  `end_to_end` is
the headline path from the BAM file: open and inflate, region query, upload, the call.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
SYNTH = os.path.join(ROOT, "rabbitvar_amd", "_build", "rvsynth")
ORACLE = os.path.join(ROOT, "oracle", "_build", "rvcpu")

METRIC = "reads/s pileup→realign→calls, 1 GPU; screened-mode speedup at ≥99.9% recall"
WORKLOAD = ["--seed", "7", "--len", "5000000", "--depth", "60", "--indel-rate", "0.000125", "--indel-max", "30",
            "--long-dels", "200", "--repeat-frac", "0.02"]
CPU_SAMPLE = (1, 2000000)

HBM_PEAK_GBS = 8000.0


def walk_alg_bytes(batch, region_bp):
    """Algorithmic HBM bytes of one pileup launch: every read's bases, qualities, CIGAR and 16 B of
    fixed fields in, the reference window in (region + 2 x 1200 bp), and the minimum per-position
    output out (coverage + 4 bases x 2 strands as int32 = 36 B)."""
    import numpy as np

    lseq = np.diff(batch.seq_off).astype(np.int64)
    ncig = np.diff(batch.cigar_off).astype(np.int64)
    return int((2 * lseq + 4 * ncig + 16).sum()) + (region_bp + 2400) + 36 * region_bp


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensure_data(d):
    if all(os.path.exists(os.path.join(d, f)) for f in ("reads.rvb", "ref.fa", "reads.bam")):
        return
    os.makedirs(d, exist_ok=True)
    log(f"generating the workload into {d}")
    subprocess.run([SYNTH, "--out", d, "--bam"] + WORKLOAD, check=True, stdout=subprocess.DEVNULL)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def end_to_end(d, eng, name, seq, steps):
    """The headline path from the BAM file: open (inflate + index), the region query (decode to
    SoA), the call (upload + device pileup + realignment + scoring + text)."""
    from rabbitvar_amd import BamFile

    L = len(seq)
    best = None
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        bam = BamFile(os.path.join(d, "reads.bam"), threads=host_cores())
        t1 = time.perf_counter()
        reads = bam.query(name, 1, L)
        t2 = time.perf_counter()
        eng.run_region(name, 1, L, reads, name, seq)
        t3 = time.perf_counter()
        bam.close()
        if best is None or t3 - t0 < best[3] - best[0]:
            best = (t0, t1, t2, t3, reads.n)
    t0, t1, t2, t3, n = best
    res = {"reads_per_s": round(n / (t3 - t0), 1), "reads": n, "s": round(t3 - t0, 4),
           "bam_open_s": round(t1 - t0, 4), "query_s": round(t2 - t1, 4), "call_s": round(t3 - t2, 4),
           "bam_bytes": os.path.getsize(os.path.join(d, "reads.bam")), "threads": host_cores(),
           "note": "best of the passes; the call includes the PCIe upload of the region's records"}
    log(f"end to end from BAM: {res['reads_per_s']:.0f} reads/s ({res['s']} s: open {res['bam_open_s']}, "
        f"query {res['query_s']}, call {res['call_s']})")
    return res


# the 20,000-region amplicon panel, sampled: 100 regions drawn from its distribution (lengths
# log-normal(7.71, 1.21) clamped to 100 bp-50 kb, depth U(500, 5000)x per region, pairs starting at
# the amplicon ends, variants at 1/2000 bp), a seeded 0.5% sample of the full list
PANEL_REGIONS = 100
PANEL = ["--seed", "77", "--len", "3000000", "--panel", str(PANEL_REGIONS), "--panel-len", "7.71", "1.21",
         "--panel-depth", "500", "5000", "--snv-rate", "0.0004", "--indel-rate", "0.0001", "--vaf", "0.005", "0.5",
         "--bam", "--no-rvb"]


def panel_bench(d, steps):
    """Region-list throughput: the engine pool (4 engines) against one engine, same regions."""
    from rabbitvar_amd import BamFile, read_fasta
    from rabbitvar_amd.engine import EnginePool

    if not os.path.exists(os.path.join(d, "reads.bam")):
        os.makedirs(d, exist_ok=True)
        log(f"generating the panel into {d}")
        subprocess.run([SYNTH, "--out", d] + PANEL, check=True, stdout=subprocess.DEVNULL)
    regions = []
    for ln in open(os.path.join(d, "panel.bed")):
        c, s0, e0, g = ln.rstrip("\n").split("\t")
        regions.append((c, int(s0), int(e0), g))
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    t0 = time.perf_counter()
    bam = BamFile(os.path.join(d, "reads.bam"), threads=16)
    open_s = time.perf_counter() - t0
    n_reads = sum(bam.query(c, s0, e0).n for c, s0, e0, _ in regions)
    res = {"workload": f"{len(regions)} amplicon regions drawn from the 20,000-region panel's distribution "
                       "(log-normal(7.71, 1.21) lengths in [100, 50000], depth U(500, 5000)x, variants 1/2000 bp, "
                       "VAF 0.005-0.5)", "regions": len(regions), "reads": n_reads,
           "region_bp": sum(e0 - s0 + 1 for _, s0, e0, _ in regions),
           "bam_bytes": os.path.getsize(os.path.join(d, "reads.bam")), "bam_open_s": round(open_s, 3)}
    for workers in (1, 4):
        pool = EnginePool(workers=workers, sample="S1", fisher=1)
        pool.run_regions(regions, bam, {name: seq})  # warmup: buffers grow to the largest region
        t0 = time.perf_counter()
        for _ in range(steps):
            pool.run_regions(regions, bam, {name: seq})
        dt = (time.perf_counter() - t0) / steps
        pool.close()
        res[f"reads_per_s_{workers}_engine" + ("s" if workers > 1 else "")] = round(n_reads / dt, 1)
        res[f"s_per_pass_{workers}"] = round(dt, 4)
    log(f"panel: {len(regions)} regions, {n_reads} reads; 1 engine {res['reads_per_s_1_engine']:.0f} reads/s, "
        f"4 engines {res['reads_per_s_4_engines']:.0f} reads/s")
    return res


# the 10 Mb, 200x/100x somatic workload: somatic SNVs and indels at 6e-5/bp (600) with VAF
# U(0.01, 0.2), germline variants shared by both samples, 2% duplicated records; a seed that no
# screen training or validation genome uses
SOMATIC = ["--seed", "2026", "--len", "10000000", "--depth", "200", "--normal-depth", "100", "--somatic-rate",
           "0.00006", "--vaf", "0.01", "0.2", "--indel-rate", "0.0001", "--dup-rate", "0.02"]


def somatic_bench(d, steps, warmup):
    """Somatic mode on the tumor/normal pair: exact and screened throughput (tumor + normal reads,
    resident in HBM), and the screened call set against the exact one."""
    import torch

    from rabbitvar_amd import read_fasta, read_rvb, screen
    from rabbitvar_amd.engine import DeviceReads, Engine

    if not os.path.exists(os.path.join(d, "normal.rvb")):
        os.makedirs(d, exist_ok=True)
        log(f"generating the tumor/normal pair into {d}")
        subprocess.run([SYNTH, "--out", d] + SOMATIC, check=True, stdout=subprocess.DEVNULL)
    name, seq = read_fasta(os.path.join(d, "ref.fa"))
    tumor, normal = read_rvb(os.path.join(d, "reads.rvb")), read_rvb(os.path.join(d, "normal.rvb"))
    n_reads = tumor.n + normal.n
    dtum, dnor = DeviceReads(tumor), DeviceReads(normal)
    eng = Engine(sample="T|N", fisher=1, removeDuplicatedReads=1)

    def timed():
        out = None
        for _ in range(warmup):
            out = eng.run_region_somatic(name, 1, len(seq), dtum, dnor, name, seq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = eng.run_region_somatic(name, 1, len(seq), dtum, dnor, name, seq)
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t0) / steps

    text, dt = timed()
    params, threshold, _ = screen.load_model()
    eng.set_screen(params, threshold)
    stext, sdt = timed()
    st = eng.stats()
    eng.set_screen(None, 0)
    eng.close()
    cmp_ = screen.compare(text, stext)
    res = {"workload": "10 Mb contig, tumor 200x + normal 100x 150 bp PE, somatic SNVs/indels 6e-5 at VAF 0.01-0.2, "
                       "germline SNV 1e-3 + indels 1e-4, 2% duplicates, seed 2026; -b 'T|N' --fisher -t",
           "reads": n_reads, "exact_reads_per_s": round(n_reads / dt, 1), "exact_ms": round(dt * 1e3, 2),
           "screened_reads_per_s": round(n_reads / sdt, 1), "screened_ms": round(sdt * 1e3, 2),
           "speedup": round(dt / sdt, 4), "recall": round(cmp_["recall"], 6),
           "precision": round(cmp_["precision"], 6), "identical_lines": round(cmp_["identical_lines"], 6),
           "calls": cmp_["exact_calls"], "candidates": st["screen_candidates"], "skipped": st["screen_skipped"],
           "host_detail_ms_last_step": dict(zip(["first_sample", "second_sample", "ref_passes", "group_by_position",
                                                 "print"], [round(x, 2) for x in st["detail_ms"][:5]]))}
    log(f"somatic: {n_reads} reads, exact {res['exact_reads_per_s']:.0f} reads/s, screened "
        f"{res['screened_reads_per_s']:.0f} reads/s ({res['speedup']}x), recall {res['recall']}, "
        f"precision {res['precision']}")
    return res


def cpu_baseline(d, L):
    """The CPU restatement of the path, timed on this host: (1) every core over the contig's 10 kb
    --auto_resize tiles (a BED list, regions on threads as the reference's OpenMP loop), (2) one core
    over one 2 Mb region (-R is single-threaded in the reference)."""
    from rabbitvar_amd.regions import auto_resize

    if not os.path.exists(ORACLE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    threads = host_cores()
    tiles = [r for row in auto_resize([[("chrS", 1, L, "chrS")]]) for r in row]
    bed = os.path.join(d, "tiles.bed")
    with open(bed, "w") as f:
        for c, s0, e0, g in tiles:
            f.write(f"{c}\t{s0}\t{e0}\t{g}\n")
    log(f"timing the CPU path: {len(tiles)} tiles of the whole contig on {threads} threads")
    r = subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1",
                        "-i", bed, "-c", "1", "-S", "2", "-E", "3", "-g", "4", "--fisher", "--th", str(threads),
                        "--time", "--out", os.path.join(d, "cpu_tiles.txt")], check=True, capture_output=True,
                       text=True)
    tt = json.loads(r.stdout.strip().splitlines()[-1])
    s, e = CPU_SAMPLE
    log(f"timing the CPU path on chrS:{s}-{e}, one thread")
    r = subprocess.run([ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, "reads.rvb"), "-N", "S1",
                        "-R", f"chrS:{s}-{e}", "--fisher", "--time", "--out", os.path.join(d, "cpu_sample.txt")],
                       check=True, capture_output=True, text=True)
    t1 = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"CPU: {tt['reads_per_s']:.0f} reads/s on {threads} threads, {t1['reads_per_s']:.0f} reads/s on one")
    return {"value": round(tt["reads_per_s"], 1), "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"the whole 5 Mb contig as {len(tiles)} 10 kb --auto_resize tiles on {threads} threads "
                      f"({tt['reads']} region reads, {tt['seconds']:.1f} s; CPU restatement of the path)",
            "cpu": cpu_model(),
            "single_thread": {"value": round(t1["reads_per_s"], 1), "unit": "reads/s", "cores": 1,
                              "sample": f"chrS:{s}-{e}, one region ({t1['reads']} reads, {t1['seconds']:.1f} s)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--data", default=os.environ.get("RV_BENCH_DATA", "/tmp/rv_bench_5mb_60x_bam"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-panel", action="store_true")
    ap.add_argument("--panel-data", default=os.environ.get("RV_BENCH_PANEL", "/tmp/rv_bench_panel_c4"))
    ap.add_argument("--no-somatic", action="store_true")
    ap.add_argument("--somatic-data", default=os.environ.get("RV_BENCH_SOMATIC", "/tmp/rv_bench_somatic_10mb"))
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the calling path has no multi-GPU mode here (asked for {a.gpus}); "
              "regions would be sharded across GPUs by the caller.", file=sys.stderr)
        sys.exit(2)

    import torch

    sys.path.insert(0, ROOT)
    from rabbitvar_amd import read_fasta, read_rvb
    from rabbitvar_amd.engine import DeviceReads, Engine

    ensure_data(a.data)
    name, seq = read_fasta(os.path.join(a.data, "ref.fa"))
    batch = read_rvb(os.path.join(a.data, "reads.rvb"))
    L = len(seq)
    log(f"{batch.n} reads over {L} bp; uploading to HBM")
    dev = DeviceReads(batch)
    eng = Engine(sample="S1", fisher=1)

    def step():
        return eng.run_region(name, 1, L, dev, name, seq)

    for i in range(a.warmup):
        t0 = time.perf_counter()
        text = step()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f} s, {len(text.splitlines())} calls")
    walk_ms = []
    stage = {k: 0.0 for k in ("prep_ms", "walk_ms", "anchor_ms", "score_ms", "reduce_ms", "realign_ms", "order_ms",
                              "format_ms", "wall_ms", "call_ms")}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        text = step()
        st = eng.stats()
        walk_ms.append(st["walk_ms"])
        for k in stage:
            stage[k] += st[k]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    osort = st["order_sorted"]
    log(f"timed {a.steps} steps in {dt:.3f} s (nonInsertionVariants order by {'slot sort' if osort & 1 else 'map replay'},"
        f" alignedVariants order by {('device' if osort & 4 else 'host') + ' per-growth sort' if osort & 2 else 'map replay'})")
    reads_per_s = batch.n * a.steps / dt

    alg_bytes = walk_alg_bytes(batch, L)
    avg_walk_s = sum(walk_ms) / len(walk_ms) / 1e3
    achieved = alg_bytes / avg_walk_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "walk_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    res = {
        "metric": METRIC,
        "value": round(reads_per_s, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic",
        "config": {"workload": "5 Mb contig, 60x 150 bp PE, SNV 1e-3, indels 1-30 bp at 1/8000, 200 deletions "
                               "50-300 bp, 2% tandem repeats; one region chrS:1-5000000, --fisher",
                   "reads": batch.n, "region_bp": L},
        "stages_ms_per_step": {k: round(v / a.steps, 3) for k, v in stage.items()},
        "host_detail_ms_last_step": dict(zip(["rl_prefetch", "rl_mnp", "rl_del", "rl_ins", "rl_lgins", "record_list",
                                              "key_sort", "order_wait", "aligned_order", "rank_sort", "reduce_events",
                                              "reduce_clips", "nopass", "row_miss_fetch", "n_nopass", "n_row_miss_fetches"],
                                             [round(x, 2) for x in st["detail_ms"]])),
        "realign_device": {"kernel": "k_find35", "pairs": st["realign_device_pairs"],
                           "ms": round(st["realign_device_ms"], 4)},
        "aligned_order_split": {"operations": st["order_ops"], "list_ms": round(st["order_ops_ms"], 3),
                                "device_ms": round(st["order_device_ms"], 3)},
        "timeline_ms_last_step": dict(zip(["pileup", "side_threads", "realign", "record_list", "score", "format",
                                           "order", "released", "p_ref", "p_prep", "p_walk", "p_copy", "p_wait_release",
                                           "p_reduce", "line_offsets", "lines_and_calls"],
                                          [round(x, 2) for x in st["timeline_ms"]])),
        "pileup_records_bytes": st["event_bytes"],
        "roofline": {"kernel": "k_walk", "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": round(avg_walk_s * 1e3, 4)},
    }
    # screened mode: the realigner skips the candidates the trained network scores below its
    # threshold; the same workload, timed the same way, and its call set against the exact one
    from rabbitvar_amd import screen

    params, threshold, _ = screen.load_model()
    eng.set_screen(params, threshold)
    for i in range(a.warmup):
        step()
    sc_stage = {k: 0.0 for k in ("realign_ms", "wall_ms")}
    skipped = cands = 0
    scr_ms = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        stext = step()
        st = eng.stats()
        for k in sc_stage:
            sc_stage[k] += st[k]
        skipped, cands = st["screen_skipped"], st["screen_candidates"]
        scr_ms += st["screen_ms"]
    torch.cuda.synchronize()
    sdt = time.perf_counter() - t0
    eng.set_screen(None, 0)
    res["end_to_end"] = end_to_end(a.data, eng, name, seq, 2)
    cmp_ = screen.compare(text, stext)
    res["screened"] = {"value": round(batch.n * a.steps / sdt, 1), "unit": "reads/s",
                       "speedup": round(dt / sdt, 4), "threshold": threshold,
                       "recall": round(cmp_["recall"], 6), "precision": round(cmp_["precision"], 6),
                       "identical_lines": round(cmp_["identical_lines"], 6), "calls": cmp_["exact_calls"],
                       "candidates": cands, "skipped": skipped, "screen_ms": round(scr_ms / a.steps, 3),
                       "screen_list_ms": round(st["screen_list_ms"], 3),
                       "stages_ms_per_step": {k: round(v / a.steps, 3) for k, v in sc_stage.items()},
                       "realign_detail_ms_last_step": dict(zip(["rl_prefetch", "rl_mnp", "rl_del", "rl_ins",
                                                                "rl_lgins", "record_list"],
                                                               [round(x, 2) for x in st["detail_ms"][:6]]))}
    log(f"screened: {res['screened']['value']} reads/s ({res['screened']['speedup']}x), recall "
        f"{cmp_['recall']:.6f}, precision {cmp_['precision']:.6f}, skipped {skipped}/{cands}")
    eng.close()
    if not a.no_panel:
        res["panel"] = panel_bench(a.panel_data, max(1, min(a.steps, 3)))
    if not a.no_somatic:
        res["somatic"] = somatic_bench(a.somatic_data, max(1, min(a.steps, 3)), 1)
    if not a.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(a.data, L)
        cb = res["cpu_baseline"]
        res["vs_cpu"] = {"all_cores": round(res["value"] / cb["value"], 2),
                         "one_core": round(res["value"] / cb["single_thread"]["value"], 2),
                         "end_to_end_all_cores": round(res["end_to_end"]["reads_per_s"] / cb["value"], 2)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
