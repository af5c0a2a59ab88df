"""Benchmark: bit-exact GCUPS of the MI355X SW-Gotoh path on the
uneven-length 50,000-sequence DNA workload (persistent batched schedule vs the
plain one-launch-per-block kernel), with the CPU oracle timed beside it.

A step is one reference Block of the 50k workload decomposed as the
reference does for P = 16 ranks (RangePartitioner, 3125 x 3125 pairs,
~7e12 DP cells); the steps cycle through rank 0's off-diagonal checkerboard
blocks.  Sequences are resident in HBM before timing; each step is one call
of the drop-in host entry point (swg_compute_block), so the timed region
includes the host-side planning, the kernels and the copy of the block's
distances into pinned host memory.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU issue peak: CUs x SIMDs x lanes x clock = 3.93e13 int32 lane-operations/s;
# packed int16 (v_pk_*_i16) does two 16-bit operations per lane-instruction.
PEAK_OPS = {"int32": 256 * 4 * 16 * 2.4e9, "int16x2": 2 * 256 * 4 * 16 * 2.4e9}
# Algorithmic operations per DP cell: 24 for the recurrence with the carried
# traceback statistics, 9 for its score-only subset with the first-maximum
# update (the score pass; the reverse pass tracks where the maximum is
# reached instead); the global variant carries one more word (lead
# and trailing gap run), 30.  Mode = the arithmetic the kernel runs in.
ALG_OPS = {"score16": (9, "int16x2"), "lps16": (9, "int16x2"), "reverse": (9, "int16x2"), "rerun_capture": (24, "int32"), "rerun_track": (24, "int32"),
           "rect64": (24, "int32"),
           "one_pass": (24, "int32"), "lpp16": (24, "int32"), "global": (30, "int32")}
# VALU instructions per DP cell in each kernel's inner loop (4 columns x R rows
# unrolled), counted by hand from this build's gfx950 assembly of
# swg_kernels.hip (hipcc -S: the loop body's VALU instructions over the cells
# it evaluates); a packed instruction counts once for the two cells it
# evaluates.  The global kernel's figure is for its variant without the
# aligned-region word (protein PercentSimilarity; with it, 35.9).  Reported
# for information.
# The reverse pass evaluates a data-dependent band (the DP cells that can lie
# on an optimal path back from the maximum), so its work is the cells it
# evaluates, not the prefix rectangles it decides.
EVALUATED_IS_WORK = {"reverse"}
VALU_PER_CELL = {"score16": 4.757, "reverse": 5.672, "rerun_capture": 21.797, "rerun_track": 25.25, "one_pass": 23.375,
                 "global": 20.906}
# the score pass by (max3 form, strip rows): loop body plus a quarter of the
# first-maximum fold (once per 16 columns).  The max3 form spends 7.5 packed
# instructions on the 9 operations of a cell pair (its three-way maxima are
# one instruction each), so its roofline fraction at 9 operations per cell
# can reach 9 / 7.5 = 1.2 of the issue rate.
VALU_PER_CELL_SCORE16 = {(True, 40): 3.966, (True, 32): 3.981, (True, 24): 4.007, (False, 40): 4.716,
                         (False, 32): 4.731, (False, 24): 4.757}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# kernel variants (SWG_SCHED_* values of swg_block_config.variant)
VNAME = {1: "plain", 2: "int32 one-pass", 3: "int16 score pass + sorted int32 re-run", 4: "16 lanes per pair",
         5: "int16 score pass with 16 lanes per pair + sorted int32 re-run"}
TUNING = {0: "fixed", 1: "autotuner trial", 2: "autotuner choice", 3: "fastest so far"}


def cfg_name(cfg):
    """One line for a block's launch configuration (swg_last_block_config)."""
    v = cfg["variant"]
    shape = ("score strip %d, re-run strip %d" % (cfg["strip_rows"], cfg["rerun_strip_rows"]) if v == 3
             else "%d rows per lane" % cfg["strip_rows"] if v == 4 else "strip %d" % cfg["strip_rows"])
    return "%s, %s, %d pairs per wave; %s" % (VNAME[v], shape, cfg["pairs_per_wave"], TUNING[cfg["tuning"]])


def kernels_source_hash():
    """First 16 hex digits of the SHA-256 of the HIP kernel source: the HBM
    traffic in profiles/traffic.json holds for the source it was measured on."""
    import hashlib
    with open(os.path.join(ROOT, "csharp_amd", "csrc", "swg_kernels.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed blocks; the first blocks also time each kernel variant for the autotuner")
    ap.add_argument("--world", type=int, default=16, help="reference rank count used for the block decomposition")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--variant-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--strip", type=int, default=0, help="0 = autotuned strip height")
    ap.add_argument("--score-strip", type=int, default=0, help="int16 score-pass strip rows (0 = autotuned)")
    ap.add_argument("--reverse-strip", type=int, default=0, help="reverse-pass strip rows (0 = the default)")
    ap.add_argument("--stream-chunk-mib", type=int, default=0,
                    help="c4: device output per streamed row chunk in MiB (0 = the library's default, 64)")
    ap.add_argument("--score-max3", type=int, default=1, choices=[0, 1],
                    help="1: the score pass's max3 form where the set allows it (the default); 0: its plain form")
    ap.add_argument("--full-pairs-per-block", type=int, default=200,
                    help="c3full: seeded pairs per block checked against the oracle")
    ap.add_argument("--schedule", default="auto", choices=["auto", "i32", "i16", "lpp16", "lps16", "plain"],
                    help="force one kernel variant for the timed steps (auto = autotuned)")
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c2nw", "c3", "c3full", "c4", "c5thin"],
                    help="c3 = the headline block benchmark; c1/c2/c4 run that workload end to end with its parity "
                         "check; c3full = every block of the 50,000-sequence matrix at P = --world; c5thin = one "
                         "and four serving queries against the panel as thin blocks, with per-kernel times")
    ap.add_argument("--distance", default="pid", choices=["pid", "kimura2", "jc"],
                    help="c3 distance function: PercentIdentity (the headline), Kimura2 or JukesCantor")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (Poisson arrivals) measurement")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-rates", default="10,30,100,300,1000")
    ap.add_argument("--serve-schedule", default="auto", choices=["auto", "i16", "lpp16", "lps16"],
                    help="serving batch kernels: auto (the library's policy: the int16 16-lanes-per-pair score "
                         "pass where the alphabet allows, else 16 lanes per pair for batches too small to fill "
                         "the GPU), "
                         "i16 (the two-pass schedule for every batch), lpp16 (the int32 16-lanes-per-pair kernel "
                         "always) or lps16 (the int16 16-lanes-per-pair score pass always)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "launches": 3}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(across GPUs, each rank would compute its blocks of the reference's P x P block "
              "checkerboard on its own GPU, one process per GPU)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import csharp_amd as ca
    from csharp_amd import workloads as WL

    if args.serve_only:
        print(json.dumps({"serving": serve_bench(args)}), flush=True)
        return
    if args.workload == "c3full":
        print(json.dumps(full_c3_run(args)), flush=True)
        return
    if args.workload == "c5thin":
        print(json.dumps(thin_blocks(args)), flush=True)
        return
    if args.workload != "c3":
        print(json.dumps(workload_run(args)), flush=True)
        return

    t0 = time.time()
    seqs = WL.sequences("c3")
    M, o, e = WL.scoring("c3")
    n = len(seqs)
    L = seqs.lengths
    log("workload c3: %d sequences, mean length %.1f (generated in %.1fs)" % (n, L.mean(), time.time() - t0))

    ranges = ca.range_partition(n, args.world)
    owned = [j for j in range(args.world) if j != 0 and ca.block_is_computed(0, j)]
    blocks = [(ranges[0], ranges[j]) for j in owned]
    max_rows = max(r[1] - r[0] + 1 for r, _ in blocks)
    max_cols = max(c[1] - c[0] + 1 for _, c in blocks)

    torch.cuda.init()
    # distances land in pinned host memory through the drop-in host entry
    # point (swg_compute_block: kernels, then the D2H copy), every step
    host_bufs = [torch.empty((max_rows, max_cols), dtype=torch.int16, pin_memory=True).numpy() for _ in range(2)]
    DIST = {"pid": (ca.DIST_PERCENT_IDENTITY, "PercentIdentity"), "kimura2": (ca.DIST_KIMURA2, "Kimura2"),
            "jc": (ca.DIST_JUKES_CANTOR, "JukesCantor")}
    dist_fn, dist_name = DIST[args.distance]
    al = ca.Aligner(M, o, e, alignment_type=ca.ALIGN_NUCLEIC, distance_fn=dist_fn, tdistance=ca.TDIST_INT16,
                    schedule=ca.SCHED_PERSISTENT)
    al.set_sequences(seqs)
    SCHED = {"auto": ca.SCHED_PERSISTENT, "i32": ca.SCHED_PERSISTENT_I32, "i16": ca.SCHED_PERSISTENT_I16,
             "lpp16": ca.SCHED_PERSISTENT_LPP16, "lps16": ca.SCHED_PERSISTENT_LPS16, "plain": ca.SCHED_PLAIN}
    al.set_schedule(SCHED[args.schedule], args.strip)
    al.set_score_strip(args.score_strip)
    al.set_reverse_strip(args.reverse_strip)
    al.set_score_max3(bool(args.score_max3))

    def run_block(k):
        (r0, r1), (c0, c1) = blocks[k % len(blocks)]
        out = host_bufs[k % 2][: r1 - r0 + 1, : c1 - c0 + 1]
        _, _, pairs = al.compute_block(r0, r1, c0, c1, False, out=out)
        ms, cells, _, _ = al.last_kernel_stats()
        return pairs, cells, ms, al.last_block_config(), out

    # warm-up: graph capture, and the autotuner's trials (each a real block,
    # timed on its own); if the search needs more blocks than --warmup, the
    # warm-up goes on until the first block that runs the autotuner's choice
    warm = []
    k = 0
    while k < args.warmup or (warm and warm[-1]["tuning"] == 1 and k < args.warmup + 12):
        p, c, ms, cfg, _ = run_block(k)
        warm.append(dict(cfg, gcups=round(c / ms / 1e6, 1), kernel_ms=round(ms, 1)))
        log("warmup block %d: %.3e cells, kernel %.1f ms -> %.1f GCUPS (%s)" % (k, c, ms, c / ms / 1e6, cfg_name(cfg)))
        k += 1
    warm_run = k

    tot_cells, tot_kms, tot_pairs = 0.0, 0.0, 0
    # per-kernel times of every timed block: in the library's kernel-timing
    # mode the block's graph also runs a one-thread stamp kernel (device wall
    # clock) before its kernels and after each, on the stream they run on
    ksum = []
    al.set_kernel_timing(True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        p, c, ms, cfg, out = run_block(warm_run + k)
        tot_cells += c; tot_kms += ms; tot_pairs += p
        kt = al.kernel_times()
        if len(ksum) != len(kt) or any(a[0] != b[0] for a, b in zip(ksum, kt)):
            ksum = [[name, 0.0, 0.0, 0.0, 0] for name, _, _, _ in kt]
        for acc, (name, kms, kc, krun) in zip(ksum, kt):
            acc[1] += kms; acc[2] += kc; acc[3] += krun; acc[4] += 1
        log("step %d: %d pairs %.3e cells kernel %.1f ms -> %.1f GCUPS (%s)" % (
            k, p, c, ms, c / ms / 1e6, cfg_name(cfg)))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    al.set_kernel_timing(False)
    gcups = tot_cells / wall / 1e9
    kernel_gcups = tot_cells / (tot_kms / 1e3) / 1e9
    last_block = blocks[(warm_run + args.steps - 1) % len(blocks)]
    result = out.copy()  # the last timed block's distances, as they arrived in host memory

    # average per-kernel time over the timed blocks, for the roofline of the
    # dominant kernel
    ktimes = [(name, ms / n, kc / n, krun / n) for name, ms, kc, krun, n in ksum if n]
    for name, kms, kc, krun in ktimes:
        log("kernel %s: %.2f ms, %.3e cells (%.3e evaluated) per timed block" % (name, kms, kc, krun))

    # the same blocks into a device buffer (no D2H copy), for comparison
    dev_out = torch.empty(max_rows * max_cols, dtype=torch.int16, device="cuda")
    dc, dms = 0.0, 0.0
    td = time.perf_counter()
    for k in range(min(args.steps, 3)):
        (r0, r1), (c0, c1) = blocks[(warm_run + k) % len(blocks)]
        al.compute_block_device(r0, r1, c0, c1, False, dev_out.data_ptr(), c1 - c0 + 1)
        al.synchronize()
        dc += al.last_kernel_stats()[1]
    device_entry_gcups = dc / (time.perf_counter() - td) / 1e9
    del dev_out

    # each forced variant on one block (kernel-only GCUPS), for the record
    variants = {}
    if args.variant_steps > 0:
        for sched in (ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16):
            al.set_schedule(sched, 0)
            vc, vms = 0.0, 0.0
            for k in range(args.variant_steps):
                p, c, ms, _, _ = run_block(warm_run + k)
                vc += c; vms += ms
            variants[VNAME[sched]] = round(vc / (vms / 1e3) / 1e9, 2)
            log("variant %s: %.1f GCUPS kernel-only" % (VNAME[sched], variants[VNAME[sched]]))
        al.set_schedule(ca.SCHED_PERSISTENT, 0)

    # plain schedule: one launch per block, natural order, fixed strip
    plain_cells, plain_wall, plain_kms = 0.0, 0.0, 0.0
    if args.plain_steps > 0:
        al.set_schedule(ca.SCHED_PLAIN, 16)
        tp = time.perf_counter()
        for k in range(args.plain_steps):
            p, c, ms, _, _ = run_block(warm_run + k)
            plain_cells += c; plain_kms += ms
            log("plain step %d: kernel %.1f ms -> %.1f GCUPS" % (k, ms, c / ms / 1e6))
        plain_wall = time.perf_counter() - tp
    plain_gcups = plain_cells / plain_wall / 1e9 if plain_wall > 0 else None

    # CPU baseline (oracle, literal restatement of the reference aligner) on a
    # bounded random sample of the last timed block; the same sample checks
    # the GPU distances bit-exactly
    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as O
        host = O.cpu_info()
        threads = host["usable"]  # one worker per host CPU this process may use
        (r0, r1), (c0, c1) = last_block
        rng = np.random.default_rng(0)

        def sample(k):
            a = rng.integers(r0, r1 + 1, size=k)
            b = rng.integers(c0, c1 + 1, size=k)
            return a, b

        As, Bs, Ds = [], [], []
        cells, dt, chunk = 0.0, 0.0, 512
        while dt < args.cpu_seconds:
            a, b = sample(chunk)
            tc = time.perf_counter()
            d, _ = O.sw_pairs_mt(seqs.residues, seqs.offsets, a, b, M, o, e, dist_fn=dist_fn, threads=threads)
            dt += time.perf_counter() - tc
            cells += float((L[a].astype(np.float64) * L[b]).sum())
            As.append(a); Bs.append(b); Ds.append(d)
            chunk = min(chunk * 2, 16384)
        A = np.concatenate(As); B = np.concatenate(Bs); D = np.concatenate(Ds)
        cpu = {"value": cells / dt / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
               "sample": "%d random pairs of the last timed block (%.3e cells, %.1f s)" % (len(A), cells, dt),
               "host_nproc": host["nproc"], "host_affinity_cpus": host["affinity"],
               "host_cgroup_cpus": host["cgroup_quota_cpus"]}
        got = result[A - r0, B - c0]
        parity = {"checked_pairs": int(len(A)), "mismatches": int((got != D).sum())}
        log("cpu baseline: %.2f GCUPS on %d threads; parity %s" % (cpu["value"], threads, parity))

    al.close()
    serving = None if args.no_serve else serve_bench(args)

    # roofline of the dominant kernel: its algorithmic cells x algorithmic
    # operations per cell / its event time, against the VALU peak of its mode
    dom = max(ktimes, key=lambda k: k[1])
    dname, dms, dcells, drun = dom
    w, mode = ALG_OPS.get(dname, (None, "int32"))
    achieved = (drun if dname in EVALUATED_IS_WORK else dcells) * w / (dms / 1e3) / 1e12 if w else None
    peak = PEAK_OPS[mode] / 1e12
    # HBM bytes per launch of that kernel: a committed constant from
    # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (copy-kernel calibrated,
    # profiles/traffic.json), used only while the kernel source is the one it
    # was measured on
    traffic, traffic_src = None, None
    try:
        key = "%s|c3_block_p%d" % (dname, args.world)
        ent = json.load(open(os.path.join(ROOT, "profiles", "traffic.json"))).get(key, {})
        if ent.get("kernels_sha16") == kernels_source_hash():
            traffic = ent.get("traffic_bytes")
            traffic_src = "profiles/traffic.json: rocprofv3 counters of this kernel source (%s)" % ent["kernels_sha16"]
        elif ent:
            traffic_src = "not measured for this kernel source (profiles/traffic.json is for %s)" % (
                ent.get("kernels_sha16"),)
    except (OSError, ValueError):
        traffic = None
    line = {
        "metric": "bit-exact GCUPS (SW-Gotoh all-pairs, persistent batched schedule)",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "dtype_detail": "score and reverse passes in packed int16 (saturation-checked), statistics re-runs in int32; "
                        "exact integer arithmetic throughout",
        "data": "synthetic",
        "config": {
            "workload": "uneven-length DNA all-pairs: 50,000 sequences (60%% N(450,50), 40%% N(1450,100)), "
                        "EDNAFULL 16/4, %s Int16; step = one %dx%d reference block at P=%d"
                        % (dist_name, max_rows, max_cols, args.world),
            "sequences": n,
            "pairs_per_step": tot_pairs // max(args.steps, 1),
            "cells_per_step": tot_cells / max(args.steps, 1),
            "output": "Int16 distances in pinned host memory through swg_compute_block (D2H included)",
            "score_pass_form": "max3" if args.score_max3 else "plain",
        },
        "kernel_gcups": round(kernel_gcups, 2),
        "device_entry_gcups": round(device_entry_gcups, 2),
        "launch_config": cfg,
        "variant": cfg_name(cfg),
        "autotune_warmup": warm,
        "warmup_blocks_run": warm_run,
        "variants_kernel_gcups": variants,
        "plain_gcups": round(plain_gcups, 2) if plain_gcups else None,
        "speedup_vs_plain": round(gcups / plain_gcups, 3) if plain_gcups else None,
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3) if achieved else None,
            "peak": round(peak, 3),
            "unit": "T %s VALU ops/s (%d per DP cell)" % (mode, w) if w else None,
            "frac": round(achieved / peak, 4) if achieved else None,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": dname,
            "kernel_ms": round(dms, 2),
            "cells": dcells,
            "cells_evaluated": drun,
            "valu_instr_per_cell": (VALU_PER_CELL_SCORE16.get((bool(args.score_max3), cfg.get("strip_rows")))
                                    if dname == "score16" else VALU_PER_CELL.get(dname)),
            "note": ("the max3 form of the score pass issues 7.5 packed instructions for the 9 operations of a "
                     "cell pair (three-way maxima in one instruction): at full VALU issue its fraction is 1.2"
                     if dname == "score16" and args.score_max3 else None),
        },
        "kernels": [{"kernel": k[0], "ms": round(k[1], 2), "cells": k[2], "cells_evaluated": k[3],
                     "frac": (round((k[3] if k[0] in EVALUATED_IS_WORK else k[2]) * ALG_OPS[k[0]][0] / (k[1] / 1e3)
                                    / PEAK_OPS[ALG_OPS[k[0]][1]], 4)
                              if k[0] in ALG_OPS and k[1] > 0 else None)}
                    for k in ktimes],
        "cpu_baseline": cpu,
        "parity": parity,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


def oracle_sample(seqs, a, b, M, o, e, nucleic, tdouble, global_=False):
    from oracle import oracle as O
    threads = O.cpu_threads()
    t = time.perf_counter()
    fn = O.nw_pairs_mt if global_ else O.sw_pairs_mt
    d, _ = fn(seqs.residues, seqs.offsets, a, b, M, o, e, nucleic=nucleic, tdist_double=tdouble, threads=threads)
    return d, time.perf_counter() - t, threads


def workload_run(args):
    """C1: 1,000 DNA all pairs -> full Double matrix (SWMPI layout), every pair
    checked against the oracle.  C2: 10,000 proteins, the full Int16 matrix
    (PercentSimilarity), sampled pairs checked; c2nw: the same with the global
    (Needleman-Wunsch) variant.  C4: rank 0 of 100,000 DNA at
    P = 4: its 25,000-row stripe, diagonal block (lower-mirror) and the three
    blocks right of it, streamed to host and written as SWMS block files."""
    import hashlib
    import tempfile
    import csharp_amd as ca
    from csharp_amd import workloads as WL

    name = args.workload
    glob = name == "c2nw"
    base = "c2" if glob else name
    cfg = WL.CONFIGS[base]
    seqs = WL.sequences(base)
    M, o, e = WL.scoring(base)
    n = len(seqs)
    L = seqs.lengths.astype(np.float64)
    nucleic = cfg["alignment_type"] == ca.ALIGN_NUCLEIC
    tdouble = cfg["tdistance"] == ca.TDIST_DOUBLE
    al = ca.Aligner(M, o, e, alignment_type=cfg["alignment_type"], tdistance=cfg["tdistance"],
                    mode=ca.MODE_NW_GLOBAL if glob else ca.MODE_SW_LOCAL)
    al.set_sequences(seqs)
    if args.stream_chunk_mib:
        al.set_stream_chunk_bytes(args.stream_chunk_mib << 20)
    if args.schedule != "auto":  # a forced schedule, for comparison with the autotuned one
        al.set_schedule({"i32": ca.SCHED_PERSISTENT_I32, "i16": ca.SCHED_PERSISTENT_I16,
                         "lpp16": ca.SCHED_PERSISTENT_LPP16, "lps16": ca.SCHED_PERSISTENT_LPS16,
                         "plain": ca.SCHED_PLAIN}[args.schedule], args.strip)
    rng = np.random.default_rng(7)
    out = {"workload": name, "sequences": n}
    if name in ("c1", "c2", "c2nw"):
        al.set_kernel_timing(True)  # a stamp kernel around each of the block's kernels
        t = time.perf_counter()
        full, pairs = al.compute_full_matrix(world=1)
        wall = time.perf_counter() - t
        cells = float((np.cumsum(L)[:-1] * L[1:]).sum())
        out.update(pairs=pairs, cells=cells, seconds=round(wall, 3), gcups=round(cells / wall / 1e9, 1),
                   sha256=hashlib.sha256(full.tobytes()).hexdigest(), launch_config=al.last_block_config())
        out["kernels"] = [{"kernel": k[0], "ms": round(k[1], 3)} for k in al.kernel_times()]
        al.set_kernel_timing(False)
        if name == "c1":
            a, b = np.tril_indices(n, -1)  # every pair, s1 = the larger index (diagonal block)
        else:
            a = rng.integers(1, n, size=20000)
            b = (rng.random(20000) * a).astype(np.int64)
        d, dt, th = oracle_sample(seqs, a, b, M, o, e, nucleic, tdouble, global_=glob)
        got = full[a, b]
        same = (got == d) | (np.isnan(got) & np.isnan(d)) if tdouble else (got == d)
        sym = bool(np.array_equal(full[a, b], full[b, a], equal_nan=True)) if tdouble else bool(
            np.array_equal(full[a, b], full[b, a]))
        out["parity"] = {"checked_pairs": int(len(a)), "mismatches": int((~same).sum()), "mirror_symmetric": sym}
        out["cpu_baseline"] = {"value": round(float((L[a] * L[b]).sum()) / dt / 1e9, 3), "unit": "GCUPS",
                               "cores": th, "kind": "port", "sample": "%d pairs" % len(a)}
    else:
        world, rank = 4, 0
        ranges = ca.range_partition(n, world)
        (r0, r1) = ranges[rank]
        tmp = tempfile.mkdtemp(prefix="c4_blocks_")
        labels = ["seq%d" % i for i in range(n)]
        hashes, tot_pairs, tot_cells = {}, 0, 0.0
        # block j's file is written by a writer thread while block j+1 is
        # computed (the library call and the file write release the GIL);
        # the timed region ends when the last file is written.  The SHA-256
        # of the files is taken afterwards, outside it.
        from concurrent.futures import ThreadPoolExecutor
        writer = ThreadPoolExecutor(max_workers=1)
        pending, paths = [], []
        t = time.perf_counter()
        samples = []
        for j in range(world):
            c0, c1 = ranges[j]
            blk, _, pairs = al.compute_block(r0, r1, c0, c1, j == rank)
            path = os.path.join(tmp, "%d-%d.bin" % (rank, j))
            pending.append(writer.submit(ca.write_swms_block, path, "c4", rank, "mi355x", "c4.fasta", rank, j,
                                         (r0, r1), (c0, c1), [labels[r0], labels[r1], labels[c0], labels[c1]], blk))
            paths.append(("%d-%d" % (rank, j), path))
            tot_pairs += pairs
            tot_cells += WL.block_cells(L, r0, r1, c0, c1, diag_lower=(j == rank))
            ka = rng.integers(r0, r1 + 1, size=1500)
            kb = rng.integers(c0, c1 + 1, size=1500)
            if j == rank:
                keep = ka != kb
                ka, kb = ka[keep], kb[keep]
                ka, kb = np.maximum(ka, kb), np.minimum(ka, kb)
            samples.append((ka, kb, blk[ka - r0, kb - c0].copy()))
            log("c4 block %d-%d: %d pairs, %.1f s elapsed" % (rank, j, pairs, time.perf_counter() - t))
            del blk
            while len(pending) > 1:  # at most one block waits for its file
                pending.pop(0).result()
        for f in pending:
            f.result()
        wall = time.perf_counter() - t
        writer.shutdown()
        for name, path in paths:
            with open(path, "rb") as fh:
                hashes[name] = hashlib.sha256(fh.read()).hexdigest()
            os.remove(path)
        os.rmdir(tmp)
        A = np.concatenate([x[0] for x in samples]); B = np.concatenate([x[1] for x in samples])
        G = np.concatenate([x[2] for x in samples])
        d, dt, th = oracle_sample(seqs, A, B, M, o, e, nucleic, tdouble)
        # after the timed region: one row chunk's worth of the second block
        # (as many rows as fit a chunk) in kernel-timing mode, for the
        # per-kernel split of a streamed chunk
        c0, c1 = ranges[1]
        chunk_rows = max(1, ((args.stream_chunk_mib or 64) << 20) // ((c1 - c0 + 1) * np.dtype(np.int16).itemsize))
        al.set_kernel_timing(True)
        _, _, cpairs = al.compute_block(r0, min(r1, r0 + chunk_rows - 1), c0, c1, False)
        ckms, ccells, _, _ = al.last_kernel_stats()
        out["chunk_kernels"] = {"rows": min(chunk_rows, r1 - r0 + 1), "cols": c1 - c0 + 1, "pairs": cpairs,
                                "cells": ccells, "kernel_ms": round(ckms, 2), "gcups": round(ccells / ckms / 1e6, 1),
                                "kernels": [{"kernel": k[0], "ms": round(k[1], 3)} for k in al.kernel_times()]}
        al.set_kernel_timing(False)
        out.update(rank=rank, world=world, pairs=tot_pairs, cells=tot_cells, seconds=round(wall, 2),
                   gcups_end_to_end=round(tot_cells / wall / 1e9, 1), block_files_sha256=hashes,
                   parity={"checked_pairs": int(len(A)), "mismatches": int((G != d).sum())},
                   cpu_baseline={"value": round(float((L[A] * L[B]).sum()) / dt / 1e9, 3), "unit": "GCUPS",
                                 "cores": th, "kind": "port", "sample": "%d pairs" % len(A)})
    al.close()
    return out


def thin_blocks(args):
    """Thin blocks of the serving workload without the server: the first
    query (395 residues), then the first four, against the 20,000-sequence
    panel through the host entry point with the library's choice for thin
    blocks (the lanes-per-pair int16 score pass and the int32 re-run), three
    times each, the last in kernel-timing mode; bit-exact against the oracle
    on 2,000 seeded panel pairs per query."""
    import csharp_amd as ca
    from csharp_amd import workloads as WL
    panel = WL.sequences("c5_panel")
    qset = WL.sequences("c5_queries")
    M, o, e = WL.scoring("c5_panel")
    rng = np.random.default_rng(7)
    out = {"workload": "c5thin", "panel": len(panel), "blocks": []}
    for nq in (1, 4):
        qs = [qset.seq(i) for i in range(nq)]
        s = ca.SequenceSet.from_strings(qs + [panel.seq(i) for i in range(len(panel))])
        al = ca.Aligner(M, o, e)
        al.set_sequences(s)
        n = len(s)
        for rep_ in range(3):
            al.set_kernel_timing(rep_ == 2)
            t = time.perf_counter()
            got, _, pairs = al.compute_block(0, nq - 1, nq, n - 1, False)
            dt = time.perf_counter() - t
        kt = al.kernel_times()
        ms, cells, _, _ = al.last_kernel_stats()
        a = np.repeat(np.arange(nq), 2000)
        b = np.concatenate([rng.choice(np.arange(nq, n), 2000, replace=False) for _ in range(nq)])
        d, _, _ = oracle_sample(s, a, b, M, o, e, True, False)
        blk = {"queries": nq, "query_lengths": [len(q) for q in qs], "pairs": int(pairs), "cells": cells,
               "host_ms": round(dt * 1e3, 2), "kernel_ms": round(ms, 2), "gcups": round(cells / ms / 1e6, 1),
               "launch_config": al.last_block_config(),
               "kernels": [{"kernel": k[0], "ms": round(k[1], 3)} for k in kt],
               "parity": {"checked_pairs": int(len(a)), "mismatches": int((got[a, b - nq] != d).sum())}}
        log("c5thin %d queries: %s" % (nq, blk))
        out["blocks"].append(blk)
        al.close()
    return out


def full_c3_run(args):
    """All 1,249,975,000 pairs of the 50,000-sequence workload as the
    reference's ranks compute them at P = --world: for every rank r its
    checkerboard blocks (r, j) (Salsa.SmithWatermanMPI/Program.cs:618-642),
    the P diagonal blocks with the lower mirror (:209-231), each through the
    host entry point into pinned host memory; a seeded sample of every block
    checked against the oracle."""
    import torch
    import csharp_amd as ca
    from csharp_amd import workloads as WL
    from oracle import oracle as O

    seqs = WL.sequences("c3")
    M, o, e = WL.scoring("c3")
    n = len(seqs)
    L = seqs.lengths.astype(np.float64)
    world = args.world
    ranges = ca.range_partition(n, world)
    todo = [(r, j) for r in range(world) for j in range(world) if ca.block_is_computed(r, j)]
    rmax = max(b - a + 1 for a, b in ranges)
    buf = torch.empty((rmax, rmax), dtype=torch.int16, pin_memory=True).numpy()
    al = ca.Aligner(M, o, e, alignment_type=ca.ALIGN_NUCLEIC, tdistance=ca.TDIST_INT16)
    al.set_sequences(seqs)
    rng = np.random.default_rng(9)
    samples = []
    tot_pairs, tot_cells = 0, 0.0
    kinds = {"diagonal": [0, 0.0, 0.0], "off-diagonal": [0, 0.0, 0.0]}  # blocks, cells, seconds
    t0 = time.perf_counter()
    for k, (r, j) in enumerate(todo):
        (r0, r1), (c0, c1) = ranges[r], ranges[j]
        diag = r == j
        out = buf[: r1 - r0 + 1, : c1 - c0 + 1]
        tb = time.perf_counter()
        _, _, pairs = al.compute_block(r0, r1, c0, c1, diag, out=out)
        dt = time.perf_counter() - tb
        cells = WL.block_cells(L, r0, r1, c0, c1, diag_lower=diag)
        kd = kinds["diagonal" if diag else "off-diagonal"]
        kd[0] += 1; kd[1] += cells; kd[2] += dt
        tot_pairs += pairs
        tot_cells += cells
        m = args.full_pairs_per_block
        if diag:
            a = rng.integers(r0 + 1, r1 + 1, size=m)
            b = r0 + (rng.random(m) * (a - r0)).astype(np.int64)
        else:
            a = rng.integers(r0, r1 + 1, size=m)
            b = rng.integers(c0, c1 + 1, size=m)
        samples.append((a, b, out[a - r0, b - c0].copy()))
        if k % 8 == 0 or diag or k == len(todo) - 1:
            log("c3full block %d/%d (%d, %d)%s: %.1f GCUPS, %.1f s elapsed" % (
                k + 1, len(todo), r, j, " diagonal" if diag else "", cells / dt / 1e9, time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    cfg = al.last_block_config()
    al.close()
    A = np.concatenate([x[0] for x in samples]); B = np.concatenate([x[1] for x in samples])
    G = np.concatenate([x[2] for x in samples])
    tc = time.perf_counter()
    D, _ = O.sw_pairs_mt(seqs.residues, seqs.offsets, A, B, M, o, e, threads=O.cpu_threads())
    log("c3full oracle sample: %d pairs in %.1f s" % (len(A), time.perf_counter() - tc))
    return {
        "workload": "c3full", "sequences": n, "world": world, "blocks": len(todo), "pairs": tot_pairs,
        "cells": tot_cells, "seconds": round(wall, 2), "gcups_end_to_end": round(tot_cells / wall / 1e9, 1),
        "per_kind": {kk: {"blocks": v[0], "cells": v[1], "seconds": round(v[2], 2),
                          "gcups": round(v[1] / v[2] / 1e9, 1) if v[2] else None} for kk, v in kinds.items()},
        "output": "Int16 distances in pinned host memory through swg_compute_block (D2H included)",
        "launch_config": cfg,
        "parity": {"checked_pairs": int(len(A)), "mismatches": int((G != D).sum())},
    }


# queries per serving batch: a deep queue fills batches of up to this many
# queries (one launch per batch)
SERVE_MAX_BATCH = 32


def serve_bench(args):
    """Serving workload: a device-resident panel of 20,000 DNA sequences
    (lengths 300-1500) and query arrivals (lengths 50-2000).  First the
    saturation rate: a burst of queries all submitted at once (closed system,
    completed queries / span).  Then Poisson arrivals at each rate, at least
    200 queries per rate so that p99 rests on a hundred-odd samples above the
    median; a rate above 90% of the saturation rate is labelled an overload
    run, whose latencies measure the queue building up over the run rather
    than a steady state."""
    import csharp_amd as ca
    from csharp_amd import workloads as WL
    from csharp_amd.serving import run_poisson

    panel = WL.sequences("c5_panel")
    qset = WL.sequences("c5_queries")
    M, o, e = WL.scoring("c5_panel")
    qmax = int(qset.lengths.max())
    sched = {"auto": ca.SCHED_PERSISTENT, "i16": ca.SCHED_PERSISTENT_I16, "lpp16": ca.SCHED_PERSISTENT_LPP16,
             "lps16": ca.SCHED_PERSISTENT_LPS16}
    srv = ca.Server(panel, M, o, e, max_query_len=qmax, max_batch=SERVE_MAX_BATCH, lanes=3, hbm_budget=4 << 30,
                    stats=False, schedule=sched[args.serve_schedule])
    # warm-up: scratch allocation and graph capture on every lane
    warm = [srv.submit(qset.seq(i)) for i in range(6)]
    for t in warm:
        srv.wait(t)
        srv.release(t)
    out = {"panel": len(panel), "max_batch": SERVE_MAX_BATCH, "lanes": 3, "schedule": args.serve_schedule,
           "results": []}
    base = 6

    def take(k):
        nonlocal base
        qs = [qset.seq((base + i) % len(qset)) for i in range(k)]
        base += k
        return qs

    burst = run_poisson(srv, take(320), 1e9, seed=1)
    sat = burst["achieved_qps"]
    out["saturation"] = {"queries": burst["queries"], "qps": sat, "gcups": burst["gcups"],
                         "p50_ms": burst["p50_ms"], "p99_ms": burst["p99_ms"],
                         "method": "320 queries submitted at once; completed queries / (last completion - first submit)"}
    log("serving saturation: %.1f q/s, %.1f GCUPS" % (sat, burst["gcups"]))
    for rate in [float(x) for x in args.serve_rates.split(",") if x]:
        r = run_poisson(srv, take(200), rate, seed=int(rate))
        r["regime"] = "steady" if rate <= 0.9 * sat else "overload"
        log("serving %.0f q/s (%s): %s" % (rate, r["regime"], r))
        out["results"].append(r)
    srv.close()
    return out


if __name__ == "__main__":
    main()
