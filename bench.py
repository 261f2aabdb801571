"""Benchmark: exact support counting on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-e2e] [--no-cpu] [--no-per-node] [--tuned]
                    [--paths fp4,mfma,packed] [--no-sweep] [--sustain-seconds S]
    python bench.py --copy-calibration MIB     (a plain device copy, for counter scaling)

A step is one pass of the hot path over one batch: the isolated count
C[i][j] = popcount(A_i AND B_j) for 65,536 prefix bitmaps x
512 item bitmaps over 2^22 transactions (i.i.d. Bernoulli(0.3) bits, seed 1),
resident in HBM before timing.  `value` is support-count Tbit-op/s of the
first of --paths (default: the fp4 matrix-core path, each bit an exact e2m1
value summed in fp32; bit-ops = rows x items x transactions).  Also reported
in the same run: the int8-MFMA path and the packed AND+v_bcnt path on the
same data, the CPU baseline (the
oracle's popcount on a 512-row sample, host threads stated), the plain
per-node count kernel's HBM fraction on two shapes (unless --no-per-node), and unless
--no-e2e the end-to-end mining time of the dense replicated workload (49,046-row base x256
= 12,555,776 transactions, support 0.85) for the plain per-node GPU engine
(one and eight host workers) and the batched MAFIA engine with and without the
learned policy, with their MFI sets compared, and the same comparison on the
two wide workloads (clickstream: 990,000 transactions at support 0.1 %; gene:
1,100 x 20,000 at support 0.12), checked against their golden MFI digests.
Unless --no-sweep: the reference microbench's eight shapes
(code/microbench/bmma_bench.cu:438-447) on both paths beside its published
RTX 3060 Ti figures.  --sustain-seconds S (default 30, 0 = off): the headline
path run back to back on the headline shape for S seconds, per-launch times
before and after the first 3 s (bmma_bench.cu:216-289), counts re-checked
afterwards.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tensorfim_amd as tf  # noqa: E402

METRIC = "support-count Tbit-op/s; end-to-end MFI seconds, batched+learned vs plain GPU"
HBM_PEAK_GBS = 8000.0
I8_PEAK_TOPS = 5000.0      # dense int8 MFMA, MI355X
FP4_PEAK_TFLOPS = 10100.0  # dense fp4 (e2m1) MFMA, MI355X
VALU_PEAK_TOPS = 39.3      # 256 CU x 64 lanes x 2.4 GHz int32 ops
# AND + popcount of 32 bit-pairs per lane takes two VALU ops: 16 bit-ops per op
VALU_BITOP_PEAK_T = VALU_PEAK_TOPS * 16
# memory-side bytes per launch of the packed kernel on this workload (FETCH_SIZE,
# same runs as TRAFFIC_BYTES; its reads are 64-byte, factor 1.00)
PACKED_TRAFFIC_BYTES = 141876016.75 * 1024.0

# HBM bytes per launch of the default int8-MFMA count kernel on this workload,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# `bench.py --no-e2e --no-cpu --paths mfma`, each scaled by the same counter's
# reading for device traffic of known size (`bench.py --copy-calibration 1024`:
# WRITE_SIZE exact; FETCH_SIZE exact for 64-byte reads, the pattern of this
# kernel); profiles/popmat_mfma_rp_65536x512x4M_traffic.md holds the readings.
# Read 44,790,465 KiB + written 1,048,576 KiB (mean over dispatches).
TRAFFIC_KERNEL = "k_popmat_mfma_rp"
# The same readings for the fp4 matrix-core kernel (`bench.py --no-e2e --no-cpu
# --paths fp4,mfma`, profiles/popmat_fp4_65536x512x4M_kernel_stats.md): its
# loads are 64-byte row segments (8 words per row per stage), read exactly by
# FETCH_SIZE (factor 1.00, as for the int8 kernel); WRITE_SIZE exact.
# Read 36,051,761 KiB + written 1,048,576 KiB (mean over dispatches).
FP4_TRAFFIC_KERNEL = "k_popmat_fp4_256"
FP4_TRAFFIC_BYTES = (36051761 + 1048576) * 1024.0
HEAD_DESC = {
    "fp4": "fp4 matrix-core path (each bit an exact e2m1 value, v_mfma_f32_32x32x64_f8f6f4, fp32 sums exact; "
           "256x256 tiles, eight waves, stages fetched straight into LDS, operands expanded in registers)",
    "mfma": "int8-MFMA path (256x128 tiles, operand bytes expanded in registers from packed bits in LDS)",
    "packed": "packed AND + v_bcnt path (VALU)",
}
TRAFFIC_BYTES = (44790465 + 1048576) * 1024.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Host CPUs a one-GPU job may use: the GPU boxes grant each GPU a share of 16
# host CPUs, whatever the machine's thread count.
HOST_SHARE_PER_GPU = 16


def host_cpu_share():
    """Threads for the CPU baselines: the CPUs this process may run on (its
    affinity mask), capped by the cgroup CPU quota if one is set, else by
    the per-GPU host share."""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(-(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    return min(n, quota) if quota else min(n, HOST_SHARE_PER_GPU)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def isolated_sample_rows(b, pitch):
    """Rows whose counts are checked against the oracle: the first and the
    last 256-row tile, the rows on both sides of every 4 GiB boundary of A,
    and 64 random rows."""
    rows = set(range(min(256, b))) | set(range(max(0, b - 256), b))
    per4g = (1 << 32) // (pitch * 8)
    for k in range(1, b // per4g + 1):
        rows |= {k * per4g - 1, min(b - 1, k * per4g)}
    rng = np.random.default_rng(12)
    rows |= set(int(r) for r in rng.choice(b, size=min(64, b), replace=False))
    return np.array(sorted(rows), dtype=np.int64)


def cpu_baseline(A_rows, B_rows, n_bits, threads, min_seconds=10.0):
    """The oracle's popcount (bmma_bench.cu cpu_ref semantics) on a sample,
    repeated until at least `min_seconds` of CPU work has been timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    reps, t0 = 0, time.perf_counter()
    while True:
        C = oracle_lib.popcount_matrix(A_rows, B_rows, threads=threads)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    bitops = reps * A_rows.shape[0] * B_rows.shape[0] * n_bits
    return C, bitops / dt / 1e12, dt, reps


def bench_isolated(ctx, args):
    b, d, n_bits = 65536, 512, 1 << 22
    pitch = tf.pitch_words(n_bits)
    log("isolated count: A %d x %d bits (%.1f GiB), B %d x %d bits" % (b, n_bits, b * pitch * 8 / 2 ** 30, d, n_bits))
    A = ctx.buffer(b * pitch * 8)
    B = ctx.buffer(d * pitch * 8)
    C = ctx.buffer(b * d * 4)
    t0 = time.perf_counter()
    ctx.fill_bernoulli(A, b, n_bits, pitch, seed=1, matrix=0, p=0.3)
    ctx.fill_bernoulli(B, d, n_bits, pitch, seed=1, matrix=1, p=0.3)
    ctx.sync()
    log("generated on device in %.2f s" % (time.perf_counter() - t0))
    bitops = float(b) * d * n_bits
    alg_bytes = (b + d) * n_bits / 8 + 4.0 * b * d
    res = {}
    sample_rows = isolated_sample_rows(b, pitch)
    B_host = B.read(np.uint64, d * pitch).reshape(d, pitch)
    A_sample = np.stack([A.read(np.uint64, pitch, offset=int(r) * pitch * 8) for r in sample_rows])
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    t0 = time.perf_counter()
    C_exp = oracle_lib.popcount_matrix(A_sample, B_host, threads=host_cpu_share())
    log("oracle counts of %d sampled rows x %d items in %.1f s" % (len(sample_rows), d, time.perf_counter() - t0))
    for path in args.paths:
        split = 8 if path in ("mfma", "fp4") else 1
        ms = ctx.time_popcount_matrix(A, B, C, b, d, pitch, path=path, split_k=split, warmup=args.warmup,
                                      reps=args.steps)
        # exactness on the sampled rows, all 512 columns each: both 256-row
        # edge tiles, the rows around every 4 GiB boundary of A, 64 random rows
        got = np.stack([C.read(np.int32, d, offset=int(r) * d * 4) for r in sample_rows])
        ok = bool((got == C_exp).all())
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.popcount_matrix(A, B, C, b, d, pitch, path=path, split_k=split)
        ctx.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        tbit = bitops / (ms * 1e-3) / 1e12
        res[path] = {"kernel_ms": ms, "wall_ms_per_step": wall_ms, "tbitops": tbit, "exact_on_sample": ok,
                     "split_k": split}
        if path == "mfma":
            res[path]["roofline"] = {"bound": "mfma", "achieved_tops": 2 * tbit, "peak_tops": I8_PEAK_TOPS,
                                     "frac": 2 * tbit / I8_PEAK_TOPS,
                                     "memory_gbs": TRAFFIC_BYTES / (ms * 1e-3) / 1e9,
                                     "memory_frac": TRAFFIC_BYTES / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        elif path == "fp4":
            # one multiply-add of two e2m1 values per bit-op: 2 flops
            res[path]["roofline"] = {"bound": "mfma", "achieved_tflops": 2 * tbit, "peak_tflops": FP4_PEAK_TFLOPS,
                                     "frac": 2 * tbit / FP4_PEAK_TFLOPS}
        else:
            res[path]["roofline"] = {"bound": "valu", "achieved_tbitops": tbit, "peak_tbitops": VALU_BITOP_PEAK_T,
                                     "frac": tbit / VALU_BITOP_PEAK_T,
                                     "memory_gbs": PACKED_TRAFFIC_BYTES / (ms * 1e-3) / 1e9,
                                     "memory_frac": PACKED_TRAFFIC_BYTES / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log("%-6s %.3f ms/launch  %.1f Tbit-op/s  exact=%s" % (path, ms, tbit, ok))
        res[path]["checked_rows"] = int(len(sample_rows))
        if not ok:
            raise SystemExit("support counts differ from the oracle on sampled rows (%s path)" % path)
        if args.tuned:
            # the library's tuner (split_k = 0): its choice for this shape and its time
            tile, tsplit = ctx.popcount_matrix_plan(A, B, C, b, d, pitch, path=path)
            tms = ctx.time_popcount_matrix(A, B, C, b, d, pitch, path=path, split_k=0, warmup=1, reps=args.steps)
            got = np.stack([C.read(np.int32, d, offset=int(r) * d * 4) for r in sample_rows])
            res[path]["tuned"] = {"tile": tile, "split_k": tsplit, "kernel_ms": tms,
                                  "tbitops": bitops / (tms * 1e-3) / 1e12, "exact_on_sample": bool((got == C_exp).all())}
            log("%-6s tuned: tile %d split %d  %.3f ms" % (path, tile, tsplit, tms))
    if args.sustain_seconds > 0 and args.paths[0] in ("fp4", "mfma"):
        res[args.paths[0]]["sustain"] = sustain(ctx, A, B, C, b, d, pitch, args.sustain_seconds, sample_rows, C_exp,
                                                args.paths[0])
    for x in (A, B, C):
        x.free()
    # the CPU baseline's sample: the 512 distinct rows of the two edge tiles
    edge = np.concatenate([np.flatnonzero(sample_rows < 256), np.flatnonzero(sample_rows >= b - 256)])
    return res, bitops, alg_bytes, B_host, A_sample[edge], n_bits


def sustain(ctx, A, B, C, b, d, pitch, seconds, sample_rows, C_exp, path):
    """The headline matrix-core path (split-k 8) back to back for `seconds` of wall time,
    each launch timed with HIP events; mean time in the first 3 s against
    the rest, as the reference's throttle probe reports it
    (bmma_bench.cu:216-289).  The counts are re-checked afterwards."""
    log("sustain: %s path, %.0f s ..." % (path, seconds))
    bitops = float(b) * d * pitch * 64
    ms_list, walls = [], []
    t0 = time.perf_counter()
    next_log = 5.0
    while True:
        wall = time.perf_counter() - t0
        if wall >= seconds:
            break
        ms_list.append(ctx.time_popcount_matrix(A, B, C, b, d, pitch, path=path, split_k=8, warmup=0, reps=1))
        walls.append(time.perf_counter() - t0)
        if walls[-1] >= next_log:
            log("  %.0f s: %d launches, last %.2f ms" % (walls[-1], len(ms_list), ms_list[-1]))
            next_log += 5.0
    got = np.stack([C.read(np.int32, d, offset=int(r) * d * 4) for r in sample_rows])
    early = [m for m, w in zip(ms_list, walls) if w <= 3.0]
    late = [m for m, w in zip(ms_list, walls) if w > 3.0]
    out = {"seconds": walls[-1] if walls else 0.0, "launches": len(ms_list),
           "mean_ms": float(np.mean(ms_list)), "max_ms": float(np.max(ms_list)), "min_ms": float(np.min(ms_list)),
           "tbitops_mean": bitops / (float(np.mean(ms_list)) * 1e-3) / 1e12,
           "exact_after": bool((got == C_exp).all())}
    if early and late:
        out["first3s_ms"] = float(np.mean(early))
        out["after3s_ms"] = float(np.mean(late))
        out["ratio"] = out["after3s_ms"] / out["first3s_ms"]
    log("  sustain: %d launches, mean %.2f ms (%.0f Tbit-op/s), after/first 3 s %.3f, exact=%s" % (
        out["launches"], out["mean_ms"], out["tbitops_mean"], out.get("ratio", float("nan")), out["exact_after"]))
    if not out["exact_after"]:
        raise SystemExit("support counts differ from the oracle after the sustained run")
    return out


# The reference microbench's sweep (code/microbench/bmma_bench.cu:438-447):
# (transactions N, items d, prefix rows b) and its published RTX 3060 Ti
# Tbit-op/s (code/microbench/sweep_results_v2.txt:15-70): naive BMMA, smem
# BMMA, bitmap loop; None where the sweep skipped the shape.
SWEEP = [((1 << 20, 256, 128), (12.7, 8.6, 3.7)), ((1 << 20, 256, 1024), (55.9, 42.7, 1.6)),
         ((1 << 22, 256, 1024), (46.8, 41.7, 1.1)), ((1 << 22, 1024, 1024), (5.9, 53.0, 1.0)),
         ((1 << 24, 1024, 1024), (5.4, 53.3, 0.9)), ((1 << 24, 4096, 128), None),
         ((1 << 26, 256, 128), (12.7, 8.3, 1.3)), ((1 << 20, 16384, 1024), (6.0, 56.2, 1.3))]


def bench_sweep(ctx, args):
    """Both paths on each sweep shape with the library's tuned tile and split
    (split_k = 0), Bernoulli(0.3) bits; counts checked against the oracle on
    the first, middle and last rows over the first and last 64 items."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    out = []
    for (n_bits, d, b), pub in SWEEP:
        pitch = tf.pitch_words(n_bits)
        A, B, C = ctx.buffer(b * pitch * 8), ctx.buffer(d * pitch * 8), ctx.buffer(b * d * 4)
        ctx.fill_bernoulli(A, b, n_bits, pitch, seed=7, matrix=0, p=0.3)
        ctx.fill_bernoulli(B, d, n_bits, pitch, seed=7, matrix=1, p=0.3)
        rows = sorted({0, b // 2, b - 1})
        cols = sorted(set(range(min(64, d))) | set(range(max(0, d - 64), d)))
        Ar = np.stack([A.read(np.uint64, pitch, offset=r * pitch * 8) for r in rows])
        Bc = np.stack([B.read(np.uint64, pitch, offset=c * pitch * 8) for c in cols])
        exp = oracle_lib.popcount_matrix(Ar, Bc, threads=host_cpu_share())
        bitops = float(b) * d * n_bits
        ent = {"transactions": n_bits, "items": d, "rows": b}
        if pub:
            ent["published_rtx3060ti_tbitops"] = {"bmma_naive": pub[0], "bmma_smem": pub[1], "bitmap": pub[2]}
        for path in ("fp4", "mfma", "packed"):
            tile, split = ctx.popcount_matrix_plan(A, B, C, b, d, pitch, path=path)
            ms = ctx.time_popcount_matrix(A, B, C, b, d, pitch, path=path, split_k=0, warmup=1, reps=max(3, args.steps // 2))
            got = np.stack([C.read(np.int32, d, offset=r * d * 4)[cols] for r in rows])
            ok = bool((got == exp).all())
            ent[path] = {"kernel_ms": ms, "tbitops": bitops / (ms * 1e-3) / 1e12, "tile": tile, "split_k": split,
                         "exact_on_sample": ok}
            if not ok:
                raise SystemExit("sweep %s: %s path differs from the oracle" % (ent, path))
        ent["mfma"]["i8_frac"] = 2 * ent["mfma"]["tbitops"] / I8_PEAK_TOPS
        ent["fp4"]["fp4_frac"] = 2 * ent["fp4"]["tbitops"] / FP4_PEAK_TFLOPS
        log("sweep N=%d d=%d b=%d: fp4 %.0f Tbit-op/s (split %d), int8 %.0f (tile %d split %d), packed %.0f" % (
            n_bits, d, b, ent["fp4"]["tbitops"], ent["fp4"]["split_k"], ent["mfma"]["tbitops"], ent["mfma"]["tile"],
            ent["mfma"]["split_k"], ent["packed"]["tbitops"]))
        out.append(ent)
        for x in (A, B, C):
            x.free()
    return out


def bench_per_node(ctx, args):
    """The plain per-node path's kernel (tfim_count_node: prefix AND + count
    against every survivor) on one node of 3 prefix items, timed with HIP
    events over back-to-back launches.  Two shapes: the dense replicated
    workload's (39 survivor bitmaps x 12,555,776 transactions, 61 MB, which
    the 256 MB Infinity Cache can hold) and 512 bitmaps x 2^23 transactions
    (512 MB, which it cannot).  Algorithmic bytes per node = (k + survivors)
    bitmaps x N/8 + 4 bytes per survivor; the roof is HBM."""
    out = {}
    rng = np.random.default_rng(7)
    for name, d, n_bits in (("dense_replicated_39x12555776", 39, 12555776), ("large_512x8388608", 512, 1 << 23)):
        words = tf.pitch_words(n_bits)
        rows = rng.integers(0, 2 ** 63, size=(d, words), dtype=np.int64).astype(np.uint64)
        rows[:, n_bits // 64:] = 0          # zero pad past the last transaction (n_bits % 64 == 0 here)
        ctx.load_bitmaps(rows, n_bits)
        items = [0, 1, 2]
        got = ctx.count_node(items)
        m = rows[0] & rows[1] & rows[2]
        cols = range(d) if d <= 64 else list(range(8)) + [d - 1]
        ok = all(int(got[e]) == int(np.bitwise_count(m & rows[e]).sum()) for e in cols)
        ms = ctx.time_count_node(items, warmup=5, reps=max(50, 20 * args.steps))
        alg = (len(items) + d) * n_bits / 8 + 4 * d
        gbs = alg / (ms * 1e-3) / 1e9
        out[name] = {"kernel_ms": ms, "algorithmic_bytes": alg, "achieved_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                     "tbitops": d * n_bits / (ms * 1e-3) / 1e12, "exact_on_sample": ok}
        log("per-node %s: %.4f ms  %.0f GB/s (%.0f%% of HBM)  exact=%s" % (name, ms, gbs, 100 * gbs / HBM_PEAK_GBS,
                                                                          ok))
        del rows
        if not ok:
            raise SystemExit("per-node counts differ from numpy's popcount (%s)" % name)
    return out


def run_engine(data, engine, extra, out):
    t0 = time.perf_counter()
    path, st = tf.mine(data, 0.85, engine=engine, out=out, extra=["--replicate", "256"] + extra, timeout=1800)
    wall = time.perf_counter() - t0
    return path, st, wall


def bench_e2e(args):
    tmp = tempfile.mkdtemp(prefix="tfim_bench_")
    data = os.path.join(tmp, "dense_replicated_base.txt")
    subprocess.run([tf.binary("tfim_gen"), "dense-replicated", data], check=True)
    out = {}
    sets = {}
    # batched+learned and batched MAFIA; the reference TensorFIM engine's
    # enumeration in waves, autotuned and with every wave on the matrix cores
    # (its split-k matrix form, on the fp4 and on the int8 matrix cores); the
    # plain per-node engine
    # (the reference's split-k ablation, BMMA_SPLITK_OFF, kernel_bmma.cu:351-368
    # and README.md:80: the same matrix-core waves with one word range per tile)
    nodes = os.path.join(tmp, "dense_replicated.nodes")
    runs = [("batched_learned", "mafia", ["--learned"]), ("batched", "mafia", []),
            ("wave_autotuned", "wave", []), ("wave_mfma_splitk", "wave", ["--wave-kernel", "4"]),
            ("wave_mfma_int8_splitk", "wave", ["--wave-kernel", "4", "--mfma-format", "int8"]),
            ("wave_mfma_splitk_off", "wave", ["--wave-kernel", "4", "--split-k", "off"]),
            ("plain_w8", "plain", ["--workers", "8", "--dump-nodes", nodes]), ("plain", "plain", [])]
    for name, engine, extra in runs:
        log("dense replicated e2e: %s ..." % name)
        path, st, wall = run_engine(data, engine, extra, os.path.join(tmp, name + ".res"))
        sets[name] = tf.load_results(path)
        out[name] = {"wall_s": wall, "computation_s": float(st["computation_time(s)"]),
                     "rows": int(st["rows"]), "waves": int(st["waves"]), "mfis": int(st["MFIsNumber"]),
                     "kernel_ms": float(st.get("kernel_ms", 0)), "counted_bitops": float(st["counted_bitops"])}
        for k in ("hut_rows", "hut_success", "model_lookaheads", "mlp_train_steps", "graph_launches", "lds_waves",
                  "l2_waves", "mfma_waves", "fused_waves", "wave_size", "spec_hits"):
            if k in st:
                out[name][k] = int(st[k])
        for k in ("learn_time(s)", "model_setup_time(s)", "process_time(s)", "build_time(s)"):
            if k in st:
                out[name][k.replace("(s)", "_s")] = float(st[k])
        if out[name]["kernel_ms"] > 0:
            # the batched count kernels against the AND + popcount VALU roof
            tb = out[name]["counted_bitops"] / (out[name]["kernel_ms"] * 1e-3) / 1e12
            out[name]["count_kernels_tbitops"] = tb
            out[name]["count_kernels_valu_frac"] = tb / VALU_BITOP_PEAK_T
        log("  %s: %.2f s wall, %.3f s computation, %d rows" % (name, wall, out[name]["computation_s"],
                                                                out[name]["rows"]))
    out["identical_mfi_sets"] = all(sets[n] == sets["plain"] for n, _, _ in runs)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import mfi_digest
    with open(os.path.join(ROOT, "tests", "golden", "synthetic", "dense-replicated.json")) as f:
        out["matches_golden"] = mfi_digest(sets["batched_learned"]) == json.load(f)["digest"]
    out["splitk_off_slowdown"] = out["wave_mfma_splitk_off"]["computation_s"] / out["wave_mfma_splitk"]["computation_s"]
    if not args.no_cpu:
        out["cpu_engine"] = cpu_engine_baseline(data, nodes, args)
    best_plain = min(out["plain"]["computation_s"], out["plain_w8"]["computation_s"])
    out["speedup_computation_vs_plain_1worker"] = out["plain"]["computation_s"] / out["batched_learned"]["computation_s"]
    out["speedup_computation_vs_best_plain"] = best_plain / out["batched_learned"]["computation_s"]
    out["speedup_wall_vs_plain_1worker"] = out["plain"]["wall_s"] / out["batched_learned"]["wall_s"]
    if not out["identical_mfi_sets"] or not out["matches_golden"]:
        raise SystemExit("a batched MFI set differs from the plain engine or from the golden set")
    return out


def bench_e2e_wide(args):
    """End to end on the two wide workloads, each checked against its golden
    MFI digest (tests/golden/synthetic): clickstream (990,000 transactions,
    41,270 items, support 0.1 %: 1,758 survivors) and gene (1,100 x 20,000,
    support 0.12: 3,876 survivors), batched+learned and batched MAFIA next to
    the plain per-node engine."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import mfi_digest
    tmp = tempfile.mkdtemp(prefix="tfim_bench_wide_")
    res = {}
    for name, support, runs in (
            ("clickstream", 0.001, [("batched_learned", "mafia", ["--learned"]), ("batched", "mafia", []),
                                    ("plain_w8", "plain", ["--workers", "8", "--dump-nodes", "NODES"])]),
            ("gene", 0.12, [("batched_learned", "mafia", ["--learned"]), ("batched", "mafia", []),
                            ("plain_w8", "plain", ["--workers", "8", "--dump-nodes", "NODES"]),
                            ("plain", "plain", [])])):
        with open(os.path.join(ROOT, "tests", "golden", "synthetic", name + ".json")) as f:
            fx = json.load(f)
        data = os.path.join(tmp, name + ".txt")
        nodes = os.path.join(tmp, name + ".nodes")
        subprocess.run([tf.binary("tfim_gen"), name, data], check=True)
        out = {}
        for run, engine, extra in runs:
            extra = [nodes if x == "NODES" else x for x in extra]
            log("%s e2e: %s ..." % (name, run))
            t0 = time.perf_counter()
            path, st = tf.mine(data, support, engine=engine, out=os.path.join(tmp, name + run + ".res"), extra=extra,
                               timeout=600)
            wall = time.perf_counter() - t0
            got = tf.load_results(path)
            out[run] = {"wall_s": wall, "computation_s": float(st["computation_time(s)"]), "rows": int(st["rows"]),
                        "waves": int(st["waves"]), "mfis": len(got), "kernel_ms": float(st.get("kernel_ms", 0)),
                        "counted_bitops": float(st["counted_bitops"]),
                        "matches_golden": mfi_digest(got) == fx["digest"]}
            for k in ("l2_waves", "mfma_waves", "fused_waves", "wave_size", "mlp_train_steps", "hut_rows",
                      "hut_success"):
                if k in st:
                    out[run][k] = int(st[k])
            for k in ("learn_time(s)", "model_setup_time(s)", "process_time(s)"):
                if k in st:
                    out[run][k.replace("(s)", "_s")] = float(st[k])
            if out[run]["kernel_ms"] > 0:
                # listed (row, column) pairs' bit-ops over the count kernels' time
                out[run]["count_kernels_tbitops"] = out[run]["counted_bitops"] / (out[run]["kernel_ms"] * 1e-3) / 1e12
            log("  %s: %.3f s computation, %d rows, golden=%s" % (run, out[run]["computation_s"], out[run]["rows"],
                                                                  out[run]["matches_golden"]))
            if not out[run]["matches_golden"]:
                raise SystemExit("%s %s: MFI set differs from the golden set" % (name, run))
        best_plain = min(out[r]["computation_s"] for r in out if r.startswith("plain"))
        out["speedup_computation_vs_best_plain"] = best_plain / out["batched_learned"]["computation_s"]
        if not args.no_cpu:
            # gene's whole DFS (19,044 nodes) is timed; clickstream's 94,602
            # nodes (1,758 survivors x 989,270 bits) are sampled
            out["cpu_engine"] = cpu_engine_baseline(data, nodes, args, support=support, replicate=0,
                                                    per_thread=20000 if name == "gene" else 250)
        res[name] = out
    return res


def cpu_engine_baseline(data, nodes_file, args, support=0.85, replicate=256, per_thread=1000):
    """The reference bitmap engine's logic on the host cores: the worker's
    per-node step of the oracle's DFS (kernel_bitmap.cu:853-894 semantics,
    64-bit popcount counter, children built as the worker builds them) on a
    random sample of the plain GPU run's nodes (same enumeration; the node
    list is that run's), `per_thread` nodes per thread (at least 1,000) at 1
    thread and at the host share, extrapolated to every node of the list."""
    exe = os.path.join(ROOT, "oracle", "_build", "tfim_oracle")
    res = []
    nthreads = args.cpu_threads or host_cpu_share()
    for threads in (1, nthreads):
        k = max(1000, per_thread * threads)
        log("cpu engine baseline: %d threads, %d random DFS nodes ..." % (threads, k))
        cmd = [exe, data, repr(float(support)), "--threads", str(threads), "--time-nodes", nodes_file,
               "--sample", str(k), "--seed", "11"]
        if replicate > 1:
            cmd += ["--replicate", str(replicate)]
        r = subprocess.run(cmd, capture_output=True, text=True, check=True)
        st = dict(ln.split(":", 1) for ln in r.stdout.splitlines() if ":" in ln)
        total, timed, secs = int(st["nodes_total"]), int(st["nodes_timed"]), float(st["seconds"])
        full = timed >= total
        res.append({"threads": threads, "nodes_timed": timed, "seconds": secs,
                    "extrapolated_computation_s": float(st["extrapolated_s"]), "nodes_total": total, "kind": "port",
                    "cpu_model": cpu_model(),
                    "note": "every node of the DFS" if full else
                            "random sample of the DFS nodes, extrapolated to all of them"})
        log("  %.2f s for %d of %d nodes -> %.2f s" % (secs, timed, total, res[-1]["extrapolated_computation_s"]))
    return res


def copy_calibration(mib):
    """Plain device traffic of known size (torch): a copy of `mib` MiB and a
    read-only sum over the same `mib` MiB, three times each -- the reference
    points that scale the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * 2 ** 20 // 4
    src = torch.ones(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    for _ in range(3):
        s = src.sum()
    torch.cuda.synchronize()
    log("copied %d MiB x3, summed %d MiB x3 (sum %d)" % (mib, mib, int(s.item())))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--paths", default="fp4,mfma,packed")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-node", action="store_true")
    ap.add_argument("--tuned", action="store_true", help="also time the tuner's tile/split choice")
    ap.add_argument("--no-sweep", action="store_true", help="skip the reference's eight microbench shapes")
    ap.add_argument("--sustain-seconds", type=float, default=30.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    args.paths = [p for p in args.paths.split(",") if p]
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (transaction or subtree "
              "sharding is a design note); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    ctx = tf.Context(0)
    info = ctx.device_info()
    log("device:", info)
    res, bitops, alg_bytes, B_host, A_sample, n_bits = bench_isolated(ctx, args)
    per_node = None if args.no_per_node else bench_per_node(ctx, args)
    sweep = None if args.no_sweep else bench_sweep(ctx, args)
    ctx.close()
    head_path = args.paths[0]
    head = res[head_path]

    threads = args.cpu_threads or host_cpu_share()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline: oracle popcount, %d threads, 512 rows x 512 items x 2^22 ..." % threads)
        # 512 distinct device-generated rows: the two 256-row edge tiles of A
        _, cpu_tbit, cpu_s, reps = cpu_baseline(A_sample, B_host, n_bits, threads)
        log("cpu: %.3f Tbit-op/s (%d passes, %.1f s)" % (cpu_tbit, reps, cpu_s))
        cpu = {"value": cpu_tbit, "unit": "Tbit-op/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(), "hardware_threads": os.cpu_count(),
               "sample": "oracle AND-popcount of 512 distinct rows (both 256-row edge tiles of A) x 512 items x "
                         "2^22 transactions, %d passes (%.1f s) on %d threads (the host share of one GPU)"
                         % (reps, cpu_s, threads)}

    e2e = None if args.no_e2e else bench_e2e(args)
    e2e_wide = None if args.no_e2e else bench_e2e_wide(args)

    kernel_s = head["kernel_ms"] * 1e-3
    if head_path == "fp4":
        # one e2m1 multiply-add per bit-op = 2 flops
        achieved = 2.0 * bitops / kernel_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP4_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP4_PEAK_TFLOPS, "traffic": FP4_TRAFFIC_BYTES, "kernel": FP4_TRAFFIC_KERNEL,
                "algorithmic_bytes": alg_bytes}
    elif head_path == "mfma":
        achieved = 2.0 * bitops / kernel_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": I8_PEAK_TOPS, "unit": "TFLOP/s",
                "frac": achieved / I8_PEAK_TOPS, "traffic": TRAFFIC_BYTES, "kernel": TRAFFIC_KERNEL,
                "algorithmic_bytes": alg_bytes}
    else:
        achieved = alg_bytes / kernel_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    line = {
        "metric": METRIC,
        "value": head["tbitops"],
        "unit": "Tbit-op/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp4": "fp4 e2m1 x e2m1 -> fp32 (operands 0, 0.5, 1, 2; every sum an exact integer)",
                  "mfma": "i8"}.get(head_path, "u32"),
        "data": "synthetic (Bernoulli(0.3) bits generated on the device, seed 1)",
        "config": {"workload": "isolated support count: 65,536 prefix bitmaps x 512 item bitmaps x 2^22 "
                               "transactions, split-k 8, %s" % HEAD_DESC[head_path],
                   "rows": 65536, "items": 512, "transactions": 1 << 22},
        "roofline": roof,
        "cpu_baseline": cpu,
        "paths": res,
        "device": info,
    }
    if per_node is not None:
        line["per_node_kernel"] = per_node
    if sweep is not None:
        line["sweep"] = sweep
    if e2e is not None:
        line["e2e_dense_replicated"] = e2e
    if e2e_wide is not None:
        line["e2e_clickstream"] = e2e_wide["clickstream"]
        line["e2e_gene"] = e2e_wide["gene"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
