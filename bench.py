#!/usr/bin/env python3
"""Benchmark of the MI355X MCZI counting and subtraction path.

One JSON line. The headline is MC dataset-presence counting of 32 datasets in
one segmented batch (9.6 Gbp, the largest single-GPU MC workload):
  20 Mbp iid genome (seed 4); 32 variants with 0.1% SNPs, each sequenced as
  150 bp reads at 15x with 0.5% substitutions; k=31, m=21, x=5 (present in
  more than 5 datasets), 1024 partitions.
The other workloads are reported beside it, each with the same fields:
  c2: 100 Mbp genome (seed 2) + 2% in 5 kbp repeats x10, 150 bp reads at 30x,
      1% substitutions; k31 m21 x5; FASTA-path counting;
  c3: 50 Mbp genome (seed 3), lognormal long reads (median 8 kbp, sigma 0.8,
      [500, 100000]) at 20x with 8% errors (4/2/2 sub/ins/del) and 150 bp
      reads at 10x in one batch; k31 m21 x3;
  c1: 1 Mbp genome (seed 1), 200k 150 bp reads, 0.5% substitutions; k31 m21
      x1, 64 partitions, KFF (KMC-style single pass);
  zi: served subtraction, index = canonical k-mers of a 100 Mbp genome
      (seed 5); 10k requests of 150 bp reads, log-uniform in [10, 1e6] reads
      (half from the indexed genome with 0.5% substitutions, half from an
      unrelated genome), at k=31 (m=19) and k=63 (m=23).
A step = one pass of the counting path over the whole read set, resident in
HBM before timing starts: minimizer scan + super-k-mer records -> two-pass
partitioning into buckets -> per-partition LDS counting with HBM spill ->
count > x -> output order. Each MC workload is autotuned first (LDS table
size, table fill and so the partition count, scan blocks per CU; untimed).
`plain` is the per-file GPU path: with several input files (c3: the long-read
and short-read files; c4: the 32 dataset files) each file is counted as a job
of its own (scan, partitioning and count of its tiles, every k-mer kept) and
the per-file (k-mer, count) pairs are merged on the device (sorted, summed,
count > x); a single-file workload scans one 128 Mi-base chunk per launch.
`cpu_baseline` is the oracle's restatement of the reference CPU path timed
on a stated sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the factor a plain
    copy kernel of known size reads on the same counter), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            return json.load(f)[workload]["kernels"][kernel]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def zi_traffic(k: int, lookups: int):
    """HBM bytes of the stream's lookup kernels: the bytes per lookup that
    the committed PMC summary measured on a request stream of the same
    shapes (FETCH_SIZE and WRITE_SIZE over every dispatch, scaled like
    measured_traffic), times the stream's lookups; None without it."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            per = json.load(f)["zi_request_stream"]["kernels"][f"k_zi_lookup_k{k}"]["bytes_per_lookup"]
        return per * lookups
    except (OSError, KeyError, ValueError):
        return None


def decode_range(words, b0, b1):
    """Packed-seq codes (one byte per base) of bases [b0, b1)."""
    import numpy as np
    w0 = b0 // 16
    w = words[w0:(b1 + 15) // 16 + 1].view(np.uint8)
    b = np.unpackbits(w[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
    c = (b[:, :, 0] | (b[:, :, 1] << 1)).reshape(-1).astype(np.uint8)
    return c[b0 - 16 * w0:b1 - 16 * w0]


def concat_reads(a, b):
    """Two packed read sets in one segmented buffer (word-aligned; the gap
    between them is a fragment too short for any k-mer)."""
    import numpy as np
    na = int(a.frag_off[-1])
    wa = (na + 15) // 16
    pad = wa * 16
    words = np.concatenate([a.words[:wa], b.words])
    off = np.concatenate([a.frag_off, [pad] if pad != na else [], b.frag_off[1:] + pad]).astype(np.uint64)
    return words, off


# ---- MC workloads ------------------------------------------------------------
def make_c1(M):
    r = M.SynthReads(M.SynthSpec(genome_len=1_000_000, genome_seed=1, read_seed=101, n_reads=200_000, read_len=150,
                                 sub_rate=0.005))
    return r, r.words, r.frag_off, None, None


def make_c2(M):
    r = M.SynthReads(M.SynthSpec(genome_len=100_000_000, genome_seed=2, read_seed=102, n_reads=20_000_000,
                                 read_len=150, sub_rate=0.01, repeat_fraction=0.02, repeat_len=5000,
                                 repeat_copies=10))
    return r, r.words, r.frag_off, None, None


def make_c3(M):
    import numpy as np
    long = M.SynthReads(M.SynthSpec(genome_len=50_000_000, genome_seed=3, read_seed=103, long_bases=1_000_000_000,
                                    long_median=8000, long_sigma=0.8, long_min=500, long_max=100_000,
                                    sub_rate=0.04, ins_rate=0.02, del_rate=0.02))
    short = M.SynthReads(M.SynthSpec(genome_len=50_000_000, genome_seed=3, read_seed=203, n_reads=3_333_334,
                                     read_len=150, sub_rate=0.005))
    words, off = concat_reads(long, short)
    # two inputs: the long-read file (and the word-alignment gap) and the short-read file
    n_long = len(off) - 1 - (len(short.frag_off) - 1)
    marks = np.array([0, n_long, len(off) - 1], dtype=np.uint64)
    return (long, short), words, off, None, marks


def make_c4(M):
    import numpy as np
    r = M.SynthReads(M.SynthSpec(genome_len=20_000_000, genome_seed=4, read_seed=104, n_reads=2_000_000,
                                 read_len=150, sub_rate=0.005, n_datasets=32, snp_rate=0.001))
    marks = np.arange(0, 32 * 2_000_000 + 1, 2_000_000, dtype=np.uint64)  # one input per dataset file
    return r, r.words, r.frag_off, r.dataset, marks


# name: (maker, k, m, x, partitions, mode, order, autotune, description)
MC = {
    "c4_fof_32_datasets_20mbp": (make_c4, 31, 21, 5, 1024, "datasets", "encoded", True,
                                 "fof dataset presence, 32 datasets in one segmented batch (autotuned)"),
    "c2_short_reads_100mbp": (make_c2, 31, 21, 5, 1024, "single", "encoded", True,
                              "FASTA-path counting, every super-k-mer staged (autotuned)"),
    "c3_long_and_short_reads_50mbp": (make_c3, 31, 21, 3, 1024, "single", "encoded", True,
                                      "long and short reads in one batch (autotuned)"),
    "c1_small_1mbp": (make_c1, 31, 21, 1, 64, "single", "kff", True, "KMC-style KFF counting (autotuned)"),
}


def cpu_baseline_mc(M, name, reads, words, off, ds, k, m, x, parts, seconds_target=12.0):
    """The oracle's restatement of the reference CPU path on a bounded sample
    of the same input, on every host core where the reference path is
    threaded (Rayon), else one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O
    cores = os.cpu_count() or 1
    n_frag = len(off) - 1
    if name.startswith("c4"):
        # fof: the first reads of each of the 32 datasets, on every host core
        # (datasets in order, each dataset's reads split among the threads)
        per = 4000
        while True:
            parts_c, offs, dss = [], [0], []
            for d in range(32):
                f0 = d * 2_000_000
                c = decode_range(words, int(off[f0]), int(off[f0 + per]))
                parts_c.append(c)
                offs.extend((np.asarray(off[f0 + 1:f0 + per + 1], dtype=np.int64) - int(off[f0]) + offs[-1]).tolist())
                dss.extend([d] * per)
            codes = np.concatenate(parts_c)
            o = np.array(offs, dtype=np.uint64)
            t0 = time.perf_counter()
            O.count_datasets_mt(codes, o, np.array(dss, dtype=np.uint32), 32, k, m, x, parts, cores)
            dt = time.perf_counter() - t0
            if dt > seconds_target / 4 or per >= 200_000:
                break
            per = min(200_000, int(per * max(2.0, seconds_target / max(dt, 1e-3))))
        return {"value": int(o[-1]) / dt, "unit": "bases/s", "cores": cores, "kind": "port",
                "sample": f"first {per} reads of each of the 32 datasets ({int(o[-1])} bases), fof dataset "
                          f"presence with the dataset-minimizer pass, {cores} threads, {dt:.1f} s"}
    if name.startswith("c1"):
        codes = decode_range(words, 0, int(off[-1]))
        t0 = time.perf_counter()
        O.kff_u8_kmc(codes, off, k, m, x, parts)
        dt = time.perf_counter() - t0
        return {"value": int(off[-1]) / dt, "unit": "bases/s", "cores": 1, "kind": "port",
                "sample": f"the whole read set ({int(off[-1])} bases), KMC-style KFF path, 1 thread, {dt:.1f} s"}
    # start near 3e7 bases, then grow the prefix toward seconds_target
    take = max(1, min(n_frag, int(np.searchsorted(np.asarray(off), 30_000_000))))
    while True:
        codes = decode_range(words, 0, int(off[take]))
        o = np.asarray(off[:take + 1], dtype=np.uint64)
        t0 = time.perf_counter()
        O.count_two_pass_mt(codes, o, k, m, x, parts, cores)
        dt = time.perf_counter() - t0
        if dt > seconds_target / 4 or take >= n_frag:
            break
        take = min(n_frag, int(take * max(2.0, seconds_target / max(dt, 1e-3))))
    return {"value": int(o[-1]) / dt, "unit": "bases/s", "cores": cores, "kind": "port",
            "sample": f"first {take} fragments ({int(o[-1])} bases) of the same input, two-pass FASTA path, "
                      f"{cores} threads, {dt:.1f} s"}


def mc_workload(M, name, steps, warmup, cpu=True, plain=True):
    import numpy as np
    maker, k, m, x, parts, mode_s, order_s, tune, desc = MC[name]
    mode = {"datasets": M.MODE_DATASETS, "single": M.MODE_SINGLE_PASS}[mode_s]
    order = M.ORDER_ENCODED if order_s == "encoded" else M.ORDER_KFF_U8
    t0 = time.perf_counter()
    reads, words, off, ds, marks = maker(M)
    gen_s = time.perf_counter() - t0
    nb = int(off[-1])
    print(f"[bench] {name}: generated {nb} bases in {len(off) - 1} fragments ({gen_s:.1f} s)",
          file=sys.stderr, flush=True)
    t_up = time.perf_counter()
    batch = M.Batch.from_packed(words, off, ds, marks)
    upload_s = time.perf_counter() - t_up
    cfg = M.CounterConfig(k, m, x, parts)
    print(f"[bench] {name}: uploaded ({upload_s:.2f} s)", file=sys.stderr, flush=True)
    tuned = None
    if tune:
        t_tune = time.perf_counter()
        tuned = batch.autotune(cfg, mode, M.LAUNCH_BATCHED, order)
        print(f"[bench] {name}: autotuned in {time.perf_counter() - t_tune:.1f} s, {tuned:.0f} k-mer instances "
              f"per partition", file=sys.stderr, flush=True)
    for i in range(warmup):
        st = batch.count(cfg, mode, M.LAUNCH_BATCHED, order)
        print(f"[bench] {name} warmup {i}: {st.ms_total + st.ms_sort:.1f} ms device, out {st.n_out}",
              file=sys.stderr, flush=True)
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = batch.count(cfg, mode, M.LAUNCH_BATCHED, order)
        dev_ms.append(st.ms_total + st.ms_sort)
    wall = time.perf_counter() - t0
    last = st.as_dict()
    res = {"workload": name, "description": desc, "k": k, "m": m, "x": x, "partition_count": parts,
           "bases": nb, "fragments": int(len(off) - 1), "generate_s": gen_s,
           "value": nb * steps / wall, "unit": "bases/s", "ms_per_step": 1e3 * wall / steps,
           "device_ms_per_step": sum(dev_ms) / len(dev_ms), "upload_ms": upload_s * 1e3,
           "pcie_inclusive_value": nb / (wall / steps + upload_s), "autotuned_instances_per_partition": tuned,
           "tuned": {"lds_slots": st.lds_slots, "internal_partitions": st.internal_partitions,
                     "scan_blocks_per_cu": st.scan_blocks_per_cu}}
    # the two-pass mode (the reference's pass 1 run on the device, filtering staging)
    two = M.MODE_DATASETS_TWO_PASS if mode == M.MODE_DATASETS else M.MODE_TWO_PASS
    st2 = batch.count(cfg, two, M.LAUNCH_BATCHED, order)
    res["two_pass"] = {"device_ms": st2.ms_total + st2.ms_sort + st2.ms_pass1, "pass1_ms": st2.ms_pass1,
                       "unique_minimizers": st2.unique_minimizers, "minimizers_retained": st2.minimizers_retained,
                       "n_out": st2.n_out, "same_output": st2.n_out == last["n_out"]}
    print(f"[bench] {name}: {res['ms_per_step']:.1f} ms per step; two-pass {res['two_pass']['device_ms']:.1f} ms",
          file=sys.stderr, flush=True)
    # per-kernel device times: one eager run with HIP events between kernels
    os.environ["MCZI_NO_GRAPH"] = "1"
    prof = batch.count(cfg, mode, M.LAUNCH_BATCHED, order).as_dict()
    if plain:
        batch.count(cfg, mode, M.LAUNCH_PLAIN, order)
        tp = time.perf_counter()
        for _ in range(2):
            stp = batch.count(cfg, mode, M.LAUNCH_PLAIN, order)
        res["plain_value"] = nb * 2 / (time.perf_counter() - tp)
        res["batched_vs_plain"] = res["value"] / res["plain_value"]
        res["plain_launch"] = ("each input file counted on its own (scan, partitioning, count of every k-mer), "
                               "the per-file counts merged on the device" if marks is not None
                               else "one scan launch per 128 Mi bases, one count")
        assert stp.n_out == last["n_out"], "plain and batched counts differ"
    del os.environ["MCZI_NO_GRAPH"]
    R = prof["n_records"]
    P = prof["internal_partitions"]
    # algorithmic (compulsory) bytes per kernel group, each operand once:
    #  k_scan: packed reads in (2 bits/base), 16-byte records out
    #  partitioning: every record read once and written once into its bucket
    #  k_count: partitioned records + partition bounds in, (k-mer, count, partition) out
    kernels = {"k_scan": (nb / 4 + 16 * R, prof["ms_scan1"]),
               "partitioning": (32 * R, prof["ms_scan2"]),
               "k_count": (16 * R + 16 * P + 16 * prof["n_out"], prof["ms_count"])}
    dom = max(kernels, key=lambda n: kernels[n][1])
    alg, ms = kernels[dom]
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    res["phase_ms"] = {n: prof[n] for n in ("ms_scan1", "ms_scan2", "ms_count", "ms_sort", "ms_total")}
    res["kernel_groups"] = {n: {"algorithmic_bytes": a, "ms": t, "GBs": a / (t * 1e-3) / 1e9 if t > 0 else 0.0}
                            for n, (a, t) in kernels.items()}
    res["stats"] = {n: last[n] for n in ("n_windows", "n_pieces", "n_records", "n_distinct", "n_overflow", "n_out",
                                         "internal_partitions", "partition_bytes")}
    res["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(name, dom),
                       "algorithmic_bytes": alg, "kernel_ms": ms}
    del batch
    print(f"[bench] {name}: plain {res.get('plain_value', 0):.3g} bases/s; CPU baseline next",
          file=sys.stderr, flush=True)
    res["cpu_baseline"] = cpu_baseline_mc(M, name, reads, words, off, ds, k, m, x, parts) if cpu else None
    return res


# ---- ZI: the served request stream ----------------------------------------------
def zi_stream(M, k, n_requests=10_000, pool_reads=1 << 24, seed=5, cpu=True, lo=10, hi=1_000_000):
    """ZI served subtraction of a request stream: read counts log-uniform in
    [lo, hi] (rounded up to a multiple of 8), each request a slice of a
    seeded pool of 150 bp reads alternating between the indexed genome and an
    unrelated one (0.5% substitutions), starting at a seeded read. Requests
    are generated before timing. Measures the plain path (one request at a
    time), the served path (all submitted at once, continuous batching) and
    Poisson arrivals at 1.5x the plain request rate (latency p50 / p99)."""
    import ctypes as C

    import numpy as np

    G = 100_000_000
    t_gen = time.perf_counter()
    ga = M.synth_genome(G, seed)
    gb = M.synth_genome(G, seed + 1000)
    codes = np.unpackbits(ga.view(np.uint8)[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
    codes = (codes[:, :, 0] | (codes[:, :, 1] << 1)).reshape(-1)[:G]
    text = b">index\n" + np.frombuffer(b"ACTG", dtype=np.uint8)[codes].tobytes() + b"\n"
    rng = np.random.default_rng(seed + k)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_requests)).astype(np.int64)
    counts = np.minimum(((counts + 7) // 8) * 8, pool_reads)
    pool = M.synth_reads_from([(ga, G), (gb, G)], pool_reads, 150, 0.005, seed + 1)
    starts = rng.integers(0, (pool_reads - counts) // 8 + 1) * 8
    off_all = np.arange(0, 150 * (int(counts.max()) + 1), 150, dtype=np.uint64)
    words_ptr = pool.words.ctypes.data
    gen_s = time.perf_counter() - t_gen
    m = 19 if k == 31 else 23
    h = C.c_void_p()
    t0 = time.perf_counter()
    M._check(M._lib.mczi_zi_server_create_from_text(text, len(text), k, m, 96 << 30, 3, C.byref(h)))
    build_s = time.perf_counter() - t0
    del text

    def submit(i):
        rid = C.c_uint64()
        M._check(M._lib.mczi_zi_server_submit_packed(h, words_ptr + int(starts[i]) * 150 // 4,
                                                      off_all.ctypes.data, int(counts[i]), 0, C.byref(rid)))
        return rid.value

    def poll(rid):
        p = C.POINTER(M._ZiResult)()
        M._check(M._lib.mczi_zi_server_poll(h, rid, 1, C.byref(p)))
        r = p.contents
        out = (r.n_windows, r.n_absent, r.latency_ms, r.queue_ms)
        M._lib.mczi_zi_result_free(p)
        return out

    def lookup_stats():
        n, ms = C.c_uint64(), C.c_double()
        M._check(M._lib.mczi_zi_server_lookup_stats(h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    print(f"[bench] zi k={k}: {n_requests} requests, {int(counts.sum())} reads; index built in {build_s:.1f} s",
          file=sys.stderr, flush=True)
    try:
        poll(submit(0))  # warm-up
        l0, k0 = lookup_stats()
        t0 = time.perf_counter()
        plain = []
        for i in range(n_requests):
            plain.append(poll(submit(i)))
            if (i + 1) % 2000 == 0:
                print(f"[bench] zi k={k}: plain {i + 1} requests, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        plain_s = time.perf_counter() - t0
        l1, k1 = lookup_stats()
        lookups = sum(x[0] for x in plain)
        print(f"[bench] zi k={k}: plain {plain_s:.1f} s, {lookups} lookups", file=sys.stderr, flush=True)
        # served, closed loop: requests are submitted while at most
        # `window_bytes` of request reads are outstanding, results collected
        # in order (continuous batching behind a bounded client queue)
        t0 = time.perf_counter()
        window_bytes = 1 << 30
        burst, queue, nxt, outstanding = [], [], 0, 0
        while len(burst) < n_requests:
            while nxt < n_requests and (not queue or outstanding + int(counts[nxt]) * 150 // 4 <= window_bytes):
                queue.append((submit(nxt), int(counts[nxt]) * 150 // 4))
                outstanding += queue[-1][1]
                nxt += 1
            rid, nbytes = queue.pop(0)
            burst.append(poll(rid))
            outstanding -= nbytes
            if len(burst) % 2000 == 0:
                print(f"[bench] zi k={k}: served {len(burst)} results, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        burst_s = time.perf_counter() - t0
        assert [x[1] for x in burst] == [x[1] for x in plain], "served results differ from plain"
        print(f"[bench] zi k={k}: served {burst_s:.1f} s", file=sys.stderr, flush=True)
        rate = 1.5 * n_requests / plain_s
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
        ids = [None] * n_requests

        def feeder():
            tb = time.perf_counter()
            for i in range(n_requests):
                d = arrivals[i] - (time.perf_counter() - tb)
                if d > 0:
                    time.sleep(d)
                ids[i] = submit(i)

        th = threading.Thread(target=feeder)
        t0 = time.perf_counter()
        th.start()
        res = []
        for i in range(n_requests):
            while ids[i] is None:
                time.sleep(0.0005)
            res.append(poll(ids[i]))
            if (i + 1) % 2000 == 0:
                print(f"[bench] zi k={k}: Poisson {i + 1} results, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        th.join()
        open_s = time.perf_counter() - t0
        lat = np.array([x[2] for x in res])
        assert [x[1] for x in res] == [x[1] for x in plain], "Poisson-served results differ from plain"
    finally:
        M._lib.mczi_zi_server_destroy(h)
    # roofline: each lookup gathers one 64-byte line of the index (the
    # compulsory access of a hashed probe); lookup kernel time from HIP events
    kernel_ms = k1 - k0
    achieved = (l1 - l0) * 64 / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    out = {
        "k": k, "m": m, "requests": n_requests, "reads": int(counts.sum()), "lookups": int(lookups),
        "request_reads": {"min": int(counts.min()), "median": float(np.median(counts)), "max": int(counts.max())},
        "generate_s": gen_s, "index_build_s": build_s,
        "plain_lookups_per_s": lookups / plain_s,
        "served_lookups_per_s": lookups / burst_s,
        "served_vs_plain": plain_s / burst_s,
        "served_client_window_bytes": window_bytes,
        "poisson_rate_per_s": rate,
        "poisson_lookups_per_s": lookups / open_s,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "plain_p50_ms": float(np.percentile([x[2] for x in plain], 50)),
        "plain_p99_ms": float(np.percentile([x[2] for x in plain], 99)),
        "lookup_kernel_lookups_per_s": (l1 - l0) / (kernel_ms * 1e-3) if kernel_ms > 0 else 0.0,
        "roofline": {"bound": "hbm", "kernel": "k_zi_lookup", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": zi_traffic(k, l1 - l0),
                     "algorithmic_bytes": (l1 - l0) * 64, "kernel_ms": kernel_ms,
                     "note": "64 B per lookup (one index line); lookups/s ceiling 1.25e11; traffic = the "
                             "measured HBM bytes per lookup of the committed PMC summary x lookups"},
        "cpu_baseline": None,
    }
    if cpu:
        # ZI's one-thread query loop (zi.rs:152-175) over the same index, on the
        # first requests of the stream up to ~5e6 lookups
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        take, tot = 0, 0
        while take < n_requests and tot < 5_000_000:
            tot += int(counts[take]) * (150 - k + 1)
            take += 1
        qc = np.concatenate([decode_range(pool.words, int(starts[i]) * 150, int(starts[i] + counts[i]) * 150)
                             for i in range(take)])
        qo = np.arange(0, qc.size + 1, 150, dtype=np.uint64)
        print(f"[bench] zi k={k}: CPU baseline (one-thread index build, then lookups)", file=sys.stderr, flush=True)
        nl, na, b_s, q_s = O.zi_lookup_timed(codes, np.array([0, G], dtype=np.uint64), qc, qo, k)
        out["cpu_baseline"] = {"value": nl / q_s, "unit": "lookups/s", "cores": 1, "kind": "port",
                               "sample": f"first {take} requests of the stream ({nl} lookups), hash-set index "
                                         f"of the same 100 Mbp genome (build {b_s:.1f} s, not counted), "
                                         f"1 thread, {q_s:.1f} s"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4_fof_32_datasets_20mbp", choices=sorted(MC),
                    help="the MC workload of the headline value")
    ap.add_argument("--only", choices=("headline", "mc", "zi"), default=None,
                    help="headline: only the headline workload; mc: every MC workload; zi: only ZI")
    ap.add_argument("--zi-requests", type=int, default=10_000)
    ap.add_argument("--zi-pool-reads", type=int, default=1 << 24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(sharding over several GPUs is design-only in this build)", file=sys.stderr)
        return 2

    import mczi_amd as M
    device = M.device_info()
    copy_bytes = 2 << 30
    copy_ms = M.copy_probe(copy_bytes, 10)
    cpu = not args.no_cpu_baseline

    line = {"metric": "MC bases/s + ZI k-mer lookups/s, p99; batched vs per-file GPU, bit-exact",
            "value": None, "unit": "bases/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u64", "data": "synthetic (seeded SplitMix64 genomes and reads, resident in HBM)",
            "config": None, "roofline": None, "cpu_baseline": None,
            "host": host_info(), "device": device,
            "copy_probe": {"bytes": copy_bytes, "ms": copy_ms,
                           "read_plus_write_GBs": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9}}
    # The ZI stream runs first, in the fresh process: its served path is bound by
    # device-to-host copies into the pinned memory the server takes when it is
    # created, and measured 15-20% slower when it followed the MC workloads.
    if args.only is None or args.only == "zi":
        line["zi"] = {}
        for k in (31, 63):
            line["zi"][f"k{k}"] = zi_stream(M, k, args.zi_requests, args.zi_pool_reads, cpu=cpu)
            print(f"[bench] zi k{k}: {json.dumps({x: line['zi'][f'k{k}'][x] for x in ('served_lookups_per_s', 'served_vs_plain', 'p99_ms')})}",
                  file=sys.stderr, flush=True)
        if line["value"] is None:
            z = line["zi"]["k31"]
            line["value"], line["unit"] = z["served_lookups_per_s"], "lookups/s"
            line["config"] = {"workload": "zi_served_100mbp_index_k31", "requests": z["requests"],
                              "reads": z["reads"]}
            line["roofline"] = z["roofline"]
            line["cpu_baseline"] = z["cpu_baseline"]
    if args.only != "zi":
        head = mc_workload(M, args.workload, args.steps, args.warmup, cpu, not args.no_plain)
        line["value"], line["unit"] = head["value"], "bases/s"
        line["ms_per_step"] = head["ms_per_step"]
        line["config"] = {"workload": head["workload"], "k": head["k"], "m": head["m"], "x": head["x"],
                          "partition_count": head["partition_count"], "bases": head["bases"],
                          "fragments": head["fragments"], "path": head["description"],
                          "launch": "batched: scan, partitioning and count in one hipGraph"}
        line["roofline"] = head["roofline"]
        line["cpu_baseline"] = head["cpu_baseline"]
        line["headline"] = head
        print(f"[bench] headline: {json.dumps({k2: head[k2] for k2 in ('value', 'ms_per_step')})}",
              file=sys.stderr, flush=True)
        if args.only is None or args.only == "mc":
            line["workloads"] = {}
            for name in MC:
                if name == args.workload:
                    continue
                line["workloads"][name] = mc_workload(M, name, args.steps, args.warmup, cpu, not args.no_plain)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
