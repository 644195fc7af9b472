#!/usr/bin/env python3
"""Benchmark of the MI355X MCZI counting path.

Default workload (MC abundance mode, short reads at scale): random iid genome
of 100 Mbp (seed 2) plus 2% of it duplicated as 5 kbp repeats at 10 copies
each, 150 bp reads at 30x (20M reads, 3 Gbp), 1% substitutions, random
strand; k=31, m=21, x=5, partition count 1024, FASTA-path (two-pass) counting.
A step = one pass of the counting path over the whole read set, which is
resident in HBM before timing starts: minimizer scan + super-k-mer records
-> partitioning (radix sort by partition) -> LDS counting with HBM spill ->
count > x filter -> sort by encoded k-mer. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (synth spec kwargs, k, m, x, partitions, mode)
    "c2_short_reads_100mbp": (dict(genome_len=100_000_000, genome_seed=2, read_seed=102, n_reads=20_000_000,
                                   read_len=150, sub_rate=0.01, repeat_fraction=0.02, repeat_len=5000,
                                   repeat_copies=10), 31, 21, 5, 1024, 0),
    "c1_small_1mbp": (dict(genome_len=1_000_000, genome_seed=1, read_seed=101, n_reads=200_000, read_len=150,
                           sub_rate=0.005), 31, 21, 1, 64, 1),
}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the factor a plain
    copy kernel of known size reads on the same counter), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d[workload]["kernels"][kernel]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def decode_prefix(reads, n_frag):
    """Packed-seq codes (one per base) of the first n_frag reads, and offsets."""
    import numpy as np
    off = reads.frag_off[: n_frag + 1].copy()
    nb = int(off[-1])
    w = reads.words[: (nb + 15) // 16 + 1].view(np.uint8)
    b = np.unpackbits(w[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
    c = (b[:, :, 0] | (b[:, :, 1] << 1)).reshape(-1)[:nb].astype(np.uint8)
    return c, off


def cpu_baseline(reads, k, m, x, parts, mode, seconds_target=10.0, max_reads=4_000_000):
    """The oracle's restatement of the reference CPU path on a bounded prefix
    of the same read set: the two-pass count on up to 16 host threads (the
    reference runs it under Rayon on all cores), or the KFF path on one
    thread. Returns bases/s, the threads used and the sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    threads = max(1, min(16, os.cpu_count() or 1)) if mode == 0 else 1

    def run(codes, off):
        if mode == 0:
            O.count_two_pass_mt(codes, off, k, m, x, parts, threads)
        else:
            O.kff_u8_kmc(codes, off, k, m, x, parts)

    n_frag = int(len(reads.frag_off) - 1)
    take = min(n_frag, 20_000 * threads)
    codes, off = decode_prefix(reads, take)
    t0 = time.perf_counter()
    run(codes, off)
    dt = time.perf_counter() - t0
    # grow the sample to ~seconds_target of wall time
    take2 = min(n_frag, max_reads, max(take, int(take * seconds_target / max(dt, 1e-3))))
    if take2 > take:
        del codes, off
        codes, off = decode_prefix(reads, take2)
        t0 = time.perf_counter()
        run(codes, off)
        dt = time.perf_counter() - t0
        take = take2
    rate = int(off[-1]) / max(dt, 1e-9)
    return {"value": rate, "unit": "bases/s", "cores": threads, "kind": "port",
            "sample": f"first {take} reads ({int(off[-1])} bases) of the same read set, "
                      f"{'two-pass FASTA path' if mode == 0 else 'KMC-style KFF path'}, {threads} threads, "
                      f"{dt:.1f} s"}


def many_inputs(M, n_inputs: int = 4000, reads_per_input: int = 50, steps: int = 3):
    """Many small inputs in one packed batch (reads of a 1 Mbp random genome, 150 bp
    reads at 0.5% substitutions; k31 m21 x1 P64): the batched launch scans
    them all at once, the plain launch one input per launch. Both count the
    same k-mers; returns bases/s of each and their ratio."""
    import numpy as np
    r = M.SynthReads(M.SynthSpec(genome_len=1_000_000, genome_seed=1, read_seed=101,
                                 n_reads=n_inputs * reads_per_input, read_len=150, sub_rate=0.005))
    marks = np.arange(0, n_inputs * reads_per_input + 1, reads_per_input, dtype=np.uint64)
    b = M.Batch.from_packed(r.words, r.frag_off, input_frag=marks)
    cfg = M.CounterConfig(31, 21, 1, 64)
    res = {"inputs": n_inputs, "reads_per_input": reads_per_input, "bases": r.n_bases}
    outs = []
    for name, launch in (("batched", M.LAUNCH_BATCHED), ("plain", M.LAUNCH_PLAIN)):
        st = b.count(cfg, M.MODE_TWO_PASS, launch)  # warm-up (plan, graph)
        t0 = time.perf_counter()
        for _ in range(steps):
            st = b.count(cfg, M.MODE_TWO_PASS, launch)
        res[name + "_bases_per_s"] = r.n_bases * steps / (time.perf_counter() - t0)
        outs.append(st.n_out)
    assert outs[0] == outs[1], "batched and plain counts differ"
    res["batched_vs_plain"] = res["batched_bases_per_s"] / res["plain_bases_per_s"]
    return res


def zi_serving(M, n_requests: int, k: int, seed: int = 5):
    """ZI served subtraction: index = canonical k-mers of a 100 Mbp genome,
    requests of 150 bp reads (half from the indexed genome, half from an
    unrelated genome, 0.5% substitutions), read counts log-uniform in
    [10, 1e6] rounded up to a multiple of 8. Measures the plain path (one
    request at a time) and the served path (continuous batching) in lookups/s
    and request latency."""
    import ctypes as C
    import threading

    import numpy as np

    G = 100_000_000
    ga = M.synth_genome(G, seed)
    gb = M.synth_genome(G, seed + 1000)
    # index FASTA text of genome A
    codes = np.unpackbits(ga.view(np.uint8)[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
    codes = (codes[:, :, 0] | (codes[:, :, 1] << 1)).reshape(-1)[:G]
    text = b">index\n" + np.frombuffer(b"ACTG", dtype=np.uint8)[codes].tobytes() + b"\n"
    del codes
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(10), np.log(1e6), size=n_requests)).astype(np.int64)
    counts = ((counts + 7) // 8) * 8
    reads = M.synth_reads_from([(ga, G), (gb, G)], int(counts.sum()), 150, 0.005, seed + 1)
    starts = np.concatenate([[0], np.cumsum(counts)])
    reqs = []
    for i in range(n_requests):
        b0, b1 = int(starts[i]) * 150, int(starts[i + 1]) * 150
        w = np.ascontiguousarray(reads.words[b0 // 16:(b1 + 15) // 16 + 1])
        off = np.arange(0, (b1 - b0) + 1, 150, dtype=np.uint64)
        reqs.append((w, off))

    h = C.c_void_p()
    t0 = time.perf_counter()
    M._check(M._lib.mczi_zi_server_create_from_text(text, len(text), k, 0, 96 << 30, 3, C.byref(h)))
    build_s = time.perf_counter() - t0

    def submit(i):
        w, off = reqs[i]
        rid = C.c_uint64()
        M._check(M._lib.mczi_zi_server_submit_packed(h, w.ctypes.data, off.ctypes.data, off.size - 1, 0,
                                                      C.byref(rid)))
        return rid.value

    def poll(rid):
        p = C.POINTER(M._ZiResult)()
        M._check(M._lib.mczi_zi_server_poll(h, rid, 1, C.byref(p)))
        r = p.contents
        out = (r.n_windows, r.n_absent, r.latency_ms, r.queue_ms)
        M._lib.mczi_zi_result_free(p)
        return out

    try:
        poll(submit(0))  # warm-up
        # plain: one request at a time
        t0 = time.perf_counter()
        plain = [poll(submit(i)) for i in range(n_requests)]
        plain_s = time.perf_counter() - t0
        lookups = sum(x[0] for x in plain)
        # served, closed loop: every request submitted at once
        t0 = time.perf_counter()
        ids = [submit(i) for i in range(n_requests)]
        burst = [poll(r) for r in ids]
        burst_s = time.perf_counter() - t0
        assert [x[1] for x in burst] == [x[1] for x in plain], "served results differ from plain"
        # served, open loop: Poisson arrivals at 1.5x the plain request rate
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
        th.join()
        open_s = time.perf_counter() - t0
        lat = np.array([x[2] for x in res])
    finally:
        M._lib.mczi_zi_server_destroy(h)
    return {
        "k": k, "requests": n_requests, "reads": int(counts.sum()), "lookups": int(lookups),
        "index_build_s": build_s,
        "plain_lookups_per_s": lookups / plain_s,
        "served_lookups_per_s": lookups / burst_s,
        "served_vs_plain": plain_s / burst_s,
        "poisson_rate_per_s": rate,
        "poisson_lookups_per_s": lookups / open_s,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "plain_p50_ms": float(np.percentile([x[2] for x in plain], 50)),
        "plain_p99_ms": float(np.percentile([x[2] for x in plain], 99)),
    }


def end_to_end(M, workload: str, baseline: bool):
    """The whole count -> simplitig FASTA path from a FASTA file on disk: host
    ingest (parse + 2-bit pack), upload, counting, download and the simplitig
    writer, each leg timed, then the library's one-call path. With baseline,
    the oracle's one-thread parser and simplitig writer (restatements of the
    reference's) run on the same file and k-mers, and the two simplitig files
    are compared byte for byte."""
    import shutil
    import tempfile
    spec_kw, k, m, x, parts, mode = WORKLOADS[workload]
    tmp = tempfile.mkdtemp(prefix="mczi_e2e_")
    fa, out, out_o = (os.path.join(tmp, n) for n in ("reads.fa", "mczi.fa", "oracle.fa"))
    try:
        reads = M.SynthReads(M.SynthSpec(**spec_kw))
        reads.write_fasta(fa)
        nb = int(reads.n_bases)
        del reads
        M.device_info()
        cfg = M.CounterConfig(k, m, x, parts)
        t0 = time.perf_counter()
        batch = M.Batch.from_files([fa])
        t1 = time.perf_counter()
        batch.count(cfg)
        t2 = time.perf_counter()
        arr = batch.download()
        t3 = time.perf_counter()
        M.write_fasta_path(out, arr, k)
        t4 = time.perf_counter()
        del batch
        t5 = time.perf_counter()
        M.count_inputs_to_fasta_path([fa], cfg, out)
        t6 = time.perf_counter()
        res = {"workload": workload, "fasta_bytes": os.path.getsize(fa), "bases": nb, "kmers_out": int(arr.size),
               "simplitig_bytes": os.path.getsize(out),
               "ingest_upload_s": t1 - t0, "count_s": t2 - t1, "download_s": t3 - t2,
               "simplitig_write_s": t4 - t3, "one_call_total_s": t6 - t5,
               "ingest_bases_per_s": nb / (t1 - t0), "simplitig_kmers_per_s": arr.size / (t4 - t3)}
        if baseline:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as O
            ta = time.perf_counter()
            codes, off = O.parse_fastx(fa, k)
            tb = time.perf_counter()
            del codes, off
            O.write_simplitigs(out_o, arr, k)
            tc = time.perf_counter()
            with open(out, "rb") as f1, open(out_o, "rb") as f2:
                same = True
                while same:
                    a, b = f1.read(1 << 24), f2.read(1 << 24)
                    same = a == b
                    if not a:
                        break
            res["cpu_baseline"] = {"parse_s": tb - ta, "simplitig_s": tc - tb, "cores": 1, "kind": "port",
                                   "sample": "the same FASTA file and k-mer set, oracle parser and writer"}
            res["simplitig_identical_to_oracle"] = bool(same)
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2_short_reads_100mbp", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--zi-requests", type=int, default=48)
    ap.add_argument("--no-zi", action="store_true")
    ap.add_argument("--zi-only", action="store_true", help="run only the ZI serving measurement")
    ap.add_argument("--zi-k63-only", action="store_true", help="run only the ZI serving measurement at k=63")
    ap.add_argument("--e2e", action="store_true",
                    help="run only the end-to-end count -> simplitig FASTA measurement (file on disk)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    import mczi_amd as M

    if args.zi_only:
        print(json.dumps({"zi": zi_serving(M, args.zi_requests, 31)}), flush=True)
        return 0
    if args.zi_k63_only:
        print(json.dumps({"zi_k63": zi_serving(M, args.zi_requests, 63)}), flush=True)
        return 0
    if args.e2e:
        print(json.dumps({"e2e": end_to_end(M, args.workload, not args.no_cpu_baseline)}), flush=True)
        return 0

    spec_kw, k, m, x, parts, mode = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    reads = M.SynthReads(M.SynthSpec(**spec_kw))
    print(f"[bench] generated {reads.n_bases} bases in {len(reads.frag_off) - 1} reads "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    device = M.device_info()  # HIP runtime start-up happens here, outside the upload timing
    t_up = time.perf_counter()
    batch = M.Batch.from_packed(reads.words, reads.frag_off)
    upload_s = time.perf_counter() - t_up  # host -> HBM copy of the packed reads (pageable source)
    cfg = M.CounterConfig(k, m, x, parts)
    order = M.ORDER_ENCODED if mode == 0 else M.ORDER_KFF_U8

    for i in range(args.warmup):
        st = batch.count(cfg, mode, M.LAUNCH_BATCHED, order)
        print(f"[bench] warmup {i}: {st.ms_total + st.ms_sort:.1f} ms device, out {st.n_out}",
              file=sys.stderr, flush=True)

    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = batch.count(cfg, mode, M.LAUNCH_BATCHED, order)
        dev_ms.append(st.ms_total + st.ms_sort)
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = reads.n_bases * args.steps / wall
    last = st.as_dict()

    # per-kernel device times: one eager run with HIP events between kernels
    os.environ["MCZI_NO_GRAPH"] = "1"
    prof = batch.count(cfg, mode, M.LAUNCH_BATCHED, order).as_dict()
    plain_value = None
    if not args.no_plain:
        batch.count(cfg, mode, M.LAUNCH_PLAIN, order)
        tp = time.perf_counter()
        for _ in range(2):
            batch.count(cfg, mode, M.LAUNCH_PLAIN, order)
        plain_value = reads.n_bases * 2 / (time.perf_counter() - tp)
    del os.environ["MCZI_NO_GRAPH"]
    many = None if args.no_plain else many_inputs(M)

    # plain device copy of 2 GiB: achievable HBM bandwidth on this box, and
    # the reference dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE
    copy_bytes = 2 << 30
    copy_ms = M.copy_probe(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9

    nb = reads.n_bases
    R = prof["n_records"]
    P = prof["internal_partitions"]
    out_bytes = prof["n_out"] * 16
    # algorithmic (compulsory) bytes per kernel group, each operand once:
    #  k_scan: packed reads in (2 bits/base); records (16 B) + partition ids (4 B) out
    #  partition sort: (partition id, record) pairs in and out once
    #  k_count: partitioned records + partition bounds in; (k-mer, count, partition) out
    kernels = {
        "k_scan": (nb / 4 + 20 * R, prof["ms_scan1"]),
        "partition_sort": (40 * R, prof["ms_scan2"]),
        "k_count": (16 * R + 16 * P + out_bytes, prof["ms_count"]),
    }
    dom = max(kernels, key=lambda n: kernels[n][1])
    alg, ms = kernels[dom]
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = measured_traffic(args.workload, dom)

    line = {
        "metric": "MC bases/s + ZI k-mer lookups/s, p99; batched vs per-file GPU, bit-exact",
        "value": value,
        "unit": "bases/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded SplitMix64 genome and reads, resident in HBM)",
        "config": {"workload": args.workload, "k": k, "m": m, "x": x, "partition_count": parts,
                   "bases": nb, "reads": int(len(reads.frag_off) - 1),
                   "path": "two-pass FASTA counting" if mode == 0 else "KMC-style KFF counting",
                   "launch": "batched (persistent kernels, hipGraph)"},
        "device_ms_per_step": sum(dev_ms) / len(dev_ms),
        "upload_ms": upload_s * 1e3,
        "pcie_inclusive_value": nb / (ms_per_step * 1e-3 + upload_s),
        "plain_value": plain_value,
        "batched_vs_plain": (value / plain_value) if plain_value else None,
        "many_inputs": many,
        "phase_ms": {k2: prof[k2] for k2 in ("ms_scan1", "ms_scan2", "ms_count", "ms_sort", "ms_total")},
        "stats": {k2: last[k2] for k2 in ("n_windows", "n_pieces", "n_records", "n_distinct", "n_overflow",
                                           "n_out", "internal_partitions")},
        "device": device,
        "copy_probe": {"bytes": copy_bytes, "ms": copy_ms, "read_plus_write_GBs": copy_gbs},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg, "kernel_ms": ms},
        "cpu_baseline": None,
    }
    if not args.no_zi:
        del batch
        line["zi"] = zi_serving(M, args.zi_requests, 31)
        print(f"[bench] zi: {line['zi']}", file=sys.stderr, flush=True)
        line["zi_k63"] = zi_serving(M, args.zi_requests, 63)
        print(f"[bench] zi k63: {line['zi_k63']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(reads, k, m, x, parts, mode)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
