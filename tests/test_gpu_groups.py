"""Counts beyond one device pass (MI355X only): the partitions are split into
groups by key hash and each group is its own pass over the resident batch
(the reference instead spills every partition's super-k-mers to disk and
counts them one by one, src/lib.rs:2803-2845). Forced group counts
(MCZI_GROUP_PASSES) and a small device budget (MCZI_DEVICE_BUDGET, which
picks the group count itself) give the same k-mers, counts, dataset
presence and KFF bytes as one pass, and as the oracle."""
import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import same_counted

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reads():
    spec = M.SynthSpec(genome_len=2_000_000, n_reads=60_000, read_len=150, sub_rate=0.01, n_datasets=40,
                       genome_seed=7, read_seed=8)
    return M.SynthReads(spec)


def _count(reads, cfg, mode, order=M.ORDER_ENCODED, datasets=None, env=None, monkeypatch=None):
    for key in ("MCZI_GROUP_PASSES", "MCZI_DEVICE_BUDGET"):
        monkeypatch.delenv(key, raising=False)
    for key, v in (env or {}).items():
        monkeypatch.setenv(key, str(v))
    ds = datasets if datasets is not None else None
    b = M.Batch.from_packed(reads.words, reads.frag_off, ds)
    try:
        st = b.count(cfg, mode=mode, order=order)
        out = b.download()
    finally:
        b.close()
        for key in (env or {}):
            monkeypatch.delenv(key, raising=False)
    return out, st


@pytest.mark.parametrize("k,m,x,mode", [(31, 21, 1, M.MODE_SINGLE_PASS), (31, 21, 0, M.MODE_TWO_PASS),
                                        (21, 11, 2, M.MODE_SINGLE_PASS), (63, 31, 1, M.MODE_SINGLE_PASS)])
def test_group_passes_equal_one_pass(reads, monkeypatch, k, m, x, mode):
    cfg = M.CounterConfig(k=k, minimizer=m, threshold=x)
    one, st1 = _count(reads, cfg, mode, monkeypatch=monkeypatch)
    assert st1.partition_groups == 1 and one.size > 1000
    for g in (2, 7):
        got, st = _count(reads, cfg, mode, env={"MCZI_GROUP_PASSES": g}, monkeypatch=monkeypatch)
        print(f"k={k} x={x} mode={mode} groups={g}: {got.size} k-mers", flush=True)
        assert st.partition_groups == g
        assert np.array_equal(got, one)
        assert st.n_distinct == st1.n_distinct and st.n_records == st1.n_records
        assert st.partition_bytes == st1.partition_bytes


def test_small_budget_picks_groups(reads, monkeypatch):
    cfg = M.CounterConfig(k=31, minimizer=21, threshold=1)
    one, _ = _count(reads, cfg, M.MODE_SINGLE_PASS, monkeypatch=monkeypatch)
    got, st = _count(reads, cfg, M.MODE_SINGLE_PASS, env={"MCZI_DEVICE_BUDGET": 48 << 20}, monkeypatch=monkeypatch)
    print(f"budget 48 MiB: {st.partition_groups} groups", flush=True)
    assert st.partition_groups > 1
    assert np.array_equal(got, one)
    # and against the oracle's exact counts
    codes = reads.codes()
    want = O.count_bruteforce(codes, reads.frag_off, 31, 1)
    assert same_counted(got, want)


def test_group_passes_kff_order_and_datasets(reads, monkeypatch):
    cfg = M.CounterConfig(k=31, minimizer=21, threshold=1, partition_count=64)
    one, _ = _count(reads, cfg, M.MODE_SINGLE_PASS, order=M.ORDER_KFF_U8, monkeypatch=monkeypatch)
    got, _ = _count(reads, cfg, M.MODE_SINGLE_PASS, order=M.ORDER_KFF_U8, env={"MCZI_GROUP_PASSES": 5},
                    monkeypatch=monkeypatch)
    assert np.array_equal(got, one)  # same records in the same KFF order
    # dataset presence: 8 datasets (one table pass) and 40 (groups of 32)
    for nd in (8, 40):
        ds = (reads.dataset % nd).astype(np.uint32)
        c = M.CounterConfig(k=31, minimizer=21, threshold=2)
        one, _ = _count(reads, c, M.MODE_DATASETS, datasets=ds, monkeypatch=monkeypatch)
        got, st = _count(reads, c, M.MODE_DATASETS, datasets=ds, env={"MCZI_GROUP_PASSES": 3},
                         monkeypatch=monkeypatch)
        assert st.partition_groups == 3 and one.size > 100
        assert np.array_equal(got, one), nd


@pytest.mark.parametrize("mode", [M.MODE_SINGLE_PASS, M.MODE_DATASETS])
def test_per_file_plain_with_group_passes(reads, monkeypatch, mode):
    """The plain launch's per-file path with every file counted in forced
    group passes gives the one-pass batched result (counts; fof with one
    dataset per input, 40 datasets)."""
    n = len(reads.frag_off) - 1
    ds = reads.dataset if mode == M.MODE_DATASETS else None
    per = n // 40
    assert per * 40 == n and np.array_equal(reads.dataset[::per], np.arange(40))
    marks = np.arange(0, n + 1, per, dtype=np.uint64)
    cfg = M.CounterConfig(k=31, minimizer=21, threshold=3)
    one, _ = _count(reads, cfg, mode, datasets=ds, monkeypatch=monkeypatch)
    for key in ("MCZI_GROUP_PASSES", "MCZI_DEVICE_BUDGET"):
        monkeypatch.delenv(key, raising=False)
    monkeypatch.setenv("MCZI_GROUP_PASSES", "3")
    b = M.Batch.from_packed(reads.words, reads.frag_off, ds, input_frag=marks)
    try:
        st = b.count(cfg, mode=mode, launch=M.LAUNCH_PLAIN)
        got = b.download()
    finally:
        b.close()
        monkeypatch.delenv("MCZI_GROUP_PASSES", raising=False)
    print(f"per-file plain mode={mode}: {got.size} k-mers", flush=True)
    assert np.array_equal(got, one) and st.n_out == one.size
