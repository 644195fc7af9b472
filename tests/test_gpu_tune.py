"""The autotuner and its cache (MI355X only): tuning never changes a result,
every LDS table size the tuner can pick counts exactly the same, and a tuned
configuration is recorded in the cache and used by a later batch of the same
shape."""
import numpy as np
import pytest

import mczi_amd as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reads():
    return M.SynthReads(M.SynthSpec(genome_len=3_000_000, n_reads=40_000, read_len=150, sub_rate=0.01,
                                    genome_seed=9, read_seed=19, n_datasets=5))


def _key_lines(path):
    return [l.split("\t") for l in path.read_text().splitlines() if l.strip()] if path.exists() else []


@pytest.mark.parametrize("mode", [M.MODE_SINGLE_PASS, M.MODE_DATASETS])
def test_table_sizes_and_tuning_give_the_same_counts(reads, tmp_path, monkeypatch, mode):
    cache = tmp_path / "tune.tsv"
    monkeypatch.setenv("MCZI_TUNE_CACHE", str(cache))
    cfg = M.CounterConfig(k=31, minimizer=21, threshold=1)
    ds = reads.dataset if mode == M.MODE_DATASETS else None
    b = M.Batch.from_packed(reads.words, reads.frag_off, ds)
    try:
        st = b.count(cfg, mode=mode)
        want = b.download()
        assert st.lds_slots == 2048 and want.size > 1000
        tuned = b.autotune(cfg, mode=mode)
        st = b.count(cfg, mode=mode)
        assert np.array_equal(b.download(), want)
        assert tuned > 0 and st.lds_slots in (1024, 2048, 4096) and st.scan_blocks_per_cu >= 1
    finally:
        b.close()
    lines = _key_lines(cache)
    assert len(lines) == 1 and int(lines[0][1]) in (10, 11, 12)
    key = lines[0][0]
    # every table size, through cache entries of this configuration
    for ls, fill, bpc in ((10, 512, 2), (12, 16384, 3), (11, 1024, 0)):
        with open(cache, "a") as f:
            f.write(f"{key}\t{ls}\t{fill}\t{bpc}\t1.0\n")
        b = M.Batch.from_packed(reads.words, reads.frag_off, ds)
        try:
            st = b.count(cfg, mode=mode)
            assert st.lds_slots == 1 << ls
            if bpc:
                assert st.scan_blocks_per_cu == bpc
            assert np.array_equal(b.download(), want), ls
        finally:
            b.close()
