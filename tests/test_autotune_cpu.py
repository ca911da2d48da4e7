"""Host logic of the launch-shape autotuner (no GPU): the candidate set,
the selection rule, the outcome check and the cache."""
import pytest

from kgmt_amd import autotune as T
from kgmt_amd import workloads as W


def test_candidates_cover_the_grid_default_first():
    c = T.candidates(256)
    assert len(c) == len(T.SHAPES) * len(T.PROP_PER_CU)
    assert c[0] == {"propagate_block": 256, "lanes_per_child": 1, "propagate_grid": 1024, "chunk_grid": 1024}
    assert len({tuple(sorted(d.items())) for d in c}) == len(c)
    assert {d["propagate_block"] for d in c} == {128, 256, 512}
    assert {d["lanes_per_child"] for d in c} == {1, 2, 4}
    # lanes > 1 only with 256 threads (the shapes the library builds)
    assert all(d["propagate_block"] == 256 for d in c if d["lanes_per_child"] > 1)
    second = T.chunk_candidates(c[3], 256)
    assert second[0] == c[3] and len(second) == len(T.CHUNK_PER_CU)
    assert {d["chunk_grid"] for d in second} == {256, 512, 1024}


def test_search_picks_lowest_median_and_checks_outcomes():
    cands = [{"propagate_block": b, "lanes_per_child": 1, "propagate_grid": 1, "chunk_grid": 1} for b in (128, 256, 512)]
    cost = {128: [3.0, 1.0, 9.0, 2.0], 256: [1.0, 1.5, 1.5, 1.5], 512: [5.0, 0.5, 4.0, 4.0]}
    calls = {}

    def run(shape):
        b = shape["propagate_block"]
        i = calls.get(b, 0)
        calls[b] = i + 1
        return cost[b][i], "same"

    out = T.search(cands, run, reps=3)
    # medians over the three timed runs (the first call is the warm-up)
    assert out["launch"]["propagate_block"] == 256 and out["seconds"] == 1.5
    assert [t["seconds"] for t in out["trials"]] == [2.0, 1.5, 4.0]
    assert out["trials"][0]["launch"]["propagate_block"] == 128

    def bad(shape):
        return 1.0, shape["propagate_block"]

    with pytest.raises(RuntimeError):
        T.search(cands, bad, reps=1)
    with pytest.raises(ValueError):
        T.search([], bad)


def test_signature_and_cache(tmp_path):
    qs = [W.make_query("di6d", s) for s in range(3)]
    cfgs = [q.cfg for q in qs]
    a = T.signature(cfgs, [20, 20, 20])
    assert a == T.signature(cfgs, [20, 20, 20])
    assert a != T.signature(cfgs, [20, 20, 21])
    assert a != T.signature([c.replace(capacity=1 << 12) for c in cfgs], [20, 20, 20])
    path = str(tmp_path / "sub" / "launch.json")
    assert T.load_cache(path) == {}
    T.store_cache("k1", {"launch": {"propagate_block": 128}}, path)
    T.store_cache("k2", {"launch": {"propagate_block": 512}}, path)
    d = T.load_cache(path)
    assert d["k1"]["launch"]["propagate_block"] == 128 and d["k2"]["launch"]["propagate_block"] == 512
