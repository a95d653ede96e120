"""The autotuner's choice is deterministic for a given set of timings (the
first minimum in candidate order), and the recorded table is applied per
depth (nearest recorded depth, the lower one on a tie).  No reference
counterpart: the reference hard-codes its launch shapes."""
import random

from its_amd import _capi, autotune


def test_choose_first_minimum_on_ties():
    cands = [(2.0, "a"), (1.0, "b"), (3.0, "c"), (1.0, "d"), (1.0, "e")]
    assert autotune.choose(cands) == (1.0, "b")
    # the same timings in the same candidate order always give the same pick
    for _ in range(5):
        assert autotune.choose(list(cands)) == (1.0, "b")
    # a strictly smaller time wins wherever it sits
    rng = random.Random(0)
    for _ in range(50):
        ts = [rng.choice([1.0, 2.0, 3.0]) for _ in range(8)]
        c = [(t, k) for k, t in enumerate(ts)]
        assert autotune.choose(c) == (min(ts), ts.index(min(ts)))


def test_table_applies_per_depth():
    table = autotune.load()
    keys = (_capi.TUNE_DENSE_CELLS, _capi.TUNE_EVAL_BLOCKS_PER_CU, _capi.TUNE_EVAL_ILP, _capi.TUNE_EVAL_QUEUE,
            _capi.TUNE_COARSE_VERTS)
    saved = {k: _capi.tune_get(k) for k in keys}
    try:
        assert set(table) >= {5, 6, 7, 8}
        for depth in (4, 5, 6, 7, 8, 9, 12):
            d = autotune.apply(depth, table)
            assert d == min(table, key=lambda k: (abs(k - depth), k))
            t = table[d]
            assert _capi.tune_get(_capi.TUNE_DENSE_CELLS) == t["dense_cells"]
            assert _capi.tune_get(_capi.TUNE_EVAL_BLOCKS_PER_CU) == t["eval_blocks_per_cu"]
            assert _capi.tune_get(_capi.TUNE_EVAL_ILP) == t["eval_ilp"]
            assert _capi.tune_get(_capi.TUNE_EVAL_QUEUE) == t["eval_queue"]
            assert _capi.tune_get(_capi.TUNE_COARSE_VERTS) == t["coarse_verts"]
        # a depth half way between two recorded depths takes the lower one
        t2 = {5: table[5], 7: table[7]}
        assert autotune.apply(6, t2) == 5
    finally:
        for k, v in saved.items():
            _capi.tune_set(k, v)
