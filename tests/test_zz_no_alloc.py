"""misfit_flags_noalloc: insertion flags staged in the library's static
device array instead of buffers allocated per call.  Two calls in a row, the
second with fewer pairs, must each equal misfit_flags on the same pairs and
the oracle's adjacency; a call past the array's capacity is refused.  Run on
an R-MAT graph (sorted rows) and on a graph whose rows are not ascending
(binary searches through the sorted row copy), with half of the pairs
existing edges so both flag values occur.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs(engine_mod, rng, ni, nl, count):
    n = len(ni) - 1
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, count, int(rng.integers(1 << 30)))
    # plus uniform pairs, a few repeated and a few reversed
    a = rng.integers(0, n, count // 4).astype(np.int32)
    b = rng.integers(0, n, count // 4).astype(np.int32)
    keep = a != b
    u = np.concatenate([u, a[keep], v[:50]]).astype(np.int32)
    v = np.concatenate([v, b[keep], u[:50]]).astype(np.int32)
    return u, v


@pytest.mark.parametrize("shuffled_rows", [False, True])
def test_flags_noalloc_equals_flags_twice(oracle, engine_mod, shuffled_rows):
    rng = np.random.default_rng(21 + shuffled_rows)
    ni, nl = engine_mod.gen_rmat(14, 16, 5)
    nl = nl.copy()
    if shuffled_rows:
        for x in range(0, len(ni) - 1, 7):
            rng.shuffle(nl[ni[x]:ni[x + 1]])
    g = oracle.Graph.from_csr(ni, nl)
    with engine_mod.Engine(0) as eng:
        eng.load_csr(ni, nl)
        for count in (40000, 3000):
            u, v = _pairs(engine_mod, rng, ni, nl, count)
            print("flags_noalloc: shuffled=%d, %d pairs" % (shuffled_rows, len(u)), flush=True)
            got = eng.flags_noalloc(u, v)
            want = eng.flags(u, v)
            assert got.dtype == np.uint8 and len(got) == len(u)
            assert np.array_equal(got, want)
            assert np.array_equal(got, g.flags(u, v))
            assert 0 < int(got.sum()) < len(u)


def test_flags_noalloc_limits(engine_mod):
    ni, nl = engine_mod.gen_rmat(10, 8, 2)
    with engine_mod.Engine(0) as eng:
        eng.load_csr(ni, nl)
        assert len(eng.flags_noalloc(np.zeros(0, np.int32), np.zeros(0, np.int32))) == 0
        n = len(ni) - 1
        big = engine_mod.NOALLOC_MAX_PAIRS + 1
        u = np.zeros(big, np.int32)
        v = np.ones(big, np.int32)
        with pytest.raises(engine_mod.MisfitError) as ei:
            eng.flags_noalloc(u, v)
        assert "MISFIT_NOALLOC_MAX_PAIRS" in str(ei.value)
        with pytest.raises(engine_mod.MisfitError):
            eng.flags_noalloc(np.array([0], np.int32), np.array([n], np.int32))
        # the largest count the array holds still works after the refusals
        full = engine_mod.NOALLOC_MAX_PAIRS
        rng = np.random.default_rng(3)
        u = rng.integers(0, n, full).astype(np.int32)
        v = ((u + 1 + rng.integers(0, n - 1, full)) % n).astype(np.int32)
        assert np.array_equal(eng.flags_noalloc(u, v), eng.flags(u, v))
