"""Host side of the synthetic CSR replay: the calibration tool's file
contract, the column synthesis and the numpy restatement of the replay PPE
that the GPU test compares against."""
import os

import numpy as np
import pytest

from lsmps_amd import replay as R
import replay_ref as RR


def _profile(n, seed=3):
    rng = np.random.default_rng(seed)
    row_len = rng.integers(5, 40, n).astype(np.int32)
    t = rng.choice(4, n, p=[0.5, 0.2, 0.2, 0.1]).astype(np.uint8)
    bad = rng.random(n) < 0.08
    return row_len, t, bad


def test_contract_round_trip_and_checks(tmp_path):
    n = 500
    row_len, t, bad = _profile(n)
    tl = [{"step": s, "injected": 10, "deleted": 9, "cnl_rebuild": int(s % 4 == 0), "ppe_iterations": 12}
          for s in range(6)]
    d = str(tmp_path / "w")
    R.write_contract(d, row_len, t, bad, timeline=tl)
    c = R.load_contract(d, column_seed=5)
    assert c["columns"].startswith("synthesized")
    assert c["csr_row_ptr"].dtype == np.int64 and c["csr_col_idx"].dtype == np.int32
    assert c["csr_col_idx"].shape[0] == row_len.sum() and c["timeline"] == tl
    assert c["csr_col_idx"].min() >= 0 and c["csr_col_idx"].max() < n
    # Materialized columns are used as written.
    col = c["csr_col_idx"][::-1].copy()
    R.write_contract(d, row_len, t, bad, col=col)
    assert np.array_equal(R.load_contract(d)["csr_col_idx"], col)
    # Wrong dtypes, shapes and values are refused.
    np.save(os.path.join(d, "csr_row_ptr.npy"), np.cumsum(np.r_[0, row_len]).astype(np.int32))
    with pytest.raises(ValueError, match="int64"):
        R.load_contract(d)
    R.write_contract(d, row_len, t, bad, col=col)
    np.save(os.path.join(d, "neighbor_row_length.npy"), (row_len + 1).astype(np.int32))
    with pytest.raises(ValueError, match="disagrees"):
        R.load_contract(d)
    R.write_contract(d, row_len, np.full(n, 7, np.uint8), bad, col=col)
    with pytest.raises(ValueError, match="0..4"):
        R.load_contract(d)
    R.write_contract(d, row_len, t, bad, col=np.full(col.shape, n, np.int32))
    with pytest.raises(ValueError, match="outside"):
        R.load_contract(d)


def test_synthesized_columns_follow_the_tool_rule():
    """Offsets around the row are rint(N(0, max(8, sqrt n))), wrapped mod n
    (python/workload_calibration.py:255-262)."""
    n = 40000
    row_len = np.full(n, 30, np.int32)
    col = R.synthesize_columns(row_len, 11)
    rows = np.repeat(np.arange(n), 30)
    off = (col.astype(np.int64) - rows + n // 2) % n - n // 2
    sigma = max(8.0, np.sqrt(n))
    assert abs(off.mean()) < 0.05 * sigma
    assert abs(off.std() / sigma - 1.0) < 0.02
    assert np.array_equal(off, np.rint(off))


def test_hash_restatement_pins():
    """splitmix64: the finaliser of 0x9E3779B97F4A7C15 is the first output of
    SplitMix64 seeded with 0 (Steele, Lea, Flood 2014); the weights are
    symmetric, inside [0.5, 1.5) / l0^2, and x* inside [-1, 1)."""
    assert int(RR.mix64(np.uint64(0x9E3779B97F4A7C15))) == 0xE220A8397B1DCDAF
    lo = np.arange(1000)
    hi = lo + 17
    w = RR.replay_weight(5, lo, hi, 0.01)
    assert np.all((w >= 0.5e4) & (w < 1.5e4)) and np.unique(w).size == w.size
    x = RR.xstar_random(3, np.arange(100000))
    assert x.min() >= -1 and x.max() < 1 and abs(x.mean()) < 0.01


def test_replay_ppe_restatement_structure():
    """Closed form of the graph Laplacian: every row but row 0 sums to zero
    unless its diagonal was reset; row 0 is the identity; b = A x*."""
    n = 400
    row_len, t, bad = _profile(n, 9)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(row_len, out=rp[1:])
    col = R.synthesize_columns(row_len, 2)
    A = RR.replay_ppe(rp, col, t, 7, 4, 0.01)
    m = A["rev"].size
    assert A["row"][1] == 1 and A["val"][0] == 1.0
    sums = np.add.reduceat(A["val"], A["row"][:-1])
    single = np.diff(A["row"]) == 1
    assert np.allclose(sums[1:][~single[1:]], 0.0, atol=1e-9 * 1e4)
    assert np.all(np.isin(RR.TYPE_MAP[t[A["rev"]]], (0, 3)))
    assert m == int(np.isin(t, (0, 1)).sum())
