"""Shared helpers for the parity tests."""
import numpy as np


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.uint32)
    b = np.ascontiguousarray(b, np.float32).view(np.uint32)
    return a == b


def assert_nearest_equal(leaf, toi, ref_leaf, ref_t, what=""):
    """Bit-exact (leaf, toi) agreement; misses are leaf -1 and toi +inf on both sides."""
    leaf = np.asarray(leaf).astype(np.int64)
    ok = (leaf == ref_leaf) & same_bits(toi, ref_t)
    if not ok.all():
        bad = np.where(~ok)[0]
        i = bad[:8]
        raise AssertionError(f"{what}: {len(bad)} / {len(ok)} queries differ; first {i.tolist()} "
                             f"got {leaf[i].tolist()} {np.asarray(toi)[i].tolist()} "
                             f"want {ref_leaf[i].tolist()} {ref_t[i].tolist()}")


def csr_equal(off, leaf, toi, ref_off, ref_leaf, ref_t=None, what=""):
    off = np.asarray(off).astype(np.uint64)
    assert np.array_equal(off, ref_off.astype(np.uint64)), f"{what}: per-query counts differ"
    assert np.array_equal(np.asarray(leaf).astype(np.int64), ref_leaf), f"{what}: ids or order differ"
    if ref_t is not None:
        assert same_bits(toi, ref_t).all(), f"{what}: toi bits differ"
