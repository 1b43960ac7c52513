"""Shared checks for the parity tests."""
from __future__ import annotations

import numpy as np


def expected_starts(keys: np.ndarray, M: int) -> np.ndarray:
    """Bucket-start table of a counting sort: exclusive scan of the per-bucket
    counts with the sentinel N in entry M (SimulationEngine.h:229-282)."""
    counts = np.bincount(keys.astype(np.int64), minlength=M)
    assert counts.shape[0] == M, "key outside [0, M)"
    out = np.zeros(M + 1, dtype=np.uint64)
    out[1:] = np.cumsum(counts)
    return out.astype(np.uint32)


def check_buckets(starts: np.ndarray, ids: np.ndarray, keys_by_id: np.ndarray, id_offset: int = 0):
    """Every particle listed in [starts[b], starts[b+1]) has key b, and ids is a
    permutation (the validator of SimulationEngine.h:229-282)."""
    n = ids.shape[0]
    local = ids.astype(np.int64) - id_offset
    assert np.array_equal(np.sort(local), np.arange(n)), "bucket ids are not a permutation"
    bucket_of_slot = np.searchsorted(starts.astype(np.int64), np.arange(n), side="right") - 1
    assert np.array_equal(keys_by_id[local], bucket_of_slot), "a particle sits in the wrong bucket"


def bucket_multisets_equal(starts_a, ids_a, starts_b, ids_b) -> bool:
    """Same table and, per bucket, the same set of particle ids (order inside a bucket free)."""
    if not np.array_equal(starts_a, starts_b):
        return False
    n = ids_a.shape[0]
    bucket = np.searchsorted(starts_a.astype(np.int64), np.arange(n), side="right") - 1
    ka = np.lexsort((ids_a, bucket))
    kb = np.lexsort((ids_b, bucket))
    return bool(np.array_equal(ids_a[ka], ids_b[kb]))


def max_rel(a, b, floor: float = 0.0) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def max_norm_rel(a, b) -> float:
    """max |a - b| / max |b| over all entries (infinity-norm relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def point_triangle_distances(p, tris):
    """FP64 distance from each point to each triangle: to the plane where the
    projection falls inside the triangle, else to the nearest edge (a formulation
    independent of the closest-point region tests of SPHFunctions.cpp:156-201)."""
    t = tris.astype(np.float64).reshape(-1, 3, 3)
    a, b, c = t[:, 0], t[:, 1], t[:, 2]
    n = np.cross(b - a, c - a)
    P = p.astype(np.float64)[:, None, :]
    d_plane = np.abs(((P - a) * n).sum(-1)) / np.linalg.norm(n, axis=-1)
    q = P - (((P - a) * n).sum(-1) / (n * n).sum(-1))[..., None] * n
    inside = np.ones(d_plane.shape, bool)
    for u, v in ((a, b), (b, c), (c, a)):
        inside &= (np.cross(v - u, q - u) * n).sum(-1) >= 0
    def seg(u, v):
        w = v - u
        s = np.clip(((P - u) * w).sum(-1) / (w * w).sum(-1), 0.0, 1.0)
        return np.linalg.norm(P - (u + s[..., None] * w), axis=-1)
    d_edge = np.minimum(np.minimum(seg(a, b), seg(b, c)), seg(c, a))
    return np.where(inside, np.minimum(d_plane, d_edge), d_edge)
