"""Constraints::fill collision rows (Constraints.cpp:417-470) from proximity records.

The library's CSR must equal a direct restatement of the reference loop (rows in
record order; (3,1): -nor1 on verts2[0]; (2,2): -nor2 w2[j] on verts2[0..1];
(1,3): -nor2 w2[j] on verts2[0..2]; others skipped; EOL vertices skip the row),
exactly.  Host-only: runs without a GPU."""
import numpy as np
import scipy.sparse as sp

import eolc


def restated(recs, N, eol=None):
    rows = []
    for r in recs:
        c1, c2 = r.count1, r.count2
        if (c1, c2) == (3, 1):
            vs, nor, w = [r.verts2[0]], list(r.nor1), [1.0]
        elif (c1, c2) == (2, 2):
            vs, nor, w = [r.verts2[0], r.verts2[1]], list(r.nor2), [r.weights2[0], r.weights2[1]]
        elif (c1, c2) == (1, 3):
            vs, nor, w = [r.verts2[j] for j in range(3)], list(r.nor2), [r.weights2[j] for j in range(3)]
        else:
            continue
        if eol is not None and any(eol[v] for v in vs):
            continue
        row = {}
        for v, wj in zip(vs, w):
            for k in range(3):
                row[3 * v + k] = row.get(3 * v + k, 0.0) + (-nor[k] * wj if (c1, c2) != (3, 1) else -nor[k])
        rows.append(row)
    A = sp.lil_matrix((len(rows), 3 * N))
    for i, row in enumerate(rows):
        for c, v in row.items():
            A[i, c] = v
    return A.tocsr()


def make_records(seed, n, N):
    rng = np.random.default_rng(seed)
    recs = []
    kinds = [(3, 1), (2, 2), (1, 3), (1, 1), (3, 3)]
    for i in range(n):
        r = eolc.CollisionRec()
        r.count1, r.count2 = kinds[rng.integers(len(kinds))]
        for k in range(3):
            r.nor1[k], r.nor2[k] = rng.normal(), rng.normal()
            r.verts2[k] = int(rng.integers(N))
            r.weights2[k] = rng.random()
        recs.append(r)
    return recs


def test_collision_rows_match_restatement():
    N = 50
    recs = make_records(1, 200, N)
    A = eolc.collision_constraints(recs, N)
    R = restated(recs, N)
    assert A.shape == R.shape and A.shape[0] > 0
    assert (abs(A - R)).max() == 0.0
    # columns increasing within every row
    for i in range(A.shape[0]):
        cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
        assert np.all(np.diff(cols) > 0)


def test_collision_rows_skip_eol_and_merge_duplicates():
    N = 20
    recs = make_records(2, 100, N)
    eol = np.zeros(N, dtype=np.uint8)
    eol[::3] = 1
    A = eolc.collision_constraints(recs, N, eol=eol)
    R = restated(recs, N, eol=eol)
    assert A.shape == R.shape
    assert (abs(A - R)).max() == 0.0
    # a (2,2) record whose two vertices coincide: one merged column triple
    r = eolc.CollisionRec()
    r.count1, r.count2 = 2, 2
    r.verts2[0] = r.verts2[1] = 4
    r.weights2[0], r.weights2[1] = 0.25, 0.5
    r.nor2[0], r.nor2[1], r.nor2[2] = 1.0, 2.0, 3.0
    A1 = eolc.collision_constraints([r], N).toarray()
    assert np.count_nonzero(A1) == 3
    assert np.allclose(A1[0, 12:15], [-0.75, -1.5, -2.25])


def test_collision_rows_empty_and_bad_vertex():
    assert eolc.collision_constraints([], 10).shape == (0, 30)
    r = eolc.CollisionRec()
    r.count1, r.count2 = 3, 1
    r.verts2[0] = 11
    try:
        eolc.collision_constraints([r], 10)
        raise AssertionError("out-of-range vertex accepted")
    except eolc.EolcError as e:
        assert e.code == eolc.ERR_INVALID
