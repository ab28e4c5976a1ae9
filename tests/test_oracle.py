"""Pins the CPU oracle (oracle/rtt_oracle.cpp) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement (dropped
term, transposed operand, wrong index or physical layout) fails: brute-force
dense products, closed forms of the stencil workloads, the tile-format known
answer of the reference's design doc, the reference's own tail planner and
Matrix Market reader compiled from its sources (oracle/_ref), and exact
rational arithmetic for the double-double truth.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle


def rand_csr(rng, m, n, density, dup_rate=0.0, unsorted=False, explicit_zero=0.0):
    rows, cols, vals = [], [], []
    for i in range(m):
        k = rng.binomial(n, density) if n else 0
        cs = list(rng.choice(n, size=k, replace=False)) if k else []
        if dup_rate and cs:
            cs += [c for c in cs if rng.random() < dup_rate]
        if not unsorted:
            cs.sort()
        else:
            rng.shuffle(cs)
        for c in cs:
            rows.append(i)
            cols.append(int(c))
            vals.append(0.0 if rng.random() < explicit_zero else rng.uniform(-1, 1))
    rp = np.zeros(m + 1, np.int32)
    for r in rows:
        rp[r + 1] += 1
    rp = np.cumsum(rp).astype(np.int32)
    return m, n, rp, np.array(cols, np.int32), np.array(vals, np.float64)


def dense(a):
    m, n, rp, ci, v = a
    d = np.zeros((m, n))
    p = np.zeros((m, n), np.int64)
    for i in range(m):
        for e in range(rp[i], rp[i + 1]):
            d[i, ci[e]] += v[e]
            p[i, ci[e]] = 1
    return d, p


@pytest.mark.parametrize("seed", range(6))
def test_spgemm_matches_dense_product(seed):
    rng = np.random.default_rng(seed)
    m, k, n = rng.integers(1, 40, size=3)
    a = rand_csr(rng, m, k, rng.uniform(0.02, 0.4), dup_rate=0.2, explicit_zero=0.1)
    b = rand_csr(rng, k, n, rng.uniform(0.02, 0.4), dup_rate=0.2, explicit_zero=0.1)
    c = oracle.spgemm(a, b)
    da, pa = dense(a)
    db, pb = dense(b)
    want_pat = (pa @ pb) > 0
    got = np.zeros((m, n))
    got_pat = np.zeros((m, n), bool)
    for i in range(m):
        cols = c["col"][c["rowptr"][i]:c["rowptr"][i + 1]]
        assert np.all(np.diff(cols) > 0), "columns ascending and unique"
        got_pat[i, cols] = True
        got[i, cols] = c["val"][c["rowptr"][i]:c["rowptr"][i + 1]]
    assert np.array_equal(got_pat, want_pat)
    np.testing.assert_allclose(got, da @ db, rtol=1e-12, atol=1e-14)
    # Transposed operand must not pass.
    if m == n == k and not np.allclose(da @ db, da @ db.T):
        assert not np.allclose(got, da @ db.T, rtol=1e-12, atol=1e-14)


def test_double_double_is_exact_on_cancellation():
    # Terms that cancel catastrophically in plain FP64.
    a = (1, 4, np.array([0, 4], np.int32), np.array([0, 1, 2, 3], np.int32),
         np.array([1e16, 1.0, -1e16, 3.0]))
    b = (4, 1, np.array([0, 1, 2, 3, 4], np.int32), np.zeros(4, np.int32),
         np.array([1.0, 1e-3, 1.0, 7e-5]))
    c = oracle.spgemm(a, b)
    exact = sum(Fraction(x) * Fraction(y) for x, y in zip(a[4], b[4]))
    got = Fraction(c["hi"][0]) + Fraction(c["lo"][0])
    assert abs(got - exact) <= abs(exact) * Fraction(1, 10**28)
    assert c["absprod"][0] == pytest.approx(2e16 + 1e-3 + 21e-5)


def test_closed_forms_laplace_and_stencil():
    import rtt_amd
    a = rtt_amd.gen_laplace2d(64, 0.1, 20260101)
    A = (a.m, a.n, a.rowptr, a.col, a.val)
    assert a.nnz == 20224
    assert oracle.nnz_cub(*A[:4], aat=False) == 100104
    c = oracle.spgemm(A, A, with_dd=False)
    assert c["nnz"] == 51972
    for g in (3, 5, 8):
        s = rtt_amd.gen_stencil27(g, 0.5, 1.5, 7)
        S = (s.m, s.n, s.rowptr, s.col, s.val)
        assert s.nnz == (3 * g - 2) ** 3
        assert oracle.nnz_cub(*S[:4], aat=True) == (9 * g - 10) ** 3
        t = oracle.transpose(*S)
        c = oracle.spgemm(S, (s.n, s.m) + t, with_dd=False)
        assert c["nnz"] == (5 * g - 6) ** 3


def test_spgemm_structure_matches_scipy():
    sp = pytest.importorskip("scipy.sparse")
    rng = np.random.default_rng(11)
    a = rand_csr(rng, 300, 300, 0.02)
    ones = (a[0], a[1], a[2], a[3], np.ones_like(a[4]))
    c = oracle.spgemm(ones, ones, with_dd=False)
    s = sp.csr_matrix((ones[4], ones[3], ones[2]), shape=(300, 300))
    p = (s @ s).tocsr()
    p.sort_indices()
    assert np.array_equal(c["rowptr"], p.indptr)
    assert np.array_equal(c["col"], p.indices)
    np.testing.assert_array_equal(c["val"], p.data)


def test_nnz_cub_definition():
    rng = np.random.default_rng(3)
    a = rand_csr(rng, 50, 50, 0.1, dup_rate=0.3)
    m, n, rp, ci, v = a
    want_aa = sum(rp[c + 1] - rp[c] for c in ci)
    deg = np.bincount(ci, minlength=n)
    want_aat = int(sum(deg[c] for c in ci))
    assert oracle.nnz_cub(m, n, rp, ci, aat=False) == want_aa
    assert oracle.nnz_cub(m, n, rp, ci, aat=True) == want_aat


# ---- tile format -------------------------------------------------------------

def test_tile_known_answer_from_design_doc():
    # The reordered 16x8 example (docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md
    # section 10): active rows 0..3 hold (0,0)=a, (1,1)=c, (2,4)=b, (3,5)=d.
    a_, b_, c_, d_ = 1.5, 2.5, 3.5, 4.5
    rp = np.array([0, 1, 2, 3, 4], np.int32)
    ci = np.array([0, 1, 4, 5], np.int32)
    v = np.array([a_, c_, b_, d_])
    t = oracle.hybrid_tiles(4, 8, rp, ci, v, 8, 4, False)
    assert list(t["tile_row_ptr"]) == [0, 2]
    assert list(t["tile_col_idx"]) == [0, 1]
    assert list(t["value_offsets"]) == [0, 2, 4]
    assert [int(x) for x in t["masks"]] == [0x21, 0x2100]
    assert list(t["values"]) == [a_, c_, b_, d_]


def test_dense_threshold_boundary_and_duplicates():
    # One 8x4 tile with 23 then 24 unique positions.
    for unique, dense in ((23, False), (24, True)):
        pos = list(range(unique))
        rows = [p // 4 for p in pos]
        cols = [p % 4 for p in pos]
        rp = np.zeros(9, np.int32)
        for r in rows:
            rp[r + 1] += 1
        rp = np.cumsum(rp).astype(np.int32)
        v = np.arange(1, unique + 1, dtype=float)
        t = oracle.hybrid_tiles(8, 4, rp, np.array(cols, np.int32), v, 8, 4, False)
        assert t["dense_tiles"] == int(dense)
        span = t["value_offsets"][1] - t["value_offsets"][0]
        assert span == (32 if dense else unique)
        if dense:
            assert list(t["values"][:unique]) == list(v) and not t["values"][unique:].any()
    # Duplicates are summed; a zero-sum duplicate keeps its mask bit.
    rp = np.array([0, 3, 3, 3, 3, 3, 3, 3, 3], np.int32)
    ci = np.array([1, 1, 2], np.int32)
    v = np.array([2.0, -2.0, 5.0])
    t = oracle.hybrid_tiles(8, 4, rp, ci, v, 8, 4, False)
    assert int(t["masks"][0]) == 0b110
    assert list(t["values"]) == [0.0, 5.0]


def test_b_layout_is_column_major():
    # B 4x8: entry (k, c) sits at physical c*4 + k.
    rp = np.array([0, 1, 2, 3, 4], np.int32)
    ci = np.array([7, 0, 3, 5], np.int32)
    v = np.array([1.0, 2.0, 3.0, 4.0])
    t = oracle.hybrid_tiles(4, 8, rp, ci, v, 4, 8, True, build_csc=True)
    want = (1 << (7 * 4 + 0)) | (1 << (0 * 4 + 1)) | (1 << (3 * 4 + 2)) | (1 << (5 * 4 + 3))
    assert int(t["masks"][0]) == want
    assert list(t["values"]) == [2.0, 3.0, 4.0, 1.0]
    assert list(t["tile_col_ptr"]) == [0, 1] and list(t["csc_tile_ids"]) == [0]


def test_b_tiles_of_transpose_equal_a_tiles():
    # B = X^T in 4x8 column-major tiles is A = X in 8x4 row-major tiles with
    # the tile grid transposed: same masks and payloads.
    rng = np.random.default_rng(5)
    x = rand_csr(rng, 37, 29, 0.15, dup_rate=0.2)
    a = oracle.hybrid_tiles(*x, 8, 4, False, build_csc=True)
    xt = oracle.transpose(*x)
    b = oracle.hybrid_tiles(x[1], x[0], *xt, 4, 8, True, build_csc=True)
    assert a["num_tiles"] == b["num_tiles"]
    # B tile (K, J) <-> A tile (J, K)
    a_map = {}
    for r in range(a["tile_row_count"]):
        for t in range(a["tile_row_ptr"][r], a["tile_row_ptr"][r + 1]):
            o, e = a["value_offsets"][t], a["value_offsets"][t + 1]
            a_map[(r, a["tile_col_idx"][t])] = (int(a["masks"][t]), list(a["values"][o:e]))
    for k in range(b["tile_row_count"]):
        for t in range(b["tile_row_ptr"][k], b["tile_row_ptr"][k + 1]):
            o, e = b["value_offsets"][t], b["value_offsets"][t + 1]
            assert a_map[(b["tile_col_idx"][t], k)] == (int(b["masks"][t]), list(b["values"][o:e]))


@pytest.mark.parametrize("seed", range(3))
def test_c_tiles_roundtrip_through_tile2csr(seed):
    rng = np.random.default_rng(100 + seed)
    x = rand_csr(rng, 45, 70, 0.2)
    m, n, rp, ci, v = x
    ct = oracle.c_tiles(m, n, rp.astype(np.int64), ci, v)
    back = oracle.tile2csr(m, n, ct["tilem"], ct["tile_ptr"], ct["tile_col"], ct["tile_nnz"],
                           ct["tile_csr_ptr"], ct["tile_csr_col"], ct["tile_csr_val"])
    assert back is not None
    assert np.array_equal(back[0], rp) and np.array_equal(back[1], ci)
    assert np.array_equal(back[2], v)
    # Mask bits and row offsets agree.
    for t in range(ct["num_tiles"]):
        mk = int(ct["tile_mask"][t])
        for r in range(8):
            below = mk & ((1 << (8 * r)) - 1)
            assert ct["tile_csr_ptr"][8 * t + r] == bin(below).count("1")
    # Row restore with a permutation of the active prefix.
    perm = rng.permutation(m).astype(np.int32)
    back = oracle.tile2csr(m, n, ct["tilem"], ct["tile_ptr"], ct["tile_col"], ct["tile_nnz"],
                           ct["tile_csr_ptr"], ct["tile_csr_col"], ct["tile_csr_val"],
                           row_new_to_old=perm, permutation_size=m)
    for new in range(m):
        old = perm[new]
        assert np.array_equal(back[1][back[0][old]:back[0][old + 1]], ci[rp[new]:rp[new + 1]])
    # A repeated id is rejected; a non-empty row outside the prefix is rejected.
    bad = perm.copy()
    bad[1] = bad[0]
    assert oracle.tile2csr(m, n, ct["tilem"], ct["tile_ptr"], ct["tile_col"], ct["tile_nnz"],
                           ct["tile_csr_ptr"], ct["tile_csr_col"], ct["tile_csr_val"],
                           row_new_to_old=bad, permutation_size=m) is None
    last = int(np.nonzero(np.diff(rp))[0].max())
    assert oracle.tile2csr(m, n, ct["tilem"], ct["tile_ptr"], ct["tile_col"], ct["tile_nnz"],
                           ct["tile_csr_ptr"], ct["tile_csr_col"], ct["tile_csr_val"],
                           row_new_to_old=perm[:last], permutation_size=last) is None


# ---- tail-queue policy against the reference's own planner ----------------------

# Checked by file, not by loading it: collecting this module in a GPU test
# process must not map the checker library.
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="reference-source checker (oracle/_ref) not built")


@needs_ref
def test_tail_chunk_count_matches_reference():
    ref = oracle.ref_lib()
    for scans in (0, 1, 2, 3, 7, 64, 2048, 100000):
        for matches in (0, 1, 5, 100, 2047, 2048, 2049, 3000, 50000):
            if matches > scans and scans > 1:
                continue
            for thr, tgt, mx in ((16384.0, 8192.0, 8), (256.0, 256.0, 8), (100.0, 7.0, 32)):
                assert oracle.tail_chunk_count(scans, matches, thr, tgt, mx) == \
                    ref.ref_tail_chunk_count(scans, matches, thr, tgt, mx)
    # FP32 admission: work exactly at the threshold is not admitted.
    assert oracle.tail_chunk_count(16384, 0) == 0
    assert oracle.tail_chunk_count(16385, 0) == 3


@needs_ref
@pytest.mark.parametrize("seed", range(4))
def test_tail_plan_matches_reference(seed):
    rng = np.random.default_rng(seed)
    n = 5000
    scans = rng.integers(0, 40000, n)
    matches = (scans * rng.uniform(0, 1, n)).astype(np.int64)
    out_nnz = rng.integers(0, 65, n)
    for kw in (dict(), dict(max_heavy_fraction=1.0), dict(record_capacity=3),
               dict(workspace_limit=1024, max_heavy_fraction=1.0), dict(enabled=False),
               dict(threshold=-1.0)):
        a = oracle.plan_tail_queue(scans, matches, out_nnz, **kw)
        b = oracle.plan_tail_queue(scans, matches, out_nnz, use_reference=True, **kw)
        assert a["reason"] == b["reason"] and a["heavy"] == b["heavy"]
        assert a["chunks"] == b["chunks"] and a["workspace"] == b["workspace"]
        assert np.array_equal(a["heavy_ids"], b["heavy_ids"])
        assert np.array_equal(a["heavy_chunks"], b["heavy_chunks"])


# ---- Matrix Market reader against the reference's reader -------------------------

def write_mtx(path, kind, sym, entries, m, n):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {kind} {sym}\n% comment\n")
        f.write(f"{m} {n} {len(entries)}\n")
        for (i, j, v) in entries:
            if kind == "pattern":
                f.write(f"{i + 1} {j + 1}\n")
            elif kind == "integer":
                f.write(f"{i + 1} {j + 1} {int(v)}\n")
            else:
                f.write(f"{i + 1} {j + 1} {v!r}\n")


@needs_ref
@pytest.mark.parametrize("kind,sym", [("real", "general"), ("real", "symmetric"),
                                      ("pattern", "general"), ("integer", "symmetric")])
def test_mmio_reader_matches_reference(tmp_path, kind, sym):
    rng = np.random.default_rng(9)
    m = n = 30
    ent = []
    for _ in range(120):
        i, j = int(rng.integers(m)), int(rng.integers(n))
        if sym == "symmetric" and j > i:
            i, j = j, i
        ent.append((i, j, float(rng.integers(-9, 9)) if kind == "integer" else rng.uniform(-1, 1)))
    ent += ent[:5]  # duplicates stay as separate entries
    p = str(tmp_path / "a.mtx")
    write_mtx(p, kind, sym, ent, m, n)
    got = oracle.mmio_read(p)
    ref = oracle.ref_mmio_read(p)
    for g, r in zip(got, ref):
        assert np.array_equal(np.asarray(g), np.asarray(r))


# ---- committed vectors of the reference's functions ----------------------------
# tests/golden/reference_vectors.json was written by
# tests/golden/make_reference_vectors.py from the reference sources compiled
# into oracle/_ref; these checks need no reference build.

def golden():
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        return json.load(f)


def test_tail_chunk_count_matches_reference_vectors():
    g = golden()["tail_chunk_count"]
    assert len(g["rows"]) > 100
    for scans, matches, thr, tgt, mx, want in g["rows"]:
        assert oracle.tail_chunk_count(int(scans), int(matches), thr, tgt, int(mx)) == want


def test_tail_plan_matches_reference_vectors():
    for p in golden()["plan_tail_queue"]:
        got = oracle.plan_tail_queue(np.array(p["scans"]), np.array(p["matches"]),
                                     np.array(p["output_nnz"]), **p["settings"])
        for k in ("reason", "heavy", "chunks", "workspace"):
            assert got[k] == p[k], (p["settings"], k)
        assert got["heavy_ids"].tolist() == p["heavy_ids"]
        assert got["heavy_chunks"].tolist() == p["heavy_chunks"]


# ---- legacy 16x16 tile CPU SpGEMM -------------------------------------------------

def _laplace(g, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    n = g * g
    rows, cols, vals = [], [], []
    for i in range(g):
        for j in range(g):
            r = i * g + j
            rows.append(r); cols.append(r); vals.append(4.0)
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < g and 0 <= jj < g:
                    rows.append(r); cols.append(ii * g + jj); vals.append(-1.0 + rng.uniform(-0.1, 0.1))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A


def _tup(A):
    return (A.shape[0], A.shape[1], A.indptr.astype(np.int32), A.indices.astype(np.int32),
            A.data.astype(np.float64))


@pytest.mark.parametrize("case", ["laplace64_aa", "random_rect", "banded_aat", "partial_tiles"])
def test_legacy_tile_spgemm_equals_the_spa(case):
    """The legacy 16x16 tile algorithm restated from src/spgemm-cpu.h takes
    each C entry's products in increasing k, like the row SPA, so on inputs
    without duplicates both give the same C bit for bit; the 64^2 Laplacian
    A^2 (the C1 configuration) has 51,972 entries (closed form)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    if case == "laplace64_aa":
        A = _laplace(64, 20260101); B = A
    elif case == "random_rect":
        A = sp.random(150, 90, density=0.06, random_state=4, format="csr")
        B = sp.random(90, 200, density=0.05, random_state=5, format="csr")
    elif case == "banded_aat":
        A = sp.diags([rng.uniform(-1, 1, 300 - abs(o)) for o in range(-7, 8)], list(range(-7, 8)),
                     format="csr")
        B = A.T.tocsr()
    else:
        A = sp.random(37, 45, density=0.2, random_state=6, format="csr")
        B = sp.random(45, 23, density=0.2, random_state=7, format="csr")
    A.sort_indices(); B.sort_indices()
    rp, col, val = oracle.legacy_tile_spgemm(_tup(A), _tup(B))
    ref = oracle.spgemm(_tup(A), _tup(B), with_dd=False)
    assert np.array_equal(rp, ref["rowptr"])
    assert np.array_equal(col, ref["col"])
    assert np.array_equal(val.view(np.uint64), ref["val"].view(np.uint64))
    if case == "laplace64_aa":
        assert rp[-1] == 51972


def test_legacy_tile_spgemm_sums_duplicates_in_the_tile_build():
    """Duplicates are summed when the tiles are built, so with duplicates the
    legacy C agrees with the SPA's within rounding, structure exact."""
    import scipy.sparse as sp
    A = sp.random(60, 60, density=0.1, random_state=8, format="coo")
    r = np.concatenate([A.row, A.row[:20]]); c = np.concatenate([A.col, A.col[:20]])
    v = np.concatenate([A.data, np.linspace(-1, 1, 20)])
    order = np.lexsort((c, r))
    rp = np.zeros(61, np.int32)
    np.add.at(rp, r + 1, 1)
    a = (60, 60, np.cumsum(rp).astype(np.int32), c[order].astype(np.int32), v[order])
    lrp, lcol, lval = oracle.legacy_tile_spgemm(a, a)
    ref = oracle.spgemm(a, a)
    assert np.array_equal(lrp, ref["rowptr"]) and np.array_equal(lcol, ref["col"])
    err = oracle.value_errors(lval, ref["hi"], ref["lo"], ref["absprod"])
    assert err["fwd_max"] <= 1e-14
