"""GPU parity of general C = A*B (an independent B CSR, the reference's
--b B.mtx, src/main.cu:1554-1566) and of the values-only B update
(docs/BGRF_V1_PPOPP_ALGORITHM_AND_DATA_FORMAT_CN.md section 9.10).

Structure bit-exact against the oracle, FP64 values within the forward-error
bound of test_gpu_parity.  A values-only update must give the same B tiles
(payload included) bit for bit as a fresh rebuild from the same values; C
from either is checked against the oracle (C's values are sums the numeric
kernel accumulates with LDS atomics, so their last bit may differ between
calls).
"""
import numpy as np
import pytest

import oracle
from test_gpu_parity import FWD_TOL, as_tuple, check_c, csr_from, oracle_tiles_equal, sub_rows

pytestmark = pytest.mark.gpu


def product(rt, pa, b_csr, a, positive=False):
    src = rt.upload_csr(b_csr)
    pb = rt.rebuild_b_general(pa, src)
    assert pb.last["path"] == 3
    tc, stats = rt.tilespgemm(pa, pb)
    ref = oracle.spgemm(as_tuple(a), as_tuple(b_csr))
    check_c(rt, tc, ref, positive)
    assert stats["output_nnz"] == ref["nnz"]
    return src, pb, tc, ref


def nnz_cub_ab_truth(a, b):
    rowlen = np.diff(b.rowptr.astype(np.int64))
    return int(rowlen[a.col].sum()) if a.nnz else 0


@pytest.mark.parametrize("seed", range(4))
def test_general_ab_random_rectangular(gpu, seed):
    rng = np.random.default_rng(100 + seed)
    m, k, n = (int(x) for x in rng.integers(20, 600, size=3))
    a = gpu.gen_random(m, k, int(rng.integers(1, 10)), 7 * seed + 1)
    b = gpu.gen_random(k, n, int(rng.integers(1, 12)), 7 * seed + 2)
    pa = gpu.prepare_identity_a(a)
    src, pb, _, _ = product(gpu, pa, b, a)
    assert gpu.nnz_cub_ab(pa, src) == nnz_cub_ab_truth(a, b)
    # B's tiles are the oracle's 4x8 column-major tiles of B's active rows.
    _, ai, _ = pa.info()
    want = oracle.hybrid_tiles(*sub_rows(as_tuple(b), ai), 4, 8, True, 24, build_csc=True)
    oracle_tiles_equal(pb.tiles(), want, csc=True)


def test_general_ab_unsorted_duplicates_and_dense_tiles(gpu):
    rng = np.random.default_rng(7)
    k, n = 96, 130
    rows, cols, vals = [], [], []
    for i in range(k):
        if i % 9 == 4:
            cs = list(range(16, 24))          # rows filling whole B tiles (dense)
        else:
            cs = list(rng.choice(n, size=int(rng.integers(0, 10)), replace=False))
        if cs and i % 3 == 0:
            cs.append(cs[0])                  # duplicate coordinate
        cs = list(rng.permutation(cs))        # unsorted row
        rows += [i] * len(cs)
        cols += cs
        vals += list(rng.uniform(-1, 1, size=len(cs)))
    b = csr_from(k, n, rows, cols, vals)   # groups rows, keeps each row's order
    a = gpu.gen_random(70, k, 6, 11)
    pa = gpu.prepare_identity_a(a)
    product(gpu, pa, b, a)


def test_general_ab_shape_errors(gpu):
    a = gpu.gen_random(30, 40, 3, 1)
    b = gpu.gen_random(41, 10, 3, 2)
    pa = gpu.prepare_identity_a(a)
    src = gpu.upload_csr(b)
    with pytest.raises(gpu.RttError):
        gpu.rebuild_b_general(pa, src)
    with pytest.raises(gpu.RttError):
        gpu.nnz_cub_ab(pa, src)


def test_general_ab_through_external_order(gpu):
    rng = np.random.default_rng(3)
    m, k, n = 200, 150, 90
    a = gpu.gen_random(m, k, 5, 21)
    b = gpu.gen_random(k, n, 4, 22)
    rnew = rng.permutation(m).astype(np.int32)
    inew = rng.permutation(k).astype(np.int32)
    pa = gpu.prepare_external_a(a, rnew, inew)
    src = gpu.upload_csr(b)
    pb = gpu.rebuild_b_general(pa, src)
    tc, _ = gpu.tilespgemm(pa, pb)
    ar, _, ident = pa.info()
    assert not ident
    csr = gpu.tile2csr_restore_rows(tc, pa.row_new_to_old(), ar)
    ref = oracle.spgemm(as_tuple(a), as_tuple(b))
    assert np.array_equal(csr.rowptr.astype(np.int64), ref["rowptr"])
    assert np.array_equal(csr.col, ref["col"])
    err = oracle.value_errors(csr.val, ref["hi"], ref["lo"], ref["absprod"])
    assert err["fwd_max"] <= FWD_TOL, err


@pytest.mark.parametrize("policy", ["always-clear", "noclear", "safe-auto"])
def test_values_update_equals_rebuild(gpu, policy):
    a = gpu.gen_random(300, 260, 6, 31)
    b = gpu.gen_random(260, 180, 7, 32)
    pa = gpu.prepare_identity_a(a)
    src = gpu.upload_csr(b)
    pb = gpu.rebuild_b_general(pa, src)
    rng = np.random.default_rng(5)
    for it in range(3):
        new_vals = rng.uniform(-2, 2, size=b.nnz)
        src.set_values(new_vals)
        d = gpu.update_b_values(pa, src, pb, policy)
        assert pb.last["path"] == 4
        want = {"always-clear": "policy-always-clear", "noclear": "eligible-explicit-noclear",
                "safe-auto": "eligible-safe-auto"}[policy]
        assert d["reason_name"] == want
        assert d["clear_payload"] == (1 if policy == "always-clear" else 0)
        tc, _ = gpu.tilespgemm(pa, pb)
        b2 = gpu.CsrMatrix(b.m, b.n, b.rowptr, b.col, new_vals)
        fresh = gpu.rebuild_b_general(pa, gpu.upload_csr(b2))
        # B's tile view (values included) equals the fresh build bit for bit.
        oracle_tiles_equal(pb.tiles(), fresh.tiles(), csc=True)
        tc2, _ = gpu.tilespgemm(pa, fresh)
        assert np.array_equal(tc.tile_csr_Col, tc2.tile_csr_Col)
        ref = oracle.spgemm(as_tuple(a), as_tuple(b2))
        check_c(gpu, tc, ref)


def test_values_update_with_duplicates_and_unsorted_rows(gpu):
    """Duplicates: safe-auto falls back to clearing (reason
    active-source-duplicates); the update still sums each coordinate's
    entries in input order, like the build, and values arrive in the caller's
    (unsorted) entry order."""
    rng = np.random.default_rng(9)
    k, n = 120, 77
    rp = [0]
    cols, vals = [], []
    for i in range(k):
        cs = list(rng.choice(n, size=int(rng.integers(0, 8)), replace=False))
        if cs and i % 4 == 1:
            cs += [cs[-1], cs[0]]
        cs = list(rng.permutation(cs))
        cols += cs
        vals += list(rng.uniform(-1, 1, size=len(cs)))
        rp.append(len(cols))
    b = gpu.CsrMatrix(k, n, np.array(rp, np.int32), np.array(cols, np.int32),
                      np.array(vals, np.float64))
    a = gpu.gen_random(90, k, 5, 41)
    pa = gpu.prepare_identity_a(a)
    src = gpu.upload_csr(b)
    pb = gpu.rebuild_b_general(pa, src)
    new_vals = rng.uniform(-3, 3, size=b.nnz)
    src.set_values(new_vals)
    d = gpu.update_b_values(pa, src, pb, "safe-auto")
    assert d["reason_name"] == "active-source-duplicates"
    assert d["clear_payload"] == 1 and d["fallback"] == 1
    tc, _ = gpu.tilespgemm(pa, pb)
    b2 = gpu.CsrMatrix(k, n, b.rowptr, b.col, new_vals)
    fresh = gpu.rebuild_b_general(pa, gpu.upload_csr(b2))
    oracle_tiles_equal(pb.tiles(), fresh.tiles(), csc=True)
    check_c(gpu, tc, oracle.spgemm(as_tuple(a), as_tuple(b2)))


def test_values_update_on_another_csr_rebuilds(gpu):
    a = gpu.gen_random(100, 80, 4, 51)
    b = gpu.gen_random(80, 60, 4, 52)
    c = gpu.gen_random(80, 60, 5, 53)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b_general(pa, gpu.upload_csr(b))
    other = gpu.upload_csr(c)
    d = gpu.update_b_values(pa, other, pb, "safe-auto")
    assert d["reason_name"] == "invalid-dynamic-b-state"
    assert pb.last["path"] == 3
    tc, _ = gpu.tilespgemm(pa, pb)
    check_c(gpu, tc, oracle.spgemm(as_tuple(a), as_tuple(c)))
    # An unknown policy is an error, before any rebuild.
    assert gpu.load().rtt_update_b_values(pa.handle, other.handle, 7, pb.handle, None,
                                          None) == 1


# ---- AMG Galerkin product R * (A * P) (a small instance of the batched
#      AMG workload's shape: 7-point Poisson with coefficient jumps,
#      aggregates of 2x2x2, one Jacobi smoothing step, R = P^T) -------------

def poisson7_jumps(g, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    kc = 10.0 ** rng.uniform(-1, 1, size=g ** 3)
    idx = np.arange(g ** 3).reshape(g, g, g)
    rows, cols, vals = [], [], []
    diag = np.zeros(g ** 3)
    for axis in range(3):
        for sgn in (-1, 1):
            nb = np.roll(idx, -sgn, axis=axis)
            valid = np.ones_like(idx, bool)
            sl = [slice(None)] * 3
            sl[axis] = -1 if sgn == 1 else 0
            valid[tuple(sl)] = False
            i, j = idx[valid], nb[valid]
            w = 0.5 * (kc[i] + kc[j])
            rows += list(i)
            cols += list(j)
            vals += list(-w)
            diag[i] += w
            # boundary faces add the cell's own coefficient (Dirichlet)
            diag[idx[~valid]] += kc[idx[~valid]]
    rows += list(range(g ** 3))
    cols += list(range(g ** 3))
    vals += list(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(g ** 3, g ** 3))


def prolongator(A, g, omega=2.0 / 3.0):
    import scipy.sparse as sp
    idx = np.arange(g ** 3).reshape(g, g, g)
    x, y, z = np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing="ij")
    h = g // 2
    agg = ((x // 2) * h + (y // 2)) * h + (z // 2)
    T = sp.csr_matrix((np.ones(g ** 3), (idx.ravel(), agg.ravel())), shape=(g ** 3, h ** 3))
    Dinv = sp.diags(1.0 / A.diagonal())
    return (T - omega * (Dinv @ A) @ T).tocsr()


def to_rtt(rt, M):
    M = M.tocsr()
    M.sort_indices()
    return rt.CsrMatrix(M.shape[0], M.shape[1], M.indptr.astype(np.int32),
                        M.indices.astype(np.int32), M.data.astype(np.float64))


@pytest.mark.parametrize("g", [8, 12])
def test_amg_galerkin_rap(gpu, g):
    A = poisson7_jumps(g, 60 + g)
    P = prolongator(A, g)
    a, p, r = to_rtt(gpu, A), to_rtt(gpu, P), to_rtt(gpu, P.T)
    pa = gpu.prepare_identity_a(a)
    _, _, tc_ap, ref_ap = product(gpu, pa, p, a)
    ap = gpu.tile2csr(tc_ap, canonical=True)
    pr = gpu.prepare_identity_a(r)
    _, _, tc_rap, ref_rap = product(gpu, pr, ap, r)
    rap = gpu.tile2csr(tc_rap, canonical=True)
    # RAP of a symmetric A is symmetric in structure; values up to rounding.
    import scipy.sparse as sp
    S = sp.csr_matrix((rap.val, rap.col, rap.rowptr), shape=(rap.m, rap.n))
    assert (S != S.T).nnz == 0 or abs(S - S.T).max() <= 1e-12 * abs(S).max()
    # A smoothing-weight change is a values-only update of P.
    srcp = gpu.upload_csr(p)
    pb = gpu.rebuild_b_general(pa, srcp)
    P2 = prolongator(A, g, omega=0.5)
    p2 = to_rtt(gpu, P2)
    assert np.array_equal(p2.col, p.col) and np.array_equal(p2.rowptr, p.rowptr)
    srcp.set_values(p2.val)
    d = gpu.update_b_values(pa, srcp, pb, "safe-auto")
    assert d["reason_name"] == "eligible-safe-auto"
    tc2, _ = gpu.tilespgemm(pa, pb)
    check_c(gpu, tc2, oracle.spgemm(as_tuple(a), as_tuple(p2)))


# ---- block-structured FEM-like A^2 (4x4 dense blocks on a random 3-D mesh,
#      log-uniform magnitudes over six decades, random signs) -------------

def fem_blocks(nodes, seed):
    from scipy.spatial import cKDTree
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    pts = rng.uniform(size=(nodes, 3))
    _, nb = cKDTree(pts).query(pts, k=7)          # self + 6 nearest
    pairs = set()
    for i in range(nodes):
        for j in nb[i]:
            pairs.add((i, int(j)))
            pairs.add((int(j), i))                   # symmetrised
    bi = np.array([p[0] for p in pairs])
    bj = np.array([p[1] for p in pairs])
    nblk = len(bi)
    mag = 10.0 ** rng.uniform(-3, 3, size=(nblk, 4, 4))
    val = mag * rng.choice([-1.0, 1.0], size=(nblk, 4, 4))
    r = (4 * bi[:, None, None] + np.arange(4)[None, :, None]).repeat(4, axis=2)
    c = (4 * bj[:, None, None] + np.arange(4)[None, None, :]).repeat(4, axis=1)
    return sp.csr_matrix((val.ravel(), (r.ravel(), c.ravel())), shape=(4 * nodes, 4 * nodes))


def test_fem_block_matrix_square(gpu):
    """Dense 4x4 blocks fill whole 8x4 A tiles and 4x8 B tiles (dense
    payloads), and the magnitudes span 1e-3..1e3 with random signs, so the
    values are checked with the forward-error bound."""
    A = fem_blocks(1500, 77)
    a = to_rtt(gpu, A)
    pa = gpu.prepare_identity_a(a)
    assert pa.tiles()["dense_tiles"] > 0
    pb = gpu.rebuild_b(pa, 0)
    tc, _ = gpu.tilespgemm(pa, pb)
    ref = oracle.spgemm(as_tuple(a), as_tuple(a))
    check_c(gpu, tc, ref)


# ---- a batch of banded A^2 problems of uneven sizes ----------------------

def banded(n, bw, per_row, rng):
    import scipy.sparse as sp
    rows, cols = [], []
    for i in range(n):
        lo, hi = max(0, i - bw), min(n, i + bw + 1)
        k = min(per_row, hi - lo)
        cs = set(rng.choice(np.arange(lo, hi), size=k, replace=False).tolist()) | {i}
        rows += [i] * len(cs)
        cols += sorted(cs)
    return sp.csr_matrix((rng.uniform(-1, 1, size=len(rows)), (rows, cols)), shape=(n, n))


def test_banded_batch(gpu):
    """Uneven banded problems (log-normal sizes, bandwidth 8..64, 5..30
    entries per row inside the band) solved one after another on one
    handle set each, as the batched workload's problems are."""
    rng = np.random.default_rng(2026)
    sizes = np.clip(np.round(np.exp(rng.normal(np.log(3000), 0.8, size=6))), 1000, 20000)
    for n in sizes.astype(int):
        A = banded(int(n), int(rng.integers(8, 65)), int(rng.integers(5, 31)), rng)
        a = to_rtt(gpu, A)
        pa = gpu.prepare_identity_a(a)
        pb = gpu.rebuild_b(pa, 0)
        tc, _ = gpu.tilespgemm(pa, pb)
        check_c(gpu, tc, oracle.spgemm(as_tuple(a), as_tuple(a)))


def test_general_ab_empty_operands(gpu):
    """An empty B (or an empty A) gives an empty C of the right shape, and a
    values-only update of an empty B is a no-op."""
    import numpy as np
    a = gpu.gen_random(50, 40, 4, 71)
    empty_b = gpu.CsrMatrix(40, 30, np.zeros(41, np.int32), np.zeros(0, np.int32), np.zeros(0))
    pa = gpu.prepare_identity_a(a)
    src = gpu.upload_csr(empty_b)
    pb = gpu.rebuild_b_general(pa, src)
    tc, stats = gpu.tilespgemm(pa, pb)
    assert stats["output_nnz"] == 0 and tc.nnz == 0 and (tc.m, tc.n) == (50, 30)
    src.set_values(np.zeros(0))
    d = gpu.update_b_values(pa, src, pb, "safe-auto")
    assert pb.last["path"] == 4 and d["reason_name"] == "eligible-safe-auto"
    assert gpu.nnz_cub_ab(pa, src) == 0
    empty_a = gpu.CsrMatrix(20, 40, np.zeros(21, np.int32), np.zeros(0, np.int32), np.zeros(0))
    pe = gpu.prepare_identity_a(empty_a)
    b = gpu.gen_random(40, 30, 3, 72)
    tc2, stats2 = gpu.tilespgemm(pe, gpu.rebuild_b_general(pe, gpu.upload_csr(b)))
    assert stats2["output_nnz"] == 0 and (tc2.m, tc2.n) == (20, 30)
