"""GPU parity of the tile-pair numeric path (FP64 matrix cores, one
v_mfma_f64_4x4x4f64 per (A tile, B tile) pair) against the CPU oracle:
structure bit-exact, values within the 1e-14 forward-error bound, and the
same bits on every call (each C tile sums its pairs in increasing K)."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import as_tuple, check_c, oracle_product

pytestmark = pytest.mark.gpu


def tile_run(rt, a, aat, dense_threshold=24, positive=False, **kw):
    pa = rt.prepare_identity_a(a, dense_threshold)
    pb = rt.rebuild_b(pa, aat)
    tc, stats = rt.tilespgemm(pa, pb, rt.schedule(numeric_path="tile", **kw))
    assert stats["numeric_path"] == rt.NUMERIC_PATHS["tile"]
    ref = oracle_product(a, aat)
    err = check_c(rt, tc, ref, positive)
    assert stats["output_nnz"] == ref["nnz"]
    return pa, pb, tc, stats, err


CASES = [
    ("laplace64", lambda rt: rt.gen_laplace2d(64, 0.1, 20260101), 0, False),
    ("stencil5", lambda rt: rt.gen_stencil27(5, 0.5, 1.5, 7), 1, True),
    ("stencil12", lambda rt: rt.gen_stencil27(12, 0.5, 1.5, 8), 1, True),
    ("random_aa", lambda rt: rt.gen_random(300, 300, 9, 31), 0, False),
    ("random_aat", lambda rt: rt.gen_random(257, 190, 7, 32), 1, False),
    ("fem_morton", lambda rt: rt.gen_fem_blocks(1500, 6, True, 77), 0, False),
    ("fem_draw_order", lambda rt: rt.gen_fem_blocks(600, 6, False, 78), 0, False),
    ("rmat12", lambda rt: rt.gen_rmat(12, 16, seed=5), 0, False),
    ("rmat12_aat", lambda rt: rt.gen_rmat(12, 16, seed=6), 1, False),
]


@pytest.mark.parametrize("name,make,aat,positive", CASES, ids=[c[0] for c in CASES])
def test_tile_path_matches_oracle(gpu, name, make, aat, positive):
    print(f"case {name}", flush=True)
    tile_run(gpu, make(gpu), aat, positive=positive)


def test_tile_path_rows_beyond_the_lds_staging(gpu):
    """Tile rows whose A tiles (more than 128) or A payload (more than 512
    values) do not fit the kernel's LDS staging read A from HBM instead; rows
    of both kinds and rows that fit run in one call."""
    import scipy.sparse as sp
    n = 4000
    wide = gpu.gen_random(16, n, 300, 51).to_scipy()              # ~900 A tiles per tile row
    deep = gpu.gen_random(16, 480, 100, 52).to_scipy()            # <= 120 A tiles, ~800 values
    deep = sp.hstack([deep, sp.csr_matrix((16, n - 480))])
    small = gpu.gen_random(32, n, 4, 53).to_scipy()
    a = sp.vstack([wide, deep, small]).tocsr()
    a.sort_indices()
    m = gpu.CsrMatrix(a.shape[0], a.shape[1], a.indptr.astype(np.int32),
                      a.indices.astype(np.int32), a.data.astype(np.float64))
    at = gpu.prepare_identity_a(m).tiles()
    per_row = np.diff(at["tile_row_ptr"])
    pay = np.diff(at["value_offsets"][at["tile_row_ptr"]])
    assert per_row.max() > 128 and ((per_row <= 128) & (pay > 512)).any()
    b = gpu.gen_random(4000, 300, 3, 54)
    pa = gpu.prepare_identity_a(m)
    pb = gpu.rebuild_b_general(pa, gpu.upload_csr(b))
    tc, _ = gpu.tilespgemm(pa, pb, gpu.schedule(numeric_path="tile"))
    check_c(gpu, tc, oracle.spgemm(as_tuple(m), as_tuple(b)))
    tile_run(gpu, m, 1)


@pytest.mark.parametrize("dense_threshold", [1, 24, 32])
def test_tile_path_dense_thresholds(gpu, dense_threshold):
    """Dense (zero-filled, 32-slot) and sparse payloads decode the same."""
    a = gpu.gen_random(200, 200, 20, 90 + dense_threshold)
    tile_run(gpu, a, 0, dense_threshold=dense_threshold)
    tile_run(gpu, a, 1, dense_threshold=dense_threshold)


def test_tile_path_general_ab_rectangular(gpu):
    a = gpu.gen_random(333, 170, 6, 41)
    b = gpu.gen_random(170, 251, 5, 42)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b_general(pa, gpu.upload_csr(b))
    tc, stats = gpu.tilespgemm(pa, pb, gpu.schedule(numeric_path="tile"))
    ref = oracle.spgemm(as_tuple(a), as_tuple(b))
    check_c(gpu, tc, ref)


def test_tile_path_is_bit_reproducible(gpu):
    a = gpu.gen_fem_blocks(2000, 6, True, 5)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 0)
    s = gpu.schedule(numeric_path="tile")
    t1, _ = gpu.tilespgemm(pa, pb, s)
    t2, _ = gpu.tilespgemm(pa, pb, s)
    assert np.array_equal(t1.tile_csr_Value.view(np.uint64), t2.tile_csr_Value.view(np.uint64))
    r = gpu.gen_rmat(13, 16, seed=9)
    pr = gpu.prepare_identity_a(r)
    pbr = gpu.rebuild_b(pr, 0)
    u1, _ = gpu.tilespgemm(pr, pbr, s)
    u2, _ = gpu.tilespgemm(pr, pbr, s)
    assert np.array_equal(u1.tile_csr_Value.view(np.uint64), u2.tile_csr_Value.view(np.uint64))


def test_auto_path_follows_tile_fill(gpu):
    """AUTO takes the matrix cores for the block-FEM matrix (16..32 stored
    values per A tile) and the LDS-accumulator kernels for the stencil
    (about 6 per tile); both give the oracle's C."""
    f = gpu.gen_fem_blocks(800, 6, True, 3)
    pa = gpu.prepare_identity_a(f)
    tc, st = gpu.tilespgemm(pa, gpu.rebuild_b(pa, 0))
    assert st["numeric_path"] == gpu.NUMERIC_PATHS["tile"]
    check_c(gpu, tc, oracle_product(f, 0))
    s = gpu.gen_stencil27(8, 0.5, 1.5, 3)
    ps = gpu.prepare_identity_a(s)
    tc2, st2 = gpu.tilespgemm(ps, gpu.rebuild_b(ps, 1))
    assert st2["numeric_path"] == gpu.NUMERIC_PATHS["vector"]
    check_c(gpu, tc2, oracle_product(s, 1), positive=True)


def test_tile_path_values_update(gpu):
    """A values-only B update keeps B's tile columns (structure and payload
    offsets) and the tile path reads the new values."""
    a = gpu.gen_random(180, 150, 6, 61)
    b = gpu.gen_random(150, 140, 5, 62)
    pa = gpu.prepare_identity_a(a)
    src = gpu.upload_csr(b)
    pb = gpu.rebuild_b_general(pa, src)
    s = gpu.schedule(numeric_path="tile")
    gpu.tilespgemm(pa, pb, s)
    b2 = gpu.CsrMatrix(b.m, b.n, b.rowptr, b.col, b.val * -2.5 + 0.25)
    src.set_values(b2.val)
    gpu.update_b_values(pa, src, pb, "safe-auto")
    tc, _ = gpu.tilespgemm(pa, pb, s)
    check_c(gpu, tc, oracle.spgemm(as_tuple(a), as_tuple(b2)))


# ---- split precision (Ozaki scheme on BF16 matrix cores) --------------------

SPLIT_FWD_TOL = 1e-12


def split_bound_ok(rt, a, b, tc, ref):
    """Checks every entry of C against the split form's error bound: each
    operand entry keeps 64 bits below its row (A) / column (B) maximum, so
    |c - c*| <= 2^-61 (maxrow|a|_i (P_A |B|)_ij + maxcol|b|_j (|A| P_B)_ij)
    plus FP64 rounding of the recombination, 2^-50 (|A||B|)_ij.  Returns the
    value errors (oracle.value_errors) for the caller's own tolerance."""
    import scipy.sparse as sp
    A, B = a.to_scipy().tocsr(), b.to_scipy().tocsr()
    absA, absB = abs(A), abs(B)
    PA, PB = absA.copy(), absB.copy()
    PA.data[:] = 1.0
    PB.data[:] = 1.0
    ra = np.asarray(absA.max(axis=1).todense()).ravel()
    rb = np.asarray(absB.max(axis=0).todense()).ravel()
    t1 = (sp.diags(ra) @ (PA @ absB)).tocsr()
    t2 = ((absA @ PB) @ sp.diags(rb)).tocsr()
    ab = (absA @ absB).tocsr()
    for M in (t1, t2, ab):
        M.sort_indices()
    csr = rt.tile2csr(tc, canonical=True)
    assert np.array_equal(t1.indices, csr.col) and np.array_equal(ab.indices, ref["col"])
    err = np.abs((csr.val - ref["hi"]) - ref["lo"])
    bound = 2.0 ** -61 * (t1.data + t2.data) + 2.0 ** -50 * ab.data
    worst = np.max(err / bound) if err.size else 0.0
    assert worst <= 1.0, worst
    return oracle.value_errors(tc.tile_csr_Value, *(
        oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], v)["tile_csr_val"]
        for v in (ref["hi"], ref["lo"], ref["absprod"])))


def split_run(rt, a, aat):
    pa = rt.prepare_identity_a(a)
    pb = rt.rebuild_b(pa, aat)
    tc, stats = rt.tilespgemm(pa, pb, rt.schedule(precision="split-bf16"))
    assert stats["numeric_path"] == rt.NUMERIC_PATHS["tile"]
    assert stats["precision"] == rt.PRECISIONS["split-bf16"]
    ref = oracle_product(a, aat)
    # structure is exact whatever the arithmetic
    ct = oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], ref["val"])
    assert np.array_equal(tc.tile_columnidx, ct["tile_col"])
    assert np.array_equal(tc.tile_nnz, ct["tile_nnz"])
    assert np.array_equal(tc.tile_csr_Col, ct["tile_csr_col"])
    b = a
    if aat:
        b = rt.csr_transpose(a)
    return pa, pb, tc, split_bound_ok(rt, a, b, tc, ref)


SPLIT_CASES = [
    ("fem_morton", lambda rt: rt.gen_fem_blocks(1500, 6, True, 77), 0, True),
    ("fem_draw_order", lambda rt: rt.gen_fem_blocks(600, 6, False, 78), 0, True),
    ("laplace64", lambda rt: rt.gen_laplace2d(64, 0.1, 20260101), 0, True),
    ("stencil8", lambda rt: rt.gen_stencil27(8, 0.5, 1.5, 8), 1, True),
    ("random_aa", lambda rt: rt.gen_random(300, 300, 9, 31), 0, False),
    ("rmat12", lambda rt: rt.gen_rmat(12, 16, seed=5), 0, False),
]


@pytest.mark.parametrize("name,make,aat,bounded_spread", SPLIT_CASES,
                         ids=[c[0] for c in SPLIT_CASES])
def test_split_precision_within_its_bound(gpu, name, make, aat, bounded_spread):
    """Every entry within the slicing bound; on inputs whose magnitudes span
    at most six decades per tile row / column (block FEM, Laplacian,
    stencil) that is a forward error of at most 1e-12.  Uniform [-1, 1]
    values (random, R-MAT) have unbounded spread, so only the bound holds."""
    print(f"case {name}", flush=True)
    _, _, _, err = split_run(gpu, make(gpu), aat)
    print(name, err, flush=True)
    if bounded_spread:
        assert err["fwd_max"] <= SPLIT_FWD_TOL, err


def test_split_precision_is_bit_reproducible(gpu):
    a = gpu.gen_fem_blocks(1200, 6, True, 19)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 0)
    s = gpu.schedule(precision="split-bf16")
    t1, _ = gpu.tilespgemm(pa, pb, s)
    t2, _ = gpu.tilespgemm(pa, pb, s)
    assert np.array_equal(t1.tile_csr_Value.view(np.uint64), t2.tile_csr_Value.view(np.uint64))


def test_split_precision_rejects_the_vector_path(gpu):
    a = gpu.gen_random(40, 40, 3, 1)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 0)
    with pytest.raises(gpu.RttError):
        gpu.tilespgemm(pa, pb, gpu.schedule(numeric_path="vector", precision="split-bf16"))


def test_auto_routes_rows_of_one_call_to_both_paths(gpu):
    """A block-diagonal batch of a block-FEM matrix (16..32 stored values per
    A tile) and a sparse random one (about 2): AUTO sends the FEM tile rows
    to the tile-pair kernel and the others to the LDS-accumulator kernels in
    the same call, and C is the oracle's."""
    f = gpu.gen_fem_blocks(700, 6, True, 21)
    r = gpu.gen_random(3000, 3000, 5, 22)
    big, ro, co = gpu.block_diag([f, r, f], 8)
    pa = gpu.prepare_identity_a(big)
    pb = gpu.rebuild_b(pa, 0)
    tc, st = gpu.tilespgemm(pa, pb)
    assert st["numeric_path"] == gpu.NUMERIC_PATHS["auto"]
    assert 0 < st["mfma_pairs"] < st["pair_count"]
    check_c(gpu, tc, oracle_product(big, 0))
    # the vector rows' neighbours in AA^T (general transpose path) as well
    pb2 = gpu.rebuild_b(pa, 1)
    tc2, st2 = gpu.tilespgemm(pa, pb2)
    assert st2["numeric_path"] == gpu.NUMERIC_PATHS["auto"]
    check_c(gpu, tc2, oracle_product(big, 1))


def test_split_precision_on_wide_spread_operands(gpu):
    """The Galerkin product R (A P) of a Poisson operator with 10^+-3
    coefficient jumps: A P's tile columns span far more than 20 binades, so
    the split form multiplies those batches on the FP64 cores, and the
    forward error stays at most 1e-12 everywhere (the slices alone would
    lose up to 2e-10 there)."""
    g = 10
    A = gpu.gen_poisson7_jumps(g, 3.0, 91)
    P = gpu.gen_aggregation_prolongator(A, g, 2.0 / 3.0)
    R = gpu.csr_transpose(P)
    ap = oracle.spgemm(as_tuple(A), as_tuple(P))
    AP = gpu.CsrMatrix(ap["m"], ap["n"], ap["rowptr"].astype(np.int32), ap["col"], ap["val"])
    pr = gpu.prepare_identity_a(R)
    pb = gpu.rebuild_b_general(pr, gpu.upload_csr(AP))
    tc, st = gpu.tilespgemm(pr, pb, gpu.schedule(precision="split-bf16"))
    ref = oracle.spgemm(as_tuple(R), as_tuple(AP))
    err = split_bound_ok(gpu, R, AP, tc, ref)
    print("R(AP) split", err, flush=True)
    assert err["fwd_max"] <= SPLIT_FWD_TOL, err
