"""Full-size correctness of the benchmark's other configurations, in the form
the benchmark runs them:

  * block-FEM A*A at 500k nodes (2M rows, the benchmark's seed) on the FP64
    matrix cores and in the split BF16 form;
  * 512 banded A^2 problems in one block-diagonal call;
  * R*(A*P) over 64 AMG hierarchies in two block-diagonal calls.

C is too large to download whole, so each product is checked through what
is exact at any size: the nnz of every row against the oracle's structure
pass (the restated reference SPA), and every row sum against the float64
product with the all-ones vector (linearity), within the forward bound
scaled by |A| (|B| 1).  Row blocks spread over the matrix are compared with
the oracle entry by entry."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import as_tuple, check_c

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

SEED = 20260101
ROW_SUM_TOL = 1e-12   # of (|A| (|B| 1))_i: the per-entry forward bound summed over the row


def struct(m):
    return (m.m, m.n, m.rowptr, m.col)


def check_rows(gpu, pa, a, b, row_order=None):
    """Every row's nnz against the oracle structure pass and every row sum
    against A (B 1); returns nnz(C)."""
    print("row stats", flush=True)
    nnz, sums = gpu.last_c_row_stats(pa)
    print("oracle structure pass", flush=True)
    want = np.diff(oracle.spgemm_row_counts(struct(a), struct(b)))
    assert np.array_equal(nnz[:a.m], want)
    A, B = a.to_scipy().tocsr(), b.to_scipy().tocsr()
    one = np.ones(b.n)
    ref = A @ (B @ one)
    scale = abs(A) @ (abs(B) @ one)
    assert np.all(np.abs(sums[:a.m] - ref) <= ROW_SUM_TOL * scale + 1e-300)
    return int(want.sum())


def rows_of(a, r0, rows):
    rows = min(rows, a.m - r0)
    rp = (a.rowptr[r0:r0 + rows + 1] - a.rowptr[r0]).astype(np.int32)
    sl = slice(int(a.rowptr[r0]), int(a.rowptr[r0 + rows]))
    return (rows, a.n, rp, a.col[sl], a.val[sl])


def check_block(gpu, pa, a, b, r0, rows, col0=0, fwd_tol=None):
    """Rows r0 .. r0+rows of C against the oracle's product of those rows of
    A with B (structure bit-exact; values within the FP64 bound, or within
    fwd_tol for the split form)."""
    sub = rows_of(a, r0, rows)
    ref = oracle.spgemm(sub, as_tuple(b))
    tc = gpu.batch_extract(pa, r0, col0, sub[0], b.n)
    if fwd_tol is None:
        check_c(gpu, tc, ref)
        return
    ct = oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], ref["val"])
    assert np.array_equal(tc.tile_columnidx, ct["tile_col"])
    assert np.array_equal(tc.tile_nnz, ct["tile_nnz"])
    assert np.array_equal(tc.tile_csr_Col, ct["tile_csr_col"])
    hi, lo, ab = (oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], ref[k])
                  ["tile_csr_val"] for k in ("hi", "lo", "absprod"))
    err = oracle.value_errors(tc.tile_csr_Value, hi, lo, ab)
    assert err["fwd_max"] <= fwd_tol, err


@pytest.fixture(scope="module")
def fem(gpu):
    return gpu.gen_fem_blocks(500000, 6, True, SEED)


@pytest.mark.parametrize("precision", ["fp64", "split-bf16"])
def test_block_fem_500k_at_full_size(gpu, fem, precision):
    print(f"block FEM 500k, {precision}", flush=True)
    pa = gpu.prepare_identity_a(fem)
    pb = gpu.rebuild_b(pa, 0)
    _, st = gpu.tilespgemm(pa, pb, gpu.schedule(materialize_output=False, numeric_path="tile",
                                                precision=precision))
    assert st["numeric_path"] == gpu.NUMERIC_PATHS["tile"]
    assert st["mfma_pairs"] > 0
    total = check_rows(gpu, pa, fem, fem)
    assert st["output_nnz"] == total
    for r0 in (0, fem.m // 2, fem.m - 2048):
        print(f"oracle block at row {r0}", flush=True)
        check_block(gpu, pa, fem, fem, r0, 2048,
                    fwd_tol=None if precision == "fp64" else 1e-12)


def test_banded_batch_512_in_one_call(gpu):
    """The 512-problem banded batch of the benchmark (same shapes and seeds)
    as one block-diagonal call on the default (AUTO) path: per-row checks
    over the whole batch, and four whole problems against the oracle."""
    rows, bw, pr = gpu.gen_banded_batch_shapes(512, seed=SEED)
    mats = [gpu.gen_banded(int(rows[p]), int(bw[p]), int(pr[p]), SEED + p) for p in range(512)]
    big, ro, co = gpu.block_diag(mats, 8)
    print(f"batch: {big.m} rows, {big.nnz} nnz", flush=True)
    pa = gpu.prepare_identity_a(big)
    pb = gpu.rebuild_b(pa, 0)
    _, st = gpu.tilespgemm(pa, pb, gpu.schedule(materialize_output=False))
    total = check_rows(gpu, pa, big, big)
    assert st["output_nnz"] == total
    for p in (0, 1, 255, 511):
        print(f"oracle problem {p} ({mats[p].m} rows)", flush=True)
        tc = gpu.batch_extract(pa, ro[p], co[p], mats[p].m, mats[p].n)
        check_c(gpu, tc, oracle.spgemm(as_tuple(mats[p]), as_tuple(mats[p])))


def test_amg_batch_64_galerkin_products(gpu):
    """R*(A*P) over the benchmark's 64 AMG hierarchies (g 32..96), as two
    block-diagonal calls: A*P checked row by row against the oracle, then
    R times the GPU's A*P (downloaded as CSR) row by row, and two whole
    hierarchies against the oracle's chain."""
    gs = np.random.default_rng(SEED).integers(32, 97, 64)
    A = [gpu.gen_poisson7_jumps(int(g), 3.0, SEED + p) for p, g in enumerate(gs)]
    P = [gpu.gen_aggregation_prolongator(a, int(g), 2.0 / 3.0) for a, g in zip(A, gs)]
    R = [gpu.csr_transpose(p) for p in P]
    bA, roA, coA = gpu.block_diag(A, 8)
    bP, roP, coP = gpu.block_diag(P, 8)
    bR, roR, coR = gpu.block_diag(R, 8)
    assert np.array_equal(coA, roP) and np.array_equal(coP, roR)
    print(f"A*P: {bA.m} rows", flush=True)
    pa = gpu.prepare_identity_a(bA)
    tc, st = gpu.tilespgemm(pa, gpu.rebuild_b_general(pa, gpu.upload_csr(bP)))
    assert st["output_nnz"] == check_rows(gpu, pa, bA, bP)
    ap = gpu.tile2csr(tc, canonical=True)
    print(f"R*(A*P): {bR.m} rows", flush=True)
    pr = gpu.prepare_identity_a(bR)
    _, st2 = gpu.tilespgemm(pr, gpu.rebuild_b_general(pr, gpu.upload_csr(ap)),
                            gpu.schedule(materialize_output=False))
    assert st2["output_nnz"] == check_rows(gpu, pr, bR, ap)
    for p in (0, 63):
        print(f"oracle hierarchy {p} (g = {gs[p]})", flush=True)
        ap_ref = oracle.spgemm(as_tuple(A[p]), as_tuple(P[p]))
        ap_t = (ap_ref["m"], ap_ref["n"], ap_ref["rowptr"].astype(np.int32), ap_ref["col"],
                ap_ref["val"])
        rap = oracle.spgemm(as_tuple(R[p]), ap_t)
        tc_p = gpu.batch_extract(pr, roR[p], coP[p], R[p].m, P[p].n)
        ct = oracle.c_tiles(rap["m"], rap["n"], rap["rowptr"], rap["col"], rap["val"])
        assert np.array_equal(tc_p.tile_columnidx, ct["tile_col"])
        assert np.array_equal(tc_p.tile_nnz, ct["tile_nnz"])
        got = gpu.tile2csr(tc_p, canonical=True).to_scipy()
        As, Ps, Rs = A[p].to_scipy(), P[p].to_scipy(), R[p].to_scipy()
        want = (Rs @ (As @ Ps)).tocsr()
        scale = (abs(Rs) @ abs(As) @ abs(Ps)).tocsr()
        diff = abs(got - want).tocsr()
        assert (diff - scale.multiply(1e-13).tocsr()).max() <= 0.0
