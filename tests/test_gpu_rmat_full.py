"""Full-size correctness of the R-MAT 2^20 C = A*A product the benchmark
measures (scale 20, 16 draws per row, seed 20260101): nnz(C) is about 9.7e9,
so C's value offsets pass 2^31 and 2^32, and in the identity layout C has
about 2.1e9 tiles, near INT_MAX -- the regime where the reference's int32
output fails (src/super16_symbolic.cuh:2966-2971).

C is far too large to download, so it is checked through what is exact at
any size:
  * nnz of every row against the oracle's structure pass over the whole
    matrix (the restated reference SPA, orc_spgemm_count);
  * C tiles of every tile row against the same structure pass on the
    tile-grouped operands (rows of A grouped by 8, columns of B by 8);
  * every row sum against A (A 1) in float64 (linearity), within 1e-11 of
    |A| (|A| 1);
  * full oracle comparisons (structure bit-exact, values within the 1e-14
    forward bound) of the hub tile row and of row blocks whose value offsets
    lie past 2^31 and past 2^32.
Both the identity layout and the BGRF-v1 layout (rows permuted) are run."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import check_c

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

SCALE = 20
BLOCK = 512


@pytest.fixture(scope="module")
def rmat(gpu):
    return gpu.gen_rmat(SCALE, 16, 0.57, 0.19, 0.19, 20260101)


def grouped_counts(a, row_order):
    """C tiles per tile row of A*A, rows of A taken in row_order (new -> old):
    the structure pass on (A's rows grouped by 8) x (A's columns grouped by 8)."""
    import scipy.sparse as sp
    A = a.to_scipy().tocsr()
    P = A[row_order] if row_order is not None else A
    m = A.shape[0]
    G = sp.csr_matrix((np.ones(m), (np.arange(m) // 8, np.arange(m))),
                      shape=((m + 7) // 8, m))
    H = sp.csr_matrix((np.ones(A.shape[1]), (np.arange(A.shape[1]), np.arange(A.shape[1]) // 8)),
                      shape=(A.shape[1], (A.shape[1] + 7) // 8))
    A8 = (G @ abs(P)).tocsr()
    B8 = (abs(A) @ H).tocsr()
    for M in (A8, B8):
        M.sort_indices()
    t = (A8.shape[0], A8.shape[1], A8.indptr.astype(np.int32), A8.indices.astype(np.int32))
    u = (B8.shape[0], B8.shape[1], B8.indptr.astype(np.int32), B8.indices.astype(np.int32))
    return np.diff(oracle.spgemm_row_counts(t, u))


def check_block(gpu, a, pa, r0, rows, row_order):
    """Rows r0 .. r0+rows of C (tiled order) against the oracle's product of
    the same rows of A (through row_order) with A."""
    rows = min(rows, a.m - r0)
    old = row_order[r0:r0 + rows] if row_order is not None else np.arange(r0, r0 + rows)
    rp = np.zeros(rows + 1, np.int64)
    lens = (a.rowptr[old + 1] - a.rowptr[old]).astype(np.int64)
    rp[1:] = np.cumsum(lens)
    idx = np.concatenate([np.arange(a.rowptr[o], a.rowptr[o + 1]) for o in old])
    sub = (rows, a.n, rp.astype(np.int32), a.col[idx], a.val[idx])
    ref = oracle.spgemm(sub, (a.m, a.n, a.rowptr, a.col, a.val))
    tc = gpu.batch_extract(pa, r0, 0, rows, a.n)
    check_c(gpu, tc, ref)


def check_full(gpu, a, pa, row_order):
    print("row stats", flush=True)
    nnz, sums = gpu.last_c_row_stats(pa)
    old = row_order if row_order is not None else np.arange(a.m)
    print("oracle structure pass", flush=True)
    want = np.diff(oracle.spgemm_row_counts(
        (a.m, a.n, a.rowptr, a.col), (a.m, a.n, a.rowptr, a.col)))
    assert np.array_equal(nnz, want[old])
    total = int(nnz.sum())
    assert total > 2 ** 32
    # Row sums: C 1 = A (A 1).
    A = a.to_scipy().tocsr()
    one = np.ones(a.n)
    ref = A @ (A @ one)
    scale = abs(A) @ (abs(A) @ one)
    assert np.all(np.abs(sums - ref[old]) <= 1e-11 * scale[old] + 1e-300)
    # Tile counts per tile row.
    print("tile counts", flush=True)
    tp = gpu.last_c_tile_ptr(pa)
    tiles = grouped_counts(a, row_order)
    assert np.array_equal(np.diff(tp), tiles[:len(tp) - 1])
    # Oracle blocks: the hub tile row and blocks past the 2^31 / 2^32 offsets.
    prefix = np.concatenate([[0], np.cumsum(nnz)])
    starts = [0]
    for mark in (2 ** 31, 2 ** 32):
        r = int(np.searchsorted(prefix, mark, side="right")) - 1
        starts.append(r // 8 * 8)
    for r0 in starts:
        print(f"oracle block at row {r0} (offset {prefix[r0]})", flush=True)
        check_block(gpu, a, pa, r0, BLOCK if r0 else 8, row_order)
    return total, int(tp[-1])


def test_rmat20_identity_layout_at_full_size(gpu, rmat):
    pa = gpu.prepare_identity_a(rmat)
    pb = gpu.rebuild_b(pa, 0)
    _, st = gpu.tilespgemm(pa, pb, gpu.schedule(materialize_output=False))
    total, tiles = check_full(gpu, rmat, pa, None)
    assert st["output_nnz"] == total and st["output_tiles"] == tiles
    assert tiles > 2 ** 30
    print(f"identity: nnz(C) {total}, C tiles {tiles}", flush=True)


def test_rmat20_bgrf_layout_at_full_size(gpu, rmat):
    pa = gpu.prepare_bgrf_a(rmat)
    perm = pa.row_new_to_old()
    assert perm is not None
    pb = gpu.rebuild_b(pa, 0)
    _, st = gpu.tilespgemm(pa, pb, gpu.schedule(materialize_output=False))
    total, tiles = check_full(gpu, rmat, pa, np.asarray(perm, np.int64))
    assert st["output_nnz"] == total and st["output_tiles"] == tiles
    print(f"bgrf: nnz(C) {total}, C tiles {tiles}", flush=True)
