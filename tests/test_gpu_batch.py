"""GPU parity of the batched form: many products in one call over the
block-diagonal concatenation (every problem 8-aligned, so no tile mixes two
problems), each problem's C extracted and compared with the oracle's product
of that problem alone; the autotuner's choice gives the same C."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import as_tuple, check_c

pytestmark = pytest.mark.gpu


def banded_batch(rt, problems, median, seed):
    rows, bw, pr = rt.gen_banded_batch_shapes(problems, median_rows=median, sigma=0.8,
                                              min_rows=20, max_rows=4 * median, seed=seed)
    return [rt.gen_banded(int(rows[p]), int(bw[p]), int(pr[p]), seed + p)
            for p in range(problems)]


@pytest.mark.parametrize("path", ["auto", "vector", "tile"])
def test_banded_batch_in_one_call(gpu, path):
    mats = banded_batch(gpu, 12, 300, 7)
    big, ro, co = gpu.block_diag(mats, 8)
    pa = gpu.prepare_identity_a(big)
    pb = gpu.rebuild_b(pa, 0)
    _, st = gpu.tilespgemm(pa, pb, gpu.schedule(materialize_output=False, numeric_path=path))
    assert gpu.nnz_cub(pa, 0) == sum(
        oracle.nnz_cub(m.m, m.n, m.rowptr, m.col, False) for m in mats)
    total = 0
    for p, m in enumerate(mats):
        tc = gpu.batch_extract(pa, ro[p], co[p], m.m, m.n)
        ref = oracle.spgemm(as_tuple(m), as_tuple(m))
        check_c(gpu, tc, ref)
        total += ref["nnz"]
    assert st["output_nnz"] == total


def test_amg_galerkin_batch(gpu):
    """R*(A*P) for three hierarchies in two batched calls (A*P, then R times
    it), every problem against the oracle's chain."""
    gs = [6, 8, 9]
    A = [gpu.gen_poisson7_jumps(g, 3.0, 40 + g) for g in gs]
    P = [gpu.gen_aggregation_prolongator(a, g, 2.0 / 3.0) for a, g in zip(A, gs)]
    R = [gpu.csr_transpose(p) for p in P]
    bA, roA, coA = gpu.block_diag(A, 8)
    bP, roP, coP = gpu.block_diag(P, 8)
    bR, roR, coR = gpu.block_diag(R, 8)
    assert np.array_equal(coA, roP) and np.array_equal(coP, roR)
    pa = gpu.prepare_identity_a(bA)
    tc, _ = gpu.tilespgemm(pa, gpu.rebuild_b_general(pa, gpu.upload_csr(bP)))
    ap = gpu.tile2csr(tc, canonical=True)
    pr = gpu.prepare_identity_a(bR)
    gpu.tilespgemm(pr, gpu.rebuild_b_general(pr, gpu.upload_csr(ap)),
                   gpu.schedule(materialize_output=False))
    for p in range(len(gs)):
        # Exact structure against the oracle's chain; values against
        # R (A P) in float64 within 1e-13 of (|R||A||P|)_ij, the forward
        # bound of two FP64 products (the GPU's A*P feeds the second one).
        ap_ref = oracle.spgemm(as_tuple(A[p]), as_tuple(P[p]))
        ap_t = (ap_ref["m"], ap_ref["n"], ap_ref["rowptr"].astype(np.int32), ap_ref["col"],
                ap_ref["val"])
        rap_ref = oracle.spgemm(as_tuple(R[p]), ap_t)
        tc_p = gpu.batch_extract(pr, roR[p], coP[p], R[p].m, P[p].n)
        ct = oracle.c_tiles(rap_ref["m"], rap_ref["n"], rap_ref["rowptr"], rap_ref["col"],
                            rap_ref["val"])
        assert np.array_equal(tc_p.tile_columnidx, ct["tile_col"])
        assert np.array_equal(tc_p.tile_nnz, ct["tile_nnz"])
        got = gpu.tile2csr(tc_p, canonical=True).to_scipy()
        As, Ps, Rs = A[p].to_scipy(), P[p].to_scipy(), R[p].to_scipy()
        want = (Rs @ (As @ Ps)).tocsr()
        scale = (abs(Rs) @ abs(As) @ abs(Ps)).tocsr()
        diff = abs(got - want).tocsr()
        bound = scale.multiply(1e-13).tocsr()
        assert (diff - bound).max() <= 0.0


def test_autotuner_choice_gives_the_same_product(gpu):
    mats = banded_batch(gpu, 8, 400, 11)
    big, ro, co = gpu.block_diag(mats, 8)
    pa = gpu.prepare_identity_a(big)
    pb = gpu.rebuild_b(pa, 0)
    sched, table, chosen = gpu.autotune(pa, pb, repeats=1)
    assert len(table) >= 5 and 0 <= chosen < len(table)
    assert all(t["ms"] > 0 for t in table)
    assert table[chosen]["ms"] == min(t["ms"] for t in table)
    sched.materialize_output = 1
    tc, _ = gpu.tilespgemm(pa, pb, sched)
    check_c(gpu, tc, oracle.spgemm(as_tuple(big), as_tuple(big)))


def test_batch_extract_rejects_unaligned_blocks(gpu):
    a = gpu.gen_random(40, 40, 3, 5)
    pa = gpu.prepare_identity_a(a)
    gpu.tilespgemm(pa, gpu.rebuild_b(pa, 0))
    with pytest.raises(gpu.RttError):
        gpu.batch_extract(pa, 3, 0, 10, 10)
    with pytest.raises(gpu.RttError):
        gpu.batch_extract(pa, 0, 0, 41, 40)
