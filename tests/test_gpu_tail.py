"""GPU parity of the tail queue of the tile-pair path (adaptive-tail): the
heavy C tiles the device admits, their chunk counts and the gate decision
against the oracle's restatement of the reference planner
(orc_plan_tail_queue, pinned against dmma_plan_tile_tail_queue), and C --
with the heavy tiles split into chunks and reduced in order -- against the
oracle's product, the same bits on every call."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import as_tuple, check_c

pytestmark = pytest.mark.gpu


def merge_counts(pa, pb, tc):
    """Per C tile: scans (merge steps of A row I's and B column J's K lists)
    and matches, from the tile structures (the reference's counting,
    src/dmma_spgemm.h:2748-2761: scans = |A row| + |B col| - matches)."""
    at, bt = pa.tiles(), pb.tiles()
    a_ptr, a_k = at["tile_row_ptr"], at["tile_col_idx"]
    b_ptr, b_k = bt["tile_col_ptr"], bt["tile_row_idx"]
    scans, matches = [], []
    for i in range(tc.tilem):
        ka = a_k[a_ptr[i]:a_ptr[i + 1]]
        for t in range(tc.tile_ptr[i], tc.tile_ptr[i + 1]):
            j = tc.tile_columnidx[t]
            kb = b_k[b_ptr[j]:b_ptr[j + 1]]
            m = len(np.intersect1d(ka, kb, assume_unique=True))
            matches.append(m)
            scans.append(len(ka) + len(kb) - m)
    nnz = np.diff(tc.tile_nnz)
    return np.array(scans), np.array(matches), nnz


def dense_ish(rt, m, per_row, seed):
    return rt.gen_random(m, m, per_row, seed)


def run_tail(rt, a, aat, precision="fp64", **kw):
    pa = rt.prepare_identity_a(a)
    pb = rt.rebuild_b(pa, aat)
    sched = rt.schedule(numeric_path="tile", precision=precision, **kw)
    tc, st = rt.tilespgemm(pa, pb, sched)
    return pa, pb, tc, st, sched


@pytest.mark.parametrize("aat", [0, 1])
def test_forced_split_matches_planner_and_oracle(gpu, aat):
    a = dense_ish(gpu, 512, 180, 3 + aat)
    pa, pb, tc, st, sc = run_tail(gpu, a, aat, split_threshold=500.0, split_chunk_target=256.0,
                                  split_max_chunks=8, force_tail_split=1)
    scans, matches, nnz = merge_counts(pa, pb, tc)
    plan = oracle.plan_tail_queue(scans, matches, nnz, threshold=500.0, target=256.0,
                                  max_chunks=8, max_heavy_fraction=1.0,
                                  record_capacity=len(scans))
    assert plan["reason"] == 1 and st["tail_reason"] == 1
    rec = gpu.last_tail_records(pa)
    order = np.argsort(rec["task"])
    assert np.array_equal(rec["task"][order], plan["heavy_ids"])
    assert np.array_equal(rec["chunks"][order], plan["heavy_chunks"])
    assert np.array_equal(rec["scans"][order], scans[plan["heavy_ids"]])
    assert np.array_equal(rec["matches"][order], matches[plan["heavy_ids"]])
    assert st["tail_records"] == plan["heavy"] and st["heavy_chunks"] == plan["chunks"]
    ref = oracle.spgemm(as_tuple(a), as_tuple(gpu.csr_transpose(a) if aat else a))
    check_c(gpu, tc, ref)
    # chunk partials are summed in chunk order: the same bits again
    tc2, _ = gpu.tilespgemm(pa, pb, sc)
    assert np.array_equal(tc.tile_csr_Value.view(np.uint64), tc2.tile_csr_Value.view(np.uint64))


def test_gate_reasons(gpu):
    """The gate declines the split -- every tile then runs whole -- when no
    tile clears the threshold (4), when the admitted share passes
    max_heavy_fraction (6), or when the records outgrow their capacity (5);
    C is the oracle's in every case."""
    a = dense_ish(gpu, 256, 100, 9)
    ref = oracle.spgemm(as_tuple(a), as_tuple(a))
    pa, pb, tc, st, _ = run_tail(gpu, a, 0)                       # default threshold 16384
    assert st["tail_reason"] == 4 and st["heavy_chunks"] == 0
    check_c(gpu, tc, ref)
    pa, pb, tc, st, _ = run_tail(gpu, a, 0, split_threshold=100.0, split_chunk_target=64.0)
    scans, matches, nnz = merge_counts(pa, pb, tc)
    cap = int(np.floor(len(scans) * 0.01)) + 1
    plan = oracle.plan_tail_queue(scans, matches, nnz, threshold=100.0, target=64.0,
                                  record_capacity=cap)
    assert st["tail_reason"] == plan["reason"] and plan["reason"] in (5, 6)
    assert st["heavy_chunks"] == 0
    check_c(gpu, tc, ref)
    pa, pb, tc, st, _ = run_tail(gpu, a, 0, split_threshold=100.0, split_chunk_target=64.0,
                                 force_tail_split=1, tail_record_capacity=3)
    assert st["tail_reason"] == 5
    check_c(gpu, tc, ref)


def test_tail_split_in_split_precision(gpu):
    """The chunk kernel runs the split BF16 product too: the split tiles stay
    within the split form's bound on the block-FEM matrix."""
    from test_gpu_tile import split_bound_ok
    a = gpu.gen_fem_blocks(900, 6, True, 17)
    pa, pb, tc, st, _ = run_tail(gpu, a, 0, precision="split-bf16", split_threshold=40.0,
                                 split_chunk_target=16.0, force_tail_split=1)
    assert st["tail_reason"] == 1 and st["heavy_chunks"] > 0
    ref = oracle.spgemm(as_tuple(a), as_tuple(a))
    err = split_bound_ok(gpu, a, a, tc, ref)
    assert err["fwd_max"] <= 1e-12


def test_no_tail_outside_adaptive_tail(gpu):
    a = dense_ish(gpu, 256, 100, 12)
    _, _, tc, st, _ = run_tail(gpu, a, 0, lb_mode="dynamic", split_threshold=100.0,
                               force_tail_split=1)
    assert st["tail_reason"] == 0 and st["heavy_chunks"] == 0
    check_c(gpu, tc, oracle.spgemm(as_tuple(a), as_tuple(a)))
