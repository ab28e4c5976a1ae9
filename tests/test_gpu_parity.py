"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Structure (tile arrays, masks, offsets, nnz, CSR after tile2csr) must be
bit-exact.  FP64 values: |c - c*| / (|A||B|)_ij <= 1e-14 everywhere (the
forward-error bound of a 27..700-term FP64 sum is below that), and
|c - c*| / |c*| <= 1e-14 on inputs without cancellation (positive values);
c* is the oracle's double-double sum.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-14
REL_TOL = 1e-14


def as_tuple(a):
    return (a.m, a.n, a.rowptr, a.col, a.val)


def sub_rows(t, rows):
    m, n, rp, ci, v = t
    rp = np.asarray(rp, np.int64)
    rp2 = np.zeros(rows + 1, np.int64)
    rp2[:] = rp[:rows + 1]
    e = int(rp2[rows])
    return rows, n, rp2.astype(np.int32), ci[:e], v[:e]


def oracle_tiles_equal(got, want, csc):
    for k in ("tile_row_count", "tile_col_count", "num_tiles", "payload_size", "dense_tiles",
              "sparse_tiles"):
        assert got[k] == want[k], k
    for k in ("tile_row_ptr", "tile_col_idx", "value_offsets", "masks"):
        assert np.array_equal(got[k], want[k]), k
    np.testing.assert_array_equal(got["values"], want["values"])
    if csc:
        for k in ("tile_col_ptr", "tile_row_idx", "csc_tile_ids"):
            assert np.array_equal(got[k], want[k]), k


def check_a_tiles(rt, a, pa, dense_threshold=24):
    ar, ai, _ = pa.info()
    m, n, rp, ci, v = as_tuple(a)
    want = oracle.hybrid_tiles(*sub_rows((m, ai, rp, ci, v), ar), 8, 4, False, dense_threshold)
    oracle_tiles_equal(pa.tiles(), want, csc=False)


def check_b_tiles(rt, a, pa, pb, aat, dense_threshold=24):
    ar, ai, _ = pa.info()
    m, n, rp, ci, v = as_tuple(a)
    if aat:
        trp, tci, tv = oracle.transpose(m, n, rp, ci, v)
        src = (n, m, trp, tci, tv)
    else:
        src = (m, n, rp, ci, v)
    want = oracle.hybrid_tiles(*sub_rows(src, ai), 4, 8, True, dense_threshold, build_csc=True)
    oracle_tiles_equal(pb.tiles(), want, csc=True)


def oracle_product(a, aat):
    A = as_tuple(a)
    if aat:
        trp, tci, tv = oracle.transpose(*A)
        B = (a.n, a.m, trp, tci, tv)
    else:
        B = A
    return oracle.spgemm(A, B)


def check_c(rt, tc, ref, positive=False):
    ct = oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], ref["val"])
    assert tc.numtile == ct["num_tiles"]
    assert tc.nnz == ref["nnz"]
    # C has A's tile-row count, trimmed to the active row prefix
    # (dmma_copy_output_to_host, src/dmma_spgemm.h:6409-6414); no oracle tile
    # lies beyond it.
    assert np.array_equal(tc.tile_ptr, ct["tile_ptr"][:tc.tilem + 1])
    assert np.all(ct["tile_ptr"][tc.tilem:] == ct["num_tiles"])
    assert np.array_equal(tc.tile_columnidx, ct["tile_col"])
    assert np.array_equal(tc.tile_nnz, ct["tile_nnz"])
    assert np.array_equal(tc.tile_csr_Ptr, ct["tile_csr_ptr"])
    assert np.array_equal(tc.tile_csr_Col, ct["tile_csr_col"])
    # values are stored in the same order as the oracle's tile order
    hi = oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], ref["hi"])["tile_csr_val"]
    lo = oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], ref["lo"])["tile_csr_val"]
    ab = oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"],
                        ref["absprod"])["tile_csr_val"]
    err = oracle.value_errors(tc.tile_csr_Value, hi, lo, ab)
    assert err["fwd_max"] <= FWD_TOL, err
    if positive:
        assert err["rel_max"] <= REL_TOL, err
    # CSR after tile2csr equals the oracle CSR.
    csr = rt.tile2csr(tc, canonical=True)
    assert np.array_equal(csr.rowptr.astype(np.int64), ref["rowptr"])
    assert np.array_equal(csr.col, ref["col"])
    return err


def run(rt, a, aat, positive=False, dense_threshold=24, sched=None):
    pa = rt.prepare_identity_a(a, dense_threshold)
    check_a_tiles(rt, a, pa, dense_threshold)
    pb = rt.rebuild_b(pa, aat)
    check_b_tiles(rt, a, pa, pb, aat, dense_threshold)
    tc, stats = rt.tilespgemm(pa, pb, sched)
    ref = oracle_product(a, aat)
    err = check_c(rt, tc, ref, positive)
    assert stats["output_nnz"] == ref["nnz"]
    assert rt.nnz_cub(pa, aat) == oracle.nnz_cub(a.m, a.n, a.rowptr, a.col, aat)
    return pa, pb, tc, stats, err


def csr_from(m, n, rows, cols, vals):
    import rtt_amd
    order = np.lexsort((np.arange(len(rows)), rows))
    rows, cols, vals = np.asarray(rows)[order], np.asarray(cols)[order], np.asarray(vals)[order]
    rp = np.zeros(m + 1, np.int64)
    np.add.at(rp, np.asarray(rows, np.int64) + 1, 1)
    return rtt_amd.CsrMatrix(m, n, np.cumsum(rp).astype(np.int32), cols.astype(np.int32),
                             vals.astype(np.float64))


# ---------------------------------------------------------------------------

@pytest.mark.parametrize("aat", [0, 1])
def test_laplace_64(gpu, aat):
    a = gpu.gen_laplace2d(64, 0.1, 20260101)
    _, _, tc, _, err = run(gpu, a, aat)
    assert tc.nnz == 51972


@pytest.mark.parametrize("aat", [0, 1])
@pytest.mark.parametrize("g", [5, 12])
def test_stencil27(gpu, aat, g):
    a = gpu.gen_stencil27(g, 0.5, 1.5, 3)
    _, _, tc, _, _ = run(gpu, a, aat, positive=True)
    if aat:
        assert tc.nnz == (5 * g - 6) ** 3


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("aat", [0, 1])
def test_random(gpu, seed, aat):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(50, 700))
    n = m if not aat else int(rng.integers(30, 700))
    a = gpu.gen_random(m, n, int(rng.integers(1, 12)), seed)
    run(gpu, a, aat)


@pytest.mark.parametrize("aat", [0, 1])
def test_unsorted_duplicates_explicit_zeros(gpu, aat):
    rng = np.random.default_rng(42)
    m = n = 203
    rows, cols, vals = [], [], []
    for i in range(m):
        cs = list(rng.choice(n, size=int(rng.integers(0, 9)), replace=False))
        cs += cs[:2]          # duplicates (including a zero-sum pair below)
        rng.shuffle(cs)       # unsorted rows
        for c in cs:
            rows.append(i)
            cols.append(int(c))
            vals.append(0.0 if rng.random() < 0.1 else float(rng.uniform(-1, 1)))
    rows += [7, 7]
    cols += [3, 3]
    vals += [1.25, -1.25]     # zero-sum duplicate keeps its structural slot
    a = csr_from(m, n, rows, cols, vals)
    order = np.argsort(rng.random(a.nnz))  # shuffle within rows again
    for i in range(m):
        s, e = a.rowptr[i], a.rowptr[i + 1]
        idx = np.arange(s, e)[np.argsort(order[s:e])]
        a.col[s:e], a.val[s:e] = a.col[idx].copy(), a.val[idx].copy()
    pa, pb, tc, stats, _ = run(gpu, a, aat)
    assert pa.stats["rows_were_sorted"] == 0


@pytest.mark.parametrize("aat", [0, 1])
def test_dense_tiles_and_threshold(gpu, aat):
    # 4x4 dense blocks (block-FEM-like): fully dense 8x4 / 4x8 tiles.
    rng = np.random.default_rng(7)
    nb = 40
    rows, cols, vals = [], [], []
    for bi in range(nb):
        for bj in {bi, (bi + 1) % nb, int(rng.integers(nb))}:
            for r in range(4):
                for c in range(4):
                    rows.append(4 * bi + r)
                    cols.append(4 * bj + c)
                    vals.append(float(rng.uniform(0.5, 1.5)))
    a = csr_from(4 * nb, 4 * nb, rows, cols, vals)
    for thr in (24, 1, 32):
        pa, pb, tc, _, _ = run(gpu, a, aat, positive=True, dense_threshold=thr)
        t = pa.tiles()
        assert t["dense_tiles"] > 0 if thr <= 16 else True


# LDS budgets of the VECTOR numeric instances, per C tile row: C tiles, C
# values, A values (the one-wave TinyRow, the workgroup SmallRow, MediumRow
# and LargeRow, first fit; larger rows go to the slab kernels).
ROW_TIERS = [("tiny", 16, 128, 64), ("small", 128, 1024, 256), ("medium", 64, 2048, 256),
             ("large", 512, 4096, 2048)]


def row_tiers(a, b):
    """Tile rows of A*B per numeric instance, from the scipy product."""
    A, B = a.to_scipy().tocsr(), b.tocsr()
    C = (A @ B).tocoo()
    T = (A.shape[0] + 7) // 8
    w = C.shape[1] // 8 + 1
    keys = np.unique((C.row // 8).astype(np.int64) * w + C.col // 8)
    nt = np.bincount(keys // w, minlength=T)
    nz = np.bincount(C.row // 8, minlength=T)
    al = np.bincount(A.tocoo().row // 8, minlength=T)
    out = {}
    for r in range(T):
        name = next((t for t, mt, mz, ma in ROW_TIERS
                     if nt[r] <= mt and nz[r] <= mz and al[r] <= ma), "slab")
        out[name] = out.get(name, 0) + 1
    return out


@pytest.mark.parametrize("aat", [0, 1])
def test_vector_row_tiers_in_one_call(gpu, aat):
    # Five bands of 2048 rows each, with 2, 6, 12, 16 and 25 entries per
    # row within +-8, 24, 40, 48 and 150 of the diagonal: C tile rows of
    # every LDS instance (the one-wave TinyRow, SmallRow, MediumRow,
    # LargeRow) in the same VECTOR call, against the oracle.
    rng = np.random.default_rng(7)
    n = 10240
    bands = [(2, 8), (6, 24), (12, 40), (16, 48), (25, 150)]
    rows, cols = [], []
    for i in range(n):
        k, w = bands[i // 2048]
        c = np.union1d(np.clip(i + rng.integers(-w, w + 1, k), 0, n - 1), [i])
        rows += [i] * len(c)
        cols += list(c)
    a = csr_from(n, n, rows, cols, rng.uniform(0.5, 1.5, len(rows)))
    s = a.to_scipy().tocsr()
    tiers = row_tiers(a, s.T if aat else s)
    assert all(tiers.get(t[0], 0) > 0 for t in ROW_TIERS), tiers
    run(gpu, a, aat, positive=True, sched=gpu.schedule(numeric_path="vector"))


@pytest.mark.parametrize("aat", [0, 1])
def test_wide_rows_take_the_slab_paths(gpu, aat):
    # An arrow matrix: dense first rows/columns give C tile rows with more
    # tiles than the wave hash holds and more nonzeros than the LDS
    # accumulator, exercising both workgroup slab kernels.
    n = 6000
    rows, cols, vals = [], [], []
    rng = np.random.default_rng(1)
    for i in range(n):
        rows.append(i)
        cols.append(i)
        vals.append(2.0)
        if i < 3:
            for j in range(0, n, 2):
                rows.append(i)
                cols.append(j)
                vals.append(float(rng.uniform(0.5, 1.5)))
        else:
            rows.append(i)
            cols.append(i % 3)
            vals.append(float(rng.uniform(0.5, 1.5)))
    a = csr_from(n, n, rows, cols, vals)
    # merge duplicate (i, i) entries for rows < 3 are fine: duplicates allowed
    _, _, tc, stats, _ = run(gpu, a, aat, positive=True,
                             sched=gpu.schedule(numeric_path="vector"))
    assert stats["heavy_rows"] > 0


@pytest.mark.parametrize("aat", [0, 1])
def test_slab_row_beyond_the_window_walk(gpu, aat):
    # The first eight rows alternate over 65536 columns, so the first tile
    # row holds 16384 A tiles: more than the slab numeric pass walks in LDS
    # windows, so its products go to C as HBM atomics. The other rows hold
    # two entries each and stay small. In A*A^T the first rows meet over
    # 32768 columns; values from {0.5, 1, 1.5} keep every such sum exact in
    # FP64 whatever its order, so the check stays at the usual tolerance and
    # a lost or repeated product still shows.
    n = 65536
    rng = np.random.default_rng(5)
    rows, cols = [], []
    for r in range(8):
        c = np.arange(r % 2, n, 2)
        rows.append(np.full(len(c), r))
        cols.append(c)
    i = np.arange(8, n)
    rows += [i, i]
    cols += [i, (7 * i) % n]
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    keep = np.unique(rows.astype(np.int64) * n + cols, return_index=True)[1]
    rows, cols = rows[keep], cols[keep]
    a = csr_from(n, n, rows, cols, rng.choice([0.5, 1.0, 1.5], len(rows)))
    # The first tile row's A tiles hold 16 values each, which AUTO would send
    # to the tile-pair kernel: the LDS-accumulator path is what is tested here.
    pa, _, _, stats, _ = run(gpu, a, aat, positive=True,
                             sched=gpu.schedule(numeric_path="vector"))
    assert stats["heavy_rows"] > 0
    assert pa.tiles()["tile_row_ptr"][1] > 8192


@pytest.mark.parametrize("scale,aat", [(12, 0), (14, 0), (12, 1)])
def test_rmat_power_law(gpu, scale, aat):
    # R-MAT rows follow a power law: hub rows with thousands of entries give
    # C tile rows wider than the LDS tables, so many rows take the slab
    # paths, and B rows mix heavy and light tiles of every class.
    a = gpu.gen_rmat(scale, 16, 0.57, 0.19, 0.19, 20260101)
    _, _, _, stats, _ = run(gpu, a, aat, sched=gpu.schedule(numeric_path="vector"))
    assert stats["heavy_rows"] > 0
    # AUTO (hub tile rows dense enough take the tile-pair kernel) gives the
    # oracle's C as well
    run(gpu, a, aat)


def test_repeat_calls_are_deterministic_on_slab_rows(gpu):
    # Slab rows share per-workgroup slabs across the rows a workgroup takes;
    # the structure of C must not depend on which rows those are.
    a = gpu.gen_rmat(14, 16, 0.57, 0.19, 0.19, 7)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 0)
    vec = gpu.schedule(numeric_path="vector")
    tc1, s1 = gpu.tilespgemm(pa, pb, vec)
    pb = gpu.rebuild_b(pa, 0, pb)
    tc2, s2 = gpu.tilespgemm(pa, pb, vec)
    assert s1["heavy_rows"] > 0
    assert (s1["output_tiles"], s1["output_nnz"]) == (s2["output_tiles"], s2["output_nnz"])
    for k in ("tile_ptr", "tile_columnidx", "tile_nnz", "tile_csr_Ptr", "tile_csr_Col"):
        assert np.array_equal(getattr(tc1, k), getattr(tc2, k)), k
    # FP64 sums in another order: equal to rounding.
    np.testing.assert_allclose(tc1.tile_csr_Value, tc2.tile_csr_Value, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("lb_mode", ["static", "dynamic", "adaptive-tail"])
def test_lb_modes_agree_on_slab_rows(gpu, lb_mode):
    """--lb-mode changes only which workgroup takes which slab row (round
    robin, a work queue in list order, or the queue heaviest row first):
    C's structure is identical and its values match the oracle."""
    a = gpu.gen_rmat(13, 16, 0.57, 0.19, 0.19, 11)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 0)
    ref = oracle_product(a, 0)
    # The first call on a handle learns how much the slab rows' tile lists
    # need; later calls stage them between the symbolic passes.
    for _ in range(2):
        tc, stats = gpu.tilespgemm(pa, pb, gpu.schedule(lb_mode=lb_mode, numeric_path="vector"))
        assert stats["heavy_rows"] > 0
        check_c(gpu, tc, ref)


def test_numeric_timeline_trace(gpu, tmp_path, monkeypatch):
    """DMMA_TRACE_FILE turns on the sampled timeline of the numeric kernel:
    the file has the DMTL0001 header, one slot per wave of every 2^shift-th
    workgroup from the phase on, each with end >= start > 0; the product
    is unchanged, and without the variable no file is written."""
    a = gpu.gen_stencil27(12, 0.5, 1.5, 3)
    path = tmp_path / "trace.bin"
    monkeypatch.setenv("DMMA_TRACE_FILE", str(path))
    monkeypatch.setenv("DMMA_TRACE_SHIFT", "2")
    monkeypatch.setenv("DMMA_TRACE_PHASE", "1")
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 1)
    tc, _ = gpu.tilespgemm(pa, pb)
    t = gpu.read_timeline(str(path))
    R = (a.m + 7) // 8
    assert (t["version"], t["waves_per_workgroup"], t["shift"], t["phase"]) == (1, 4, 2, 1)
    assert t["workgroups"] == R and t["output_tiles"] == tc.numtile
    assert t["slots"] == ((R - 1 - 1) // 4 + 1) * 4
    assert np.all(t["start_ns"] > 0) and np.all(t["end_ns"] >= t["start_ns"])
    assert t["compute_units"] > 0
    check_c(gpu, tc, oracle_product(a, 1), positive=True)
    monkeypatch.delenv("DMMA_TRACE_FILE")
    path.unlink()
    gpu.tilespgemm(pa, pb)
    assert not path.exists()


def test_rectangular_aat_and_empty_shapes(gpu):
    a = gpu.gen_random(37, 300, 5, 9)
    run(gpu, a, 1)
    # rows beyond the active prefix and empty trailing columns
    a = csr_from(64, 80, [0, 3, 17], [2, 40, 5], [1.0, 2.0, 3.0])
    run(gpu, a, 1)
    run(gpu, csr_from(80, 80, [0, 3, 17], [2, 40, 5], [1.0, 2.0, 3.0]), 0)
    # nonzero shape, no entries; and 0 x 0
    for m, n in ((16, 16), (0, 0)):
        e = csr_from(m, n, [], [], [])
        pa = gpu.prepare_identity_a(e)
        pb = gpu.rebuild_b(pa, 1)
        tc, stats = gpu.tilespgemm(pa, pb)
        assert tc.numtile == 0 and tc.nnz == 0


def test_external_orders_restore_rows(gpu):
    rng = np.random.default_rng(17)
    a = gpu.gen_random(300, 300, 6, 5)
    rp_ = rng.permutation(300).astype(np.int32)
    ip_ = rng.permutation(300).astype(np.int32)
    for aat in (0, 1):
        pa = gpu.prepare_external_a(a, rp_, ip_)
        pb = gpu.rebuild_b(pa, aat)
        tc, _ = gpu.tilespgemm(pa, pb)
        ar, ai, ident = pa.info()
        assert not ident
        csr = gpu.tile2csr_restore_rows(tc, pa.row_new_to_old(), ar)
        ref = oracle_product(a, aat)
        assert np.array_equal(csr.rowptr.astype(np.int64), ref["rowptr"])
        assert np.array_equal(csr.col, ref["col"])
        err = oracle.value_errors(csr.val, ref["hi"], ref["lo"], ref["absprod"])
        assert err["fwd_max"] <= FWD_TOL, err


@pytest.mark.parametrize("kind", ["random", "rmat12"])
def test_bgrf_layout_restores_to_the_oracle_product(gpu, kind):
    """The BGRF-v1 layout changes only where rows and K indices sit: after
    tile2csr_restore_rows, C equals the oracle's in both modes, and A's tile
    count is the plan's."""
    a = (gpu.gen_random(400, 400, 7, 13) if kind == "random"
         else gpu.gen_rmat(12, 16, 0.57, 0.19, 0.19, 8))
    for aat in (0, 1):
        pa = gpu.prepare_bgrf_a(a)
        assert pa.tiles()["num_tiles"] == pa.bgrf["tiles_final"]
        assert pa.bgrf["tiles_final"] <= pa.bgrf["tiles_identity"]
        pb = gpu.rebuild_b(pa, aat)
        tc, _ = gpu.tilespgemm(pa, pb)
        ar, _, _ = pa.info()
        perm = pa.row_new_to_old()
        csr = gpu.tile2csr(tc) if perm is None else gpu.tile2csr_restore_rows(tc, perm, ar)
        ref = oracle_product(a, aat)
        assert np.array_equal(csr.rowptr.astype(np.int64), ref["rowptr"])
        assert np.array_equal(csr.col, ref["col"])
        err = oracle.value_errors(csr.val, ref["hi"], ref["lo"], ref["absprod"])
        assert err["fwd_max"] <= FWD_TOL, err


def test_design_doc_tile_example_through_external_order(gpu):
    # 16x8 A with entries (0,0)=a, (1,4)=b, (8,1)=c, (9,5)=d and the row
    # permutation 0->0, 8->1, 1->2, 9->3 (docs/...CN.md section 10).
    a_, b_, c_, d_ = 1.5, 2.5, 3.5, 4.5
    a = csr_from(16, 8, [0, 1, 8, 9], [0, 4, 1, 5], [a_, b_, c_, d_])
    row_new_to_old = np.array([0, 8, 1, 9] + [r for r in range(16) if r not in (0, 8, 1, 9)],
                              np.int32)
    pa = gpu.prepare_external_a(a, row_new_to_old, np.arange(8, dtype=np.int32))
    t = pa.tiles()
    assert pa.info()[0] == 4
    assert list(t["tile_row_ptr"]) == [0, 2]
    assert list(t["tile_col_idx"]) == [0, 1]
    assert list(t["value_offsets"]) == [0, 2, 4]
    assert [int(x) for x in t["masks"]] == [0x21, 0x2100]
    assert list(t["values"]) == [a_, c_, b_, d_]


@pytest.mark.parametrize("aat", [0, 1])
def test_rocsparse_control_agrees_with_oracle(gpu, aat):
    """The vendor control computes the same product: rocSPARSE's nnz(C)
    equals the oracle's and the sum of its values matches the oracle's
    double-double sum (positive inputs, so no cancellation) within 1e-12."""
    a = gpu.gen_stencil27(12, 0.5, 1.5, 3)
    ref = oracle_product(a, aat)
    r = gpu.control_rocsparse(a, aat, warmup=1, iterations=2)
    assert r["nnz_c"] == ref["nnz"]
    want = float(np.sum(ref["hi"]) + np.sum(ref["lo"]))
    assert abs(r["checksum"] - want) <= 1e-12 * abs(want)
    assert r["ms"] > 0


def test_repeat_calls_reuse_workspace(gpu):
    a = gpu.gen_stencil27(10, 0.5, 1.5, 2)
    pa = gpu.prepare_identity_a(a)
    pb = None
    outs = []
    for _ in range(3):
        pb = gpu.rebuild_b(pa, 1, pb)
        tc, _ = gpu.tilespgemm(pa, pb)
        outs.append(tc)
    for tc in outs[1:]:
        assert np.array_equal(tc.tile_columnidx, outs[0].tile_columnidx)
        np.testing.assert_allclose(tc.tile_csr_Value, outs[0].tile_csr_Value, rtol=1e-14)


def test_full_size_stencil27_128_aat_properties(gpu):
    """The benchmark workload at full size (2,097,152 rows, 55.7M nnz):
    closed-form nnz(C) and nnzCub, a row-sum (linearity) check of all values
    and an oracle comparison on a block of rows."""
    g = 128
    a = gpu.gen_stencil27(g, 0.5, 1.5, 20260101)
    pa = gpu.prepare_identity_a(a)
    pb = gpu.rebuild_b(pa, 1)
    assert gpu.nnz_cub(pa, 1) == 1142 ** 3
    tc, stats = gpu.tilespgemm(pa, pb)
    assert tc.nnz == 634 ** 3 == stats["output_nnz"]
    c = gpu.last_c(pa)
    # Row sums of C = A (A^T 1), computed independently in float64.
    s = a.to_scipy()
    want = s @ (s.T @ np.ones(a.m))
    csr = gpu.tile2csr(tc, canonical=True)
    got = np.add.reduceat(csr.val, csr.rowptr[:-1]) if csr.nnz else np.zeros(a.m)
    np.testing.assert_allclose(got, want, rtol=1e-13)
    # Oracle on a block of 4096 rows in the middle of the grid.
    r0, r1 = 1_000_000, 1_004_096
    sub = (r1 - r0, a.n, (a.rowptr[r0:r1 + 1] - a.rowptr[r0]).astype(np.int32),
           a.col[a.rowptr[r0]:a.rowptr[r1]], a.val[a.rowptr[r0]:a.rowptr[r1]])
    trp, tci, tv = oracle.transpose(a.m, a.n, a.rowptr, a.col, a.val)
    ref = oracle.spgemm(sub, (a.n, a.m, trp, tci, tv))
    assert np.array_equal(csr.rowptr[r0:r1 + 1] - csr.rowptr[r0], ref["rowptr"])
    seg = slice(int(csr.rowptr[r0]), int(csr.rowptr[r1]))
    assert np.array_equal(csr.col[seg], ref["col"])
    err = oracle.value_errors(csr.val[seg], ref["hi"], ref["lo"], ref["absprod"])
    assert err["rel_max"] <= REL_TOL, err
