"""CPU tests of the shipped library's host side: the C ABI loads and exports
every declared symbol, and the host functions (Matrix Market reader,
tile2csr, order files, generators) agree with the oracle.  No GPU call."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import rtt_amd


def test_library_exports_every_header_symbol():
    lib = rtt_amd.load()
    names = rtt_amd.abi_symbols()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", rtt_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_control_library_exports():
    """The rocSPARSE control library is built beside the SpGEMM library and
    exports its two entry points (loading it needs no GPU)."""
    assert os.path.exists(rtt_amd.CONTROL_PATH)
    out = subprocess.run(["nm", "-D", "--defined-only", rtt_amd.CONTROL_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert {"rtt_control_rocsparse", "rtt_control_last_error"} <= exported


def test_smatrix_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "rtt_spgemm.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(rtt_smatrix),'
        ' offsetof(rtt_smatrix, value), offsetof(rtt_smatrix, tile_ptr),'
        ' offsetof(rtt_smatrix, numtile), offsetof(rtt_smatrix, csc_tile_rowidx),'
        ' sizeof(rtt_stats));return 0;}\n')
    exe = tmp_path / "layout"
    inc = os.path.dirname(rtt_amd.HEADER_PATH)
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    S = rtt_amd.SMatrix
    assert got == [C.sizeof(S), S.value.offset, S.tile_ptr.offset, S.numtile.offset,
                   S.csc_tile_rowidx.offset, C.sizeof(rtt_amd.Stats)]


def write(path, text):
    with open(path, "w") as f:
        f.write(text)


@pytest.mark.parametrize("kind", ["general", "symmetric", "pattern"])
def test_read_mtx_matches_oracle(tmp_path, kind):
    rng = np.random.default_rng(4)
    p = str(tmp_path / "a.mtx")
    ent = []
    for _ in range(200):
        i, j = int(rng.integers(40)), int(rng.integers(40))
        if kind == "symmetric" and j > i:
            i, j = j, i
        ent.append((i, j, rng.uniform(-1, 1)))
    ent += ent[:7]
    field = "pattern" if kind == "pattern" else "real"
    sym = "symmetric" if kind == "symmetric" else "general"
    lines = [f"%%MatrixMarket matrix coordinate {field} {sym}", "% c", f"40 40 {len(ent)}"]
    lines += [f"{i + 1} {j + 1}" if field == "pattern" else f"{i + 1} {j + 1} {v!r}"
              for i, j, v in ent]
    write(p, "\n".join(lines) + "\n")
    a = rtt_amd.read_mtx(p)
    m, n, rp, ci, v = oracle.mmio_read(p)
    assert (a.m, a.n) == (m, n)
    assert np.array_equal(a.rowptr, rp) and np.array_equal(a.col, ci)
    assert np.array_equal(a.val, v)
    r = rtt_amd.read_mtx(p, reference_values=True)
    assert np.array_equal(r.val, np.arange(r.nnz) % 10)


def test_write_then_read_roundtrip(tmp_path):
    a = rtt_amd.gen_laplace2d(16, 0.1, 3)
    p = str(tmp_path / "l.mtx")
    rtt_amd.write_mtx(p, a)
    b = rtt_amd.read_mtx(p)
    assert np.array_equal(a.rowptr, b.rowptr) and np.array_equal(a.col, b.col)
    assert np.array_equal(a.val, b.val)


def test_generators_closed_forms_and_determinism():
    a = rtt_amd.gen_laplace2d(64, 0.1, 20260101)
    assert a.nnz == 20224 and a.m == 4096
    diag = a.val[np.repeat(np.arange(a.m), np.diff(a.rowptr)) == a.col]
    assert np.all(diag == 4.0)
    off = a.val[np.repeat(np.arange(a.m), np.diff(a.rowptr)) != a.col]
    assert off.min() >= -1.1 and off.max() <= -0.9
    s = rtt_amd.gen_stencil27(16, 0.5, 1.5, 1)
    assert s.nnz == 46 ** 3 and s.val.min() >= 0.5 and s.val.max() < 1.5
    for x in (a, s):
        for i in range(x.m):
            assert np.all(np.diff(x.col[x.rowptr[i]:x.rowptr[i + 1]]) > 0)
    s2 = rtt_amd.gen_stencil27(16, 0.5, 1.5, 1)
    assert np.array_equal(s.val, s2.val)
    r = rtt_amd.gen_rmat(10, 16, seed=5)
    r2 = rtt_amd.gen_rmat(10, 16, seed=5)
    assert r.m == 1024 and 0 < r.nnz <= 16 * 1024
    assert np.array_equal(r.col, r2.col) and np.array_equal(r.val, r2.val)
    deg = np.diff(r.rowptr)
    assert deg.max() > 8 * deg.mean()   # power-law skew


def test_tile2csr_matches_oracle_including_malformed_bounds():
    rng = np.random.default_rng(8)
    m, n = 50, 61
    tilem = (m + 7) // 8
    tiles_per_row = rng.integers(0, 5, tilem)
    tile_ptr = np.concatenate([[0], np.cumsum(tiles_per_row)]).astype(np.int32)
    T = int(tile_ptr[-1])
    cols, nnzs, ptrs, lcols = [], [0], [], []
    for tr in range(tilem):
        picks = np.sort(rng.choice((n + 7) // 8, size=tiles_per_row[tr], replace=False))
        for J in picks:
            cols.append(int(J))
            k = int(rng.integers(0, 20))
            nnzs.append(nnzs[-1] + k)
            ptrs += list(np.sort(rng.integers(0, k + 3, 8)))   # may exceed the tile
            lcols += list(rng.integers(0, 9, k))               # 8 is out of range
    tc = rtt_amd.TileC(m, n, tilem, (n + 7) // 8, T, nnzs[-1], tile_ptr,
                       np.array(cols, np.int32), np.array(nnzs, np.int32),
                       np.array(ptrs, np.uint8), np.array(lcols, np.uint8),
                       rng.uniform(-1, 1, nnzs[-1]))
    for canonical in (True, False):
        got = rtt_amd.tile2csr(tc, canonical)
        want = oracle.tile2csr(m, n, tilem, tc.tile_ptr, tc.tile_columnidx, tc.tile_nnz,
                               tc.tile_csr_Ptr, tc.tile_csr_Col, tc.tile_csr_Value,
                               canonical=canonical)
        assert np.array_equal(got.rowptr.astype(np.int64), want[0])
        if canonical:
            assert np.array_equal(got.col, want[1]) and np.array_equal(got.val, want[2])
        else:
            for i in range(m):
                s, e = got.rowptr[i], got.rowptr[i + 1]
                assert np.array_equal(got.col[s:e], want[1][s:e])
    perm = rng.permutation(m).astype(np.int32)
    got = rtt_amd.tile2csr_restore_rows(tc, perm, m)
    want = oracle.tile2csr(m, n, tilem, tc.tile_ptr, tc.tile_columnidx, tc.tile_nnz,
                           tc.tile_csr_Ptr, tc.tile_csr_Col, tc.tile_csr_Value,
                           row_new_to_old=perm, permutation_size=m)
    assert np.array_equal(got.rowptr.astype(np.int64), want[0])
    assert np.array_equal(got.col, want[1])
    bad = perm.copy()
    bad[3] = bad[4]
    with pytest.raises(rtt_amd.RttError):
        rtt_amd.tile2csr_restore_rows(tc, bad, m)


def test_order_file_parser(tmp_path):
    p = str(tmp_path / "o.txt")
    write(p, "2 0\n1 3\n")
    assert list(rtt_amd.read_order_file(p, 4)) == [2, 0, 1, 3]
    for text in ("2 0 1", "2 0 1 3 4", "2 0 1 1", "2 0 1 7", "2 0 x 1", "-1 0 1 2"):
        write(p, text)
        with pytest.raises(rtt_amd.RttError):
            rtt_amd.read_order_file(p, 4)


def test_gpu_entry_points_fail_loudly_without_init():
    a = rtt_amd.gen_laplace2d(8, 0.1, 1)
    with pytest.raises(rtt_amd.RttError):
        rtt_amd.prepare_identity_a(a)


def test_b_values_clear_policy_matches_reference():
    """rtt_choose_b_values_clear against the reference's own
    dmma_choose_b_values_clear (src/dmma_b_values_clear_policy.h:91-160),
    compiled from its header into oracle/_ref: every policy value (valid and
    invalid) times every combination of the eight preconditions, and every
    reason and policy name."""
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref/libref.so was not built (no reference tree)")
    import itertools
    lib = rtt_amd.load()
    for policy in (-1, 0, 1, 2, 3):
        for flags in itertools.product((0, 1), repeat=8):
            got = rtt_amd.choose_b_values_clear(policy, *flags)
            want = oracle.ref_choose_b_values_clear(policy, *flags)
            assert (got["clear_payload"], got["skip_clear"], got["fallback"],
                    got["reason"]) == want, (policy, flags)
    ref = oracle.ref_lib()
    for reason in range(-1, 13):
        assert lib.rtt_b_values_clear_reason_name(reason) == \
            ref.ref_b_values_clear_reason_name(reason)
    for policy in range(-1, 4):
        assert lib.rtt_b_values_clear_policy_name(policy) == \
            ref.ref_b_values_clear_policy_name(policy)


def test_b_values_clear_policy_known_cases():
    """Hand cases: always-clear never skips; a valid values-only update of a
    duplicate-free B skips the clear under noclear and safe-auto; the first
    failing precondition is the reason."""
    ok = dict(values_only_update=1, dynamic_b_valid=1, has_active_duplicates=0,
              active_source_mapping_complete=1, active_source_to_payload_injective=1,
              payload_fully_initialized=1, every_payload_slot_overwritten=1,
              numeric_b_read_only=1)
    d = rtt_amd.choose_b_values_clear(0, **ok)
    assert (d["clear_payload"], d["skip_clear"], d["fallback"]) == (1, 0, 0)
    assert d["reason_name"] == "policy-always-clear"
    for pol, name in ((1, "eligible-explicit-noclear"), (2, "eligible-safe-auto")):
        d = rtt_amd.choose_b_values_clear(pol, **ok)
        assert (d["clear_payload"], d["skip_clear"], d["fallback"]) == (0, 1, 0)
        assert d["reason_name"] == name
    d = rtt_amd.choose_b_values_clear(2, **dict(ok, has_active_duplicates=1,
                                                 numeric_b_read_only=0))
    assert d["reason_name"] == "active-source-duplicates" and d["clear_payload"] == 1
    d = rtt_amd.choose_b_values_clear(1, **dict(ok, values_only_update=0, dynamic_b_valid=0))
    assert d["reason_name"] == "structure-update-legacy-rebuild"
    d = rtt_amd.choose_b_values_clear(5, **ok)
    assert d["reason_name"] == "invalid-policy"


def _permuted_tile_count(a, row_new_to_old, inner_new_to_old):
    """Occupied 8x4 tiles of A laid out by the two maps, from the oracle's
    tile builder on the explicitly permuted CSR."""
    inner_old_to_new = np.empty(a.n, np.int64)
    inner_old_to_new[inner_new_to_old] = np.arange(a.n)
    rows, cols = [], []
    for new, old in enumerate(row_new_to_old):
        cs = a.col[a.rowptr[old]:a.rowptr[old + 1]]
        rows += [new] * len(cs)
        cols += list(inner_old_to_new[cs])
    order = np.lexsort((np.asarray(cols), np.asarray(rows)))
    r = np.asarray(rows, np.int64)[order]
    c = np.asarray(cols, np.int32)[order]
    rp = np.zeros(a.m + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    t = oracle.hybrid_tiles(a.m, a.n, np.cumsum(rp).astype(np.int32), c,
                            np.ones(len(c)), 8, 4, False, 24)
    return t["num_tiles"]


@pytest.mark.parametrize("kind", ["rmat12", "stencil8", "random"])
def test_bgrf_plan_bijections_monotone_and_exact(kind):
    """BGRF-v1: both maps are bijections, the plan is deterministic, no
    stage raises the exact tile count, and the reported tile counts equal
    the oracle's tile builder on the permuted matrix."""
    a = {"rmat12": lambda: rtt_amd.gen_rmat(12, 16, 0.57, 0.19, 0.19, 5),
         "stencil8": lambda: rtt_amd.gen_stencil27(8, 0.5, 1.5, 1),
         "random": lambda: rtt_amd.gen_random(700, 500, 6, 3)}[kind]()
    r, i, st = rtt_amd.bgrf_plan(a)
    assert np.array_equal(np.sort(r), np.arange(a.m))
    assert np.array_equal(np.sort(i), np.arange(a.n))
    r2, i2, _ = rtt_amd.bgrf_plan(a)
    assert np.array_equal(r, r2) and np.array_equal(i, i2)
    assert st["tiles_after_coarse"] <= st["tiles_identity"]
    assert st["tiles_after_rows"] <= st["tiles_after_coarse"]
    assert st["tiles_final"] <= st["tiles_after_rows"]
    assert st["tiles_after_coarse"] == (st["tiles_coarse"] if st["coarse_accepted"]
                                        else st["tiles_identity"])
    assert st["tiles_identity"] == _permuted_tile_count(a, np.arange(a.m), np.arange(a.n))
    assert st["tiles_final"] == _permuted_tile_count(a, r, i)


def test_bgrf_recovers_a_scrambled_band():
    """A tridiagonal matrix under a random symmetric permutation: the
    reversed Cuthill-McKee traversal of its bipartite graph recovers a banded
    order, so the coarse stage is accepted and the tile count falls to the
    band's: 8 consecutive rows of a tridiagonal matrix touch 10 consecutive
    columns, at most four 4-wide K tiles."""
    n = 2000
    rng = np.random.default_rng(4)
    p = rng.permutation(n)
    rows, cols = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(p[i])
                cols.append(p[j])
    order = np.lexsort((cols, rows))
    r = np.asarray(rows)[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    a = rtt_amd.CsrMatrix(n, n, np.cumsum(rp).astype(np.int32),
                          np.asarray(cols, np.int32)[order], np.ones(len(rows)))
    _, _, st = rtt_amd.bgrf_plan(a)
    assert st["coarse_accepted"] == 1
    assert st["tiles_final"] < 0.5 * st["tiles_identity"]
    assert st["tiles_final"] <= 4 * ((n + 7) // 8)


def test_bgrf_plan_edge_cases():
    empty = rtt_amd.CsrMatrix(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32),
                              np.zeros(0))
    r, i, st = rtt_amd.bgrf_plan(empty)
    assert sorted(r) == list(range(5)) and sorted(i) == list(range(7))
    assert st["tiles_final"] == 0
    bad = rtt_amd.CsrMatrix(2, 2, np.array([0, 1, 2], np.int32), np.array([0, 5], np.int32),
                            np.ones(2))
    with pytest.raises(rtt_amd.RttError):
        rtt_amd.bgrf_plan(bad)


def _read_kv(path):
    with open(path) as f:
        rows = [l.strip().split(",", 1) for l in f][1:]
    return {k: v for k, v in rows}


def test_reorder_dumps_match_the_plan(tmp_path):
    """The on-disk formats of the reference's reorder diagnostics: the
    permutation CSV (header, old->new, and the inverse check column equal to
    old_id), the key/value stats, the pattern .mtx of the reordered A (every
    entry, 1-based), and the heatmap grid (exactly bins^2 cells, both sides
    summing to the entry count)."""
    a = rtt_amd.gen_rmat(10, 8, 0.57, 0.19, 0.19, 3)
    r, i, st = rtt_amd.bgrf_plan(a)
    prefix = str(tmp_path / "P")
    rtt_amd.dump_reorder(prefix, a, r, i, st)
    lines = open(prefix + "_permutations.csv").read().splitlines()
    assert lines[0] == "axis,old_id,new_id,new_to_old_at_new_id"
    body = [l.split(",") for l in lines[1:]]
    assert len(body) == a.m + a.n
    rows = [b for b in body if b[0] == "row"]
    inner = [b for b in body if b[0] == "inner"]
    assert all(b[1] == b[3] for b in body)
    assert [int(b[2]) for b in rows] == list(np.argsort(r))
    assert [int(b[2]) for b in inner] == list(np.argsort(i))
    kv = _read_kv(prefix + "_unified_stats.csv")
    assert kv["algorithm"] == "bgrf-v1"
    assert int(kv["num_tiles"]) == st["tiles_final"]
    assert int(kv["coarse_candidate_accepted"]) == st["coarse_accepted"]
    assert int(kv["sparse_tiles"]) + int(kv["dense_tiles"]) == st["tiles_final"]
    mtx = open(prefix + "_A_reordered.mtx").read().splitlines()
    assert mtx[0] == "%%MatrixMarket matrix coordinate pattern general"
    assert mtx[2] == f"{a.m} {a.n} {a.nnz}"
    pairs = sorted(tuple(int(x) for x in l.split()) for l in mtx[3:])
    r_o2n, i_o2n = np.argsort(r), np.argsort(i)
    want = sorted((int(r_o2n[row]) + 1, int(i_o2n[c]) + 1)
                  for row in range(a.m) for c in a.col[a.rowptr[row]:a.rowptr[row + 1]])
    assert pairs == want
    rtt_amd.dump_reorder_heatmap(prefix, a, r, i, 16, st)
    grid = np.loadtxt(prefix + "_heatmap.csv", delimiter=",", skiprows=1, dtype=np.int64)
    assert grid.shape == (256, 4)
    assert grid[:, 2].sum() == a.nnz and grid[:, 3].sum() == a.nnz
    meta = _read_kv(prefix + "_heatmap_meta.csv")
    assert int(meta["bins"]) == 16 and int(meta["nnz"]) == a.nnz
    # The identity baseline (no maps, no stats) and a non-bijective map.
    rtt_amd.dump_reorder(prefix + "_id", a)
    assert _read_kv(prefix + "_id_unified_stats.csv")["algorithm"] == "identity_baseline"
    with pytest.raises(rtt_amd.RttError):
        rtt_amd.dump_reorder(prefix, a, np.zeros(a.m, np.int32), i, st)


def test_b_values_clear_policy_matches_reference_vectors():
    """rtt_choose_b_values_clear against the committed outputs of the
    reference's dmma_choose_b_values_clear (tests/golden/reference_vectors.json,
    written from the reference sources compiled into oracle/_ref): every
    policy value times every combination of the eight preconditions, and the
    reason and policy names."""
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        g = json.load(f)["b_values_clear"]
    cols = g["columns"]
    assert len(g["rows"]) == 5 * 256
    lib = rtt_amd.load()
    for row in g["rows"]:
        policy, flags, want = row[0], row[1:9], tuple(row[9:])
        got = rtt_amd.choose_b_values_clear(policy, **dict(zip(cols[1:9], flags)))
        assert (got["clear_payload"], got["skip_clear"], got["fallback"],
                got["reason"]) == want, (policy, flags)
    for r, name in g["names"]["reason"].items():
        assert lib.rtt_b_values_clear_reason_name(int(r)).decode() == name
    for p, name in g["names"]["policy"].items():
        assert lib.rtt_b_values_clear_policy_name(int(p)).decode() == name
