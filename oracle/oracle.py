"""CPU oracle for the tiled FP64 SpGEMM path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/liboracle.so`` (a C++ restatement of the
reference's CPU semantics, see ``rtt_oracle.cpp`` for the file:line each
function follows) and of ``oracle/_ref/libref.so`` (reference sources compiled
where they lie: the tail-queue planner and the Matrix Market reader).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
use this module, and only as the checker / CPU baseline.  The shipped library
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

_lib = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        i64, i32, vp = C.c_int64, C.c_int, C.c_void_p
        _lib.orc_mmio_info.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32),
                                       C.POINTER(i64), C.POINTER(i32)]
        _lib.orc_mmio_expanded_nnz.argtypes = [C.c_char_p]
        _lib.orc_mmio_expanded_nnz.restype = i64
        _lib.orc_mmio_read.argtypes = [C.c_char_p, vp, vp, vp]
        _lib.orc_nnz_cub.argtypes = [i32, i32, vp, vp, i32]
        _lib.orc_nnz_cub.restype = C.c_uint64
        _lib.orc_transpose.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp]
        _lib.orc_spgemm_count.argtypes = [i32, i32, vp, vp, i32, i32, vp, vp, vp]
        _lib.orc_spgemm_count.restype = i64
        _lib.orc_legacy_tile_spgemm.argtypes = [i32, i32, vp, vp, vp, i32, i32, vp, vp, vp, vp,
                                                vp, vp]
        _lib.orc_legacy_tile_spgemm.restype = i64
        _lib.orc_spgemm_fill.argtypes = [i32, i32, vp, vp, vp, i32, i32, vp, vp, vp, vp, vp,
                                         vp, vp, vp, vp]
        _lib.orc_hybrid_tiles_count.argtypes = [i32, i32, vp, vp, i32, i32, i32, i32, vp]
        _lib.orc_hybrid_tiles_fill.argtypes = [i32, i32, vp, vp, vp, i32, i32, i32, i32, i32,
                                               vp, vp, vp, vp, vp, vp, vp, vp]
        _lib.orc_c_tiles_count.argtypes = [i32, i32, vp, vp, vp]
        _lib.orc_c_tiles_count.restype = i64
        _lib.orc_c_tiles_fill.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        _lib.orc_tile2csr_count.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, i32, vp]
        _lib.orc_tile2csr_count.restype = i64
        _lib.orc_tile2csr_fill.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, i32,
                                           vp, vp, vp]
        _lib.orc_tail_chunk_count.argtypes = [i32, i32, C.c_double, C.c_double, i32,
                                              C.c_double, C.c_double]
        _lib.orc_plan_tail_queue.argtypes = [i32, vp, vp, vp, i32, C.c_double, C.c_double, i32,
                                             C.c_double, i32, C.c_uint64, i32, vp, vp, vp]
    return _lib


def ref_available() -> bool:
    """True when the reference-source checker was built (without loading it)."""
    return os.path.exists(REF_LIB)


def ref_lib() -> C.CDLL | None:
    """Reference-source checker, or None when it could not be built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = C.CDLL(REF_LIB)
        i32, vp = C.c_int, C.c_void_p
        _ref.ref_tail_chunk_count.argtypes = [i32, i32, C.c_double, C.c_double, i32]
        _ref.ref_plan_tail_queue.argtypes = [i32, vp, vp, vp, i32, C.c_double, C.c_double, i32,
                                             C.c_double, i32, C.c_uint64, i32, vp, vp, vp]
        _ref.ref_mmio_info.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32),
                                       C.POINTER(i32), C.POINTER(i32)]
        _ref.ref_mmio_data.argtypes = [C.c_char_p, vp, vp, vp]
        _ref.ref_choose_b_values_clear.argtypes = [i32] * 9 + [vp]
        _ref.ref_choose_b_values_clear.restype = None
        _ref.ref_b_values_clear_policy_name.argtypes = [i32]
        _ref.ref_b_values_clear_policy_name.restype = C.c_char_p
        _ref.ref_b_values_clear_reason_name.argtypes = [i32]
        _ref.ref_b_values_clear_reason_name.restype = C.c_char_p
    return _ref


def ref_choose_b_values_clear(*args) -> tuple | None:
    """The reference's dmma_choose_b_values_clear
    (src/dmma_b_values_clear_policy.h:91-160) compiled from its header:
    (clear_payload, skip_clear, fallback, reason), or None without _ref/."""
    r = ref_lib()
    if r is None:
        return None
    out = np.zeros(4, np.int32)
    r.ref_choose_b_values_clear(*[int(x) for x in args], _p(out))
    return tuple(int(x) for x in out)


def _p(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- Matrix Market ---------------------------------------------------------

def mmio_read(path: str):
    """(m, n, rowptr, col, val) with mmio_allinone semantics."""
    m, n, sym = C.c_int(), C.c_int(), C.c_int()
    nnz_file = C.c_int64()
    rc = lib().orc_mmio_info(path.encode(), C.byref(m), C.byref(n), C.byref(nnz_file),
                             C.byref(sym))
    if rc != 0:
        raise ValueError(f"oracle mmio_info failed ({rc})")
    nnz = lib().orc_mmio_expanded_nnz(path.encode())
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(max(nnz, 1), np.int32)
    vv = np.zeros(max(nnz, 1), np.float64)
    if lib().orc_mmio_read(path.encode(), _p(rp), _p(ci), _p(vv)) != 0:
        raise ValueError("oracle mmio_read failed")
    return m.value, n.value, rp, ci[:nnz], vv[:nnz]


def ref_mmio_read(path: str):
    """The reference reader (compiled from its own sources), or None."""
    r = ref_lib()
    if r is None:
        return None
    m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    if r.ref_mmio_info(path.encode(), C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym)) != 0:
        raise ValueError("reference mmio_info failed")
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    vv = np.zeros(max(nnz.value, 1), np.float64)
    if r.ref_mmio_data(path.encode(), _p(rp), _p(ci), _p(vv)) != 0:
        raise ValueError("reference mmio_data failed")
    return m.value, n.value, rp, ci[:nnz.value], vv[:nnz.value]


# ---- SpGEMM ------------------------------------------------------------------

def nnz_cub(m, n, rp, ci, aat: bool) -> int:
    rp, ci = _i32(rp), _i32(ci)
    return int(lib().orc_nnz_cub(m, n, _p(rp), _p(ci), int(aat)))


def transpose(m, n, rp, ci, val):
    rp, ci, val = _i32(rp), _i32(ci), _f64(val)
    nnz = int(rp[m]) if m > 0 else 0
    trp = np.zeros(n + 1, np.int32)
    tci = np.zeros(max(nnz, 1), np.int32)
    tv = np.zeros(max(nnz, 1), np.float64)
    lib().orc_transpose(m, n, _p(rp), _p(ci), _p(val), _p(trp), _p(tci), _p(tv))
    return trp, tci[:nnz], tv[:nnz]


def spgemm_row_counts(a, b) -> np.ndarray:
    """Row pointer (int64, m + 1) of C = A*B's structure: the structure pass
    of the row SPA alone (orc_spgemm_count), for products too large to fill."""
    ma, na, rpa, cia = a[:4]
    mb, nb, rpb, cib = b[:4]
    rpa, cia, rpb, cib = _i32(rpa), _i32(cia), _i32(rpb), _i32(cib)
    rpc = np.zeros(ma + 1, np.int64)
    if lib().orc_spgemm_count(ma, na, _p(rpa), _p(cia), mb, nb, _p(rpb), _p(cib), _p(rpc)) < 0:
        raise ValueError("shape mismatch")
    return rpc


def legacy_tile_spgemm(a, b):
    """C = A*B by the reference's legacy 16x16 tile CPU algorithm
    (orc_legacy_tile_spgemm, src/spgemm-cpu.h:3-455), as CSR:
    (rowptr int64, col int32, val float64).  a, b = (m, n, rowptr, col, val)."""
    ma, na, rpa, cia, va = a
    mb, nb, rpb, cib, vb = b
    rpa, cia, va, rpb, cib, vb = _i32(rpa), _i32(cia), _f64(va), _i32(rpb), _i32(cib), _f64(vb)
    rpc = np.zeros(ma + 1, np.int64)
    f = lib().orc_legacy_tile_spgemm
    nnz = f(ma, na, _p(rpa), _p(cia), _p(va), mb, nb, _p(rpb), _p(cib), _p(vb), _p(rpc), None,
            None)
    if nnz < 0:
        raise ValueError("shape mismatch")
    cic = np.zeros(max(nnz, 1), np.int32)
    vc = np.zeros(max(nnz, 1), np.float64)
    f(ma, na, _p(rpa), _p(cia), _p(va), mb, nb, _p(rpb), _p(cib), _p(vb), _p(rpc), _p(cic),
      _p(vc))
    return rpc, cic[:nnz], vc[:nnz]


def spgemm(a, b, with_dd: bool = True):
    """C = A*B.  a, b = (m, n, rowptr, col, val).  Returns a dict with the
    reference-order values and, optionally, the double-double truth and the
    sum of |a||b| per output."""
    ma, na, rpa, cia, va = a
    mb, nb, rpb, cib, vb = b
    rpa, cia, va, rpb, cib, vb = _i32(rpa), _i32(cia), _f64(va), _i32(rpb), _i32(cib), _f64(vb)
    rpc = np.zeros(ma + 1, np.int64)
    nnz = lib().orc_spgemm_count(ma, na, _p(rpa), _p(cia), mb, nb, _p(rpb), _p(cib), _p(rpc))
    if nnz < 0:
        raise ValueError("shape mismatch")
    cic = np.zeros(max(nnz, 1), np.int32)
    vc = np.zeros(max(nnz, 1), np.float64)
    hi = np.zeros(max(nnz, 1), np.float64) if with_dd else None
    lo = np.zeros(max(nnz, 1), np.float64) if with_dd else None
    ab = np.zeros(max(nnz, 1), np.float64) if with_dd else None
    rc = lib().orc_spgemm_fill(ma, na, _p(rpa), _p(cia), _p(va), mb, nb, _p(rpb), _p(cib),
                               _p(vb), _p(rpc), _p(cic), _p(vc), _p(hi), _p(lo), _p(ab))
    if rc != 0:
        raise ValueError("oracle spgemm fill failed")
    out = {"m": ma, "n": nb, "rowptr": rpc, "col": cic[:nnz], "val": vc[:nnz], "nnz": int(nnz)}
    if with_dd:
        out.update(hi=hi[:nnz], lo=lo[:nnz], absprod=ab[:nnz])
    return out


# ---- tile formats ------------------------------------------------------------

def hybrid_tiles(m, n, rp, ci, val, tile_rows, tile_cols, col_major, dense_threshold=24,
                 build_csc=False):
    rp, ci, val = _i32(rp), _i32(ci), _f64(val)
    counts = np.zeros(4, np.int64)
    if lib().orc_hybrid_tiles_count(m, n, _p(rp), _p(ci), tile_rows, tile_cols, int(col_major),
                                    dense_threshold, _p(counts)) != 0:
        raise ValueError("invalid CSR for tiling")
    T, L = int(counts[0]), int(counts[1])
    trc = (m + tile_rows - 1) // tile_rows
    tcc = (n + tile_cols - 1) // tile_cols
    out = {"tile_row_ptr": np.zeros(trc + 1, np.int32), "tile_col_idx": np.zeros(max(T, 1), np.int32),
           "value_offsets": np.zeros(T + 1, np.int32), "masks": np.zeros(max(T, 1), np.uint32),
           "values": np.zeros(max(L, 1), np.float64)}
    csc = [None, None, None]
    if build_csc:
        out["tile_col_ptr"] = np.zeros(tcc + 1, np.int32)
        out["tile_row_idx"] = np.zeros(max(T, 1), np.int32)
        out["csc_tile_ids"] = np.zeros(max(T, 1), np.int32)
        csc = [out["tile_col_ptr"], out["tile_row_idx"], out["csc_tile_ids"]]
    lib().orc_hybrid_tiles_fill(m, n, _p(rp), _p(ci), _p(val), tile_rows, tile_cols,
                                int(col_major), dense_threshold, int(build_csc),
                                _p(out["tile_row_ptr"]), _p(out["tile_col_idx"]),
                                _p(out["value_offsets"]), _p(out["masks"]), _p(out["values"]),
                                *[_p(x) for x in csc])
    for k in ("tile_col_idx", "masks", "tile_row_idx", "csc_tile_ids"):
        if k in out:
            out[k] = out[k][:T]
    out["values"] = out["values"][:L]
    out.update(num_tiles=T, payload_size=L, dense_tiles=int(counts[2]),
               sparse_tiles=int(counts[3]), tile_row_count=trc, tile_col_count=tcc)
    return out


def c_tiles(m, n, rpc, cic, vc):
    """8x8 output tiles of a row-sorted CSR C."""
    rpc = np.ascontiguousarray(rpc, dtype=np.int64)
    cic, vc = _i32(cic), _f64(vc)
    tilem = (m + 7) // 8
    tp = np.zeros(tilem + 1, np.int64)
    T = int(lib().orc_c_tiles_count(m, n, _p(rpc), _p(cic), _p(tp)))
    nnz = int(rpc[m]) if m > 0 else 0
    out = {"tile_ptr": tp, "tile_col": np.zeros(max(T, 1), np.int32),
           "tile_nnz": np.zeros(T + 1, np.int64), "tile_csr_ptr": np.zeros(max(8 * T, 1), np.uint8),
           "tile_csr_col": np.zeros(max(nnz, 1), np.uint8),
           "tile_csr_val": np.zeros(max(nnz, 1), np.float64),
           "tile_mask": np.zeros(max(T, 1), np.uint64)}
    rc = lib().orc_c_tiles_fill(m, n, _p(rpc), _p(cic), _p(vc), _p(tp), _p(out["tile_col"]),
                                _p(out["tile_nnz"]), _p(out["tile_csr_ptr"]),
                                _p(out["tile_csr_col"]), _p(out["tile_csr_val"]),
                                _p(out["tile_mask"]))
    if rc != 0:
        raise ValueError("oracle c_tiles failed")
    for k in ("tile_col", "tile_mask"):
        out[k] = out[k][:T]
    out["tile_csr_ptr"] = out["tile_csr_ptr"][:8 * T]
    out["tile_csr_col"] = out["tile_csr_col"][:nnz]
    out["tile_csr_val"] = out["tile_csr_val"][:nnz]
    out.update(num_tiles=T, nnz=nnz, tilem=tilem)
    return out


def tile2csr(m, n, tilem, tile_ptr, tile_col, tile_nnz, tile_csr_ptr, tile_csr_col,
             tile_csr_val, row_new_to_old=None, permutation_size=0, canonical=True):
    """Returns (rowptr int64, col, val) or None when the reference would fail."""
    tile_ptr, tile_col, tile_nnz = _i32(tile_ptr), _i32(tile_col), _i32(tile_nnz)
    tcp = np.ascontiguousarray(tile_csr_ptr, dtype=np.uint8)
    tcc = np.ascontiguousarray(tile_csr_col, dtype=np.uint8)
    tcv = _f64(tile_csr_val)
    perm = None if row_new_to_old is None else _i32(row_new_to_old)
    rp = np.zeros(m + 1, np.int64)
    nnz = lib().orc_tile2csr_count(m, n, tilem, _p(tile_ptr), _p(tile_col), _p(tile_nnz),
                                   _p(tcp), _p(tcc), _p(perm), permutation_size, _p(rp))
    if nnz < 0:
        return None
    ci = np.zeros(max(nnz, 1), np.int32)
    vv = np.zeros(max(nnz, 1), np.float64)
    lib().orc_tile2csr_fill(m, n, tilem, _p(tile_ptr), _p(tile_col), _p(tile_nnz), _p(tcp),
                            _p(tcc), _p(tcv), _p(perm), permutation_size, int(canonical), _p(rp),
                            _p(ci), _p(vv))
    return rp, ci[:nnz], vv[:nnz]


# ---- tail queue ----------------------------------------------------------------

def tail_chunk_count(scans, matches, threshold=16384.0, target=8192.0, max_chunks=8):
    return int(lib().orc_tail_chunk_count(scans, matches, threshold, target, max_chunks, 1.0, 8.0))


def plan_tail_queue(scans, matches, output_nnz, enabled=True, threshold=16384.0, target=8192.0,
                    max_chunks=8, max_heavy_fraction=0.01, record_capacity=1 << 20,
                    workspace_limit=1 << 30, workspace_reduction=True, use_reference=False):
    scans, matches, output_nnz = _i32(scans), _i32(matches), _i32(output_nnz)
    n = scans.size
    out = np.zeros(4, np.int64)
    ids = np.zeros(max(n, 1), np.int32)
    chunks = np.zeros(max(n, 1), np.int32)
    f = ref_lib().ref_plan_tail_queue if use_reference else lib().orc_plan_tail_queue
    f(n, _p(scans), _p(matches), _p(output_nnz), int(enabled), threshold, target, max_chunks,
      max_heavy_fraction, record_capacity, workspace_limit, int(workspace_reduction), _p(out),
      _p(ids), _p(chunks))
    h = int(out[1])
    return {"reason": int(out[0]), "heavy": h, "chunks": int(out[2]), "workspace": int(out[3]),
            "heavy_ids": ids[:h].copy(), "heavy_chunks": chunks[:h].copy()}


# ---- error metrics -------------------------------------------------------------

def value_errors(c_vals, dd_hi, dd_lo, absprod):
    """max / median of |c - c*| / |c*| (c* != 0) and of |c - c*| / (|A||B|)."""
    c = np.asarray(c_vals, np.float64)
    err = np.abs((c - dd_hi) - dd_lo)
    truth = np.abs(dd_hi)
    nz = truth > 0
    rel = err[nz] / truth[nz] if nz.any() else np.zeros(1)
    scale = np.where(absprod > 0, absprod, 1.0)
    fwd = err / scale
    return {"rel_max": float(rel.max()) if rel.size else 0.0,
            "rel_median": float(np.median(rel)) if rel.size else 0.0,
            "fwd_max": float(fwd.max()) if fwd.size else 0.0,
            "fwd_median": float(np.median(fwd)) if fwd.size else 0.0}
