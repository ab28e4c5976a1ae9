"""Python host-side mirror of the tiled SpGEMM driver interface.

Thin ctypes layer over ``lib/librtt_spgemm.so`` (the C ABI declared in
``include/rtt_spgemm.h``).  Names follow the reference's host API:

=====================================  =========================================
reference (file:line)                  here
=====================================  =========================================
mmio_allinone  src/mmio_highlevel.h:593 read_mtx
gpu_prepare_identity_a                  prepare_identity_a
  src/gpu_dmma_tiles.h:2443
gpu_prepare_external_a  :2389           prepare_external_a
gpu_rebuild_dynamic_b_via_permuted_csr  rebuild_b(a, aat=0)
  / _transpose  :2275 / :2285           rebuild_b(a, aat=1)
gpu_compute_nnz_cub_derived  :2304      nnz_cub
dmma_tilespgemm src/dmma_spgemm.h:6490  tilespgemm
tile2csr / tile2csr_restore_rows        tile2csr / tile2csr_restore_rows
  src/tile2csr.h:232 / :337
=====================================  =========================================

Every compute call runs the HIP library; there is no CPU fallback.  Importing
works without a GPU; ``init()`` fails loudly when no device or library is
present.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "librtt_spgemm.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "rtt_spgemm.h")

RTT_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "OOM", 3: "OVERFLOW", 4: "DEVICE", 5: "UNSUPPORTED"}
LB_MODES = {"static": 0, "dynamic": 1, "adaptive-tail": 2}


class RttError(RuntimeError):
    def __init__(self, code: int, where: str, text: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(code, code)}: {text}")
        self.code = code


class SMatrix(C.Structure):
    """The legacy SMatrix layout (src/common.h:48-73)."""

    _fields_ = [
        ("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int), ("isSymmetric", C.c_int),
        ("value", C.POINTER(C.c_double)), ("columnindex", C.POINTER(C.c_int)),
        ("rowpointer", C.POINTER(C.c_int)), ("tilem", C.c_int), ("tilen", C.c_int),
        ("tile_ptr", C.POINTER(C.c_int)), ("tile_columnidx", C.POINTER(C.c_int)),
        ("tile_rowidx", C.POINTER(C.c_int)), ("tile_nnz", C.POINTER(C.c_int)),
        ("numtile", C.c_int), ("tile_csr_Value", C.POINTER(C.c_double)),
        ("tile_csr_Col", C.POINTER(C.c_ubyte)), ("tile_csr_Ptr", C.POINTER(C.c_ubyte)),
        ("mask", C.POINTER(C.c_ushort)), ("csc_tile_ptr", C.POINTER(C.c_int)),
        ("csc_tile_rowidx", C.POINTER(C.c_int)),
    ]


class DeviceTiles(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "rows", "cols", "tile_rows", "tile_cols", "tile_row_count", "tile_col_count",
        "num_tiles", "payload_size", "dense_tiles", "sparse_tiles")] + [
        ("tile_row_ptr", C.c_void_p), ("tile_col_idx", C.c_void_p),
        ("value_offsets", C.c_void_p), ("masks", C.c_void_p), ("values", C.c_void_p),
        ("tile_col_ptr", C.c_void_p), ("tile_row_idx", C.c_void_p),
        ("csc_tile_ids", C.c_void_p)]


class OfflineAStats(C.Structure):
    _fields_ = [("h2d_ms", C.c_double), ("validation_ms", C.c_double),
                ("canonicalize_ms", C.c_double), ("reorder_ms", C.c_double),
                ("tile_build_ms", C.c_double), ("total_ms", C.c_double),
                ("rows_were_sorted", C.c_int), ("transpose_src_ms", C.c_double)]


class BUpdateStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("csr_row_permute_ms", C.c_double),
                ("csr2tile_ms", C.c_double), ("csc_ms", C.c_double), ("path", C.c_int)]


class BgrfStats(C.Structure):
    _fields_ = [("tiles_identity", C.c_longlong), ("tiles_coarse", C.c_longlong),
                ("tiles_after_coarse", C.c_longlong), ("tiles_after_rows", C.c_longlong),
                ("tiles_final", C.c_longlong), ("row_fanout_before", C.c_longlong),
                ("row_fanout_after", C.c_longlong), ("inner_fanout_before", C.c_longlong),
                ("inner_fanout_after", C.c_longlong), ("coarse_accepted", C.c_int),
                ("coarse_levels", C.c_int), ("coarse_components", C.c_int),
                ("row_windows", C.c_int), ("row_windows_accepted", C.c_int),
                ("inner_windows", C.c_int), ("inner_windows_accepted", C.c_int),
                ("coarse_ms", C.c_double), ("fine_ms", C.c_double), ("plan_ms", C.c_double)]


class BValuesDecision(C.Structure):
    _fields_ = [("clear_payload", C.c_int), ("skip_clear", C.c_int), ("fallback", C.c_int),
                ("reason", C.c_int)]


B_VALUES_POLICIES = {"always-clear": 0, "noclear": 1, "safe-auto": 2}


class Schedule(C.Structure):
    _fields_ = [("lb_mode", C.c_int), ("materialize_output", C.c_int),
                ("split_threshold", C.c_double), ("split_chunk_target", C.c_double),
                ("split_max_chunks", C.c_int), ("max_heavy_fraction", C.c_double),
                ("tail_record_capacity", C.c_int), ("force_tail_split", C.c_int),
                ("numeric_path", C.c_int), ("tile_fill_threshold", C.c_double),
                ("precision", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("symbolic_ms", C.c_double), ("allocation_ms", C.c_double),
                ("numeric_ms", C.c_double), ("total_ms", C.c_double),
                ("output_copy_ms", C.c_double), ("symbolic_count_kernel_ms", C.c_double),
                ("symbolic_fill_kernel_ms", C.c_double), ("numeric_kernel_ms", C.c_double),
                ("pair_count", C.c_ulonglong), ("output_tiles", C.c_ulonglong),
                ("output_nnz", C.c_ulonglong), ("heavy_rows", C.c_ulonglong),
                ("heavy_chunks", C.c_ulonglong), ("output_materialized", C.c_int),
                ("numeric_path", C.c_int), ("b_columns_ms", C.c_double),
                ("mfma_pairs", C.c_ulonglong), ("precision", C.c_int),
                ("tail_reason", C.c_int), ("tail_records", C.c_ulonglong)]


class TuneReport(C.Structure):
    _fields_ = [("candidates", C.c_int), ("chosen", C.c_int), ("ms", C.c_double * 16),
                ("numeric_path", C.c_int * 16), ("fill", C.c_double * 16),
                ("lb_mode", C.c_int * 16)]


class DeviceC(C.Structure):
    _fields_ = [("tilem", C.c_int64), ("tilen", C.c_int64), ("num_tiles", C.c_int64),
                ("nnz", C.c_int64), ("tile_ptr", C.c_void_p), ("tile_col", C.c_void_p),
                ("tile_mask", C.c_void_p), ("tile_nnz", C.c_void_p),
                ("tile_csr_ptr", C.c_void_p), ("tile_csr_col", C.c_void_p),
                ("tile_csr_val", C.c_void_p)]


_lib = None


def load() -> C.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "rtt_last_error": (C.c_char_p, []),
        "rtt_version": (C.c_char_p, []),
        "rtt_destroy_smatrix": (None, [P(SMatrix)]),
        "rtt_read_mtx": (C.c_int, [C.c_char_p, C.c_int, P(SMatrix)]),
        "rtt_write_mtx": (C.c_int, [C.c_char_p, P(SMatrix)]),
        "rtt_init": (C.c_int, [C.c_int]),
        "rtt_sync": (C.c_int, []),
        "rtt_download_tiles": (C.c_int, [P(DeviceTiles)] + [vp] * 8),
        "rtt_prepare_identity_a": (C.c_int, [P(SMatrix), C.c_int, P(vp), P(OfflineAStats)]),
        "rtt_prepare_external_a": (C.c_int, [P(SMatrix), vp, vp, C.c_int, P(vp),
                                             P(OfflineAStats)]),
        "rtt_read_order_file": (C.c_int, [C.c_char_p, C.c_int, vp]),
        "rtt_bgrf_plan": (C.c_int, [P(SMatrix), vp, vp, P(BgrfStats)]),
        "rtt_dump_reorder": (C.c_int, [C.c_char_p, P(SMatrix), vp, vp, P(BgrfStats)]),
        "rtt_dump_reorder_heatmap": (C.c_int, [C.c_char_p, P(SMatrix), vp, vp, C.c_int,
                                               P(BgrfStats)]),
        "rtt_prepare_bgrf_a": (C.c_int, [P(SMatrix), C.c_int, P(vp), P(OfflineAStats),
                                         P(BgrfStats)]),
        "rtt_prepared_a_view": (C.c_int, [vp, P(DeviceTiles)]),
        "rtt_prepared_a_info": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "rtt_prepared_a_row_new_to_old": (vp, [vp]),
        "rtt_destroy_prepared_a": (None, [vp]),
        "rtt_rebuild_b": (C.c_int, [vp, C.c_int, P(vp), P(BUpdateStats)]),
        "rtt_dynamic_b_view": (C.c_int, [vp, P(DeviceTiles)]),
        "rtt_destroy_dynamic_b": (None, [vp]),
        "rtt_compute_nnz_cub": (C.c_int, [vp, C.c_int, P(C.c_ulonglong), P(C.c_double)]),
        "rtt_upload_csr": (C.c_int, [P(SMatrix), P(vp), P(C.c_double), P(C.c_double)]),
        "rtt_device_csr_set_values": (C.c_int, [vp, vp]),
        "rtt_destroy_device_csr": (None, [vp]),
        "rtt_rebuild_b_general": (C.c_int, [vp, vp, P(vp), P(BUpdateStats)]),
        "rtt_compute_nnz_cub_ab": (C.c_int, [vp, vp, P(C.c_ulonglong), P(C.c_double)]),
        "rtt_choose_b_values_clear": (C.c_int, [C.c_int] * 9 + [P(BValuesDecision)]),
        "rtt_b_values_clear_policy_name": (C.c_char_p, [C.c_int]),
        "rtt_b_values_clear_reason_name": (C.c_char_p, [C.c_int]),
        "rtt_update_b_values": (C.c_int, [vp, vp, C.c_int, vp, P(BUpdateStats),
                                          P(BValuesDecision)]),
        "rtt_schedule_defaults": (None, [P(Schedule)]),
        "rtt_tilespgemm": (C.c_int, [vp, vp, P(SMatrix), P(Stats), P(Schedule)]),
        "rtt_last_c_view": (C.c_int, [vp, P(DeviceC)]),
        "rtt_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
        "rtt_tile2csr": (C.c_int, [P(SMatrix), C.c_int]),
        "rtt_tile2csr_restore_rows": (C.c_int, [P(SMatrix), vp, C.c_int, C.c_int]),
        "rtt_gen_laplace2d": (C.c_int, [C.c_int, C.c_double, C.c_uint64, P(SMatrix)]),
        "rtt_gen_stencil27": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint64,
                                        P(SMatrix)]),
        "rtt_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, P(SMatrix)]),
        "rtt_gen_random": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, P(SMatrix)]),
        "rtt_gen_fem_blocks": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, P(SMatrix)]),
        "rtt_gen_banded": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, P(SMatrix)]),
        "rtt_gen_banded_batch_shapes": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int,
                                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                  C.c_uint64, vp, vp, vp]),
        "rtt_gen_poisson7_jumps": (C.c_int, [C.c_int, C.c_double, C.c_uint64, P(SMatrix)]),
        "rtt_gen_aggregation_prolongator": (C.c_int, [P(SMatrix), C.c_int, C.c_double,
                                                      P(SMatrix)]),
        "rtt_csr_transpose": (C.c_int, [P(SMatrix), P(SMatrix)]),
        "rtt_block_diag": (C.c_int, [C.c_int, vp, C.c_int, P(SMatrix), vp, vp]),
        "rtt_batch_extract": (C.c_int, [vp, P(C.c_int64), P(C.c_int64), C.c_int, C.c_int,
                                        P(SMatrix)]),
        "rtt_autotune": (C.c_int, [vp, vp, C.c_int, P(Schedule), P(TuneReport)]),
        "rtt_last_c_row_stats": (C.c_int, [vp, vp, vp]),
        "rtt_last_tail_records": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, P(C.c_int)]),
        "rtt_tile_values_noalloc": (C.c_int, [vp, vp, C.c_int, vp]),
        "rtt_stream": (C.c_int, [P(vp)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int, where: str) -> None:
    if rc != RTT_OK:
        raise RttError(rc, where, load().rtt_last_error().decode())


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


@dataclass
class CsrMatrix:
    """Host CSR (int32 row pointer / columns, float64 values)."""
    m: int
    n: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    is_symmetric: int = 0

    @property
    def nnz(self) -> int:
        return int(self.rowptr[self.m]) if self.m > 0 else 0

    def as_smatrix(self) -> SMatrix:
        self.rowptr = np.ascontiguousarray(self.rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        self.val = np.ascontiguousarray(self.val, dtype=np.float64)
        s = SMatrix()
        s.m, s.n, s.nnz, s.isSymmetric = self.m, self.n, self.nnz, self.is_symmetric
        s.rowpointer = self.rowptr.ctypes.data_as(C.POINTER(C.c_int))
        s.columnindex = self.col.ctypes.data_as(C.POINTER(C.c_int))
        s.value = self.val.ctypes.data_as(C.POINTER(C.c_double))
        return s

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.m, self.n))


def _np_copy(ptr, count: int, dtype) -> np.ndarray:
    if count <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
    addr = C.cast(ptr, C.c_void_p).value
    buf = (ct * count).from_address(addr)
    return np.array(np.ctypeslib.as_array(buf), dtype=dtype, copy=True)


def _csr_from_smatrix(s: SMatrix) -> CsrMatrix:
    out = CsrMatrix(s.m, s.n, _np_copy(s.rowpointer, s.m + 1, np.int32),
                    _np_copy(s.columnindex, s.nnz, np.int32),
                    _np_copy(s.value, s.nnz, np.float64), s.isSymmetric)
    if s.m + 1 > 0 and out.rowptr.size == 0:
        out.rowptr = np.zeros(s.m + 1, np.int32)
    return out


def _consume_csr(s: SMatrix) -> CsrMatrix:
    try:
        return _csr_from_smatrix(s)
    finally:
        load().rtt_destroy_smatrix(C.byref(s))


# ---- host utilities ----------------------------------------------------------

def read_mtx(path: str, reference_values: bool = False) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_read_mtx(path.encode(), int(reference_values), C.byref(s)), "read_mtx")
    return _consume_csr(s)


def write_mtx(path: str, a: CsrMatrix) -> None:
    s = a.as_smatrix()
    _check(load().rtt_write_mtx(path.encode(), C.byref(s)), "write_mtx")


def read_order_file(path: str, expected: int) -> np.ndarray:
    out = np.zeros(max(expected, 1), np.int32)
    _check(load().rtt_read_order_file(path.encode(), expected, _ptr(out)), "read_order_file")
    return out[:expected]


def gen_laplace2d(g: int, pert: float = 0.1, seed: int = 1) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_laplace2d(g, pert, seed, C.byref(s)), "gen_laplace2d")
    return _consume_csr(s)


def gen_stencil27(g: int, lo: float = 0.5, hi: float = 1.5, seed: int = 1) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_stencil27(g, lo, hi, seed, C.byref(s)), "gen_stencil27")
    return _consume_csr(s)


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 20260101) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(s)), "gen_rmat")
    return _consume_csr(s)


def gen_random(m: int, n: int, per_row: int, seed: int = 1) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_random(m, n, per_row, seed, C.byref(s)), "gen_random")
    return _consume_csr(s)


def gen_fem_blocks(nodes: int, neighbours: int = 6, spatial_order: bool = True,
                   seed: int = 20260101) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_fem_blocks(nodes, neighbours, int(spatial_order), seed, C.byref(s)),
           "gen_fem_blocks")
    return _consume_csr(s)


def gen_banded(n: int, bandwidth: int, per_row: int, seed: int = 1) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_banded(n, bandwidth, per_row, seed, C.byref(s)), "gen_banded")
    return _consume_csr(s)


def gen_banded_batch_shapes(problems: int, median_rows: float = 16000.0, sigma: float = 1.0,
                            min_rows: int = 1000, max_rows: int = 200000, min_bw: int = 8,
                            max_bw: int = 64, min_per_row: int = 5, max_per_row: int = 30,
                            seed: int = 20260101):
    """(rows, bandwidth, per_row) arrays of a banded batch."""
    rows = np.zeros(max(problems, 1), np.int32)
    bw = np.zeros_like(rows)
    pr = np.zeros_like(rows)
    _check(load().rtt_gen_banded_batch_shapes(problems, median_rows, sigma, min_rows, max_rows,
                                              min_bw, max_bw, min_per_row, max_per_row, seed,
                                              _ptr(rows), _ptr(bw), _ptr(pr)),
           "gen_banded_batch_shapes")
    return rows[:problems], bw[:problems], pr[:problems]


def gen_poisson7_jumps(g: int, decades: float = 3.0, seed: int = 1) -> CsrMatrix:
    s = SMatrix()
    _check(load().rtt_gen_poisson7_jumps(g, decades, seed, C.byref(s)), "gen_poisson7_jumps")
    return _consume_csr(s)


def gen_aggregation_prolongator(a: CsrMatrix, g: int, omega: float = 2.0 / 3.0) -> CsrMatrix:
    s = SMatrix()
    src = a.as_smatrix()
    _check(load().rtt_gen_aggregation_prolongator(C.byref(src), g, omega, C.byref(s)),
           "gen_aggregation_prolongator")
    return _consume_csr(s)


def csr_transpose(a: CsrMatrix) -> CsrMatrix:
    s = SMatrix()
    src = a.as_smatrix()
    _check(load().rtt_csr_transpose(C.byref(src), C.byref(s)), "csr_transpose")
    return _consume_csr(s)


def block_diag(mats, align: int = 8):
    """Block-diagonal batch of CSR matrices; returns (CsrMatrix, row_offsets, col_offsets)."""
    mats = list(mats)
    sm = [m.as_smatrix() for m in mats]
    arr = (C.POINTER(SMatrix) * max(len(sm), 1))(*[C.pointer(x) for x in sm])
    ro = np.zeros(len(mats) + 1, np.int64)
    co = np.zeros(len(mats) + 1, np.int64)
    s = SMatrix()
    _check(load().rtt_block_diag(len(mats), C.cast(arr, C.c_void_p), align, C.byref(s),
                                 _ptr(ro), _ptr(co)), "block_diag")
    return _consume_csr(s), ro, co


# ---- tile matrices -------------------------------------------------------------

@dataclass
class TileC:
    """Host tile-CSR of C (the SMatrix fields dmma_copy_output_to_host fills)."""
    m: int
    n: int
    tilem: int
    tilen: int
    numtile: int
    nnz: int
    tile_ptr: np.ndarray
    tile_columnidx: np.ndarray
    tile_nnz: np.ndarray
    tile_csr_Ptr: np.ndarray
    tile_csr_Col: np.ndarray
    tile_csr_Value: np.ndarray

    def as_smatrix(self) -> SMatrix:
        s = SMatrix()
        s.m, s.n, s.nnz, s.tilem, s.tilen, s.numtile = (
            self.m, self.n, self.nnz, self.tilem, self.tilen, self.numtile)
        for name, ct in (("tile_ptr", C.c_int), ("tile_columnidx", C.c_int),
                         ("tile_nnz", C.c_int), ("tile_csr_Ptr", C.c_ubyte),
                         ("tile_csr_Col", C.c_ubyte), ("tile_csr_Value", C.c_double)):
            arr = getattr(self, name)
            setattr(s, name, arr.ctypes.data_as(C.POINTER(ct)))
        return s


def _tilec_from_smatrix(s: SMatrix) -> TileC:
    return TileC(s.m, s.n, s.tilem, s.tilen, s.numtile, s.nnz,
                 _np_copy(s.tile_ptr, s.tilem + 1, np.int32),
                 _np_copy(s.tile_columnidx, s.numtile, np.int32),
                 _np_copy(s.tile_nnz, s.numtile + 1, np.int32),
                 _np_copy(s.tile_csr_Ptr, 8 * s.numtile, np.uint8),
                 _np_copy(s.tile_csr_Col, s.nnz, np.uint8),
                 _np_copy(s.tile_csr_Value, s.nnz, np.float64))


def _download_tiles(view: DeviceTiles) -> dict:
    T = view.num_tiles
    out = {
        "rows": view.rows, "cols": view.cols, "tile_rows": view.tile_rows,
        "tile_cols": view.tile_cols, "tile_row_count": view.tile_row_count,
        "tile_col_count": view.tile_col_count, "num_tiles": T,
        "payload_size": view.payload_size, "dense_tiles": view.dense_tiles,
        "sparse_tiles": view.sparse_tiles,
        "tile_row_ptr": np.zeros(view.tile_row_count + 1, np.int32),
        "tile_col_idx": np.zeros(max(T, 1), np.int32),
        "value_offsets": np.zeros(T + 1, np.int32),
        "masks": np.zeros(max(T, 1), np.uint32),
        "values": np.zeros(max(view.payload_size, 1), np.float64),
    }
    has_csc = bool(view.tile_col_ptr)
    if has_csc:
        out["tile_col_ptr"] = np.zeros(view.tile_col_count + 1, np.int32)
        out["tile_row_idx"] = np.zeros(max(T, 1), np.int32)
        out["csc_tile_ids"] = np.zeros(max(T, 1), np.int32)
    names = ["tile_row_ptr", "tile_col_idx", "value_offsets", "masks", "values",
             "tile_col_ptr", "tile_row_idx", "csc_tile_ids"]
    ptrs = [_ptr(out[n]) if n in out else None for n in names]
    _check(load().rtt_download_tiles(C.byref(view), *ptrs), "download_tiles")
    for n in ("tile_col_idx", "masks", "tile_row_idx", "csc_tile_ids"):
        if n in out:
            out[n] = out[n][:T]
    out["values"] = out["values"][:view.payload_size]
    return out


_initialised = False
DRIVER_PATH = os.path.join(os.path.dirname(_HERE), "bin", "rtt_spgemm")


def init(device: int = 0) -> None:
    global _initialised
    _check(load().rtt_init(device), "rtt_init")
    _initialised = True


def initialised() -> bool:
    """True once init() has opened a device in this process."""
    return _initialised


def sync() -> None:
    _check(load().rtt_sync(), "rtt_sync")


class PreparedA:
    """Owns an rtt_prepared_a handle (A in 8x4 hybrid tiles on the GPU)."""

    def __init__(self, handle: C.c_void_p, stats: OfflineAStats, csr: CsrMatrix):
        self.handle = handle
        self.stats = {f: getattr(stats, f) for f, _ in OfflineAStats._fields_}
        self.m, self.n = csr.m, csr.n

    def close(self) -> None:
        """Destroys the device handle now (also done on garbage collection)."""
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rtt_destroy_prepared_a(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def tiles(self) -> dict:
        v = DeviceTiles()
        _check(load().rtt_prepared_a_view(self.handle, C.byref(v)), "prepared_a_view")
        return _download_tiles(v)

    def info(self) -> tuple[int, int, bool]:
        ar, ai, ident = C.c_int(), C.c_int(), C.c_int()
        _check(load().rtt_prepared_a_info(self.handle, C.byref(ar), C.byref(ai),
                                          C.byref(ident)), "prepared_a_info")
        return ar.value, ai.value, bool(ident.value)

    def row_new_to_old(self):
        p = load().rtt_prepared_a_row_new_to_old(self.handle)
        return None if not p else _np_copy(p, self.m, np.int32)


def prepare_identity_a(a: CsrMatrix, dense_threshold: int = 24) -> PreparedA:
    s = a.as_smatrix()
    h = C.c_void_p()
    st = OfflineAStats()
    _check(load().rtt_prepare_identity_a(C.byref(s), dense_threshold, C.byref(h),
                                         C.byref(st)), "prepare_identity_a")
    return PreparedA(h, st, a)


def prepare_external_a(a: CsrMatrix, row_new_to_old, inner_new_to_old,
                       dense_threshold: int = 24) -> PreparedA:
    s = a.as_smatrix()
    r = np.ascontiguousarray(row_new_to_old, dtype=np.int32)
    i = np.ascontiguousarray(inner_new_to_old, dtype=np.int32)
    h = C.c_void_p()
    st = OfflineAStats()
    _check(load().rtt_prepare_external_a(C.byref(s), _ptr(r), _ptr(i), dense_threshold,
                                         C.byref(h), C.byref(st)), "prepare_external_a")
    return PreparedA(h, st, a)


def bgrf_plan(a: CsrMatrix):
    """BGRF-v1 plan (build_dmma_reorder_plan, src/dmma_reorder.h:2651):
    (row_new_to_old, inner_new_to_old, stats).  Host only."""
    s = a.as_smatrix()
    r = np.zeros(max(a.m, 1), np.int32)
    i = np.zeros(max(a.n, 1), np.int32)
    st = BgrfStats()
    _check(load().rtt_bgrf_plan(C.byref(s), _ptr(r), _ptr(i), C.byref(st)), "bgrf_plan")
    return r[:a.m], i[:a.n], {f: getattr(st, f) for f, _ in BgrfStats._fields_}


def _stats_struct(stats: dict | None):
    if stats is None:
        return None
    st = BgrfStats()
    for k, v in stats.items():
        setattr(st, k, v)
    return C.byref(st)


def dump_reorder(prefix: str, a: CsrMatrix, row_new_to_old=None, inner_new_to_old=None,
                 stats: dict | None = None) -> None:
    """<prefix>_permutations.csv, _unified_stats.csv, _A_reordered.mtx
    (dump_dmma_reorder_plan, src/dmma_reorder.h:3086)."""
    s = a.as_smatrix()
    r = None if row_new_to_old is None else np.ascontiguousarray(row_new_to_old, np.int32)
    i = None if inner_new_to_old is None else np.ascontiguousarray(inner_new_to_old, np.int32)
    _check(load().rtt_dump_reorder(prefix.encode(), C.byref(s), None if r is None else _ptr(r),
                                   None if i is None else _ptr(i), _stats_struct(stats)),
           "dump_reorder")


def dump_reorder_heatmap(prefix: str, a: CsrMatrix, row_new_to_old=None, inner_new_to_old=None,
                         bins: int = 256, stats: dict | None = None) -> None:
    """<prefix>_heatmap.csv and _heatmap_meta.csv (src/main.cu:1114-1246)."""
    s = a.as_smatrix()
    r = None if row_new_to_old is None else np.ascontiguousarray(row_new_to_old, np.int32)
    i = None if inner_new_to_old is None else np.ascontiguousarray(inner_new_to_old, np.int32)
    _check(load().rtt_dump_reorder_heatmap(prefix.encode(), C.byref(s),
                                           None if r is None else _ptr(r),
                                           None if i is None else _ptr(i), int(bins),
                                           _stats_struct(stats)), "dump_reorder_heatmap")


def read_timeline(path: str) -> dict:
    """Reads a numeric timeline trace (the DMTL0001 layout of
    dmma_write_timeline_trace, src/dmma_spgemm.h:95-164), written by the
    library when DMMA_TRACE_FILE is set: header fields and per-slot start and
    end times (ns) and compute-unit ids."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"DMTL0001":
        raise ValueError(f"{path}: not a DMTL0001 timeline")
    head = np.frombuffer(data, np.uint32, 5, 8)
    tail = np.frombuffer(data, np.uint64, 3, 28)
    slots = int(tail[2])
    off = 52
    start = np.frombuffer(data, np.uint64, slots, off)
    end = np.frombuffer(data, np.uint64, slots, off + 8 * slots)
    cu = np.frombuffer(data, np.uint32, slots, off + 16 * slots)
    return {"version": int(head[0]), "compute_units": int(head[1]),
            "waves_per_workgroup": int(head[2]), "shift": int(head[3]), "phase": int(head[4]),
            "output_tiles": int(tail[0]), "workgroups": int(tail[1]), "slots": slots,
            "start_ns": start, "end_ns": end, "cu": cu}


def prepare_bgrf_a(a: CsrMatrix, dense_threshold: int = 24) -> PreparedA:
    """A in the BGRF-v1 layout (the reference driver's default)."""
    s = a.as_smatrix()
    h = C.c_void_p()
    st = OfflineAStats()
    bs = BgrfStats()
    _check(load().rtt_prepare_bgrf_a(C.byref(s), dense_threshold, C.byref(h), C.byref(st),
                                     C.byref(bs)), "prepare_bgrf_a")
    pa = PreparedA(h, st, a)
    pa.bgrf = {f: getattr(bs, f) for f, _ in BgrfStats._fields_}
    return pa


class DynamicB:
    """Owns an rtt_dynamic_b handle (B in 4x8 hybrid tiles + tile CSC)."""

    def __init__(self):
        self.handle = C.c_void_p()
        self.last = {}

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value and _lib is not None:
            _lib.rtt_destroy_dynamic_b(self.handle)
            self.handle = C.c_void_p()

    def tiles(self) -> dict:
        v = DeviceTiles()
        _check(load().rtt_dynamic_b_view(self.handle, C.byref(v)), "dynamic_b_view")
        return _download_tiles(v)


def rebuild_b(a: PreparedA, aat: int, b: DynamicB | None = None) -> DynamicB:
    b = b or DynamicB()
    st = BUpdateStats()
    _check(load().rtt_rebuild_b(a.handle, int(aat), C.byref(b.handle), C.byref(st)),
           "rebuild_b")
    b.last = {f: getattr(st, f) for f, _ in BUpdateStats._fields_}
    return b


def nnz_cub(a: PreparedA, aat: int) -> int:
    v = C.c_ulonglong()
    ms = C.c_double()
    _check(load().rtt_compute_nnz_cub(a.handle, int(aat), C.byref(v), C.byref(ms)), "nnz_cub")
    return int(v.value)


class DeviceCsr:
    """Owns an rtt_device_csr handle: an independent B uploaded to the GPU
    (DmmaOwnedDeviceCsr, src/gpu_dmma_tiles.h:32)."""

    def __init__(self, csr: CsrMatrix):
        s = csr.as_smatrix()
        self.handle = C.c_void_p()
        h2d, val = C.c_double(), C.c_double()
        _check(load().rtt_upload_csr(C.byref(s), C.byref(self.handle), C.byref(h2d),
                                     C.byref(val)), "upload_csr")
        self.m, self.n, self.nnz = csr.m, csr.n, csr.nnz
        self.stats = {"h2d_ms": h2d.value, "validation_ms": val.value}

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value and _lib is not None:
            _lib.rtt_destroy_device_csr(self.handle)
            self.handle = C.c_void_p()

    def set_values(self, values) -> None:
        """New values, in the entry order of the CSR this handle was uploaded
        from; the structure stays."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (self.nnz,):
            raise ValueError(f"expected {self.nnz} values, got shape {v.shape}")
        _check(load().rtt_device_csr_set_values(self.handle, _ptr(v)), "device_csr_set_values")


def upload_csr(csr: CsrMatrix) -> DeviceCsr:
    """gpu_upload_csr (src/gpu_dmma_tiles.h:2259) for an independent B."""
    return DeviceCsr(csr)


def rebuild_b_general(a: PreparedA, src: DeviceCsr, b: DynamicB | None = None) -> DynamicB:
    """B for C = A*B from an independent CSR (--b, src/main.cu:1417-1420)."""
    b = b or DynamicB()
    st = BUpdateStats()
    _check(load().rtt_rebuild_b_general(a.handle, src.handle, C.byref(b.handle), C.byref(st)),
           "rebuild_b_general")
    b.last = {f: getattr(st, f) for f, _ in BUpdateStats._fields_}
    return b


def nnz_cub_ab(a: PreparedA, src: DeviceCsr) -> int:
    v = C.c_ulonglong()
    ms = C.c_double()
    _check(load().rtt_compute_nnz_cub_ab(a.handle, src.handle, C.byref(v), C.byref(ms)),
           "nnz_cub_ab")
    return int(v.value)


def choose_b_values_clear(policy: int, values_only_update=1, dynamic_b_valid=1,
                          has_active_duplicates=0, active_source_mapping_complete=1,
                          active_source_to_payload_injective=1, payload_fully_initialized=1,
                          every_payload_slot_overwritten=1, numeric_b_read_only=1) -> dict:
    """dmma_choose_b_values_clear (src/dmma_b_values_clear_policy.h:91-160)."""
    d = BValuesDecision()
    lib = load()
    _check(lib.rtt_choose_b_values_clear(
        int(policy), int(values_only_update), int(dynamic_b_valid), int(has_active_duplicates),
        int(active_source_mapping_complete), int(active_source_to_payload_injective),
        int(payload_fully_initialized), int(every_payload_slot_overwritten),
        int(numeric_b_read_only), C.byref(d)), "choose_b_values_clear")
    out = {f: getattr(d, f) for f, _ in BValuesDecision._fields_}
    out["reason_name"] = lib.rtt_b_values_clear_reason_name(d.reason).decode()
    return out


def update_b_values(a: PreparedA, src: DeviceCsr, b: DynamicB,
                    policy: str = "safe-auto") -> dict:
    """Values-only B update: same structure as the build from `src`, new
    values (DeviceCsr.set_values).  Returns the clear decision."""
    st = BUpdateStats()
    d = BValuesDecision()
    lib = load()
    _check(lib.rtt_update_b_values(a.handle, src.handle, B_VALUES_POLICIES[policy], b.handle,
                                   C.byref(st), C.byref(d)), "update_b_values")
    b.last = {f: getattr(st, f) for f, _ in BUpdateStats._fields_}
    out = {f: getattr(d, f) for f, _ in BValuesDecision._fields_}
    out["reason_name"] = lib.rtt_b_values_clear_reason_name(d.reason).decode()
    return out


NUMERIC_PATHS = {"auto": 0, "vector": 1, "tile": 2}
PRECISIONS = {"fp64": 0, "split-bf16": 1}


def schedule(lb_mode: str = "adaptive-tail", materialize_output: bool = True,
             numeric_path: str = "auto", precision: str = "fp64", **kw) -> Schedule:
    s = Schedule()
    load().rtt_schedule_defaults(C.byref(s))
    s.lb_mode = LB_MODES[lb_mode]
    s.materialize_output = int(materialize_output)
    s.numeric_path = NUMERIC_PATHS[numeric_path]
    s.precision = PRECISIONS[precision]
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def tilespgemm(a: PreparedA, b: DynamicB, sched: Schedule | None = None):
    """Runs C = A*B; returns (TileC or None, stats dict)."""
    sched = sched or schedule()
    out = SMatrix()
    st = Stats()
    _check(load().rtt_tilespgemm(a.handle, b.handle, C.byref(out), C.byref(st),
                                 C.byref(sched)), "tilespgemm")
    stats = {f: getattr(st, f) for f, _ in Stats._fields_}
    tc = None
    if sched.materialize_output:
        try:
            tc = _tilec_from_smatrix(out)
        finally:
            load().rtt_destroy_smatrix(C.byref(out))
    return tc, stats


def batch_extract(a: PreparedA, row_offset: int, col_offset: int, m: int, n: int) -> TileC:
    """Problem block (m x n at the 8-aligned offsets) of the last C on `a`."""
    out = SMatrix()
    _check(load().rtt_batch_extract(a.handle, C.byref(C.c_int64(int(row_offset))),
                                    C.byref(C.c_int64(int(col_offset))), m, n, C.byref(out)),
           "batch_extract")
    try:
        return _tilec_from_smatrix(out)
    finally:
        load().rtt_destroy_smatrix(C.byref(out))


def autotune(a: PreparedA, b: DynamicB, repeats: int = 2, sched: Schedule | None = None):
    """Fastest schedule for (a, b) among the tuner's candidates, and the table."""
    sched = sched or schedule(materialize_output=False)
    rep = TuneReport()
    _check(load().rtt_autotune(a.handle, b.handle, repeats, C.byref(sched), C.byref(rep)),
           "autotune")
    inv = {v: k for k, v in NUMERIC_PATHS.items()}
    lbinv = {v: k for k, v in LB_MODES.items()}
    table = [{"numeric_path": inv[rep.numeric_path[i]], "fill": rep.fill[i],
              "lb_mode": lbinv[rep.lb_mode[i]], "ms": rep.ms[i]} for i in range(rep.candidates)]
    return sched, table, rep.chosen


def tile_values_noalloc(a: PreparedA, b: DynamicB, precision: str, c_values_ptr: int) -> None:
    """Recomputes the last C's values into device memory at c_values_ptr
    (nnz doubles the caller owns), allocating nothing; returns after queuing."""
    _check(load().rtt_tile_values_noalloc(a.handle, b.handle, PRECISIONS[precision],
                                          C.c_void_p(int(c_values_ptr))), "tile_values_noalloc")


def stream() -> int:
    """The HIP stream (hipStream_t, as an integer) the library queues its
    work on (rtt_stream)."""
    h = C.c_void_p()
    _check(load().rtt_stream(C.byref(h)), "stream")
    return h.value or 0


def last_tail_records(a: PreparedA, max_records: int = 1 << 20) -> dict:
    """Heavy C tiles admitted by the last call's tail queue (slot order)."""
    task = np.zeros(max(max_records, 1), np.int64)
    sc, mt, ch = (np.zeros(max(max_records, 1), np.int32) for _ in range(3))
    n = C.c_int()
    _check(load().rtt_last_tail_records(a.handle, max_records, _ptr(task), _ptr(sc), _ptr(mt),
                                        _ptr(ch), C.byref(n)), "last_tail_records")
    k = n.value
    return {"task": task[:k], "scans": sc[:k], "matches": mt[:k], "chunks": ch[:k]}


def last_c_tile_ptr(a: PreparedA) -> np.ndarray:
    """C's 64-bit tile row pointer (tilem + 1) of the last product on `a`."""
    v = DeviceC()
    _check(load().rtt_last_c_view(a.handle, C.byref(v)), "last_c_view")
    out = np.zeros(v.tilem + 1, np.int64)
    _check(load().rtt_memcpy_d2h(_ptr(out), C.c_void_p(v.tile_ptr), out.nbytes), "memcpy_d2h")
    return out


def last_c_row_stats(a: PreparedA):
    """(nnz, value sum) of every row of the last C on `a`, computed on the GPU."""
    n = np.zeros(max(a.m, 1), np.int64)
    v = np.zeros(max(a.m, 1), np.float64)
    _check(load().rtt_last_c_row_stats(a.handle, _ptr(n), _ptr(v)), "last_c_row_stats")
    return n[:a.m], v[:a.m]


def last_c(a: PreparedA, with_values: bool = True) -> dict:
    """Downloads the 64-bit device C of the last tilespgemm on `a`."""
    v = DeviceC()
    _check(load().rtt_last_c_view(a.handle, C.byref(v)), "last_c_view")
    T, Z, R = v.num_tiles, v.nnz, v.tilem
    out = {"tilem": R, "tilen": v.tilen, "num_tiles": T, "nnz": Z}
    items = [("tile_ptr", v.tile_ptr, R + 1, np.int64), ("tile_col", v.tile_col, T, np.int32),
             ("tile_mask", v.tile_mask, T, np.uint64), ("tile_nnz", v.tile_nnz, T + 1, np.int64),
             ("tile_csr_ptr", v.tile_csr_ptr, 8 * T, np.uint8),
             ("tile_csr_col", v.tile_csr_col, Z, np.uint8)]
    if with_values:
        items.append(("tile_csr_val", v.tile_csr_val, Z, np.float64))
    for name, p, count, dt in items:
        arr = np.zeros(count, dt)
        if count:
            _check(load().rtt_memcpy_d2h(_ptr(arr), C.c_void_p(p), arr.nbytes), "memcpy_d2h")
        out[name] = arr
    return out


def tile2csr(tc: TileC, canonical: bool = True) -> CsrMatrix:
    s = tc.as_smatrix()
    _check(load().rtt_tile2csr(C.byref(s), int(canonical)), "tile2csr")
    try:
        return CsrMatrix(s.m, s.n, _np_copy(s.rowpointer, s.m + 1, np.int32),
                         _np_copy(s.columnindex, s.nnz, np.int32),
                         _np_copy(s.value, s.nnz, np.float64))
    finally:
        _free_csr_part(s)


def tile2csr_restore_rows(tc: TileC, row_new_to_old, permutation_size: int,
                          canonical: bool = True) -> CsrMatrix:
    s = tc.as_smatrix()
    perm = np.ascontiguousarray(row_new_to_old, dtype=np.int32)
    _check(load().rtt_tile2csr_restore_rows(C.byref(s), _ptr(perm), permutation_size,
                                            int(canonical)), "tile2csr_restore_rows")
    try:
        return CsrMatrix(s.m, s.n, _np_copy(s.rowpointer, s.m + 1, np.int32),
                         _np_copy(s.columnindex, s.nnz, np.int32),
                         _np_copy(s.value, s.nnz, np.float64))
    finally:
        _free_csr_part(s)


def _free_csr_part(s: SMatrix) -> None:
    # Only the CSR arrays belong to the library here; the tile arrays are
    # numpy-owned, so free the CSR part through a separate struct.
    t = SMatrix()
    t.rowpointer, t.columnindex, t.value = s.rowpointer, s.columnindex, s.value
    load().rtt_destroy_smatrix(C.byref(t))


def abi_symbols(header: str = HEADER_PATH) -> list[str]:
    """Function names declared in the public header."""
    import re
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(rtt_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


# ---- vendor-library control ------------------------------------------------

CONTROL_PATH = os.path.join(os.path.dirname(_HERE), "lib", "librtt_control.so")


def control_rocsparse(a: CsrMatrix, aat: int, warmup: int = 1, iterations: int = 3) -> dict:
    """C = A*A (aat 0) or A*A^T (aat 1) with rocsparse_spgemm, timed as the
    reference's cuSPARSE control (src/cusparse_core_benchmark.h:4-30).  A
    separate library (lib/librtt_control.so); raises when it is missing."""
    if not os.path.exists(CONTROL_PATH):
        raise RuntimeError(f"{CONTROL_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(CONTROL_PATH)
    lib.rtt_control_rocsparse.restype = C.c_int
    lib.rtt_control_rocsparse.argtypes = [C.POINTER(SMatrix), C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                          C.POINTER(C.c_double)]
    lib.rtt_control_last_error.restype = C.c_char_p
    s = a.as_smatrix()
    ms = C.c_double()
    nnz = C.c_longlong()
    chk = C.c_double()
    if lib.rtt_control_rocsparse(C.byref(s), aat, warmup, iterations, C.byref(ms), C.byref(nnz),
                                 C.byref(chk)) != 0:
        raise RuntimeError("rocSPARSE control: " + lib.rtt_control_last_error().decode())
    return {"ms": ms.value, "nnz_c": int(nnz.value), "checksum": chk.value}
