"""The no-allocation entry point (rtt_tile_values_noalloc) against the entry
point it varies (rtt_tilespgemm on the TILE path): the same kernels in the
same order, so C's values must agree bit for bit.  Two calls in a row, the
second on a smaller problem (the scratch array keeps the first call's
contents and must be cleared by the second), and one after a values-only B
update.  The output buffers are device memory the test owns, allocated
through the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Caller-owned device memory (hipMalloc of the runtime already loaded)."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so.7")   # the runtime the library links
        self.ptr = C.c_void_p()
        self.nbytes = max(int(nbytes), 8)
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        # NaN bytes everywhere: entries the call fails to write show up
        assert self.hip.hipMemset(self.ptr, 0xFF, C.c_size_t(self.nbytes)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def download(self, n):
        out = np.zeros(n, np.float64)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                      C.c_size_t(8 * n), 2) == 0   # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip.hipDeviceSynchronize()
            self.hip.hipFree(self.ptr)


def values_into_buffer(rt, pa, pb, precision, n):
    buf = DeviceBuffer(8 * n)
    rt.tile_values_noalloc(pa, pb, precision, buf.ptr.value)
    rt.sync()
    return buf.download(n)


def check(rt, a, aat, precision):
    pa = rt.prepare_identity_a(a)
    pb = rt.rebuild_b(pa, aat)
    _, st = rt.tilespgemm(pa, pb, rt.schedule(materialize_output=False, numeric_path="tile",
                                              precision=precision))
    want = rt.last_c(pa)["tile_csr_val"]
    got = values_into_buffer(rt, pa, pb, precision, len(want))
    assert got.size == st["output_nnz"] > 0
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    return pa, pb


@pytest.mark.parametrize("precision", ["fp64", "split-bf16"])
def test_noalloc_values_equal_the_tile_path(gpu, precision):
    check(gpu, gpu.gen_fem_blocks(1500, 6, True, 31), 0, precision)
    # second call, smaller problem
    check(gpu, gpu.gen_fem_blocks(400, 6, True, 32), 0, precision)
    # A*A^T (B's columns are A's tile rows)
    check(gpu, gpu.gen_random(300, 260, 12, 33), 1, precision)


def test_noalloc_after_a_values_update(gpu):
    rt = gpu
    a = rt.gen_random(400, 300, 10, 41)
    b = rt.gen_random(300, 350, 8, 42)
    pa = rt.prepare_identity_a(a)
    src = rt.upload_csr(b)
    pb = rt.rebuild_b_general(pa, src)
    sched = rt.schedule(materialize_output=False, numeric_path="tile")
    rt.tilespgemm(pa, pb, sched)
    src.set_values(b.val * 3.0 - 1.0)
    rt.update_b_values(pa, src, pb, "safe-auto")
    n = rt.last_c(pa, with_values=False)["nnz"]
    got = values_into_buffer(rt, pa, pb, "fp64", n)
    rt.tilespgemm(pa, pb, sched)
    want = rt.last_c(pa)["tile_csr_val"]
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_noalloc_refuses_without_a_tile_structure(gpu):
    rt = gpu
    a = rt.gen_stencil27(6, 0.5, 1.5, 3)
    pa = rt.prepare_identity_a(a)
    pb = rt.rebuild_b(pa, 1)
    rt.tilespgemm(pa, pb, rt.schedule(materialize_output=False, numeric_path="vector"))
    n = rt.last_c(pa, with_values=False)["nnz"]
    # a buffer large enough for C, so that a wrongly accepted call stays in bounds
    out = DeviceBuffer(8 * n)
    with pytest.raises(rt.RttError):
        rt.tile_values_noalloc(pa, pb, "fp64", out.ptr.value)
    rt.tilespgemm(pa, pb, rt.schedule(materialize_output=False, numeric_path="tile"))
    pb = rt.rebuild_b(pa, 1, pb)   # rebuilt since: refused as well
    with pytest.raises(rt.RttError):
        rt.tile_values_noalloc(pa, pb, "fp64", out.ptr.value)
    rt.sync()


def test_noalloc_call_replays_from_a_hip_graph(gpu):
    """The allocation-free call queues kernels only (its scratch is cleared
    by a kernel, not a memset) and does not synchronise, so it can be
    captured on the library's stream (rtt_stream) into a HIP graph.  The
    graph is replayed three times, with the program's own eager work and a
    synchronisation between replays, and every replay must give the eager
    call's values bit for bit."""
    rt = gpu
    a = rt.gen_fem_blocks(1200, 6, True, 51)
    pa = rt.prepare_identity_a(a)
    pb = rt.rebuild_b(pa, 0)
    rt.tilespgemm(pa, pb, rt.schedule(materialize_output=False, numeric_path="tile"))
    want = rt.last_c(pa)["tile_csr_val"]
    n = len(want)
    assert n > 0
    out = DeviceBuffer(8 * n)
    other = DeviceBuffer(8 * n)
    hip = out.hip
    vp = C.c_void_p
    hip.hipStreamBeginCapture.argtypes = [vp, C.c_int]
    hip.hipStreamEndCapture.argtypes = [vp, C.POINTER(vp)]
    hip.hipGraphInstantiate.argtypes = [C.POINTER(vp), vp, vp, vp, C.c_size_t]
    hip.hipGraphLaunch.argtypes = [vp, vp]
    hip.hipGraphExecDestroy.argtypes = [vp]
    hip.hipGraphDestroy.argtypes = [vp]
    stream = vp(rt.stream())
    graph, exe = vp(), vp()
    rt.sync()
    assert hip.hipStreamBeginCapture(stream, 2) == 0   # hipStreamCaptureModeRelaxed
    try:
        rt.tile_values_noalloc(pa, pb, "fp64", out.ptr.value)
    finally:
        end = hip.hipStreamEndCapture(stream, C.byref(graph))
    assert end == 0 and graph.value
    assert hip.hipGraphInstantiate(C.byref(exe), graph, None, None, 0) == 0
    try:
        for replay in range(3):
            # Eager work between the replays: the output poisoned with NaN
            # bytes, and the same call run eagerly into another buffer.
            assert out.hip.hipMemset(out.ptr, 0xFF, C.c_size_t(8 * n)) == 0
            rt.tile_values_noalloc(pa, pb, "fp64", other.ptr.value)
            rt.sync()
            assert hip.hipGraphLaunch(exe, stream) == 0
            rt.sync()
            got = out.download(n)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), replay
            assert np.array_equal(other.download(n).view(np.uint64), want.view(np.uint64))
    finally:
        hip.hipGraphExecDestroy(exe)
        hip.hipGraphDestroy(graph)
