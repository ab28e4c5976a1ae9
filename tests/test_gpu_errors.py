"""GPU-side error paths: malformed CSR input is rejected by the device
validation (validate_csr_kernel; the reference's validate_device_csr,
src/gpu_dmma_tiles.h:236-255,777-918), and a B is bound to the prepared A it
was built for.  Every case is refused with RTT_ERR_INVALID before any
product kernel runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INVALID = 1


def bad(rt, rowptr, col, m=4, n=5):
    rowptr = np.asarray(rowptr, np.int32)
    col = np.asarray(col, np.int32)
    return rt.CsrMatrix(m, n, rowptr, col, np.ones(len(col)))


@pytest.mark.parametrize("case", ["non_monotone", "column_too_large", "negative_column",
                                  "nonzero_start", "row_beyond_nnz", "negative_start"])
def test_device_validation_rejects_malformed_csr(gpu, case):
    rt = gpu
    good = dict(rowptr=[0, 2, 3, 5, 6], col=[0, 3, 1, 2, 4, 0])
    if case == "non_monotone":
        m = bad(rt, [0, 3, 2, 5, 6], good["col"])
    elif case == "column_too_large":
        m = bad(rt, good["rowptr"], [0, 3, 1, 2, 5, 0])
    elif case == "negative_column":
        m = bad(rt, good["rowptr"], [0, 3, -1, 2, 4, 0])
    elif case == "nonzero_start":
        m = bad(rt, [1, 2, 3, 5, 6], good["col"])
    elif case == "row_beyond_nnz":
        m = bad(rt, [0, 2, 9, 5, 6], good["col"])
    else:
        m = bad(rt, [0, -3, 3, 5, 6], good["col"])
    print(f"case {case}", flush=True)
    with pytest.raises(rt.RttError) as e:
        rt.upload_csr(m)
    assert e.value.code == INVALID
    # the well-formed matrix still goes through afterwards
    rt.upload_csr(bad(rt, **good))


def test_b_is_bound_to_its_prepared_a(gpu):
    rt = gpu
    a1 = rt.gen_random(64, 64, 3, 1)
    a2 = rt.gen_random(64, 64, 9, 2)        # conforming, more tiles
    p1 = rt.prepare_identity_a(a1)
    p2 = rt.prepare_identity_a(a2)
    b1 = rt.rebuild_b(p1, 0)
    with pytest.raises(rt.RttError) as e:
        rt.tilespgemm(p2, b1)
    assert e.value.code == INVALID and "another prepared A" in str(e.value)
    src = rt.upload_csr(rt.gen_random(64, 40, 4, 3))
    bg = rt.rebuild_b_general(p1, src)
    with pytest.raises(rt.RttError):
        rt.update_b_values(p2, src, bg, "safe-auto")
    # rebuilding for the other A rebinds it
    b1 = rt.rebuild_b(p2, 0, b1)
    rt.tilespgemm(p2, b1)
    with pytest.raises(rt.RttError):
        rt.tilespgemm(p1, b1)


def test_transposed_b_refuses_a_destroyed_a(gpu):
    rt = gpu
    a = rt.gen_random(48, 48, 5, 4)
    pa = rt.prepare_identity_a(a)
    pb = rt.rebuild_b(pa, 1)
    pb.tiles()
    pa.close()
    with pytest.raises(rt.RttError):
        pb.tiles()
