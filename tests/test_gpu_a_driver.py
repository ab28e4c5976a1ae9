"""The drop-in driver (bin/rtt_spgemm, the reference's src/main.cu flags and
RTT_RESULT line) on small inputs: C = A*A, C = A*A^T and general A*B with a
timed values-only B update.  nnz(C) and the result line are checked against
the oracle.

The driver runs as a child process.  A child is only started while this test
process has not opened the GPU itself (this file is collected before the
other GPU tests); otherwise the tests skip.
"""
import re
import subprocess

import numpy as np
import pytest

import oracle
import rtt_amd

pytestmark = pytest.mark.gpu


def run_driver(args, timeout=120):
    if rtt_amd.initialised():
        pytest.skip("the GPU is open in this process; no child process is started from it")
    p = subprocess.run([rtt_amd.DRIVER_PATH] + args, capture_output=True, text=True,
                       timeout=timeout)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


def parse(out):
    m = re.search(r"^C: (\d+) x (\d+), nnz (\d+), tiles (\d+)$", out, re.M)
    r = re.search(r"^RTT_RESULT (.*)$", out, re.M)
    assert m and r, out[-2000:]
    fields = dict(kv.split("=", 1) for kv in r.group(1).split())
    return tuple(int(x) for x in m.groups()), fields


@pytest.mark.parametrize("aat", [0, 1])
def test_driver_aa_and_aat(tmp_path, aat):
    a = rtt_amd.gen_stencil27(6, 0.5, 1.5, 3)
    path = str(tmp_path / "a.mtx")
    rtt_amd.write_mtx(path, a)
    out = run_driver(["-aat", str(aat), "--iterations", "3", "--warmup-iterations", "1", path])
    (m, n, nnz, _), f = parse(out)
    A = (a.m, a.n, a.rowptr, a.col, a.val)
    B = (a.n, a.m, *oracle.transpose(*A)) if aat else A
    ref = oracle.spgemm(A, B)
    assert (m, n, nnz) == (a.m, a.m if aat else a.n, ref["nnz"])
    assert f["matrix"] == path and f["mode"] == "adaptive-tail" and f["repeats"] == "3"
    assert float(f["gflops"]) > 0 and float(f["runtime_ms"]) > 0


def test_driver_general_ab_values_update(tmp_path):
    a = rtt_amd.gen_random(120, 90, 5, 61)
    b = rtt_amd.gen_random(90, 70, 6, 62)
    pa, pb = str(tmp_path / "a.mtx"), str(tmp_path / "b.mtx")
    rtt_amd.write_mtx(pa, a)
    rtt_amd.write_mtx(pb, b)
    out = run_driver(["--b", pb, "--b-update", "values", "--iterations", "2", "-aat", "1", pa])
    assert "Note: --b selects general A*B; -aat is ignored." in out
    assert re.search(r"^B_UPDATE mode=values path=4 clear=0 reason=eligible-safe-auto$", out,
                     re.M), out[-1500:]
    (m, n, nnz, _), _ = parse(out)
    ref = oracle.spgemm((a.m, a.n, a.rowptr, a.col, a.val), (b.m, b.n, b.rowptr, b.col, b.val))
    assert (m, n, nnz) == (a.m, b.n, ref["nnz"])
    bad = subprocess.run([rtt_amd.DRIVER_PATH, "--b-update", "sideways", pa],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "Usage" in bad.stdout


def test_driver_reorder_dumps(tmp_path):
    a = rtt_amd.gen_rmat(9, 8, 0.57, 0.19, 0.19, 4)
    path = str(tmp_path / "a.mtx")
    rtt_amd.write_mtx(path, a)
    prefix = str(tmp_path / "P")
    out = run_driver(["--prepare-only", "--dump-reorder-prefix", prefix,
                      "--dump-reorder-heatmap", prefix, "--heatmap-bins", "32", path])
    assert "BGRF plan" in out and f"REORDER_DUMP prefix={prefix}" in out
    lines = open(prefix + "_permutations.csv").read().splitlines()
    assert len(lines) == 1 + a.m + a.n
    stats = dict(l.split(",", 1) for l in open(prefix + "_unified_stats.csv").read().split()[1:])
    assert stats["algorithm"] == "bgrf-v1"
    grid = np.loadtxt(prefix + "_heatmap.csv", delimiter=",", skiprows=1, dtype=np.int64)
    assert grid.shape == (32 * 32, 4) and grid[:, 3].sum() == a.nnz
    r, i, st = rtt_amd.bgrf_plan(a)
    assert int(stats["num_tiles"]) == st["tiles_final"]
