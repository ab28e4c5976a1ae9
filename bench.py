"""Benchmark: tiled FP64 SpGEMM on one MI355X.

A step is one pass of the reference's timed boundary over the workload:
B rebuild + symbolic + allocate + numeric (A preparation is offline, as in
src/main.cu:2171-2318).  Inputs are resident in HBM before timing starts;
the host export of C is outside the timed region.

Workload (default): C = A*A^T for the 3D 27-point stencil on a 128^3 grid
(2,097,152 rows, 55,742,968 nnz, values U[0.5, 1.5], seed 20260101).

Prints one JSON line with the metric, the roofline of the dominant kernel
(HIP-event durations measured on the library's stream inside the timed
region) and the CPU baseline (the oracle's row SPA on a bounded row sample,
timed on this host).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "rtt-spgemm_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = ("SpGEMM GFLOP/s + roofline %, split-prec/batched vs plain FP64 at "
          "≤1e-12 rel err")
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "stencil27-128-aat": dict(kind="stencil27", g=128, aat=1),
    "stencil27-64-aat": dict(kind="stencil27", g=64, aat=1),
    "laplace2d-64-aa": dict(kind="laplace2d", g=64, aat=0),
    "rmat-20-aa": dict(kind="rmat", scale=20, aat=0),
    "rmat-18-aa": dict(kind="rmat", scale=18, aat=0),
    "fem-500k-aa": dict(kind="fem", nodes=500000, aat=0),
    "fem-100k-aa": dict(kind="fem", nodes=100000, aat=0),
}
FP64_MFMA_PEAK_TFS = 78.6
MFMA_PAIR_FLOPS = 512   # v_mfma_f64_4x4x4f64: 4 blocks x 4x4x4 x 2


def make_matrix(rt, spec):
    if spec["kind"] == "stencil27":
        return rt.gen_stencil27(spec["g"], 0.5, 1.5, 20260101), (
            f"3D 27-point stencil {spec['g']}^3, values U[0.5,1.5], seed 20260101")
    if spec["kind"] == "fem":
        return rt.gen_fem_blocks(spec["nodes"], 6, True, 20260101), (
            f"block FEM, {spec['nodes']} nodes uniform in the unit cube, symmetrised 6 nearest "
            f"neighbours + self, dense 4x4 blocks, magnitudes 10^U[-3,3] with random signs, "
            f"Morton node numbering, seed 20260101")
    if spec["kind"] == "rmat":
        return rt.gen_rmat(spec["scale"], 16, 0.57, 0.19, 0.19, 20260101), (
            f"R-MAT 2^{spec['scale']}, 16 draws per row, (a,b,c,d) = (0.57,0.19,0.19,0.05), "
            f"duplicates removed, values U[-1,1], seed 20260101")
    return rt.gen_laplace2d(spec["g"], 0.1, 20260101), (
        f"2D 5-point Laplacian {spec['g']}^2, off-diagonals -1+U[-0.1,0.1], seed 20260101")


# ---- batched workloads -----------------------------------------------------

BATCHES = {
    # 512 independent banded A^2 problems (log-normal row counts, median 16k,
    # clipped to [1k, 200k]; bandwidth U{8..64}; U{5..30} entries per row
    # inside the band plus the diagonal; values U[-1,1]).
    "banded-batch-512": dict(kind="banded", problems=512),
    "banded-batch-64": dict(kind="banded", problems=64),
    # AMG Galerkin products R*(A*P) of 7-point Poisson hierarchies on g^3
    # grids, g ~ U{32..96}, coefficient jumps 10^U[-3,3], smoothed
    # aggregation P (2x2x2 aggregates, omega = 2/3), R = P^T.
    "amg-batch-64": dict(kind="amg", problems=64),
    "amg-batch-8": dict(kind="amg", problems=8),
}


def make_batch(rt, spec, seed=20260101):
    if spec["kind"] == "banded":
        rows, bw, pr = rt.gen_banded_batch_shapes(spec["problems"], seed=seed)
        mats = [rt.gen_banded(int(rows[p]), int(bw[p]), int(pr[p]), seed + p)
                for p in range(spec["problems"])]
        return {"A": mats, "desc": f"{spec['problems']} banded problems, rows "
                f"{int(rows.min())}..{int(rows.max())} (median {int(np.median(rows))}), "
                f"sum {int(rows.sum())} rows, {sum(m.nnz for m in mats)} nnz"}
    gs = np.random.default_rng(seed).integers(32, 97, spec["problems"])
    A = [rt.gen_poisson7_jumps(int(g), 3.0, seed + p) for p, g in enumerate(gs)]
    P = [rt.gen_aggregation_prolongator(a, int(g), 2.0 / 3.0) for a, g in zip(A, gs)]
    R = [rt.csr_transpose(p) for p in P]
    return {"A": A, "P": P, "R": R, "g": gs,
            "desc": f"{spec['problems']} AMG hierarchies, g {int(gs.min())}..{int(gs.max())}, "
                    f"sum {int(sum(int(g) ** 3 for g in gs))} fine rows, "
                    f"{sum(a.nnz for a in A)} nnz(A), {sum(p.nnz for p in P)} nnz(P)"}


def timed_steps(step, steps, warmup):
    import rtt_amd as rt
    for _ in range(warmup):
        step()
    rt.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    rt.sync()
    return 1e3 * (time.perf_counter() - t0) / steps


def batch_banded(rt, data, steps, warmup, precision, sequential=True):
    """One block-diagonal call (one launch per stage over every problem's
    tiles, schedule from the autotuner) against one call per problem with
    the default schedule."""
    mats = data["A"]
    big, ro, co = rt.block_diag(mats, 8)
    pa = rt.prepare_identity_a(big)
    nnz_cub = rt.nnz_cub(pa, 0)
    pb = rt.rebuild_b(pa, 0)
    sched, table, chosen = rt.autotune(pa, pb, repeats=2)
    sched.precision = rt.PRECISIONS[precision]
    state = {}

    def batched():
        state["pb"] = rt.rebuild_b(pa, 0, state.get("pb"))
        state["st"] = rt.tilespgemm(pa, state["pb"], sched)[1]

    ms_b = timed_steps(batched, steps, warmup)
    nnz_c = state["st"]["output_nnz"]
    tuned = {"chosen": table[chosen], "candidates": table}
    if not sequential:
        return {"nnz_cub": nnz_cub, "nnz_c": nnz_c, "ms_batched": ms_b, "tuned": tuned}
    # one product per call, default schedule
    pas = [rt.prepare_identity_a(m) for m in mats]
    pbs = [None] * len(pas)
    seq_sched = rt.schedule(materialize_output=False, precision=precision)
    seq_nnz = [0]

    def sequential():
        tot = 0
        for i, p in enumerate(pas):
            pbs[i] = rt.rebuild_b(p, 0, pbs[i])
            tot += rt.tilespgemm(p, pbs[i], seq_sched)[1]["output_nnz"]
        seq_nnz[0] = tot

    ms_s = timed_steps(sequential, max(1, steps // 2), 1)
    return {"nnz_cub": nnz_cub, "nnz_c": nnz_c, "ms_batched": ms_b, "ms_sequential": ms_s,
            "nnz_c_matches_sequential": nnz_c == seq_nnz[0], "tuned": tuned}


def batch_amg(rt, data, steps, warmup, precision, sequential=True):
    """R*(A*P) over the block-diagonal hierarchies: the two products of one
    step, P and A*P resident on the device as B operands (their tile
    rebuilds are inside the step), against the same per hierarchy."""
    def prep(As, Ps, Rs):
        out = []
        for A, P, R in zip(As, Ps, Rs):
            pa = rt.prepare_identity_a(A)
            sp_ = rt.upload_csr(P)
            pr = rt.prepare_identity_a(R)
            # A*P once to obtain A*P as a CSR operand (outside the step)
            pb = rt.rebuild_b_general(pa, sp_)
            tc, _ = rt.tilespgemm(pa, pb, rt.schedule())
            ap = rt.tile2csr(tc, canonical=True)
            sap = rt.upload_csr(ap)
            out.append({"pa": pa, "sp": sp_, "pr": pr, "sap": sap, "pb": None, "pb2": None,
                        "ap": ap, "cub": rt.nnz_cub_ab(pa, sp_) + rt.nnz_cub_ab(pr, sap)})
        return out

    sched = rt.schedule(materialize_output=False, precision=precision)

    def run(items):
        st = None
        for it in items:
            it["pb"] = rt.rebuild_b_general(it["pa"], it["sp"], it["pb"])
            rt.tilespgemm(it["pa"], it["pb"], sched)
            it["pb2"] = rt.rebuild_b_general(it["pr"], it["sap"], it["pb2"])
            st = rt.tilespgemm(it["pr"], it["pb2"], sched)[1]
        return st

    bA, _, _ = rt.block_diag(data["A"], 8)
    bP, _, _ = rt.block_diag(data["P"], 8)
    bR, _, _ = rt.block_diag(data["R"], 8)
    big = prep([bA], [bP], [bR])
    ms_b = timed_steps(lambda: run(big), steps, warmup)
    res = {"nnz_cub": big[0]["cub"], "ms_batched": ms_b}
    if precision != "fp64":
        # Errors of both products of the step against the FP64 kernels, scaled
        # per entry by the product of the absolute operands (the same FP64
        # kernels on |A|, |P| and |R|, |A*P|).
        it = big[0]
        got = [rt.last_c(it["pa"])["tile_csr_val"], rt.last_c(it["pr"])["tile_csr_val"]]
        f64 = rt.schedule(materialize_output=False)
        rt.tilespgemm(it["pa"], it["pb"], f64)
        rt.tilespgemm(it["pr"], it["pb2"], f64)
        ref = [rt.last_c(it["pa"])["tile_csr_val"], rt.last_c(it["pr"])["tile_csr_val"]]

        def absm(m):
            return rt.CsrMatrix(m.m, m.n, m.rowptr, m.col, np.abs(m.val))

        ap_csr = it["ap"]
        scales = []
        for left, right in ((bA, bP), (bR, ap_csr)):
            pl = rt.prepare_identity_a(absm(left))
            rt.tilespgemm(pl, rt.rebuild_b_general(pl, rt.upload_csr(absm(right))), f64)
            scales.append(rt.last_c(pl)["tile_csr_val"])
        res["split_vs_fp64_errors"] = {
            "A*P": value_errors_vs(got[0], ref[0], scales[0]),
            "R*(A*P)": value_errors_vs(got[1], ref[1], scales[1])}
    if sequential:
        items = prep(data["A"], data["P"], data["R"])
        res["ms_sequential"] = timed_steps(lambda: run(items), max(1, steps // 2), 1)
    return res


def batch_main(args):
    import rtt_amd as rt
    spec = BATCHES[args.workload]
    rt.init(0)
    data = make_batch(rt, spec)
    if spec["kind"] == "banded":
        r = batch_banded(rt, data, args.steps, args.warmup, args.precision,
                         sequential=not args.no_sequential)
    else:
        r = batch_amg(rt, data, args.steps, args.warmup, args.precision,
                      sequential=not args.no_sequential)
    flop = 2.0 * r["nnz_cub"]
    line = {"workload": f"{args.workload}: {data['desc']}", "precision": args.precision,
            "steps": args.steps, "warmup": args.warmup, "nnz_cub": r["nnz_cub"],
            "batched": {"ms_per_step": round(r["ms_batched"], 4),
                        "value": round(flop / (r["ms_batched"] * 1e6), 3), "unit": "GFLOP/s"}}
    if "ms_sequential" in r:
        line["sequential"] = {"ms_per_step": round(r["ms_sequential"], 4),
                              "value": round(flop / (r["ms_sequential"] * 1e6), 3),
                              "unit": "GFLOP/s"}
        line["batched_over_sequential"] = round(r["ms_sequential"] / r["ms_batched"], 3)
    for k in ("nnz_c", "nnz_c_matches_sequential", "tuned", "split_vs_fp64_errors"):
        if k in r:
            line[k] = r[k]
    print(json.dumps(line), flush=True)


# ---- the split-precision form on the block-FEM matrix ----------------------

def value_errors_vs(got, ref, scale):
    """max / median of |got - ref| / |ref| (ref != 0) and of |got - ref| / scale."""
    err = np.abs(got - ref)
    nz = ref != 0
    rel = err[nz] / np.abs(ref[nz])
    fwd = err / np.where(scale > 0, scale, 1.0)
    return {"rel_max": float(rel.max()) if rel.size else 0.0,
            "rel_median": float(np.median(rel)) if rel.size else 0.0,
            "fwd_max": float(fwd.max()) if fwd.size else 0.0,
            "fwd_median": float(np.median(fwd)) if fwd.size else 0.0}


def split_form(args):
    """C = A*A on the full block-FEM matrix in plain FP64 (tile path) and in
    the split BF16 form: step times, and the split form's error per C entry
    against the FP64 result, scaled by |c| and by (|A||A|)_ij (computed by
    the same FP64 kernels on |A|).  A row sample is also scored against the
    double-double oracle (both forms)."""
    import oracle
    import rtt_amd as rt
    spec = WORKLOADS[args.workload]
    a, desc = make_matrix(rt, spec)
    rt.init(0)
    pa = rt.prepare_identity_a(a)
    nnz_cub = rt.nnz_cub(pa, 0)
    out = {"workload": f"{args.workload}: C = A*A, {desc}", "nnz_cub": nnz_cub}
    vals = {}
    for prec in ("fp64", "split-bf16"):
        sched = rt.schedule(materialize_output=False, precision=prec)
        state = {}

        def step():
            state["pb"] = rt.rebuild_b(pa, 0, state.get("pb"))
            state["st"] = rt.tilespgemm(pa, state["pb"], sched)[1]

        ms = timed_steps(step, args.steps, args.warmup)
        st = state["st"]
        vals[prec] = rt.last_c(pa)["tile_csr_val"]
        out[prec] = {"ms_per_step": round(ms, 4),
                     "value": round(2.0 * nnz_cub / (ms * 1e6), 3), "unit": "GFLOP/s",
                     "numeric_kernel_ms": round(st["numeric_kernel_ms"], 4),
                     "numeric_path": {v: k for k, v in rt.NUMERIC_PATHS.items()}[st["numeric_path"]],
                     "mfma": {"instruction": "v_mfma_f64_4x4x4f64" if prec == "fp64"
                              else "v_mfma_f32_16x16x32_bf16 (8 per pair)",
                              "pairs": st["mfma_pairs"]}}
    # (|A||A|)_ij with the FP64 kernels: the forward-error scale of every entry.
    absa = rt.CsrMatrix(a.m, a.n, a.rowptr, a.col, np.abs(a.val))
    pabs = rt.prepare_identity_a(absa)
    rt.tilespgemm(pabs, rt.rebuild_b(pabs, 0), rt.schedule(materialize_output=False))
    scale = rt.last_c(pabs)["tile_csr_val"]
    out["split_vs_fp64_errors"] = value_errors_vs(vals["split-bf16"], vals["fp64"], scale)
    out["split_vs_fp64_errors"]["entries"] = int(scale.size)
    out["split_over_fp64_speed"] = round(out["fp64"]["ms_per_step"] /
                                         out["split-bf16"]["ms_per_step"], 3)
    # Row sample against the double-double oracle.
    rows = 2048
    r0 = (a.m // 2) // 8 * 8
    sub = (rows, a.n, (a.rowptr[r0:r0 + rows + 1] - a.rowptr[r0]).astype(np.int32),
           a.col[a.rowptr[r0]:a.rowptr[r0 + rows]], a.val[a.rowptr[r0]:a.rowptr[r0 + rows]])
    ref = oracle.spgemm(sub, (a.m, a.n, a.rowptr, a.col, a.val))
    ct = lambda v: oracle.c_tiles(ref["m"], ref["n"], ref["rowptr"], ref["col"], v)["tile_csr_val"]
    hi, lo, ab = ct(ref["hi"]), ct(ref["lo"]), ct(ref["absprod"])
    sample = {}
    # The sample rows' values are one stretch of C's values: tile rows
    # r0/8 .. (r0 + rows)/8 in tile order, as the oracle's tiles of the block.
    c = rt.last_c(pa, with_values=False)
    t0, t1 = int(c["tile_ptr"][r0 // 8]), int(c["tile_ptr"][(r0 + rows) // 8])
    z0, z1 = int(c["tile_nnz"][t0]), int(c["tile_nnz"][t1])
    for prec in ("fp64", "split-bf16"):
        sample[prec] = oracle.value_errors(vals[prec][z0:z1], hi, lo, ab)
    out["oracle_sample"] = {"rows": f"{r0}..{r0 + rows - 1}", "entries": int(z1 - z0),
                            "errors": sample}
    print(json.dumps(out), flush=True)


def csr_bytes(m, nnz):
    # Compulsory CSR traffic of one operand (src/external/cusparse/main.cu:205-208).
    return 4 * (m + 1) + 12 * nnz


def cpu_baseline(a, aat, nnz_cub, budget_s=2.5, repeats=5):
    """Times the oracle's row SPA (structure + values in reference order) on a
    contiguous block of rows sized to about budget_s seconds, `repeats`
    times; reports the median.  Threads: OMP_NUM_THREADS when set (the host's
    CPU share on a GPU box), else every core."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    A = (a.m, a.n, a.rowptr, a.col, a.val)
    if aat:
        trp, tci, tv = oracle.transpose(*A)
        B = (a.n, a.m, trp, tci, tv)
    else:
        B = A
    per_row = nnz_cub / max(a.m, 1)

    def run(rows):
        r0 = max(0, a.m // 2 - rows // 2)
        r1 = min(a.m, r0 + rows)
        sub = (r1 - r0, a.n, (a.rowptr[r0:r1 + 1] - a.rowptr[r0]).astype(np.int32),
               a.col[a.rowptr[r0]:a.rowptr[r1]], a.val[a.rowptr[r0]:a.rowptr[r1]])
        products = float(np.diff(np.asarray(B[2], np.int64))[sub[3]].sum())
        t0 = time.perf_counter()
        oracle.spgemm(sub, B, with_dd=False)
        dt = time.perf_counter() - t0
        return dt, products, r1 - r0

    dt, prod, rows = run(max(1024, min(a.m, int(2e6 / max(per_row, 1)))))
    rows = int(min(a.m, rows * max(1.0, budget_s / max(dt, 1e-3))))
    runs = [run(rows) for _ in range(repeats)]
    dts = sorted(r[0] for r in runs)
    dt = dts[len(dts) // 2]
    prod, rows = runs[0][1], runs[0][2]
    return {"value": 2.0 * prod / dt / 1e9, "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"row SPA (oracle, structure + FP64 values) on {rows} contiguous rows "
                      f"({rows / a.m:.1%} of the rows, {prod:.3g} products) of the same "
                      f"workload, median of {repeats} runs ({dt:.2f} s), OpenMP {threads} "
                      f"threads"}


def c1_line(rt, steps=20, warmup=5, repeats=5):
    """The reference's CPU-runnable configuration: C = A*A on the 64^2
    5-point Laplacian (4096 rows, nnzCub 100,104).  The GPU step (B update +
    symbolic + allocate + numeric, launch-bound at this size) beside the
    cpu_baseline legs the plan names for it: the legacy 16x16 tile CPU
    SpGEMM and the row SPA, both restated in the oracle, median of
    `repeats` runs each."""
    import oracle
    a = rt.gen_laplace2d(64, 0.1, 20260101)
    pa = rt.prepare_identity_a(a)
    nnz_cub = rt.nnz_cub(pa, 0)
    sched = rt.schedule(materialize_output=False)
    state = {}

    def step():
        state["pb"] = rt.rebuild_b(pa, 0, state.get("pb"))
        state["st"] = rt.tilespgemm(pa, state["pb"], sched)[1]

    ms = timed_steps(step, steps, warmup)
    A = (a.m, a.n, a.rowptr, a.col, a.val)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)

    def med(f):
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]

    t_legacy = med(lambda: oracle.legacy_tile_spgemm(A, A))
    t_spa = med(lambda: oracle.spgemm(A, A, with_dd=False))
    gf = lambda t: round(2.0 * nnz_cub / t / 1e9, 4)
    return {"workload": "laplace2d-64-aa: C = A*A, 2D 5-point Laplacian 64^2, off-diagonals "
                        f"-1+U[-0.1,0.1], seed 20260101, {a.m} rows, {a.nnz} nnz, nnzCub "
                        f"{nnz_cub}, nnzC {state['st']['output_nnz']}",
            "gpu": {"ms_per_step": round(ms, 4), "value": gf(ms * 1e-3), "unit": "GFLOP/s",
                    "steps": steps, "warmup": warmup},
            "cpu_legacy_tile16": {"value": gf(t_legacy), "unit": "GFLOP/s", "ms": round(t_legacy * 1e3, 3),
                                  "cores": threads, "kind": "port",
                                  "sample": f"whole product, median of {repeats} runs (the "
                                            "legacy 16x16 tile CPU algorithm, oracle)"},
            "cpu_spa": {"value": gf(t_spa), "unit": "GFLOP/s", "ms": round(t_spa * 1e3, 3),
                        "cores": threads, "kind": "port",
                        "sample": f"whole product, median of {repeats} runs (row SPA, oracle)"}}


# Kernel names in the rocprofv3 counter summaries, per breakdown key.
PROFILE_KERNELS = {
    "numeric": "numeric_row_kernel<SmallRow",
    "numeric_tile": "tile_numeric_kernel<Fp64Product>",
    "symbolic_count": "symbolic_wave_kernel",
    "symbolic_fill": "symbolic_write_kernel",
}


def profile_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload: FETCH_SIZE and WRITE_SIZE (KiB per dispatch,
    from the table whose first counter columns they are), each scaled by what
    the counter reads for a device copy of known size.  None when no summary
    covers the kernel."""
    import re
    path = os.path.join(ROOT, "profiles", f"{workload}_counters.md")
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    m = re.search(r"FETCH_SIZE\s+x\s+1024\s+x\s+([0-9]+(?:\.[0-9]+)?)\s+and\s+WRITE_SIZE\s+x\s+"
                  r"1024\s+x\s+([0-9]+(?:\.[0-9]+)?)", text)
    if not m:
        return None, None
    sf, sw = float(m.group(1)), float(m.group(2))
    name = PROFILE_KERNELS.get(kernel)
    in_table = False
    for line in text.splitlines():
        cells = [c.strip(" `") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 3 and cells[0] == "kernel":
            in_table = cells[1].startswith("FETCH_SIZE") and cells[2].startswith("WRITE_SIZE")
            continue
        if in_table and len(cells) >= 3 and name and cells[0].startswith(name):
            try:
                fetch_kib, write_kib = float(cells[1]), float(cells[2])
            except ValueError:
                continue
            return fetch_kib * 1024 * sf + write_kib * 1024 * sw, os.path.relpath(path, ROOT)
    return None, None


# The precision and batching forms the metric compares with plain FP64, each
# measured in a child process (this script with its own arguments) before the
# main process touches the GPU; a failure there does not end the benchmark.
FORMS = {
    "split_block_fem": ["--workload", "fem-500k-aa", "--split-form", "--steps", "5",
                        "--warmup", "2"],
    "batched_banded": ["--workload", "banded-batch-512", "--steps", "4", "--warmup", "1"],
    "batched_amg_fp64": ["--workload", "amg-batch-64", "--steps", "3", "--warmup", "1"],
    "batched_amg_split": ["--workload", "amg-batch-64", "--steps", "3", "--warmup", "1",
                          "--precision", "split-bf16", "--no-sequential"],
}


def run_forms():
    out = {}
    for name, extra in FORMS.items():
        cmd = [sys.executable, os.path.abspath(__file__)] + extra
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        except subprocess.TimeoutExpired:
            out[name] = {"error": "timed out after 900 s"}
            continue
        lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
        if p.returncode != 0 or not lines:
            out[name] = {"error": f"exit {p.returncode}: {p.stderr[-300:]}"}
            continue
        out[name] = json.loads(lines[-1])
        print(f"form {name} done", file=sys.stderr, flush=True)
    # Headline ratios of the metric: each form's GFLOP/s over plain FP64 on
    # the same workload.
    ratios = {}
    f = out.get("split_block_fem", {})
    if "fp64" in f and "split-bf16" in f:
        ratios["split_bf16_over_fp64_block_fem"] = {
            "ratio": round(f["split-bf16"]["value"] / f["fp64"]["value"], 3),
            "split_fwd_max_err": f["split_vs_fp64_errors"]["fwd_max"],
            "split_rel_max_err": f["split_vs_fp64_errors"]["rel_max"],
            "split_rel_median_err": f["split_vs_fp64_errors"]["rel_median"]}
    b = out.get("batched_banded", {})
    if "batched_over_sequential" in b:
        ratios["batched_over_sequential_banded_512"] = b["batched_over_sequential"]
    c = out.get("batched_amg_fp64", {})
    if "batched_over_sequential" in c:
        ratios["batched_over_sequential_amg_64"] = c["batched_over_sequential"]
    d = out.get("batched_amg_split", {})
    if "batched" in c and "batched" in d:
        ratios["split_batched_over_fp64_sequential_amg_64"] = round(
            d["batched"]["value"] / c["sequential"]["value"], 3) if "sequential" in c else None
        ratios["split_batched_over_fp64_batched_amg_64"] = round(
            d["batched"]["value"] / c["batched"]["value"], 3)
    out["ratios_vs_plain_fp64"] = ratios
    return out


# The other product mode of the same benchmark configuration, reported
# beside the main line: C = A*A on the R-MAT 2^20 matrix, in the reference's
# default (BGRF-v1) layout.
SECOND_MODE = {"workload": "rmat-20-aa", "layout": "bgrf", "steps": 3, "warmup": 1}


def second_mode():
    """Runs the other mode's workload in a child process (this script with
    its own arguments) and returns the essentials of its JSON line, or an
    error entry; a failure there does not end the benchmark."""
    m = SECOND_MODE
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", m["workload"], "--layout",
           m["layout"], "--steps", str(m["steps"]), "--warmup", str(m["warmup"]),
           "--no-cpu-baseline", "--no-control", "--no-second-mode", "--no-forms"]
    try:
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return {"workload": m["workload"], "error": "timed out after 600 s"}
    lines = [l for l in p.stdout.splitlines() if l.startswith('{"metric"')]
    if p.returncode != 0 or not lines:
        return {"workload": m["workload"], "error": f"exit {p.returncode}: {p.stderr[-200:]}"}
    r = json.loads(lines[-1])
    # Most of this workload's rows take the slab kernels, which the per-kernel
    # events do not cover: its roofline is given for the whole step, as the
    # compulsory CSR bytes (A, B, C) over the step time.
    frac = r["config"]["pipeline_hbm_frac"]
    return {"workload": r["config"]["workload"], "layout": m["layout"], "value": r["value"],
            "unit": r["unit"], "ms_per_step": r["ms_per_step"], "steps": r["steps"],
            "warmup": r["warmup"],
            "roofline": {"bound": "hbm", "scope": "whole step, CSR bytes of A, B and C",
                         "achieved": round(frac * HBM_PEAK_GBS, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": frac}}


def rocsparse_control(workload):
    """rocSPARSE's SpGEMM on the same workload (the reference's vendor
    control, src/cusparse_core_benchmark.h:4-30): the buffer-size and nnz
    stages, C's allocation and the compute stage per sample, scratch reused;
    median of three after one warm-up.  A reported comparison, not the
    measured value.  It runs in a child process (this script with
    --control-only), so that a failure inside the vendor library is reported
    here instead of ending the benchmark."""
    aat = WORKLOADS[workload]["aat"]
    try:
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--workload", workload,
                            "--control-only"], capture_output=True, text=True, timeout=300)
    except subprocess.TimeoutExpired:
        return {"library": "rocSPARSE", "error": "timed out after 300 s"}
    lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
    if p.returncode != 0 or not lines:
        return {"library": "rocSPARSE", "error": f"exit {p.returncode}: {p.stderr[-200:]}"}
    r = json.loads(lines[-1])
    if "error" in r:
        return {"library": "rocSPARSE", "error": r["error"][:200]}
    return {"library": "rocSPARSE rocsparse_spgemm (CSR, FP64, int32 indices)",
            "ms": round(r["ms"], 4), "nnz_c": r["nnz_c"],
            "boundary": "buffer-size and nnz stages + C allocation + compute stage, scratch reused"
                        + ("; B = A^T built by csr2csc before timing" if aat else "")}


def finish_control(control, nnz_cub, nnz_c):
    """GFLOP/s of the control on the same flop count, and whether its nnz(C)
    equals this library's."""
    if control is None or "error" in control:
        return control
    out = dict(control)
    out["value"] = round(2.0 * nnz_cub / (control["ms"] * 1e6), 3)
    out["unit"] = "GFLOP/s"
    out["nnz_c_matches"] = control["nnz_c"] == nnz_c
    return out


def control_only(workload):
    """Child side of rocsparse_control: build the workload's A, run the
    control, print one JSON line."""
    import rtt_amd as rt
    spec = WORKLOADS[workload]
    a, _ = make_matrix(rt, spec)
    try:
        r = rt.control_rocsparse(a, spec["aat"], warmup=1, iterations=3)
    except Exception as e:  # a missing library or a vendor error status
        r = {"error": str(e)}
    print(json.dumps(r), flush=True)


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime: the plain
    copy kernel whose FETCH_SIZE / WRITE_SIZE counters scale the traffic
    counters of the SpGEMM kernels."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="stencil27-128-aat",
                    choices=sorted(WORKLOADS) + sorted(BATCHES))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lb-mode", default="adaptive-tail",
                    choices=["static", "dynamic", "adaptive-tail"],
                    help="row assignment of the slab kernels (the reference's --lb-mode)")
    ap.add_argument("--layout", default="identity", choices=["identity", "bgrf"],
                    help="A's offline layout: identity (--no-reorder) or the BGRF-v1 plan")
    ap.add_argument("--no-control", action="store_true",
                    help="skip the rocSPARSE SpGEMM control on the same workload")
    ap.add_argument("--control-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched workloads: skip the one-call-per-problem comparison")
    ap.add_argument("--split-form", action="store_true",
                    help="plain FP64 vs split-BF16 on --workload (block FEM), with errors")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the split-precision and batched forms reported beside the line")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 64^2 Laplacian line (GPU step and its CPU baselines)")
    ap.add_argument("--no-second-mode", action="store_true",
                    help="skip the R-MAT 2^20 C = A*A measurement reported beside the main line")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split-bf16"],
                    help="tile-path arithmetic: FP64 MFMA, or the Ozaki split on BF16 MFMA")
    ap.add_argument("--numeric-path", default="auto", choices=["auto", "vector", "tile"],
                    help="numeric kernels: LDS accumulator (vector), FP64 matrix cores per "
                         "tile pair (tile), or chosen by A's tile fill (auto)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (traffic-counter scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.control_only:
        control_only(args.workload)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    if args.workload in BATCHES:
        batch_main(args)
        return
    if args.split_form:
        split_form(args)
        return

    import rtt_amd as rt
    spec = WORKLOADS[args.workload]
    aat = spec["aat"]
    # The child processes (control, other mode) run before this process
    # touches the GPU.
    control = None if args.no_control else rocsparse_control(args.workload)
    other = None if args.no_second_mode else second_mode()
    forms = None if args.no_forms else run_forms()
    a, desc = make_matrix(rt, spec)
    rt.init(0)
    pa = rt.prepare_bgrf_a(a) if args.layout == "bgrf" else rt.prepare_identity_a(a)
    nnz_cub = rt.nnz_cub(pa, aat)
    sched = rt.schedule(args.lb_mode, materialize_output=False, numeric_path=args.numeric_path,
                        precision=args.precision)
    pb = None
    for _ in range(args.warmup):
        pb = rt.rebuild_b(pa, aat, pb)
        rt.tilespgemm(pa, pb, sched)
    rt.sync()
    steps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        pb = rt.rebuild_b(pa, aat, pb)
        _, st = rt.tilespgemm(pa, pb, sched)
        s1 = time.perf_counter()
        steps.append((s1 - s0, pb.last, st))
    rt.sync()
    total = time.perf_counter() - t0
    ms_per_step = 1e3 * total / args.steps
    gflops = 2.0 * nnz_cub / (ms_per_step * 1e6)

    st = steps[-1][2]
    nnz_c = st["output_nnz"]
    kern = {
        "symbolic_count": np.mean([s[2]["symbolic_count_kernel_ms"] for s in steps]),
        "symbolic_fill": np.mean([s[2]["symbolic_fill_kernel_ms"] for s in steps]),
        "numeric": np.mean([s[2]["numeric_kernel_ms"] for s in steps]),
    }
    # Algorithmic bytes of each kernel (per launch), in CSR terms:
    #  numeric        reads A and B once, writes C (values and structure;
    #                 the LDS kernels write both for the rows they hold)
    #  symbolic_fill  C structure of the rows the LDS kernels cannot hold
    #                 (none on the structured workloads)
    #  symbolic_count reads A and B structure, writes per-row counts
    #  On the tile path the structure pass writes all of C's structure and
    #  the numeric kernel reads A and B and writes C's values.
    m_b = a.n if aat else a.m
    a_struct = 4 * (a.m + 1) + 4 * a.nnz
    b_struct = 4 * (m_b + 1) + 4 * a.nnz
    tile_path = st["numeric_path"] == rt.NUMERIC_PATHS["tile"]
    alg = {
        "numeric": csr_bytes(a.m, a.nnz) + csr_bytes(m_b, a.nnz) + (
            8 * nnz_c if tile_path else csr_bytes(a.m, nnz_c)),
        "symbolic_fill": a_struct + b_struct + 4 * (a.m + 1) + 4 * nnz_c,
        "symbolic_count": a_struct + b_struct + 8 * (a.m + 1),
    }
    dom = max(kern, key=kern.get)
    achieved = alg[dom] / (kern[dom] * 1e-3) / 1e9
    traffic, traffic_src = profile_traffic(
        args.workload, "numeric_tile" if dom == "numeric" and tile_path and
        args.precision == "fp64" else dom)
    b_alg = csr_bytes(a.m, a.nnz) + csr_bytes(m_b, a.nnz) + csr_bytes(a.m, nnz_c)
    breakdown = {
        "b_update_ms": float(np.median([s[1]["total_ms"] for s in steps])),
        "symbolic_ms": float(np.median([s[2]["symbolic_ms"] for s in steps])),
        "malloc_ms": float(np.median([s[2]["allocation_ms"] for s in steps])),
        "numeric_ms": float(np.median([s[2]["numeric_ms"] for s in steps])),
        "kernel_ms": {k: round(float(v), 4) for k, v in kern.items()},
    }
    mfma = None
    if tile_path:
        flops = MFMA_PAIR_FLOPS * st["mfma_pairs"]
        mfma = {"pairs": st["mfma_pairs"], "flops": flops,
                "utilisation": round(flops / (kern["numeric"] * 1e-3) / (FP64_MFMA_PEAK_TFS * 1e12), 5),
                "useful_fraction": round(2.0 * nnz_cub / max(flops, 1), 4),
                "peak": f"{FP64_MFMA_PEAK_TFS} TFLOP/s FP64 MFMA"}
        breakdown["b_columns_ms"] = float(np.median([s[2]["b_columns_ms"] for s in steps]))
    print(json.dumps({"breakdown": breakdown, "nnz_cub": nnz_cub, "nnz_c": nnz_c,
                      "tiles_c": st["output_tiles"], "pairs": st["pair_count"], "mfma": mfma}),
          file=sys.stderr, flush=True)
    # A*A^T with identity rows: the B update copies each B tile's record from
    # A's transpose_src records, built at A's (untimed) preparation.  The reference does the corresponding transpose
    # inside its timed B rebuild, so the step is also reported with it.
    offline_tr = None
    if aat and args.layout == "identity":
        tr_ms = float(pa.stats["transpose_src_ms"])
        offline_tr = {"transpose_src_ms": round(tr_ms, 4),
                      "ms_per_step_with_it": round(ms_per_step + tr_ms, 4),
                      "value_with_it": round(2.0 * nnz_cub / ((ms_per_step + tr_ms) * 1e6), 3)}
    control = finish_control(control, nnz_cub, nnz_c)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(a, aat, nnz_cub)
    c1 = None if args.no_cpu_baseline or args.no_c1 else c1_line(rt)
    line = {
        "metric": METRIC,
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: C = A*A{'^T' if aat else ''}, {desc}, "
                               f"{a.m} rows, {a.nnz} nnz, nnzCub {nnz_cub}, nnzC {nnz_c}",
                   "precision": "plain FP64" if args.precision == "fp64" else
                                "split BF16 slices, FP64 recombination",
                   "lb_mode": args.lb_mode,
                   "numeric_path": {rt.NUMERIC_PATHS["auto"]: "auto (vector and tile rows)",
                                    rt.NUMERIC_PATHS["vector"]: "vector",
                                    rt.NUMERIC_PATHS["tile"]: "tile"}[st["numeric_path"]],
                   "layout": args.layout,
                   "pipeline_hbm_frac": round(b_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                              5)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None if traffic is None else int(traffic),
                     "algorithmic_bytes": int(alg[dom]),
                     "traffic_source": traffic_src,
                     "mfma": mfma},
        "cpu_baseline": cpu,
        "offline_transpose": offline_tr,
        "c1": c1,
        "control": control,
        "second_mode": other,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
