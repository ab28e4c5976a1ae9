"""Writes reference_vectors.json: input/output vectors of the reference's own
functions on the tiled SpGEMM path that the oracle restates, computed by the
reference sources compiled into oracle/_ref/libref.so (oracle/Makefile builds
it from the reference tree; nothing of the reference is copied here):

  * dmma_tail_chunk_count_from_counts (src/dmma_tile_tail_queue.h:151-176) on a
    grid of (scans, matches, threshold, target, max_chunks);
  * dmma_plan_tile_tail_queue (src/dmma_tile_tail_queue.h:190-292) on seeded
    task sets under the gate's settings;
  * dmma_choose_b_values_clear (src/dmma_b_values_clear_policy.h:91-160) for
    every policy value and every combination of its eight preconditions,
    with the policy and reason names.

Run from the repository root after `make -C oracle`:
    python tests/golden/make_reference_vectors.py
"""
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

PLAN_SETTINGS = [dict(), dict(max_heavy_fraction=1.0), dict(record_capacity=3),
                 dict(workspace_limit=1024, max_heavy_fraction=1.0), dict(enabled=False),
                 dict(threshold=-1.0), dict(threshold=2000.0, target=500.0, max_chunks=32,
                                            max_heavy_fraction=0.2)]


def main():
    ref = oracle.ref_lib()
    if ref is None:
        sys.exit("oracle/_ref/libref.so is not built")
    grid = []
    for scans in (0, 1, 2, 3, 7, 64, 2048, 100000):
        for matches in (0, 1, 5, 100, 2047, 2048, 2049, 3000, 50000):
            if matches > scans and scans > 1:
                continue
            for thr, tgt, mx in ((16384.0, 8192.0, 8), (256.0, 256.0, 8), (100.0, 7.0, 32),
                                 (16384.0, 8192.0, 2)):
                grid.append([scans, matches, thr, tgt, mx,
                             int(ref.ref_tail_chunk_count(scans, matches, thr, tgt, mx))])
    plans = []
    for seed in range(2):
        rng = np.random.default_rng(100 + seed)
        n = 300
        scans = rng.integers(0, 40000, n)
        matches = (scans * rng.uniform(0, 1, n)).astype(np.int64)
        out_nnz = rng.integers(0, 65, n)
        for kw in PLAN_SETTINGS:
            p = oracle.plan_tail_queue(scans, matches, out_nnz, use_reference=True, **kw)
            plans.append({"scans": scans.tolist(), "matches": matches.tolist(),
                          "output_nnz": out_nnz.tolist(), "settings": kw,
                          "reason": p["reason"], "heavy": p["heavy"], "chunks": p["chunks"],
                          "workspace": p["workspace"], "heavy_ids": p["heavy_ids"].tolist(),
                          "heavy_chunks": p["heavy_chunks"].tolist()})
    gate = []
    for policy in (-1, 0, 1, 2, 3):
        for flags in itertools.product((0, 1), repeat=8):
            gate.append([policy, *flags, *oracle.ref_choose_b_values_clear(policy, *flags)])
    names = {"reason": {str(r): ref.ref_b_values_clear_reason_name(r).decode()
                        for r in range(-1, 13)},
             "policy": {str(p): ref.ref_b_values_clear_policy_name(p).decode()
                        for p in range(-1, 4)}}
    out = {"tail_chunk_count": {"columns": ["scans", "matches", "threshold", "target",
                                            "max_chunks", "chunks"], "rows": grid},
           "plan_tail_queue": plans,
           "b_values_clear": {"columns": ["policy", "values_only_update", "dynamic_b_valid",
                                          "has_active_duplicates",
                                          "active_source_mapping_complete",
                                          "active_source_to_payload_injective",
                                          "payload_fully_initialized",
                                          "every_payload_slot_overwritten",
                                          "numeric_b_read_only", "clear_payload", "skip_clear",
                                          "fallback", "reason"],
                              "rows": gate, "names": names}}
    path = os.path.join(HERE, "reference_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path}: {len(grid)} chunk counts, {len(plans)} plans, {len(gate)} gate cases")


if __name__ == "__main__":
    main()
