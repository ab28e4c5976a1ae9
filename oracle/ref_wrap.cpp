/*
 * C-ABI wrapper around reference sources that compile with g++ as they lie
 * under the reference tree (no stand-in headers): the tail-queue admission
 * planner (src/dmma_tile_tail_queue.h), the values-only B update's clear gate
 * (src/dmma_b_values_clear_policy.h) and the Matrix Market reader of the
 * standalone tool (src/external/cusparse/mmio_highlevel.h, whose mmio_data
 * follows the same placement as src/mmio_highlevel.h:593-759).
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile and
 * used by tests to pin the oracle restatement.
 */
#include "dmma_b_values_clear_policy.h"
#include "dmma_tile_tail_queue.h"
#include "mmio_highlevel.h"

#include <cstdint>
#include <vector>

extern "C" {

// out = {clear_payload, skip_clear, fallback, reason}
void ref_choose_b_values_clear(int policy, int values_only_update, int dynamic_b_valid,
                               int has_active_duplicates, int mapping_complete, int injective,
                               int payload_fully_initialized, int every_slot_overwritten,
                               int numeric_b_read_only, int *out)
{
    const DmmaBValuesClearDecision d = dmma_choose_b_values_clear(
        static_cast<DmmaBValuesClearPolicy>(policy), values_only_update != 0,
        dynamic_b_valid != 0, has_active_duplicates != 0, mapping_complete != 0,
        injective != 0, payload_fully_initialized != 0, every_slot_overwritten != 0,
        numeric_b_read_only != 0);
    out[0] = d.clear_payload ? 1 : 0;
    out[1] = d.skip_clear ? 1 : 0;
    out[2] = d.fallback ? 1 : 0;
    out[3] = static_cast<int>(d.reason);
}

const char *ref_b_values_clear_policy_name(int policy)
{
    return dmma_b_values_clear_policy_name(static_cast<DmmaBValuesClearPolicy>(policy));
}

const char *ref_b_values_clear_reason_name(int reason)
{
    return dmma_b_values_clear_reason_name(static_cast<DmmaBValuesClearReason>(reason));
}

int ref_tail_chunk_count(int scans, int matches, double admission_threshold,
                         double chunk_target, int max_chunks)
{
    DmmaTileTailQueueExactTask task;
    task.task = 0;
    task.scans = scans;
    task.matches = matches;
    DmmaTileTailQueuePolicy policy;
    policy.enabled = true;
    policy.admission_threshold = admission_threshold;
    policy.chunk_target = chunk_target;
    policy.max_chunks = max_chunks;
    DmmaTileTailQueueCostModel model;
    return dmma_tile_tail_queue_chunk_count(task, policy, model);
}

int ref_plan_tail_queue(int n, const int *scans, const int *matches,
                        const int *output_nnz, int enabled,
                        double admission_threshold, double chunk_target,
                        int max_chunks, double max_heavy_fraction,
                        int record_capacity, uint64_t workspace_limit,
                        int workspace_reduction, int64_t *out, int *heavy_ids,
                        int *heavy_chunks)
{
    std::vector<DmmaTileTailQueueExactTask> tasks(n);
    for (int i = 0; i < n; ++i) {
        tasks[i].task = i;
        tasks[i].scans = scans[i];
        tasks[i].matches = matches[i];
        tasks[i].output_nnz = output_nnz[i];
    }
    DmmaTileTailQueuePolicy policy;
    policy.enabled = enabled != 0;
    policy.admission_threshold = admission_threshold;
    policy.chunk_target = chunk_target;
    policy.max_chunks = max_chunks;
    policy.max_heavy_fraction = max_heavy_fraction;
    policy.record_capacity = record_capacity;
    policy.workspace_limit_bytes = workspace_limit;
    policy.workspace_reduction = workspace_reduction != 0;
    DmmaTileTailQueueCostModel model;
    const DmmaTileTailQueuePlan plan =
        dmma_plan_tile_tail_queue(tasks, policy, model);
    out[0] = static_cast<int64_t>(plan.reason);
    out[1] = static_cast<int64_t>(plan.heavy_tasks);
    out[2] = static_cast<int64_t>(plan.chunks);
    out[3] = static_cast<int64_t>(plan.partial_workspace_bytes);
    for (size_t i = 0; i < plan.heavy.size(); ++i) {
        heavy_ids[i] = plan.heavy[i].task;
        heavy_chunks[i] = plan.heavy[i].chunks;
    }
    return 0;
}

int ref_mmio_info(const char *path, int *m, int *n, int *nnz, int *is_sym)
{
    return mmio_info(m, n, nnz, is_sym, const_cast<char *>(path));
}

int ref_mmio_data(const char *path, int *row_ptr, int *col_idx, double *val)
{
    return mmio_data(row_ptr, col_idx, val, const_cast<char *>(path));
}

} // extern "C"
