// The clear gate of the values-only B update, host only so that it can be
// checked exhaustively without a device.
//
// Restates dmma_choose_b_values_clear (src/dmma_b_values_clear_policy.h:91-160):
// "always clear" never skips; the other two policies skip the clear only when
// every precondition holds, and otherwise report the first one that fails,
// in a fixed order.  The payload clear is the safe default in every fallback.
#include "rtt_spgemm.h"

extern "C" {

int rtt_choose_b_values_clear(int policy, int values_only_update, int dynamic_b_valid,
                              int has_active_duplicates, int active_source_mapping_complete,
                              int active_source_to_payload_injective,
                              int payload_fully_initialized, int every_payload_slot_overwritten,
                              int numeric_b_read_only, rtt_b_values_decision *out)
{
    if (!out)
        return RTT_ERR_INVALID;
    rtt_b_values_decision d = {1, 0, 0, RTT_B_VALUES_CLEAR_POLICY_ALWAYS};
    if (policy == RTT_B_VALUES_ALWAYS_CLEAR) {
        *out = d;
        return RTT_OK;
    }
    d.fallback = 1;
    // The preconditions in the order they are checked, each with the reason
    // reported when it is the first to fail.
    const struct {
        bool holds;
        int reason;
    } gate[] = {
        {policy == RTT_B_VALUES_NOCLEAR || policy == RTT_B_VALUES_SAFE_AUTO,
         RTT_B_VALUES_FALLBACK_INVALID_POLICY},
        {values_only_update != 0, RTT_B_VALUES_FALLBACK_STRUCTURE_UPDATE},
        {dynamic_b_valid != 0, RTT_B_VALUES_FALLBACK_INVALID_STATE},
        {has_active_duplicates == 0, RTT_B_VALUES_FALLBACK_DUPLICATES},
        {active_source_mapping_complete != 0, RTT_B_VALUES_FALLBACK_INCOMPLETE_MAPPING},
        {active_source_to_payload_injective != 0, RTT_B_VALUES_FALLBACK_NONUNIQUE_MAPPING},
        {payload_fully_initialized != 0, RTT_B_VALUES_FALLBACK_UNINITIALIZED_PAYLOAD},
        {every_payload_slot_overwritten != 0, RTT_B_VALUES_FALLBACK_DENSE_HOLES},
        {numeric_b_read_only != 0, RTT_B_VALUES_FALLBACK_NUMERIC_WRITE_CONTRACT},
    };
    for (const auto &g : gate) {
        if (!g.holds) {
            d.reason = g.reason;
            *out = d;
            return RTT_OK;
        }
    }
    d.clear_payload = 0;
    d.skip_clear = 1;
    d.fallback = 0;
    d.reason = policy == RTT_B_VALUES_NOCLEAR ? RTT_B_VALUES_NOCLEAR_ELIGIBLE
                                              : RTT_B_VALUES_AUTO_ELIGIBLE;
    *out = d;
    return RTT_OK;
}

// Names as the reference prints them (src/dmma_b_values_clear_policy.h:43-88).
const char *rtt_b_values_clear_policy_name(int policy)
{
    switch (policy) {
    case RTT_B_VALUES_ALWAYS_CLEAR:
        return "always-clear";
    case RTT_B_VALUES_NOCLEAR:
        return "noclear";
    case RTT_B_VALUES_SAFE_AUTO:
        return "safe-auto";
    default:
        return "invalid";
    }
}

const char *rtt_b_values_clear_reason_name(int reason)
{
    static const char *const names[] = {
        "policy-always-clear",
        "eligible-explicit-noclear",
        "eligible-safe-auto",
        "structure-update-legacy-rebuild",
        "invalid-dynamic-b-state",
        "active-source-duplicates",
        "source-to-payload-not-injective",
        "payload-not-fully-initialized",
        "numeric-b-not-read-only",
        "invalid-policy",
        "active-source-map-incomplete",
        "dense-holes-not-overwritten-each-update",
    };
    if (reason < 0 || reason >= static_cast<int>(sizeof(names) / sizeof(names[0])))
        return "unknown";
    return names[reason];
}

} // extern "C"
