// Operator entry points not yet implemented on the device: each declines
// with OXR_EUNSUPPORTED, the reference's own decline-first contract
// (crates/oxidelake-device/src/backend.rs:129-163 defaults), so a caller
// runs its CPU path and counts a fallback.
#include "common.h"

using namespace oxr;

#define DECLINE(feature) return fail(OXR_EUNSUPPORTED, feature, "not implemented on the rocm backend yet")

extern "C" {
int oxr_agg_update_join(oxr_stream *, oxr_agg_state *, const oxr_join_table *, const oxr_batch *,
                        int32_t, const oxr_batch *, int32_t) { DECLINE("rocm.aggregate"); }
}
