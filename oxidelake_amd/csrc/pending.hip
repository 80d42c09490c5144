// Operator entry points not yet implemented on the device: each declines
// with OXR_EUNSUPPORTED, the reference's own decline-first contract
// (crates/oxidelake-device/src/backend.rs:129-163 defaults), so a caller
// runs its CPU path and counts a fallback.
#include "common.h"

using namespace oxr;

#define DECLINE(feature) return fail(OXR_EUNSUPPORTED, feature, "not implemented on the rocm backend yet")

extern "C" {
int oxr_join_build(oxr_stream *, const oxr_batch *, int32_t, oxr_join_table **) { DECLINE("rocm.join"); }
int oxr_join_table_info(const oxr_join_table *, uint64_t *, uint64_t *, int *) { DECLINE("rocm.join"); }
void oxr_join_table_free(oxr_join_table *) {}
int oxr_join_probe(oxr_stream *, const oxr_join_table *, const oxr_batch *, int32_t,
                   oxr_join_matches **, uint64_t *) { DECLINE("rocm.join"); }
void oxr_join_matches_free(oxr_join_matches *) {}
int oxr_join_gather(oxr_stream *, const oxr_batch *, const oxr_batch *, const oxr_join_table *,
                    const oxr_join_matches *, uint64_t, uint64_t, const int32_t *, int32_t,
                    oxr_batch **) { DECLINE("rocm.join"); }
int oxr_aggregate(oxr_stream *, const oxr_batch *, const oxr_agg_spec *, oxr_batch **) { DECLINE("rocm.aggregate"); }
int oxr_agg_create(oxr_stream *, const oxr_batch *, const oxr_agg_spec *, oxr_agg_state **) { DECLINE("rocm.aggregate"); }
int oxr_agg_update(oxr_stream *, oxr_agg_state *, const oxr_batch *) { DECLINE("rocm.aggregate"); }
int oxr_agg_update_join(oxr_stream *, oxr_agg_state *, const oxr_join_table *, const oxr_batch *,
                        int32_t, const oxr_batch *, int32_t) { DECLINE("rocm.aggregate"); }
int oxr_agg_finish(oxr_stream *, oxr_agg_state *, oxr_batch **) { DECLINE("rocm.aggregate"); }
void oxr_agg_free(oxr_agg_state *) {}
int oxr_vector_distance(oxr_stream *, const oxr_batch *, int32_t, const float *, int32_t, int32_t,
                        const char *, oxr_batch **) { DECLINE("rocm.vector"); }
int oxr_vector_topk(oxr_stream *, const oxr_batch *, int32_t, const float *, int32_t, int32_t,
                    int32_t, int32_t, int64_t *, float *) { DECLINE("rocm.vector"); }
}
