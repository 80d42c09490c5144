// Join handles shared by join.hip (build, probe, gather) and aggregate.hip
// (fused join + aggregation); the device view is in jtable.h.
#pragma once

#include <cstdint>

#include "common.h"
#include "jtable.h"


struct oxr_join_table {
  oxr_device *dev = nullptr;
  oxr::StreamCore *owner = nullptr;
  oxr::JTable t{};
  void *mem = nullptr;       // slot arrays
  void *rows_mem = nullptr;  // row array
  uint64_t distinct = 0;
  uint64_t nrows = 0;        // build rows with a non-NULL key
  uint64_t build_rows = 0;   // rows of the build batch
  int32_t key_type = OXR_INT64;
  uint64_t serial = 0;       // unique per table of the process (cache keys)
};

struct oxr_join_matches {
  oxr_device *dev = nullptr;
  oxr::StreamCore *owner = nullptr;
  uint64_t probe_rows = 0;
  uint64_t total = 0;
  uint64_t *offsets = nullptr;  // probe_rows + 1, device
  uint32_t *starts = nullptr;   // probe_rows, device
  void *mem = nullptr;
  const oxr_join_table *table = nullptr;
};
