// Launches of the wide passes (wide_passes.hip); layout in device_store.hpp.
#pragma once

#include <cstdint>

#include "device_store.hpp"

namespace qhamd {

// Pass R, pass L and the row-set sums of one apply (kc = 16, 32 or 64).
// `events`: null, or three hipEvent_t recorded before pass R, before pass L
// and after the sums.
void launch_wide(const WideView& V, const WideCounts& c, int kc, bool fp32, void* stream,
                 void* events);
void launch_batch_wide(const WideBatchClass& c, const double* x, double* y, void* zw, bool fp32,
                       void* stream);

}  // namespace qhamd
