// Internal interface of batches for the request server (same library).
#pragma once
#include "nrt_internal.hpp"

namespace nrt {
void batch_launch(nrt_batch* b, long long k);  // enqueue k steps, no wait
void batch_finish(nrt_batch* b);               // wait for them
hipStream_t batch_stream(nrt_batch* b);
bool batch_slot_complete(nrt_batch* b, int slot);  // host-side, no device access
// implicit.hip: steps of a single handle enqueued on its stream
void implicit_steps_async(nrt_handle* h, int64_t first, int64_t num, int64_t a0, int64_t a1,
                          int32_t trace_elem, int64_t interval, int64_t gl0, double* trace_dev);
void implicit_slot_stats(nrt_handle* h, int* status, int64_t* inner);
}  // namespace nrt
