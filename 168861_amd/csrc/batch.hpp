// Internal interface of batches for the request server (same library).
#pragma once
#include "nrt_internal.hpp"

namespace nrt {
void batch_launch(nrt_batch* b, long long k);  // enqueue k steps, no wait
// implicit: each slot takes max(1, min(kmax, elem_budget / N)) steps
void batch_launch_budget(nrt_batch* b, long long k, long long elem_budget, long long kmax);
void batch_finish(nrt_batch* b);               // wait for them
hipStream_t batch_stream(nrt_batch* b);
bool batch_slot_complete(nrt_batch* b, int slot);  // host-side, no device access
// Admission: footprint estimate, and a check against free device memory less
// `reserved` bytes (throws NRT_ERR_ADMISSION).
int64_t batch_footprint(const nrt_batch* b, int N, int64_t num_steps, int64_t trace_every);
void batch_admission_check(const nrt_batch* b, int N, int64_t num_steps, int64_t trace_every,
                           int64_t reserved);
int batch_plan_threads(const nrt_batch* b);  // batch_threads for handle_create
nrt_options batch_slot_options(const nrt_batch* b, int N);  // options a slot is planned with
// Registers a handle built elsewhere (handle_create with the batch's options and
// batch_plan_threads) as a new slot; the batch owns it from then on.
int batch_add_handle(nrt_batch* b, nrt_handle* h, int64_t num_steps, int64_t a0, int64_t a1,
                     int32_t trace_elem, int64_t trace_every);
// Results of a completed slot that is not in the launch in flight, copied on the
// caller's stream `side` without waiting for that launch; then its release.
void batch_slot_collect(nrt_batch* b, int slot, hipStream_t side, std::vector<double>* V,
                        std::vector<double>* trace, int* status, int64_t* steps_done);
void batch_slot_drop(nrt_batch* b, int slot);
// implicit.hip: steps of a single handle enqueued on its stream
void implicit_steps_async(nrt_handle* h, int64_t first, int64_t num, int64_t a0, int64_t a1,
                          int32_t trace_elem, int64_t interval, int64_t gl0, double* trace_dev);
void implicit_slot_stats(nrt_handle* h, int* status, int64_t* inner);
}  // namespace nrt
