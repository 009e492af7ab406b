// sf_batch_copy.h -- asynchronous copy-out of finished batch problems, for
// the request loop (host/sf_server.cpp): the copy of one problem's state
// runs on the batch's copy stream while the next steps of the others run
// on its compute stream.
#pragma once

#include <hip/hip_runtime.h>

struct sf_batch;

namespace sfamd {

// Doubles of host memory one copy-out of `slot` fills: the padded block of
// each of the five variables, then the problem's clock time.
long long batch_copy_doubles(const sf_batch* b, int slot);

// Enqueue the copy of `slot`'s state and time into pinned host memory
// `dst` (batch_copy_doubles(b, slot) doubles) behind the work already on
// the compute stream, then record `done`.  The problem must have reached
// its step limit (the kernels no longer touch it), and the slot must not be
// removed before `done` has completed.
int batch_copy_out_async(sf_batch* b, int slot, double* dst, hipEvent_t done);

// The unique-node SoA (5 x n0 x n1) and the time of `slot` from a completed
// copy-out.
void batch_unpack(const sf_batch* b, int slot, const double* src, double* soa, double* time);

}  // namespace sfamd
