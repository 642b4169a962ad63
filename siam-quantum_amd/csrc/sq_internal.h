// sq_internal.h - library-internal accessors between the engine and the
// device-resident SCF (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

struct sqjk_ctx;

// stream on which a J/K context queues its work
hipStream_t sqjk_stream_internal(sqjk_ctx* c);
// mixed-precision threshold set for the drop-in entry points
double sqjk_dropin_mixed_internal();
// device blocks from the library's cache (no hipFree per release, which
// would wait for the whole device)
void* sq_device_block(size_t bytes);
void sq_device_release(void* p);
