// SPDX-License-Identifier: BSD-3-Clause
//
// The kernel library (libtg_hip.so), opened on first use and never at load, as the reference opens
// its GPU drivers (include/terraingen/gpu/backend.h:50-55, src/gpu/CMakeLists.txt:3-6): a process
// without a GPU runtime loads the host library and runs its CPU work unchanged.
#pragma once

#include "tg_hip.h"

#include <string>

namespace tgh {

struct HipApi {
    decltype(&tg_hip_open) open;
    decltype(&tg_hip_device_name) device_name;
    decltype(&tg_hip_can_run) can_run;
    decltype(&tg_hip_run) run;
    decltype(&tg_hip_batch_create) batch_create;
    decltype(&tg_hip_batch_free) batch_free;
    decltype(&tg_hip_batch_stage) batch_stage;
    decltype(&tg_hip_batch_stage_async) batch_stage_async;
    decltype(&tg_hip_batch_run) batch_run;
    decltype(&tg_hip_batch_stream) batch_stream;
    decltype(&tg_hip_batch_surface) batch_surface;
    decltype(&tg_hip_batch_stage_boulders) batch_stage_boulders;
    decltype(&tg_hip_malloc) malloc;
    decltype(&tg_hip_free) free;
    decltype(&tg_hip_host_alloc) host_alloc;
    decltype(&tg_hip_host_free) host_free;
    decltype(&tg_hip_memcpy_d2h_async) memcpy_d2h_async;
    decltype(&tg_hip_event_create) event_create;
    decltype(&tg_hip_event_destroy) event_destroy;
    decltype(&tg_hip_event_record) event_record;
    decltype(&tg_hip_event_query) event_query;
    decltype(&tg_hip_stream_synchronize) stream_synchronize;
};

// The opened library; throws std::runtime_error with the loader's reason when it cannot be opened.
// Its path is $TERRAINGEN_HIP_LIBRARY, else libtg_hip.so beside this library.
const HipApi& Hip();

// Empty when the library opens, else why it does not.
std::string HipUnavailableReason();

}  // namespace tgh
