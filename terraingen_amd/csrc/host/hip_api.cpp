// SPDX-License-Identifier: BSD-3-Clause
#include "hip_api.h"

#include <dlfcn.h>

#include <cstdlib>
#include <mutex>
#include <stdexcept>

namespace tgh {

namespace {

std::string LibraryPath() {
    if (const char* env = std::getenv("TERRAINGEN_HIP_LIBRARY"); env != nullptr && *env != 0) {
        return env;
    }
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&HipUnavailableReason), &info) != 0 &&
        info.dli_fname != nullptr) {
        std::string self = info.dli_fname;
        const std::size_t slash = self.rfind('/');
        return (slash == std::string::npos ? std::string() : self.substr(0, slash + 1)) +
               "libtg_hip.so";
    }
    return "libtg_hip.so";
}

struct Loaded {
    HipApi api{};
    std::string reason;
};

const Loaded& Load() {
    static Loaded loaded;
    static std::once_flag once;
    std::call_once(once, [] {
        const std::string path = LibraryPath();
        void* handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (handle == nullptr) {
            const char* why = dlerror();
            loaded.reason = "cannot open " + path + ": " + (why != nullptr ? why : "unknown");
            return;
        }
        bool ok = true;
        auto bind = [&](auto& fn, const char* name) {
            void* sym = dlsym(handle, name);
            if (sym == nullptr) {
                ok = false;
                loaded.reason = path + " lacks " + name;
            }
            fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(sym);
        };
        HipApi& a = loaded.api;
        bind(a.open, "tg_hip_open");
        bind(a.device_name, "tg_hip_device_name");
        bind(a.can_run, "tg_hip_can_run");
        bind(a.run, "tg_hip_run");
        bind(a.batch_create, "tg_hip_batch_create");
        bind(a.batch_free, "tg_hip_batch_free");
        bind(a.batch_stage, "tg_hip_batch_stage");
        bind(a.batch_stage_async, "tg_hip_batch_stage_async");
        bind(a.batch_surface, "tg_hip_batch_surface");
        bind(a.batch_stage_boulders, "tg_hip_batch_stage_boulders");
        bind(a.batch_run, "tg_hip_batch_run");
        bind(a.batch_stream, "tg_hip_batch_stream");
        bind(a.malloc, "tg_hip_malloc");
        bind(a.free, "tg_hip_free");
        bind(a.host_alloc, "tg_hip_host_alloc");
        bind(a.host_free, "tg_hip_host_free");
        bind(a.memcpy_d2h_async, "tg_hip_memcpy_d2h_async");
        bind(a.event_create, "tg_hip_event_create");
        bind(a.event_destroy, "tg_hip_event_destroy");
        bind(a.event_record, "tg_hip_event_record");
        bind(a.event_query, "tg_hip_event_query");
        bind(a.stream_synchronize, "tg_hip_stream_synchronize");
        if (!ok) {
            loaded.api = HipApi{};
        }
    });
    return loaded;
}

}  // namespace

const HipApi& Hip() {
    const Loaded& l = Load();
    if (!l.reason.empty()) {
        throw std::runtime_error(l.reason);
    }
    return l.api;
}

std::string HipUnavailableReason() {
    return Load().reason;
}

}  // namespace tgh
