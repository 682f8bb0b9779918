// SPDX-License-Identifier: BSD-3-Clause
//
// The registry, the hip device and the strip dispatcher (backend.h). Semantics follow the
// reference's src/gpu/backend.cpp:24-200 (names, probing, default and resolution) and
// src/gpu/kernel.cpp:15-153 (dispatchable grids, FP64 cell offsets, strips of rows).

#include "backend.h"

#include "common.h"
#include "hip_api.h"

#include <atomic>
#include <cctype>
#include <cstdlib>
#include <mutex>
#include <stdexcept>

namespace tgamd {

namespace {

std::atomic<int> gDefault{-1};

// The backends in kBackendNames' order.
constexpr std::array<Backend, kBackendNames.size()> kNamed{
    Backend::Auto, Backend::Cpu, Backend::Gpu, Backend::Vulkan, Backend::Cuda, Backend::Metal,
    Backend::Hip};

// What Gpu and Auto choose from, best first.
constexpr Backend kGpuBackends[] = {Backend::Hip, Backend::Metal, Backend::Cuda, Backend::Vulkan};

bool IsDeviceBackend(Backend b) {
    return b == Backend::Vulkan || b == Backend::Cuda || b == Backend::Metal || b == Backend::Hip;
}

std::string Lower(std::string_view s) {
    std::string out(s);
    for (char& c : out) {
        c = char(std::tolower(static_cast<unsigned char>(c)));
    }
    return out;
}

BackendStatus Probe(Backend backend) {
    BackendStatus status{backend, false, {}, {}};
    if (backend == Backend::Cpu) {
        status.available = true;
        status.device = "host";
        return status;
    }
    if (backend == Backend::Hip) {
        if (const HipCompute* hip = HipCompute::Get()) {
            status.available = true;
            status.device = hip->DeviceName();
        } else {
            status.reason = HipCompute::Reason();
        }
        return status;
    }
    if (IsDeviceBackend(backend)) {
        status.reason = "this build has no " + std::string(BackendName(backend)) +
                        " device; its GPU backend is hip";
        return status;
    }
    status.reason = std::string(BackendName(backend)) + " is not a device backend";
    return status;
}

std::optional<Backend> BestGpu() {
    for (const Backend b : kGpuBackends) {
        if (Probe(b).available) {
            return b;
        }
    }
    return std::nullopt;
}

constexpr std::size_t kDefaultStripBytes = std::size_t{1} << 26;  // as the reference: 64 MiB
constexpr std::size_t kGroupSide = 16;
constexpr std::size_t kMaxGroups = 65535;

}  // namespace

std::string BackendNameList() {
    std::string list;
    for (std::size_t i = 0; i < kBackendNames.size(); ++i) {
        if (i > 0) {
            list += i + 1 == kBackendNames.size() ? " or " : ", ";
        }
        list += kBackendNames[i];
    }
    return list;
}

std::optional<Backend> BackendNamed(std::string_view name) {
    const std::string lower = Lower(name);
    for (std::size_t i = 0; i < kBackendNames.size(); ++i) {
        if (lower == kBackendNames[i]) {
            return kNamed[i];
        }
    }
    return std::nullopt;
}

std::string_view BackendName(Backend backend) {
    for (std::size_t i = 0; i < kNamed.size(); ++i) {
        if (kNamed[i] == backend) {
            return kBackendNames[i];
        }
    }
    return "unknown";
}

std::string DescribeBackend(Backend backend) {
    if (IsDeviceBackend(backend)) {
        if (const BackendStatus s = Probe(backend); s.available) {
            return std::string(BackendName(backend)) + " (" + s.device + ")";
        }
    }
    return std::string(BackendName(backend));
}

std::vector<BackendStatus> Backends() {
    return {Probe(Backend::Cpu), Probe(Backend::Vulkan), Probe(Backend::Cuda),
            Probe(Backend::Metal), Probe(Backend::Hip)};
}

Backend DefaultBackend() {
    if (const int set = gDefault.load(); set >= 0) {
        return static_cast<Backend>(set);
    }
    const char* name = std::getenv("TERRAINGEN_BACKEND");
    if (name == nullptr || *name == '\0') {
        return Backend::Cpu;
    }
    if (const std::optional<Backend> b = BackendNamed(name)) {
        return *b;
    }
    throw std::invalid_argument("TERRAINGEN_BACKEND is \"" + std::string(name) + "\"; expected " +
                                BackendNameList());
}

void SetDefaultBackend(Backend backend) {
    gDefault.store(static_cast<int>(backend));
}

Backend ResolveBackend(Backend requested) {
    if (requested == Backend::Cpu) {
        return Backend::Cpu;
    }
    if (IsDeviceBackend(requested)) {
        if (const BackendStatus s = Probe(requested); !s.available) {
            throw std::runtime_error("The " + std::string(BackendName(requested)) +
                                     " backend is unavailable: " + s.reason);
        }
        return requested;
    }
    if (requested == Backend::Gpu) {
        if (const std::optional<Backend> gpu = BestGpu()) {
            return *gpu;
        }
        std::string reasons;
        for (const Backend b : kGpuBackends) {
            reasons += "; " + std::string(BackendName(b)) + ": " + Probe(b).reason;
        }
        throw std::runtime_error("No GPU backend is available" + reasons);
    }
    return BestGpu().value_or(Backend::Cpu);  // Auto
}

// ---- The hip device ---------------------------------------------------------------------------

namespace {

struct HipShared {
    std::once_flag once;
    HipCompute* device = nullptr;
    std::string reason;
};

HipShared& Shared() {
    static HipShared shared;
    return shared;
}

}  // namespace

HipCompute* HipCompute::Get() {
    HipShared& s = Shared();
    std::call_once(s.once, [&s] {
        try {
            const tgh::HipApi& api = tgh::Hip();
            auto* device = new HipCompute();
            char err[512] = {0};
            if (api.open(0, &device->m_dev, err, sizeof err) != TG_OK) {
                delete device;
                throw std::runtime_error(err[0] != 0 ? err : "tg_hip_open failed");
            }
            const char* name = api.device_name(device->m_dev);
            device->m_name = name != nullptr ? name : "";
            s.device = device;  // never destroyed, as the reference's LazyDevice
        } catch (const std::exception& e) {
            s.reason = e.what();
        }
    });
    return s.device;
}

std::string HipCompute::Reason() {
    Get();
    return Shared().reason;
}

const std::string& HipCompute::DeviceName() const {
    return m_name;
}

bool HipCompute::CanRun(const KernelProgram& program) const {
    return program.hip != 0 && tgh::Hip().can_run(m_dev, program.hip) != 0;
}

void HipCompute::Run(const KernelProgram& program, const std::vector<DeviceBuffer>& buffers,
                     const void* pushConstants, std::size_t pushBytes, const Dispatch&) {
    if (buffers.size() < 4 || buffers[0].output == nullptr ||
        pushBytes != sizeof(tg_kernel_params)) {
        throw std::runtime_error("hip: a kernel run binds heights, xs, ys and its parameter block");
    }
    std::vector<tg_buf> extra;
    for (std::size_t i = 4; i < buffers.size(); ++i) {
        extra.push_back({buffers[i].bytes, buffers[i].size, buffers[i].count, 0});
    }
    tg_kernel_params params;
    std::memcpy(&params, pushConstants, sizeof params);
    char err[512] = {0};
    const int code = tgh::Hip().run(
        m_dev, program.hip, buffers[3].bytes, buffers[3].size,
        static_cast<float*>(buffers[0].output), static_cast<const float*>(buffers[1].bytes),
        static_cast<const float*>(buffers[2].bytes), &params, extra.data(), extra.size(), err,
        sizeof err);
    if (code != TG_OK) {
        throw std::runtime_error(err[0] != 0 ? err : "tg_hip_run failed");
    }
}

ComputeDevice* GpuDevice(Backend backend) {
    return backend == Backend::Hip ? HipCompute::Get() : nullptr;
}

// ---- RunKernel ---------------------------------------------------------------------------------

bool RunKernel(Backend backend, const KernelProgram& program, const void* kernel,
               std::size_t kernelBytes, Grid& grid, double originX, double originY,
               const std::vector<KernelBuffer>& extra, std::size_t maxStripBytes) {
    if (grid.width == 0 || grid.height == 0 || grid.values == nullptr ||
        (grid.width + kGroupSide - 1) / kGroupSide > kMaxGroups) {
        return false;
    }
    ComputeDevice* device = GpuDevice(backend);
    if (device == nullptr || !device->CanRun(program)) {
        return false;
    }
    const tgh::Grid g{grid.width, grid.height, grid.originX, grid.originY, grid.pixel};
    // Cell centres from the origin, in FP64, narrowed per column and per row.
    std::vector<float> xs(grid.width), ys(grid.height);
    for (std::size_t c = 0; c < grid.width; ++c) {
        xs[c] = static_cast<float>(g.CenterX(c) - originX);
    }
    for (std::size_t r = 0; r < grid.height; ++r) {
        ys[r] = static_cast<float>(g.CenterY(r) - originY);
    }
    tg_kernel_params params{};
    params.width = static_cast<std::uint32_t>(grid.width);
    if (grid.noData) {
        params.no_data = std::isnan(*grid.noData) ? TG_NODATA_NAN : TG_NODATA_VALUE;
        params.no_data_value = *grid.noData;
    }
    // Rows per strip: the widest per-cell buffer stays under the strip size.
    std::size_t cellBytes = sizeof(float);
    const std::size_t cells = grid.width * grid.height;
    for (const KernelBuffer& b : extra) {
        if (b.perCell != 0) {
            cellBytes = std::max(cellBytes, b.size / cells);
        }
    }
    const std::size_t strip = maxStripBytes != 0 ? maxStripBytes : kDefaultStripBytes;
    const std::size_t rowsPerStrip =
        std::clamp<std::size_t>(strip / (cellBytes * grid.width), 1, kMaxGroups * kGroupSide);
    for (std::size_t row0 = 0; row0 < grid.height; row0 += rowsPerStrip) {
        const std::size_t rows = std::min(rowsPerStrip, grid.height - row0);
        params.rows = static_cast<std::uint32_t>(rows);
        params.first_row = static_cast<std::uint32_t>(row0);
        float* heights = grid.values + row0 * grid.width;
        std::vector<DeviceBuffer> buffers{
            {heights, rows * grid.width * sizeof(float), rows * grid.width, heights},
            {xs.data(), xs.size() * sizeof(float), xs.size(), nullptr},
            {ys.data() + row0, rows * sizeof(float), rows, nullptr},
            {kernel, kernelBytes, 1, nullptr}};
        for (const KernelBuffer& b : extra) {
            if (b.perCell == 0) {
                buffers.push_back({b.bytes, b.size, b.count, nullptr});
                continue;
            }
            const std::size_t elements = b.perCell * grid.width;
            const std::size_t rowBytes = b.size / b.count * elements;
            buffers.push_back({static_cast<const std::byte*>(b.bytes) + row0 * rowBytes,
                               rows * rowBytes, rows * elements, nullptr});
        }
        Dispatch dispatch;
        dispatch.groups = {std::uint32_t((grid.width + kGroupSide - 1) / kGroupSide),
                           std::uint32_t((rows + kGroupSide - 1) / kGroupSide), 1u};
        dispatch.groupSize = {std::uint32_t(kGroupSide), std::uint32_t(kGroupSide), 1u};
        device->Run(program, buffers, &params, sizeof params, dispatch);
    }
    return true;
}

}  // namespace tgamd

// ---- C exports for hosts that are not C++ ---------------------------------------------------

extern "C" {

int tgx_backend_named(const char* name, int* backend) {
    if (name == nullptr || backend == nullptr) {
        return TG_ERR_ARGUMENT;
    }
    const std::optional<tgamd::Backend> b = tgamd::BackendNamed(name);
    if (!b) {
        return TG_ERR_ARGUMENT;
    }
    *backend = static_cast<int>(*b);
    return TG_OK;
}

size_t tgx_backend_name(int backend, char* out, size_t len) {
    const std::string name(tgamd::BackendName(static_cast<tgamd::Backend>(backend)));
    tgh::SetErr(out, len, name);
    return name.size();
}

size_t tgx_backends_report(char* out, size_t len) {
    // What `terraingen backends` prints (app/cli.cpp:192-205), default line included.
    std::string text;
    for (const tgamd::BackendStatus& s : tgamd::Backends()) {
        text += std::string(tgamd::BackendName(s.backend)) + ": " +
                (s.available ? "available on " + s.device : "unavailable: " + s.reason) + "\n";
    }
    try {
        text += "default: " + std::string(tgamd::BackendName(tgamd::DefaultBackend())) + "\n";
    } catch (const std::invalid_argument& e) {
        text += std::string("error: ") + e.what() + "\n";
    }
    tgh::SetErr(out, len, text);
    return text.size();
}

int tgx_default_backend(int* backend, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (backend == nullptr) {
            throw std::invalid_argument("tgx_default_backend: null argument");
        }
        *backend = static_cast<int>(tgamd::DefaultBackend());
    });
}

void tgx_set_default_backend(int backend) {
    tgamd::SetDefaultBackend(static_cast<tgamd::Backend>(backend));
}

int tgx_resolve_backend(int requested, int* backend, char* err, size_t errLen) {
    if (backend == nullptr) {
        tgh::SetErr(err, errLen, "tgx_resolve_backend: null argument");
        return TG_ERR_ARGUMENT;
    }
    try {
        *backend = static_cast<int>(tgamd::ResolveBackend(static_cast<tgamd::Backend>(requested)));
        return TG_OK;
    } catch (const std::exception& e) {
        tgh::SetErr(err, errLen, e.what());
        return TG_ERR_NO_DEVICE;
    }
}

int tgx_run_kernel(int backend, int program, const void* kernel, size_t kernel_bytes,
                   const tgx_grid* grid, float* values, const float* no_data, double origin_x,
                   double origin_y, const tg_buf* extra, size_t n_extra, size_t max_strip_bytes,
                   int* ran, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (grid == nullptr || ran == nullptr || (n_extra > 0 && extra == nullptr) ||
            (grid->width > 0 && grid->height > 0 && values == nullptr)) {
            throw std::invalid_argument("tgx_run_kernel: null argument");
        }
        tgamd::Grid g;
        g.width = grid->width;
        g.height = grid->height;
        g.originX = grid->origin_x;
        g.originY = grid->origin_y;
        g.pixel = grid->pixel;
        g.values = values;
        if (no_data != nullptr) {
            g.noData = *no_data;
        }
        std::vector<tgamd::KernelBuffer> buffers;
        for (size_t i = 0; i < n_extra; ++i) {
            buffers.push_back({extra[i].data, extra[i].bytes, extra[i].count, extra[i].per_cell});
        }
        tgamd::KernelProgram p;
        p.hip = program;
        *ran = tgamd::RunKernel(static_cast<tgamd::Backend>(backend), p, kernel, kernel_bytes, g,
                                origin_x, origin_y, buffers, max_strip_bytes)
                   ? 1
                   : 0;
    });
}

}  // extern "C"
