// SPDX-License-Identifier: BSD-3-Clause
//
// The backend registry, the device interface and the strip dispatcher of TerrainGen, with the hip
// backend in them: the same names, argument meanings and error behaviour as the reference's
// include/terraingen/gpu/backend.h, src/gpu/device.h and src/gpu/kernel.h, so that code written
// against those reads the same against this. Only the hip backend has a device in this build; the
// Vulkan, CUDA and Metal names parse and report themselves unavailable, as they do in a reference
// build whose drivers are absent.
#pragma once

#include "tg_hip.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <string>
#include <string_view>
#include <vector>

namespace tgamd {

// Backend (include/terraingen/gpu/backend.h:16-23) with Hip appended.
enum class Backend { Auto, Cpu, Vulkan, Gpu, Cuda, Metal, Hip };

// kBackendNames (backend.h:26-27) with "hip" appended.
inline constexpr std::array<std::string_view, 7> kBackendNames{"auto",  "cpu",   "gpu", "vulkan",
                                                               "cuda",  "metal", "hip"};

std::string BackendNameList();
// Ignores case; std::nullopt for an unknown name.
std::optional<Backend> BackendNamed(std::string_view name);
// "unknown" for an out-of-range value.
std::string_view BackendName(Backend backend);
// The name, and for an available GPU backend its device: "hip (AMD Instinct MI355X)".
std::string DescribeBackend(Backend backend);

struct BackendStatus {
    Backend backend = Backend::Cpu;
    bool available = false;
    std::string device;  // when available
    std::string reason;  // otherwise
};

// Cpu, then each device backend: Vulkan, Cuda, Metal, Hip.
std::vector<BackendStatus> Backends();

// SetDefaultBackend's choice, else $TERRAINGEN_BACKEND (std::invalid_argument for an unknown
// name), else Cpu.
Backend DefaultBackend();
void SetDefaultBackend(Backend backend);

// Gpu: the best available device backend (Hip first); Auto: the same or Cpu; a named device
// backend: itself. std::runtime_error with the reason when a requested GPU backend cannot run.
Backend ResolveBackend(Backend requested);

// DeviceBuffer, Dispatch and ComputeDevice (src/gpu/device.h:18-64).
struct DeviceBuffer {
    const void* bytes = nullptr;
    std::size_t size = 0;    // bytes
    std::size_t count = 0;   // elements the kernel sees
    void* output = nullptr;  // where the bytes return after the run; null for an input
};

struct Dispatch {
    std::array<std::uint32_t, 3> groups{1u, 1u, 1u};
    std::array<std::uint32_t, 3> groupSize{1u, 1u, 1u};
};

// KernelProgram (include/terraingen/gpu/program.h:14-19) reduced to the form this build has: the
// kernel's id in libtg_hip.so, 0 for none.
struct KernelProgram {
    int hip = 0;
};

class ComputeDevice {
public:
    virtual ~ComputeDevice() = default;
    virtual const std::string& DeviceName() const = 0;
    virtual bool CanRun(const KernelProgram& program) const = 0;
    // Binds `buffers` in order (heights, xs, ys, the kernel's parameter block, its extras) with
    // the push constants after them, runs, copies outputs back. std::runtime_error on failure.
    virtual void Run(const KernelProgram& program, const std::vector<DeviceBuffer>& buffers,
                     const void* pushConstants, std::size_t pushBytes, const Dispatch& dispatch) = 0;
};

// The hip device: libtg_hip.so opened on first use, GPU ordinal 0, never closed.
class HipCompute : public ComputeDevice {
public:
    static HipCompute* Get();        // null when unavailable
    static std::string Reason();     // empty when available
    const std::string& DeviceName() const override;
    bool CanRun(const KernelProgram& program) const override;
    void Run(const KernelProgram& program, const std::vector<DeviceBuffer>& buffers,
             const void* pushConstants, std::size_t pushBytes, const Dispatch& dispatch) override;
    tg_dev* Handle() const {
        return m_dev;
    }

private:
    HipCompute() = default;
    tg_dev* m_dev = nullptr;
    std::string m_name;
};

// GpuDevice (device.h:63-64): the device of a device backend; null otherwise or unavailable.
ComputeDevice* GpuDevice(Backend backend);

// A north-up grid of FP32 heights (the part of Raster<float> RunKernel reads).
struct Grid {
    std::size_t width = 0;
    std::size_t height = 0;
    double originX = 0.0;  // west edge
    double originY = 0.0;  // north edge
    double pixel = 1.0;
    float* values = nullptr;            // width x height, row-major from the north-west cell
    std::optional<float> noData;        // cells holding it are left alone
};

// KernelBuffer (src/gpu/kernel.h:20-27).
struct KernelBuffer {
    const void* bytes = nullptr;
    std::size_t size = 0;
    std::size_t count = 0;
    std::size_t perCell = 0;  // elements per grid cell for a buffer sliced per strip; 0: whole
};

// RunKernel (src/gpu/kernel.h:35-47): every cell holding data gets the kernel at its center
// measured from `origin` (FP64 per column and row, narrowed to FP32), in strips of rows of at
// most `maxStripBytes` per buffer (0: 64 MiB, as the reference). False, leaving `grid` untouched,
// when the backend is not a GPU one, its device is unavailable, the program has no form for it,
// or the grid is empty or wider than one dispatch reaches.
bool RunKernel(Backend backend, const KernelProgram& program, const void* kernel,
               std::size_t kernelBytes, Grid& grid, double originX, double originY,
               const std::vector<KernelBuffer>& extra = {}, std::size_t maxStripBytes = 0);

}  // namespace tgamd
