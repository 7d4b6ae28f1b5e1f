// fsi_launch.hpp -- how the library's kernels are launched, alone or batched.
//
// Every kernel body is a device function (FSI_KERNEL); fsi_launch<&body, LB,
// W>(grid, block, shared, stream, args...) runs it through a generic
// __global__ wrapper with launch bounds LB (and at most W waves per SIMD
// when W > 0), so a single launch compiles to the body with its arguments
// in the kernel argument segment, as a plain __global__ kernel would.
//
// While a thread has a LaunchRecorder installed, fsi_launch records the
// launch instead: the batched twin of the wrapper and the argument values
// packed in a std::tuple.  Instances whose recorded sequences agree kernel by
// kernel (same grids, blocks and LDS) are then run together: one launch per
// kernel with grid.z = the number of instances, block z reading its own
// instance's arguments from a device table (uniform address, scalar loads).
// Grids are 2-D (z = 1) everywhere in the library; the XCD-aware tile order
// of tile_id() is unaffected by z when it applies (block counts that are a
// multiple of 8).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <tuple>
#include <utility>
#include <vector>

#define FSI_KERNEL __device__ __forceinline__

namespace fsi {

template <class F>
struct KernelParams;
template <class... P>
struct KernelParams<void (*)(P...)> {
    using Tuple = std::tuple<P...>;
};

template <auto F, int LB, class... P>
__global__ void __launch_bounds__(LB) k_single(P... a) {
    F(a...);
}
template <auto F, int LB, int W, class... P>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(1, W))) k_single_w(P... a) {
    F(a...);
}
template <auto F, int LB, class Tuple>
__global__ void __launch_bounds__(LB) k_batch(const char* __restrict__ tab, int stride) {
    const Tuple& a = *reinterpret_cast<const Tuple*>(tab + (size_t)blockIdx.z * stride);
    std::apply(F, a);
}
template <auto F, int LB, int W, class Tuple>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(1, W))) k_batch_w(const char* __restrict__ tab,
                                                                                          int stride) {
    const Tuple& a = *reinterpret_cast<const Tuple*>(tab + (size_t)blockIdx.z * stride);
    std::apply(F, a);
}

// Launches the batched twin of a recorded kernel over `count` instances
// whose argument tuples sit `stride` bytes apart from `tab` (device memory).
using BatchedLauncher = void (*)(dim3 grid, dim3 block, size_t shared, hipStream_t s, const char* tab, int count,
                                 int stride);

struct RecordedLaunch {
    BatchedLauncher launch;
    dim3 grid, block;
    size_t shared;
    size_t offset, size;  // the argument tuple in LaunchRecorder::args
};

struct LaunchRecorder {
    std::vector<RecordedLaunch> launches;
    std::vector<char> args;
    bool overflow = false;  // a launch that cannot be batched (grid.z != 1)
    void clear() {
        launches.clear();
        args.clear();
        overflow = false;
    }
};

// the recorder of the calling thread, or null (launch directly)
extern thread_local LaunchRecorder* g_recorder;

template <auto F, int LB, int W, class Tuple>
void launch_batched(dim3 grid, dim3 block, size_t shared, hipStream_t s, const char* tab, int count, int stride) {
    const dim3 g(grid.x, grid.y, (unsigned)count);
    if constexpr (W > 0)
        k_batch_w<F, LB, W, Tuple><<<g, block, shared, s>>>(tab, stride);
    else
        k_batch<F, LB, Tuple><<<g, block, shared, s>>>(tab, stride);
}

template <auto F, int LB, int W, class... P, class... A>
inline void launch_with(std::tuple<P...>*, dim3 grid, dim3 block, size_t shared, hipStream_t s, A&&... args) {
    if (LaunchRecorder* r = g_recorder) {
        using Tuple = std::tuple<P...>;
        const Tuple t(static_cast<P>(args)...);
        if (grid.z != 1) r->overflow = true;
        const size_t off = (r->args.size() + 15) / 16 * 16;
        r->args.resize(off + sizeof(Tuple));
        std::memcpy(r->args.data() + off, &t, sizeof(Tuple));
        r->launches.push_back({&launch_batched<F, LB, W, Tuple>, grid, block, shared, off, sizeof(Tuple)});
        return;
    }
    if constexpr (W > 0)
        k_single_w<F, LB, W, P...><<<grid, block, shared, s>>>(static_cast<P>(args)...);
    else
        k_single<F, LB, P...><<<grid, block, shared, s>>>(static_cast<P>(args)...);
}

// fsi_launch<&body, LB[, W]>(grid, block, shared, stream, args...)
template <auto F, int LB, int W = 0, class... A>
inline void fsi_launch(dim3 grid, dim3 block, size_t shared, hipStream_t s, A&&... args) {
    using Tuple = typename KernelParams<decltype(F)>::Tuple;
    launch_with<F, LB, W>(static_cast<Tuple*>(nullptr), grid, block, shared, s, std::forward<A>(args)...);
}

template <auto F, int LB, int W, class... P>
hipError_t allow_dynamic_lds_with(std::tuple<P...>*, int bytes) {
    using Tuple = std::tuple<P...>;
    const void *single, *batch;
    if constexpr (W > 0) {
        single = reinterpret_cast<const void*>(&k_single_w<F, LB, W, P...>);
        batch = reinterpret_cast<const void*>(&k_batch_w<F, LB, W, Tuple>);
    } else {
        single = reinterpret_cast<const void*>(&k_single<F, LB, P...>);
        batch = reinterpret_cast<const void*>(&k_batch<F, LB, Tuple>);
    }
    hipError_t e = hipFuncSetAttribute(single, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess) e = hipFuncSetAttribute(batch, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    return e;
}

// raises the dynamic LDS limit of both wrappers of a body (above 64 KB)
template <auto F, int LB, int W = 0>
hipError_t fsi_allow_dynamic_lds(int bytes) {
    using Tuple = typename KernelParams<decltype(F)>::Tuple;
    return allow_dynamic_lds_with<F, LB, W>(static_cast<Tuple*>(nullptr), bytes);
}

}  // namespace fsi
