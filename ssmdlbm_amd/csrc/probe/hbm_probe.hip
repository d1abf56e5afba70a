// hbm_probe.hip -- streaming-bandwidth probe of the device's HBM, built into
// its own library (libssmd_probe.so) beside the solver.  bench.py measures
// with it, in the same process as the solver workloads, the rate a plain
// stream reaches on this part: the practical ceiling the lattice kernel's
// achieved bandwidth is set against.  It is not part of the solver's ABI.
//
// Kernels: 256-thread blocks, 16-byte (dwordx4) accesses, UNROLL accesses
// in flight per thread, either one pass (grid = bytes / (256 * 16 * UNROLL))
// or a persistent grid striding over the buffer.  Loads and stores are either
// cached or non-temporal (__builtin_nontemporal_load/store).
//   op 0  read:  sum of every dword, one partial per block (no elision)
//   op 1  write: a constant pattern
//   op 2  copy:  dst = src
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u32x4 ld(const u32x4 *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st(u32x4 *p, u32x4 v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// n16: number of 16-byte elements, a multiple of kThreads * kUnroll.
template <int OP, bool NT>
__global__ void __launch_bounds__(kThreads) k_stream(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                     uint64_t n16, uint32_t *__restrict__ partial) {
    const uint64_t chunk = uint64_t(kThreads) * kUnroll;
    uint32_t acc = 0;
    for (uint64_t base = uint64_t(blockIdx.x) * chunk; base < n16; base += uint64_t(gridDim.x) * chunk) {
        const uint64_t i0 = base + threadIdx.x;
        if constexpr (OP == 0) {
            u32x4 v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) v[u] = ld<NT>(src + i0 + uint64_t(u) * kThreads);
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) acc += v[u].x + v[u].y + v[u].z + v[u].w;
        } else if constexpr (OP == 1) {
            const u32x4 v = {uint32_t(i0), 1u, 2u, 3u};
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) st<NT>(dst + i0 + uint64_t(u) * kThreads, v);
        } else {
            u32x4 v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) v[u] = ld<NT>(src + i0 + uint64_t(u) * kThreads);
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) st<NT>(dst + i0 + uint64_t(u) * kThreads, v[u]);
        }
    }
    if constexpr (OP == 0) {
        __shared__ uint32_t red[kThreads];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kThreads / 2; s > 0; s >>= 1) {
            if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
    }
}

template <int OP>
void launch(bool nt, dim3 grid, hipStream_t s, const void *src, void *dst, uint64_t n16, uint32_t *partial) {
    const u32x4 *a = static_cast<const u32x4 *>(src);
    u32x4 *b = static_cast<u32x4 *>(dst);
    if (nt) hipLaunchKernelGGL((k_stream<OP, true>), grid, dim3(kThreads), 0, s, a, b, n16, partial);
    else hipLaunchKernelGGL((k_stream<OP, false>), grid, dim3(kThreads), 0, s, a, b, n16, partial);
}

}  // namespace

extern "C" {

// Largest block count a probe launch may use (size of the caller's `partial`).
uint32_t ssmd_probe_max_blocks(uint64_t bytes) {
    const uint64_t chunk = uint64_t(kThreads) * kUnroll * 16;
    return uint32_t((bytes + chunk - 1) / chunk);
}

// Mean time (ms) of `reps` launches after one warm-up.  op: 0 read, 1 write,
// 2 copy; nt: non-temporal accesses; blocks: 0 = one pass over the buffer,
// else a persistent grid of that many blocks.  bytes must be a multiple of
// 256 * 16 * 4; partial holds ssmd_probe_max_blocks(bytes) uint32 (op 0).
// Returns 0 or a hipError_t value; -1 on bad arguments.
int ssmd_probe_stream(int op, int nt, uint32_t blocks, const void *src, void *dst, uint64_t bytes, int reps,
                      uint32_t *partial, void *stream, float *ms) {
    const uint64_t chunk = uint64_t(kThreads) * kUnroll * 16;
    if (op < 0 || op > 2 || reps < 1 || !ms || bytes == 0 || bytes % chunk) return -1;
    if ((op != 1 && !src) || (op != 0 && !dst) || (op == 0 && !partial)) return -1;
    const uint32_t one_pass = uint32_t(bytes / chunk);
    const uint32_t nb = blocks ? (blocks < one_pass ? blocks : one_pass) : one_pass;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t n16 = bytes / 16;
    auto go = [&]() {
        if (op == 0) launch<0>(nt, dim3(nb), s, src, dst, n16, partial);
        else if (op == 1) launch<1>(nt, dim3(nb), s, src, dst, n16, partial);
        else launch<2>(nt, dim3(nb), s, src, dst, n16, partial);
    };
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return int(hipGetLastError());
    if (hipEventCreate(&e1) != hipSuccess) return int(hipGetLastError());
    go();
    hipError_t e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) go();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / float(reps);
    return e == hipSuccess ? 0 : int(e);
}

}  // extern "C"
