// lattice_inst.hpp -- definitions of the lattice launchers; included by one
// lattice_q<Q>_<fmt>.hip file per (lattice, storage format).
#pragma once

#include "launch.hpp"

namespace ssmd {

namespace detail {

template <int Q, class P, int TX, int TY, bool MACROS>
hipError_t lattice_one(const LatticeArgs &a, hipStream_t s) {
    const int lz = a.lz;
    const dim3 grid((a.k.nx + TX - 1) / TX, (a.k.ny + TY - 1) / TY, (a.k.nz - 2 + lz - 1) / lz);
    hipLaunchKernelGGL((k_lattice<Q, P, TX, TY, MACROS>), grid, dim3(TX, TY, 1), 0, s, a);
    return hipGetLastError();
}

template <int Q, class P, int TX, int TY, bool MACROS>
hipError_t lattice_batch_one(const JetSlot *jets, int njets, uint32_t total_tiles, uint32_t *queue, hipStream_t s) {
    static int resident = 0;  // blocks of this kernel that fit on the device at once
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess) return hipErrorInvalidDevice;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lattice_batch<Q, P, TX, TY, MACROS>, TX * TY, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        resident = cus * per_cu;
    }
    const uint32_t grid = total_tiles < uint32_t(resident) ? total_tiles : uint32_t(resident);
    hipLaunchKernelGGL((k_lattice_batch<Q, P, TX, TY, MACROS>), dim3(grid), dim3(TX, TY, 1), 0, s, jets, njets,
                       total_tiles, queue);
    return hipGetLastError();
}

}  // namespace detail

template <int Q, class P>
hipError_t launch_lattice(int tile, bool macros, const LatticeArgs &a, hipStream_t s) {
    if (macros) return detail::lattice_one<Q, P, 64, 4, true>(a, s);
    switch (tile) {
        case 0: return detail::lattice_one<Q, P, 64, 4, false>(a, s);
        case 1: return detail::lattice_one<Q, P, 128, 2, false>(a, s);
        case 2: return detail::lattice_one<Q, P, 64, 2, false>(a, s);
        case 3: return detail::lattice_one<Q, P, 128, 4, false>(a, s);
        case 4: return detail::lattice_one<Q, P, 128, 4, false>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <int Q, class P>
hipError_t launch_lattice_batch(int tile, bool macros, const JetSlot *jets, int njets, uint32_t total_tiles,
                                uint32_t *queue, hipStream_t s) {
    if (macros) return detail::lattice_batch_one<Q, P, 64, 4, true>(jets, njets, total_tiles, queue, s);
    switch (tile) {
        case 0: return detail::lattice_batch_one<Q, P, 64, 4, false>(jets, njets, total_tiles, queue, s);
        case 1: return detail::lattice_batch_one<Q, P, 128, 2, false>(jets, njets, total_tiles, queue, s);
        case 2: return detail::lattice_batch_one<Q, P, 64, 2, false>(jets, njets, total_tiles, queue, s);
        case 3: return detail::lattice_batch_one<Q, P, 128, 4, false>(jets, njets, total_tiles, queue, s);
        // the persistent kernel takes tiles from a queue: the XCD order of
        // tile 4 does not apply, its shape does
        case 4: return detail::lattice_batch_one<Q, P, 128, 4, false>(jets, njets, total_tiles, queue, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ssmd

#define SSMD_INSTANTIATE_LATTICE(Q, PT)                                                                         \
    namespace ssmd {                                                                                            \
    template hipError_t launch_lattice<Q, PT>(int, bool, const LatticeArgs &, hipStream_t);                     \
    template hipError_t launch_lattice_batch<Q, PT>(int, bool, const JetSlot *, int, uint32_t, uint32_t *,       \
                                                    hipStream_t);                                                 \
    }
