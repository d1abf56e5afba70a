// launch.hpp -- host-side launchers of the fused lattice kernel, compiled
// once per (lattice, storage format) in lattice_q*_*.hip so the heavy
// instantiations build in parallel.
#pragma once

#include "kernels.hpp"

namespace ssmd {

// Tile shapes the lattice kernels are built for (the autotuner's candidates),
// and the order tiles are dealt to the XCDs in (LatticeArgs::xcd_order).
struct TileShape {
    int tx, ty;
    int xcd_order;
};
constexpr TileShape kTiles[] = {{64, 4, 0}, {128, 2, 0}, {64, 2, 0}, {128, 4, 0}, {128, 4, 1}};
constexpr int kNumTiles = int(sizeof(kTiles) / sizeof(kTiles[0]));

// z planes per block: enough chunks that the grid holds >= ~2048 blocks on
// large boxes without letting the 4-plane halo prologue dominate a chunk.
inline int auto_lz_shape(int nx, int ny, int nz, int tx, int ty) {
    const int interior = nz - 2;
    const long cols = long((nx + tx - 1) / tx) * long((ny + ty - 1) / ty);
    int lz = interior;
    while (lz > 16 && cols * ((interior + lz - 1) / lz) < 2048) lz = (lz + 1) / 2;
    return lz < 1 ? 1 : lz;
}
inline int auto_lz(int nx, int ny, int nz, int tile) { return auto_lz_shape(nx, ny, nz, kTiles[tile].tx, kTiles[tile].ty); }

// Ensembles also take tile index kNumTiles: 128 x 4 blocks that run 64 x 8
// tiles for the members whose rows a 64-wide tile covers with fewer idle lanes
// (k_lattice_batch).  (64 x 8 tiles for every member of the 32-jet
// ensemble measured 3.26 ms per step against 2.87 ms for this choice.)
constexpr int kEnsembleMixedTile = kNumTiles;
inline bool narrow_member(int nx) { return (nx + 63) / 64 * 64 < (nx + 127) / 128 * 128; }

// One field set.  `macros` runs the tile-0 kernel that writes every moment.
template <int Q, class P>
hipError_t launch_lattice(int tile, bool macros, const LatticeArgs &a, hipStream_t s);

// Ensemble: persistent kernel over all members' tiles; grid = resident blocks;
// `queue`: two zeroed words of device memory (the kernel leaves them zeroed).
template <int Q, class P>
hipError_t launch_lattice_batch(int tile, bool macros, const JetSlot *jets, int njets, uint32_t total_tiles,
                                uint32_t *queue, hipStream_t s);

// LEARNED closure pass: tau_t = max(floor, scale * network(feat)) over n
// cells (closure.hip).
int launch_closure_update(const void *packed, const uint16_t *feat, float *tau_t, uint32_t n, float scale,
                          float floor, hipStream_t s);

}  // namespace ssmd
