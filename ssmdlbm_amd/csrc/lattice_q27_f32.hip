// Lattice launchers for D3Q27 with f32 population storage.
#include "lattice_inst.hpp"

SSMD_INSTANTIATE_LATTICE(27, ssmd::PopF32)
