// Lattice launchers for D3Q19 with f16 population storage.
#include "lattice_inst.hpp"

SSMD_INSTANTIATE_LATTICE(19, ssmd::PopF16)
