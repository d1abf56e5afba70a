// Lattice launchers for D3Q27 with f16 population storage.
#include "lattice_inst.hpp"

SSMD_INSTANTIATE_LATTICE(27, ssmd::PopF16)
