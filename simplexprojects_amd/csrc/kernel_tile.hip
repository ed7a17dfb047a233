// kernel_tile.hip -- register-resident simplex kernel (variants added below).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

int tile_variant_count(int) { return 0; }
TileVariant tile_variant(int, int) { return TileVariant{1, 8, 1, 1}; }
int tile_max_blocks_per_cu(int, int) { return 0; }
hipError_t launch_tile(const TileLaunch&, hipStream_t) { return hipErrorInvalidValue; }

}  // namespace blpg
