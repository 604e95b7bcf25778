// float64 instantiations of the fused axis sweep (fused_sweep.hpp).
// Chunk length K: 16 points (128 bytes) per lane; one lane per line for
// periods up to 16, 32 points per lane for period 2048.
#include "fused_sweep.hpp"

namespace paliwa_hip {

hipError_t launch_fused_f64(SweepParams const &p, hipStream_t stream) {
  using fused::launch_kg;
  using T = double;
  switch (p.L) {
  case 1: return launch_kg<T, 2, 1>(p, stream);
  case 2: return launch_kg<T, 4, 1>(p, stream);
  case 3: return launch_kg<T, 8, 1>(p, stream);
  case 4: return launch_kg<T, 16, 1>(p, stream);
  case 5: return launch_kg<T, 16, 2>(p, stream);
  case 6: return launch_kg<T, 16, 4>(p, stream);
  case 7: return launch_kg<T, 16, 8>(p, stream);
  case 8: return launch_kg<T, 16, 16>(p, stream);
  case 9: return launch_kg<T, 16, 32>(p, stream);
  case 10: return launch_kg<T, 16, 64>(p, stream);
  case 11: return launch_kg<T, 32, 64>(p, stream);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fused_batched_f64(SweepParams const &p, bool strided,
                                    ManyGrids grids, int64_t tiles,
                                    hipStream_t stream) {
  using fused::launch_batched_kg;
  using T = double;
  switch (p.L) {
  case 1: return launch_batched_kg<T, 2, 1>(p, strided, grids, tiles, stream);
  case 2: return launch_batched_kg<T, 4, 1>(p, strided, grids, tiles, stream);
  case 3: return launch_batched_kg<T, 8, 1>(p, strided, grids, tiles, stream);
  case 4: return launch_batched_kg<T, 16, 1>(p, strided, grids, tiles, stream);
  case 5: return launch_batched_kg<T, 16, 2>(p, strided, grids, tiles, stream);
  case 6: return launch_batched_kg<T, 16, 4>(p, strided, grids, tiles, stream);
  case 7: return launch_batched_kg<T, 16, 8>(p, strided, grids, tiles, stream);
  case 8: return launch_batched_kg<T, 16, 16>(p, strided, grids, tiles, stream);
  case 9: return launch_batched_kg<T, 16, 32>(p, strided, grids, tiles, stream);
  case 10: return launch_batched_kg<T, 16, 64>(p, strided, grids, tiles, stream);
  case 11: return launch_batched_kg<T, 32, 64>(p, strided, grids, tiles, stream);
  default: return hipErrorInvalidValue;
  }
}

} // namespace paliwa_hip
