// float32 instantiations of the fused axis sweep (fused_sweep.hpp): line
// periods 2^1 .. 2^12, chunk length from fused_chunk_of.
#include "fused_sweep.hpp"

namespace paliwa_hip {

hipError_t launch_fused_f32(SweepParams const &p, hipStream_t stream) {
  return fused::dispatch_period<float, 1, 12>(p, stream);
}

hipError_t launch_fused_batched_f32(SweepParams const &p, bool strided,
                                    ManyGrids grids, int64_t tiles,
                                    hipStream_t stream) {
  return fused::dispatch_period_batched<float, 1, 12>(p, strided, grids,
                                                       tiles, stream);
}

} // namespace paliwa_hip
