// float64 instantiations of the fused axis sweep (fused_sweep.hpp): line
// periods 2^1 .. 2^11, chunk length from fused_chunk_of.
#include "fused_sweep.hpp"

namespace paliwa_hip {

hipError_t launch_fused_f64(SweepParams const &p, hipStream_t stream) {
  return fused::dispatch_period<double, 1, 11>(p, stream);
}

hipError_t launch_fused_batched_f64(SweepParams const &p, bool strided,
                                    ManyGrids grids, int64_t tiles,
                                    hipStream_t stream) {
  return fused::dispatch_period_batched<double, 1, 11>(p, strided, grids,
                                                       tiles, stream);
}

} // namespace paliwa_hip
