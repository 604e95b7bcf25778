// int32 instantiations of the fused axis sweep (fused_sweep.hpp) for the
// learned transform's interpolet sweeps: ladder schedules only, line periods
// 2^1 .. 2^11, chunk length from fused_chunk_of.  The coefficient slots of
// each pass carry (sign, bias, shift) of the integer lifting step.
#include "fused_sweep.hpp"

namespace paliwa_hip {
namespace {

using namespace fused;

template <int K, int G, bool STRIDED>
hipError_t launch_int_layout(SweepParams const &p, hipStream_t stream) {
  using Geom = TileGeom<K, G, STRIDED, int32_t>;
  int64_t const lines = p.A * p.B;
  int64_t const tiles = (lines + Geom::TB - 1) / Geom::TB;
  size_t const bytes = sizeof(int32_t) * Geom::lds_elems(p.m, STRIDED);
  if (p.schedule != kScheduleLadderAsc && p.schedule != kScheduleLadderDesc)
    return hipErrorInvalidValue;
  bool const asc = p.schedule == kScheduleLadderAsc;
  switch (p.pattern) {
  case kPatOdd:
    return asc ? launch_one(fused_ladder<int32_t, K, G, STRIDED, kPatOdd, true>, p, bytes, tiles, Geom::NT, stream)
               : launch_one(fused_ladder<int32_t, K, G, STRIDED, kPatOdd, false>, p, bytes, tiles, Geom::NT, stream);
  case kPatOddEven:
    return asc ? launch_one(fused_ladder<int32_t, K, G, STRIDED, kPatOddEven, true>, p, bytes, tiles, Geom::NT, stream)
               : launch_one(fused_ladder<int32_t, K, G, STRIDED, kPatOddEven, false>, p, bytes, tiles, Geom::NT, stream);
  case kPatEvenOdd:
    return asc ? launch_one(fused_ladder<int32_t, K, G, STRIDED, kPatEvenOdd, true>, p, bytes, tiles, Geom::NT, stream)
               : launch_one(fused_ladder<int32_t, K, G, STRIDED, kPatEvenOdd, false>, p, bytes, tiles, Geom::NT, stream);
  default:
    return hipErrorInvalidValue;
  }
}

template <int L, int LMAX>
hipError_t dispatch_int(SweepParams const &p, hipStream_t stream) {
  if constexpr (L > LMAX) {
    return hipErrorInvalidValue;
  } else {
    if (p.L == L) {
      constexpr int K = fused_chunk_of(sizeof(int32_t), L);
      constexpr int G = (1 << L) / K;
      return p.B > 1 ? launch_int_layout<K, G, true>(p, stream)
                     : launch_int_layout<K, G, false>(p, stream);
    }
    return dispatch_int<L + 1, LMAX>(p, stream);
  }
}

} // namespace

hipError_t launch_fused_i32(SweepParams const &p, hipStream_t stream) {
  return dispatch_int<1, 11>(p, stream);
}

} // namespace paliwa_hip
