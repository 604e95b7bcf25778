// Drop-in for paliwa/paliwa_io.hpp on MI355X: dump_chunk_span_to_binary_file
// writes a grid's elements (device memory, row-major) to a headerless raw file
// in native byte order, through libpaliwa_hip.so.  The grid is any object
// with data_handle() and extents() (paliwa::GridSpan is one).  The call
// synchronizes the stream of `instance`; a file that cannot be written throws
// std::runtime_error (the reference's ofstream fails silently).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "paliwa_transform.hpp"

namespace paliwa {

template <typename ChunkType, typename ExecSpace = HipExec>
void dump_chunk_span_to_binary_file(ChunkType const span, std::string const &filename,
                                    ExecSpace instance = ExecSpace()) {
  auto const &e = span.extents();
  std::int64_t count = 1;
  for (std::size_t d = 0; d < static_cast<std::size_t>(detail::rank_of(e)); ++d)
    count *= static_cast<std::int64_t>(detail::extent_of(e, d));
  using T = detail::value_type_of<ChunkType>;
  int const rc = paliwa_dump_to_binary_file(span.data_handle(),
                                            count * static_cast<std::int64_t>(sizeof(T)),
                                            filename.c_str(), instance.stream);
  if (rc != PALIWA_OK)
    throw std::runtime_error(paliwa_last_error());
}

} // namespace paliwa
