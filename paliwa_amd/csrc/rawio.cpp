// Raw dumps of grids in device memory: the reference's
// dump_chunk_span_to_binary_file (paliwa/paliwa_io.hpp:13-23) writes the
// elements of a span in row-major order, headerless, native byte order; this
// file writes (and reads back) the same bytes, streaming through a bounded
// host buffer instead of mirroring the whole grid on the host.
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

namespace {

using namespace paliwa_hip;

constexpr int64_t kChunk = int64_t(64) << 20; // host staging per step

} // namespace

extern "C" {

int paliwa_dump_to_binary_file(const void *data, int64_t bytes,
                               const char *filename, void *stream) {
  if (filename == nullptr || bytes < 0 || (bytes > 0 && data == nullptr))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "data, bytes and filename are required");
  std::FILE *f = std::fopen(filename, "wb");
  if (f == nullptr)
    return fail(PALIWA_ERR_IO, std::string("cannot open ") + filename + " for writing");
  auto s = static_cast<hipStream_t>(stream);
  std::vector<char> host(static_cast<size_t>(std::min(bytes, kChunk)));
  char const *src = static_cast<char const *>(data);
  for (int64_t done = 0; done < bytes;) {
    int64_t const n = std::min(kChunk, bytes - done);
    if (hipMemcpyAsync(host.data(), src + done, static_cast<size_t>(n),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      std::fclose(f);
      return fail(PALIWA_ERR_HIP, "device to host copy failed");
    }
    if (std::fwrite(host.data(), 1, static_cast<size_t>(n), f) != static_cast<size_t>(n)) {
      std::fclose(f);
      return fail(PALIWA_ERR_IO, std::string("short write to ") + filename);
    }
    done += n;
  }
  if (std::fclose(f) != 0)
    return fail(PALIWA_ERR_IO, std::string("cannot close ") + filename);
  return PALIWA_OK;
}

int paliwa_load_from_binary_file(void *data, int64_t bytes,
                                 const char *filename, void *stream) {
  if (filename == nullptr || bytes < 0 || (bytes > 0 && data == nullptr))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "data, bytes and filename are required");
  std::FILE *f = std::fopen(filename, "rb");
  if (f == nullptr)
    return fail(PALIWA_ERR_IO, std::string("cannot open ") + filename + " for reading");
  if (std::fseek(f, 0, SEEK_END) != 0) {
    std::fclose(f);
    return fail(PALIWA_ERR_IO, std::string("cannot seek in ") + filename);
  }
  long const size = std::ftell(f);
  if (size != bytes) {
    std::fclose(f);
    return fail(PALIWA_ERR_SHAPE, std::string(filename) + " holds " + std::to_string(size) +
                                      " bytes, expected " + std::to_string(bytes));
  }
  std::rewind(f);
  auto s = static_cast<hipStream_t>(stream);
  std::vector<char> host(static_cast<size_t>(std::min(bytes, kChunk)));
  char *dst = static_cast<char *>(data);
  for (int64_t done = 0; done < bytes;) {
    int64_t const n = std::min(kChunk, bytes - done);
    if (std::fread(host.data(), 1, static_cast<size_t>(n), f) != static_cast<size_t>(n)) {
      std::fclose(f);
      return fail(PALIWA_ERR_IO, std::string("short read from ") + filename);
    }
    // synchronous: the staging buffer is reused for the next chunk
    if (hipMemcpyAsync(dst + done, host.data(), static_cast<size_t>(n),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      std::fclose(f);
      return fail(PALIWA_ERR_HIP, "host to device copy failed");
    }
    done += n;
  }
  std::fclose(f);
  return PALIWA_OK;
}

} // extern "C"
