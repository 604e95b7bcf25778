// Raw dumps of grids in device memory: the reference's
// dump_chunk_span_to_binary_file (paliwa/paliwa_io.hpp:13-23) writes the
// elements of a span in row-major order, headerless, native byte order; this
// file writes (and reads back) the same bytes, streaming through two pinned
// host buffers so that the device-host copy of one chunk overlaps the file
// write (or read) of the other, instead of mirroring the whole grid on the
// host.
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

namespace {

using namespace paliwa_hip;

constexpr int64_t kChunk = int64_t(16) << 20; // bytes per staging buffer

// Two staging buffers shared by every call, allocated on first use and kept
// until the process ends (pinned host memory; pageable if pinning fails).
// Calls are serialized by the mutex, and every call waits for its copies
// before it returns, so a buffer is never reused while a copy reads or fills
// it.
struct Staging {
  std::mutex mutex;
  char *buf[2] = {nullptr, nullptr};
  std::vector<char> pageable[2];
  bool ready = false;

  void prepare() {
    if (ready)
      return;
    for (int i = 0; i < 2; ++i) {
      void *p = nullptr;
      if (hipHostMalloc(&p, static_cast<size_t>(kChunk), hipHostMallocDefault) == hipSuccess &&
          p != nullptr) {
        buf[i] = static_cast<char *>(p);
      } else {
        (void)hipGetLastError();
        pageable[i].resize(static_cast<size_t>(kChunk));
        buf[i] = pageable[i].data();
      }
    }
    ready = true;
  }
};

Staging &staging() {
  static Staging s;
  return s;
}

struct Events {
  hipEvent_t e[2] = {nullptr, nullptr};
  bool ok = true;
  Events() {
    for (auto &x : e)
      ok = ok && hipEventCreateWithFlags(&x, hipEventDisableTiming) == hipSuccess;
  }
  ~Events() {
    for (auto x : e)
      if (x != nullptr)
        (void)hipEventDestroy(x);
  }
};

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
  Staging &st = staging();
  std::lock_guard<std::mutex> lock(st.mutex);
  st.prepare();
  Events ev;
  if (!ev.ok) {
    std::fclose(f);
    return fail(PALIWA_ERR_HIP, "could not create events");
  }
  char const *src = static_cast<char const *>(data);
  int64_t const chunks = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](int64_t i) {
    int64_t const off = i * kChunk, n = std::min(kChunk, bytes - off);
    int const b = static_cast<int>(i % 2);
    return hipMemcpyAsync(st.buf[b], src + off, static_cast<size_t>(n),
                          hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipEventRecord(ev.e[b], s) == hipSuccess;
  };
  int rc = PALIWA_OK;
  if (chunks > 0 && !issue(0))
    rc = fail(PALIWA_ERR_HIP, "device to host copy failed");
  for (int64_t i = 0; i < chunks && rc == PALIWA_OK; ++i) {
    // buffer (i + 1) % 2 held chunk i - 1, already written to the file
    if (i + 1 < chunks && !issue(i + 1)) {
      rc = fail(PALIWA_ERR_HIP, "device to host copy failed");
      break;
    }
    int const b = static_cast<int>(i % 2);
    if (hipEventSynchronize(ev.e[b]) != hipSuccess) {
      rc = fail(PALIWA_ERR_HIP, "device to host copy failed");
      break;
    }
    int64_t const n = std::min(kChunk, bytes - i * kChunk);
    if (std::fwrite(st.buf[b], 1, static_cast<size_t>(n), f) != static_cast<size_t>(n))
      rc = fail(PALIWA_ERR_IO, std::string("short write to ") + filename);
  }
  (void)hipStreamSynchronize(s); // no copy may still use a staging buffer
  if (std::fclose(f) != 0 && rc == PALIWA_OK)
    rc = fail(PALIWA_ERR_IO, std::string("cannot close ") + filename);
  return rc;
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
  Staging &st = staging();
  std::lock_guard<std::mutex> lock(st.mutex);
  st.prepare();
  Events ev;
  if (!ev.ok) {
    std::fclose(f);
    return fail(PALIWA_ERR_HIP, "could not create events");
  }
  char *dst = static_cast<char *>(data);
  int64_t const chunks = (bytes + kChunk - 1) / kChunk;
  int rc = PALIWA_OK;
  for (int64_t i = 0; i < chunks && rc == PALIWA_OK; ++i) {
    int const b = static_cast<int>(i % 2);
    // buffer b last fed chunk i - 2: its copy must be done before the read
    if (i >= 2 && hipEventSynchronize(ev.e[b]) != hipSuccess) {
      rc = fail(PALIWA_ERR_HIP, "host to device copy failed");
      break;
    }
    int64_t const off = i * kChunk, n = std::min(kChunk, bytes - off);
    if (std::fread(st.buf[b], 1, static_cast<size_t>(n), f) != static_cast<size_t>(n)) {
      rc = fail(PALIWA_ERR_IO, std::string("short read from ") + filename);
      break;
    }
    if (hipMemcpyAsync(dst + off, st.buf[b], static_cast<size_t>(n), hipMemcpyHostToDevice,
                       s) != hipSuccess ||
        hipEventRecord(ev.e[b], s) != hipSuccess)
      rc = fail(PALIWA_ERR_HIP, "host to device copy failed");
  }
  if (hipStreamSynchronize(s) != hipSuccess && rc == PALIWA_OK)
    rc = fail(PALIWA_ERR_HIP, "host to device copy failed");
  std::fclose(f);
  return rc;
}

} // extern "C"
