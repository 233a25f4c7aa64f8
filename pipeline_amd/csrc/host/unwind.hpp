// unwind.hpp -- host buffers handed to asynchronous copies must outlive them.
// A SyncOnUnwind declared after such buffers waits for the stream when an
// exception leaves the scope, before the unwinding frees the buffers.
#pragma once

#include <exception>

#include "dpcull_hip.h"

namespace dpcull {

class SyncOnUnwind {
 public:
  explicit SyncOnUnwind(void* stream) : m_stream(stream), m_uncaught(std::uncaught_exceptions()) {}
  ~SyncOnUnwind()
  {
    if (std::uncaught_exceptions() > m_uncaught) (void)dpc_stream_sync(m_stream);
  }
  SyncOnUnwind(const SyncOnUnwind&) = delete;
  SyncOnUnwind& operator=(const SyncOnUnwind&) = delete;

 private:
  void* m_stream;
  int m_uncaught;
};

}  // namespace dpcull
