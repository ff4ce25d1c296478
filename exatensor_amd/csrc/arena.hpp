// Device memory arena: one hipMalloc per GPU carved by a best-fit allocator
// with address-ordered free-list coalescing.
//
// Behavioural contract follows the reference's argument buffer
// (TALSH/mem_manager.cpp:544-692, get_buf_entry / free_buf_entry): a request
// larger than the arena can ever hold is DEVICE_UNABLE, a request that does
// not fit now is TRY_LATER, and both leave the arena unchanged. Unlike the
// reference's fixed binary tree (whole power-of-two entries, up to 50% waste)
// blocks are sized to the request, rounded to the alignment.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>

namespace xt {

class Arena {
 public:
  static constexpr size_t kAlign = 4096;

  // Manage [base, base+bytes). The arena does not own the memory.
  void reset(void *base, size_t bytes);
  // 0 on success (handle >= 0), TRY_LATER, or DEVICE_UNABLE.
  int alloc(size_t bytes, void **ptr, int *handle);
  // 0 on success, TALSH_INVALID_ARGS for an unknown handle.
  int release(int handle);
  size_t capacity() const { return cap_; }
  size_t free_bytes() const;
  size_t largest_free() const;
  void *base() const { return base_; }
  int live_blocks() const;

 private:
  mutable std::mutex mu_;
  char *base_ = nullptr;
  size_t cap_ = 0;
  std::map<size_t, size_t> free_;                 // offset -> size, address ordered
  std::unordered_map<int, std::pair<size_t, size_t>> used_;  // handle -> (offset, size)
  int next_handle_ = 0;
};

}  // namespace xt
