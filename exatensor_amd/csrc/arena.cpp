#include "arena.hpp"

#include "talsh.h"

namespace xt {

void Arena::reset(void *base, size_t bytes)
{
  std::lock_guard<std::mutex> g(mu_);
  base_ = static_cast<char *>(base);
  cap_ = bytes - bytes % kAlign;
  free_.clear();
  used_.clear();
  if (cap_ > 0) free_[0] = cap_;
}

int Arena::alloc(size_t bytes, void **ptr, int *handle)
{
  std::lock_guard<std::mutex> g(mu_);
  size_t need = bytes == 0 ? kAlign : (bytes + kAlign - 1) / kAlign * kAlign;
  if (need > cap_) return DEVICE_UNABLE;
  auto best = free_.end();
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= need && (best == free_.end() || it->second < best->second)) {
      best = it;
      if (it->second == need) break;
    }
  }
  if (best == free_.end()) return TRY_LATER;
  size_t off = best->first, sz = best->second;
  free_.erase(best);
  if (sz > need) free_[off + need] = sz - need;
  int h = next_handle_++;
  if (next_handle_ < 0) next_handle_ = 0;
  used_[h] = {off, need};
  *ptr = base_ + off;
  *handle = h;
  return 0;
}

int Arena::release(int handle)
{
  std::lock_guard<std::mutex> g(mu_);
  auto it = used_.find(handle);
  if (it == used_.end()) return TALSH_INVALID_ARGS;
  size_t off = it->second.first, sz = it->second.second;
  used_.erase(it);
  auto nx = free_.lower_bound(off);
  if (nx != free_.end() && off + sz == nx->first) {  // merge with the next block
    sz += nx->second;
    nx = free_.erase(nx);
  }
  if (nx != free_.begin()) {  // merge with the previous block
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) {
      pv->second += sz;
      return 0;
    }
  }
  free_[off] = sz;
  return 0;
}

size_t Arena::free_bytes() const
{
  std::lock_guard<std::mutex> g(mu_);
  size_t s = 0;
  for (auto &kv : free_) s += kv.second;
  return s;
}

size_t Arena::largest_free() const
{
  std::lock_guard<std::mutex> g(mu_);
  size_t s = 0;
  for (auto &kv : free_) s = kv.second > s ? kv.second : s;
  return s;
}

int Arena::live_blocks() const
{
  std::lock_guard<std::mutex> g(mu_);
  return (int)used_.size();
}

}  // namespace xt
