// Host-side checks of the device arena's allocator (exatensor_amd/csrc/arena.cpp)
// over a plain host buffer: best fit, TRY_LATER / DEVICE_UNABLE leaving the
// arena unchanged, fragmentation, address-ordered coalescing, alignment, and a
// randomized allocate/release sequence against a shadow model of the free
// space. The arena never touches the memory it manages, so no GPU is needed.
// Prints "ARENA_OK" on success; any failed check aborts with its line.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "arena.hpp"
#include "talsh.h"

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "arena_test: check failed at line %d: %s\n", __LINE__, #c); \
      std::abort();                                                       \
    }                                                                     \
  } while (0)

int main()
{
  using xt::Arena;
  const size_t MiB = 1 << 20;
  std::vector<char> mem(64 * MiB + 100);
  Arena a;
  a.reset(mem.data(), 64 * MiB + 100);  // capacity rounds down to the alignment
  CHECK(a.capacity() == 64 * MiB);
  CHECK(a.free_bytes() == 64 * MiB && a.largest_free() == 64 * MiB);

  // never fits: DEVICE_UNABLE, unchanged
  void *p = nullptr;
  int h = -7;
  CHECK(a.alloc(64 * MiB + 1, &p, &h) == DEVICE_UNABLE && p == nullptr && h == -7);
  CHECK(a.free_bytes() == 64 * MiB && a.live_blocks() == 0);

  // fill with eight 8 MiB blocks, laid out in address order
  void *b[8];
  int hb[8];
  for (int i = 0; i < 8; ++i) {
    CHECK(a.alloc(8 * MiB, &b[i], &hb[i]) == 0);
    CHECK(b[i] == mem.data() + i * 8 * MiB);
  }
  CHECK(a.free_bytes() == 0 && a.live_blocks() == 8);
  // full now: TRY_LATER, unchanged
  p = nullptr;
  h = -7;
  CHECK(a.alloc(1, &p, &h) == TRY_LATER && p == nullptr && h == -7);
  CHECK(a.live_blocks() == 8);

  // free every other block: 32 MiB free in four 8 MiB holes
  for (int i = 0; i < 8; i += 2) CHECK(a.release(hb[i]) == 0);
  CHECK(a.release(hb[0]) == TALSH_INVALID_ARGS);  // twice: unknown handle
  CHECK(a.free_bytes() == 32 * MiB && a.largest_free() == 8 * MiB);
  CHECK(a.alloc(16 * MiB, &p, &h) == TRY_LATER);  // fragmented: fits in total, not in one piece
  CHECK(a.free_bytes() == 32 * MiB);

  // releasing block 1 joins holes 0, 1, 2 into 24 MiB at the base
  CHECK(a.release(hb[1]) == 0);
  CHECK(a.largest_free() == 24 * MiB && a.free_bytes() == 40 * MiB);
  void *big;
  int hbig;
  CHECK(a.alloc(16 * MiB, &big, &hbig) == 0 && big == mem.data());
  // holes now: 8 MiB at 16 (the rest of the joined hole), 32 and 48 MiB; a
  // 6 MiB request takes the lowest of equally good fits, and an 8 MiB
  // request then prefers the exact fit at 32 over the 2 MiB remainder at 22
  void *s6;
  int h6;
  CHECK(a.alloc(6 * MiB, &s6, &h6) == 0 && s6 == mem.data() + 16 * MiB);
  void *s8;
  int h8;
  CHECK(a.alloc(8 * MiB, &s8, &h8) == 0 && s8 == mem.data() + 32 * MiB);  // exact fit preferred
  CHECK(a.release(h6) == 0 && a.release(h8) == 0 && a.release(hbig) == 0);
  for (int i = 3; i < 8; i += 2) CHECK(a.release(hb[i]) == 0);
  CHECK(a.free_bytes() == 64 * MiB && a.largest_free() == 64 * MiB && a.live_blocks() == 0);

  // alignment: sizes round up to 4 KiB; zero-byte requests take one unit
  void *q0, *q1;
  int h0, h1;
  CHECK(a.alloc(0, &q0, &h0) == 0 && a.alloc(1, &q1, &h1) == 0);
  CHECK((char *)q1 - (char *)q0 == (long)Arena::kAlign);
  CHECK(a.free_bytes() == 64 * MiB - 2 * Arena::kAlign);
  CHECK(a.release(h0) == 0 && a.release(h1) == 0);

  // randomized sequence against a shadow model: live blocks never overlap,
  // stay inside the arena, and the free space adds up; at the end every
  // release coalesces back to one block
  std::mt19937_64 rng(2026);
  std::map<int, std::pair<size_t, size_t>> live;  // handle -> (offset, bytes)
  size_t used = 0;
  for (int step = 0; step < 20000; ++step) {
    if (live.empty() || rng() % 3 != 0) {
      const size_t want = 1 + rng() % (3 * MiB);
      void *r;
      int hr;
      const size_t before = a.free_bytes();
      const int e = a.alloc(want, &r, &hr);
      const size_t need = (want + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign;
      if (e == 0) {
        const size_t off = (char *)r - mem.data();
        CHECK(off % Arena::kAlign == 0 && off + need <= a.capacity());
        for (auto &kv : live) CHECK(off + need <= kv.second.first || kv.second.first + kv.second.second <= off);
        live[hr] = {off, need};
        used += need;
      } else {
        CHECK(e == TRY_LATER && a.free_bytes() == before && a.largest_free() < need);
      }
    } else {
      auto it = live.begin();
      std::advance(it, rng() % live.size());
      CHECK(a.release(it->first) == 0);
      used -= it->second.second;
      live.erase(it);
    }
    CHECK(a.free_bytes() == a.capacity() - used && (int)live.size() == a.live_blocks());
  }
  for (auto &kv : live) CHECK(a.release(kv.first) == 0);
  CHECK(a.free_bytes() == a.capacity() && a.largest_free() == a.capacity());
  std::printf("ARENA_OK\n");
  return 0;
}
