// Stable LSD radix ordering of tile intersections and tile-range marking.
//
// Replaces the reference's cub::DeviceRadixSort::SortPairs (DoubleBuffer) call
// and mark_sorted_tile_ranges_kernel (renderer_cuda.cu:3215-3286, 1908-1942).
//
// Keys keep the reference layout (uint64 global_tile << 32 | float_bits(depth)).
// The sort digits are taken from a compressed, order-preserving image of the
// key: (tile << depth_bits) | (depth_bits_value - bits(near_plane)).  Every
// emitted depth lies in [near, far], and positive floats order like their bit
// patterns, so the image orders exactly like the key while needing
// bit_length(GT-1) + bit_length(bits(far) - bits(near) + 1) bits instead of
// 32 + bit_length(GT-1).  Sentinel keys (all ones) map to the all-ones image,
// which is strictly larger than any valid image.
//
// Structure (onesweep):
//   1. one histogram pass over the keys fills the digit histograms of every
//      radix pass at once;
//   2. per pass, each 512-thread workgroup takes a ticket (its tile index),
//      ranks its 8192 keys stably in LDS (each wave ranks a contiguous
//      1024-key run with 64-lane ballot matching), publishes its per-digit
//      counts and resolves its global per-digit offsets by decoupled look-back
//      over predecessor tickets, then writes the tile out digit-contiguously.
//      The look-back wait grows with the number of tiles in the chain, so the
//      tile is as large as two resident workgroups per CU allow (74 KB of LDS
//      each): 8192 keys sort about 15% faster than 4096.
// Ticket order guarantees every predecessor a workgroup waits on is already
// resident; the look-back spin is additionally bounded: a wait that gives up
// adds one to counters[VGR_COUNTER_SORT_FAILURES] instead of hanging, and every
// caller treats a nonzero count like an overflow (the order is then wrong).
// The element count is read from the device counter block, so the sequence
// never synchronizes the host and can be captured in a graph.
#include <type_traits>

#include "vgr_internal.h"

namespace vgr {
namespace {

// Look-back flag words: 0 = not published; bit 31 set = inclusive prefix in bits
// 0-30 (up to 2^31 - 1 keys); otherwise the tile's aggregate plus one (a tile
// holds at most 8192 keys, so an aggregate word is never 0 nor has bit 31).
constexpr uint32_t kFlagPrefix = 1u << 31;
constexpr uint32_t kFlagValueMask = kFlagPrefix - 1u;
constexpr uint32_t kDefaultSpinLimit = 1u << 22;
constexpr uint32_t kForceGiveUp = 0xffffffffu;  // test hook: every look-back gives up at once

struct SortKeyParams {
  int depth_bits;
  uint32_t depth_base;
  int key_bits;
};

__device__ __forceinline__ uint64_t compressed_key(uint64_t key, const SortKeyParams& p) {
  if (key == kSentinelKey) return (1ull << p.key_bits) - 1ull;
  const uint64_t tile = key >> 32;
  const uint32_t depth = static_cast<uint32_t>(key) - p.depth_base;
  return (tile << p.depth_bits) | static_cast<uint64_t>(depth);
}

__device__ __forceinline__ int64_t element_count(const int64_t* counters, int64_t capacity) {
  int64_t n = counters[VGR_COUNTER_INTERSECTIONS];
  n = n < 0 ? 0 : n;
  return n < capacity ? n : capacity;
}

__global__ __launch_bounds__(kBlockThreads) void sort_histogram_kernel(const uint64_t* __restrict__ keys,
                                                                       const int64_t* __restrict__ counters,
                                                                       int64_t capacity, SortKeyParams kp,
                                                                       int digit_bits, int passes,
                                                                       uint32_t* __restrict__ histogram) {
  __shared__ uint32_t local[16 * 256];
  const int radix = 1 << digit_bits;
  for (int i = threadIdx.x; i < passes * radix; i += kBlockThreads) local[i] = 0;
  __syncthreads();
  const int64_t n = element_count(counters, capacity);
  const uint32_t mask = static_cast<uint32_t>(radix - 1);
  // Keys are read coalesced (lane i takes key i of a 64-key run); per pass, a lane
  // adds to LDS only if it starts a run of equal digits among its wave's keys, and
  // adds the run's length.  Emission writes a splat's tiles next to each other,
  // so the high (tile) digits come in long runs that would otherwise serialize up
  // to 64 lanes on one LDS address.
  // Each wave takes four consecutive 64-key runs per iteration and issues their
  // loads together, so four loads per lane are in flight.
  constexpr int kRuns = 4;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave_id = (static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x) / kWave;
  const int64_t wave_count = static_cast<int64_t>(gridDim.x) * (kBlockThreads / kWave);
  const uint64_t above = lane == kWave - 1 ? 0ull : (~0ull << (lane + 1));
  for (int64_t base = wave_id * kRuns * kWave; base < n; base += wave_count * kRuns * kWave) {
    uint64_t ck[kRuns];
#pragma unroll
    for (int r = 0; r < kRuns; ++r) {
      const int64_t i = base + r * kWave + lane;
      ck[r] = i < n ? compressed_key(keys[i], kp) : 0ull;
    }
#pragma unroll
    for (int r = 0; r < kRuns; ++r) {
      const bool valid = base + r * kWave + lane < n;
      const uint64_t live = __ballot(valid);
      if (live == 0ull) break;
      for (int p = 0; p < passes; ++p) {
        const uint32_t digit = static_cast<uint32_t>(ck[r] >> (p * digit_bits)) & mask;
        const uint32_t before = __shfl_up(digit, 1, kWave);
        const bool head = valid && (lane == 0 || digit != before);
        const uint64_t heads = __ballot(head);
        if (head) {
          const uint64_t later = (heads & above) | ~live;  // next run head, or the first invalid lane
          const int end = later == 0ull ? kWave : __ffsll(static_cast<unsigned long long>(later)) - 1;
          atomicAdd(&local[p * radix + digit], static_cast<uint32_t>(end - lane));
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * radix; i += kBlockThreads) {
    if (local[i] != 0) atomicAdd(&histogram[i], local[i]);
  }
}

// Exclusive scan of every pass's histogram (one block, one digit per thread),
// for the per-pass digit offsets.
__global__ __launch_bounds__(kSortThreads) void sort_digit_scan_kernel(const uint32_t* __restrict__ histogram,
                                                                        int digit_bits, int passes,
                                                                        uint32_t* __restrict__ digit_base) {
  __shared__ uint32_t wave_totals[kSortThreads / kWave];
  const int radix = 1 << digit_bits;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  for (int p = 0; p < passes; ++p) {
    const uint32_t value = threadIdx.x < radix ? histogram[p * radix + threadIdx.x] : 0u;
    uint32_t inclusive = value;
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1) {
      const uint32_t t = __shfl_up(inclusive, offset, kWave);
      if (lane >= offset) inclusive += t;
    }
    if (lane == kWave - 1) wave_totals[wave] = inclusive;
    __syncthreads();
    uint32_t prefix = 0;
    for (int w = 0; w < wave; ++w) prefix += wave_totals[w];
    if (threadIdx.x < radix) digit_base[p * radix + threadIdx.x] = prefix + inclusive - value;
    __syncthreads();
  }
}

// Clears histogram, tickets and the look-back flags of the tiles the device
// count needs (graph-safe replacement for hipMemsetAsync).
__global__ __launch_bounds__(kBlockThreads) void sort_clear_kernel(const int64_t* __restrict__ counters,
                                                                   int64_t capacity, int digit_bits, int passes,
                                                                   int64_t max_tiles, int tile_keys,
                                                                   uint32_t* __restrict__ histogram,
                                                                   uint32_t* __restrict__ tickets,
                                                                   uint32_t* __restrict__ flags) {
  const int radix = 1 << digit_bits;
  const int64_t n = element_count(counters, capacity);
  int64_t tiles = (n + tile_keys - 1) / tile_keys;
  tiles = tiles < max_tiles ? tiles : max_tiles;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlockThreads;
  const int64_t first = static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x;
  if (first < static_cast<int64_t>(passes) * radix) {
    for (int64_t i = first; i < static_cast<int64_t>(passes) * radix; i += stride) histogram[i] = 0;
  }
  if (first < passes) tickets[first] = 0;
  const int64_t per_pass = tiles * radix;
  for (int p = 0; p < passes; ++p) {
    uint32_t* pass_flags = flags + static_cast<int64_t>(p) * max_tiles * radix;
    for (int64_t i = first; i < per_pass; i += stride) pass_flags[i] = 0;
  }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & (kWave - 1);
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

template <int kDigitBits, int kSortItems>
__global__ __launch_bounds__(kSortThreads) void sort_onesweep_pass_kernel(
    const uint64_t* __restrict__ keys_in, const int32_t* __restrict__ values_in, uint64_t* __restrict__ keys_out,
    int32_t* __restrict__ values_out, const int64_t* __restrict__ counters, int64_t capacity, SortKeyParams kp,
    int shift, const uint32_t* __restrict__ digit_base, uint32_t* __restrict__ flags, uint32_t* __restrict__ ticket,
    int64_t* __restrict__ failures, uint32_t spin_limit) {
  constexpr int kRadix = 1 << kDigitBits;
  constexpr int kSortTile = kSortThreads * kSortItems;
  static_assert(kRadix <= kSortThreads && kSortTile <= 65535, "one digit per thread; 16-bit offsets");
  constexpr int kWaves = kSortThreads / kWave;
  constexpr int kWaveRun = kSortTile / kWaves;  // contiguous keys ranked by one wave
  constexpr uint32_t kMask = kRadix - 1;
  // Keys and then values pass through the same LDS buffer (two phases), which
  // halves the pass's LDS footprint and doubles its occupancy.
  __shared__ uint64_t s_stage[kSortTile];
  uint64_t* const s_keys = s_stage;
  int32_t* const s_values = reinterpret_cast<int32_t*>(s_stage);
  // Per-wave digit counts, then in-place per-wave offsets (at most the tile size,
  // 8192): 16-bit entries for 9-bit digits keep two workgroups per CU in LDS.
  using Count = typename std::conditional<(kRadix > 256), uint16_t, uint32_t>::type;
  __shared__ Count s_count[kWaves][kRadix];
  __shared__ uint32_t s_local_start[kRadix];
  __shared__ int32_t s_global_start[kRadix];  // capacities are below 2^31
  __shared__ uint32_t s_ticket;

  // The grid covers the capacity; the device-side key count needs its first
  // ceil(n / tile) workgroups.  Those take the tickets (tiles in the order they
  // ask, so every predecessor of a ticket holder is running or done), and every
  // workgroup past them leaves at once, before any atomic or LDS work.
  const int64_t n = element_count(counters, capacity);
  if (static_cast<int64_t>(blockIdx.x) * kSortTile >= n) return;
  if (threadIdx.x == 0) s_ticket = atomicAdd(ticket, 1u);
  for (int i = threadIdx.x; i < kWaves * kRadix; i += kSortThreads) (&s_count[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile = s_ticket;
  const int64_t tile_start = tile * kSortTile;
  if (tile_start >= n) return;
  const int valid_count = static_cast<int>(n - tile_start < kSortTile ? n - tile_start : kSortTile);
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);

  uint64_t keys[kSortItems];
  int32_t values[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int local = wave * kWaveRun + r * kWave + lane;
    if (local < valid_count) {
      keys[r] = keys_in[tile_start + local];
      values[r] = values_in[tile_start + local];
    } else {
      keys[r] = 0;
      values[r] = 0;
    }
  }
  uint32_t digits[kSortItems];
  uint32_t ranks[kSortItems];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int local = wave * kWaveRun + r * kWave + lane;
    const bool valid = local < valid_count;
    const uint32_t digit = valid ? static_cast<uint32_t>(compressed_key(keys[r], kp) >> shift) & kMask : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kDigitBits; ++b) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t vote = __ballot(bit);
      peers &= bit ? vote : ~vote;
    }
    digits[r] = digit;
    ranks[r] = 0;
    if (valid) {
      const uint32_t below = static_cast<uint32_t>(__popcll(peers & lt));
      const uint32_t base = s_count[wave][digit];
      ranks[r] = base + below;
      if (below == 0) s_count[wave][digit] = static_cast<Count>(base + static_cast<uint32_t>(__popcll(peers)));
    }
  }
  __syncthreads();

  // Per-digit wave offsets (in place) and tile totals.
  uint32_t total = 0;
  if (threadIdx.x < kRadix) {
    uint32_t running = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = s_count[w][threadIdx.x];
      s_count[w][threadIdx.x] = static_cast<Count>(running);
      running += c;
    }
    total = running;
    // Publish this tile's aggregate (tile 0 publishes its inclusive prefix).
    const uint32_t flag = tile == 0 ? (kFlagPrefix | total) : total + 1u;
    __hip_atomic_store(&flags[tile * kRadix + threadIdx.x], flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // Exclusive scan of the digit totals (kRadix <= kSortThreads: one per thread).
  {
    __shared__ uint32_t wave_totals[kWaves];
    uint32_t inclusive = total;
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1) {
      const uint32_t t = __shfl_up(inclusive, offset, kWave);
      if (lane >= offset) inclusive += t;
    }
    if (lane == kWave - 1) wave_totals[wave] = inclusive;
    __syncthreads();
    uint32_t prefix = 0;
    for (int w = 0; w < wave; ++w) prefix += wave_totals[w];
    if (threadIdx.x < kRadix) s_local_start[threadIdx.x] = prefix + inclusive - total;
  }
  __syncthreads();

  // Phase 1: keys into LDS in block-local sorted order (positions kept for phase 2).
  uint32_t positions[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int local = wave * kWaveRun + r * kWave + lane;
    positions[r] = s_local_start[digits[r]] + s_count[wave][digits[r]] + ranks[r];
    if (local < valid_count) s_keys[positions[r]] = keys[r];
  }

  // Decoupled look-back for the global offset of every digit.
  if (threadIdx.x < kRadix) {
    uint32_t exclusive = 0;
    if (tile > 0) {
      // The flags of kWindow predecessors are loaded together (independent loads,
      // one memory round trip), then consumed in order up to the first inclusive
      // prefix; a flag not yet published restarts the window from it.
      constexpr int kWindow = 4;
      int64_t predecessor = tile - 1;
      uint32_t spins = 0;
      bool found = false;
      if (spin_limit == kForceGiveUp) {
        atomicAdd(reinterpret_cast<unsigned long long*>(failures), 1ull);
        found = true;  // skip the wait; the exclusive prefix stays incomplete
      }
      while (!found) {
        uint32_t window[kWindow];
#pragma unroll
        for (int w = 0; w < kWindow; ++w) {
          window[w] = predecessor - w >= 0
                          ? __hip_atomic_load(&flags[(predecessor - w) * kRadix + threadIdx.x], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT)
                          : 0u;
        }
        int consumed = 0;
#pragma unroll
        for (int w = 0; w < kWindow; ++w) {
          if (found || consumed < w) continue;  // stopped at an earlier entry
          const uint32_t word = window[w];
          if (word == 0u || predecessor - w < 0) continue;  // not published yet
          const bool prefix = (word & kFlagPrefix) != 0u;
          exclusive += prefix ? (word & kFlagValueMask) : word - 1u;
          consumed = w + 1;
          if (prefix || predecessor - w == 0) found = true;
        }
        if (found) break;
        predecessor -= consumed;
        if (consumed == 0) {
          if (++spins > spin_limit) {
            atomicAdd(reinterpret_cast<unsigned long long*>(failures), 1ull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __hip_atomic_store(&flags[tile * kRadix + threadIdx.x], kFlagPrefix | (exclusive + total), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_global_start[threadIdx.x] = static_cast<int32_t>(static_cast<int64_t>(digit_base[threadIdx.x]) + exclusive -
                                                       static_cast<int64_t>(s_local_start[threadIdx.x]));
  }
  __syncthreads();

  // Keys out; each thread remembers where its LDS slots go for the values.
  constexpr int kSlots = kSortTile / kSortThreads;
  int64_t destinations[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const int j = k * kSortThreads + threadIdx.x;
    destinations[k] = -1;
    if (j < valid_count) {
      const uint64_t key = s_keys[j];
      const uint32_t digit = static_cast<uint32_t>(compressed_key(key, kp) >> shift) & kMask;
      const int64_t out = static_cast<int64_t>(s_global_start[digit]) + j;
      // Always true for a consistent histogram; keeps a corrupted input in bounds.
      if (out >= 0 && out < capacity) {
        keys_out[out] = key;
        destinations[k] = out;
      }
    }
  }
  __syncthreads();  // keys read out of LDS
  // Phase 2: values through the same LDS buffer.
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int local = wave * kWaveRun + r * kWave + lane;
    if (local < valid_count) s_values[positions[r]] = values[r];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    if (destinations[k] >= 0) values_out[destinations[k]] = s_values[k * kSortThreads + threadIdx.x];
  }
}

__global__ __launch_bounds__(kBlockThreads) void ranges_kernel(const uint64_t* __restrict__ keys,
                                                               const int64_t* __restrict__ counters_in,
                                                               int64_t capacity, int64_t global_tiles,
                                                               int32_t* __restrict__ tile_starts,
                                                               int32_t* __restrict__ tile_ends,
                                                               int64_t* __restrict__ counters) {
  const int64_t n = element_count(counters_in, capacity);
  // Each lane takes kPer consecutive keys (loads issued together); the neighbours
  // across lanes come from shuffles, and lanes 0 and 63 read the one key beyond
  // their wave's run, so every key is read from memory about once.
  constexpr int kPer = 4;
  constexpr uint32_t kOutside = 0xfffffffeu;  // differs from every tile and from the sentinel tile
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave_id = (static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x) / kWave;
  const int64_t wave_count = static_cast<int64_t>(gridDim.x) * (kBlockThreads / kWave);
  int64_t ends = 0;
  for (int64_t base = wave_id * kWave * kPer; base < n; base += wave_count * kWave * kPer) {
    const int64_t first = base + static_cast<int64_t>(lane) * kPer;
    uint32_t tile[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) tile[k] = first + k < n ? static_cast<uint32_t>(keys[first + k] >> 32) : kOutside;
    uint32_t previous = __shfl_up(tile[kPer - 1], 1, kWave);
    uint32_t next = __shfl_down(tile[0], 1, kWave);
    if (lane == 0) previous = first > 0 && first < n ? static_cast<uint32_t>(keys[first - 1] >> 32) : kOutside;
    if (lane == kWave - 1) next = first + kPer < n ? static_cast<uint32_t>(keys[first + kPer] >> 32) : kOutside;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int64_t i = first + k;
      const uint32_t t = tile[k];
      if (i >= n || t >= static_cast<uint64_t>(global_tiles)) continue;
      const uint32_t before = k == 0 ? previous : tile[k - 1];
      const uint32_t after = k == kPer - 1 ? next : tile[k + 1];
      if (i == 0 || before != t) tile_starts[t] = static_cast<int32_t>(i);
      if (i + 1 == n || after != t) {
        tile_ends[t] = static_cast<int32_t>(i + 1);
        ++ends;
      }
    }
  }
  // One atomic per workgroup on the shared counter (per-wave atomics on one
  // address serialize: thousands of waves would queue on it).
  __shared__ int64_t s_ends[kBlockThreads / kWave];
  for (int offset = 32; offset > 0; offset >>= 1) ends += __shfl_xor(ends, offset, kWave);
  if (lane == 0) s_ends[threadIdx.x / kWave] = ends;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t total = 0;
    for (int w = 0; w < kBlockThreads / kWave; ++w) total += s_ends[w];
    if (total != 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(counters + VGR_COUNTER_ACTIVE_TILES),
                static_cast<unsigned long long>(total));
    }
  }
}

// Order key of a run member: its record's Gaussian ID, then the record index
// (IDs are unique within one camera's records; the index only orders records
// that a corrupted list could repeat).
__device__ __forceinline__ uint64_t tie_key(const vgr_record* __restrict__ records, int64_t visible_capacity,
                                            int32_t value) {
  const int32_t id = (value >= 0 && value < visible_capacity) ? records[value].gaussian_id : 0;
  return (static_cast<uint64_t>(static_cast<uint32_t>(id)) << 32) | static_cast<uint32_t>(value);
}

// Restores Gaussian-ID order inside runs of equal keys (same tile, same depth
// bits).  Needed only when the packed scene is stored in a spatial order, in
// which case records of a camera are not in registration-id order.  Every
// run is owned by the thread that sees its first element: insertion sort for
// short runs, in-place heap sort (O(n log n)) for runs longer than 32, such as
// the many equal depths of a fronto-parallel plane.
__global__ __launch_bounds__(kBlockThreads) void tie_order_kernel(const uint64_t* __restrict__ keys,
                                                                  int32_t* __restrict__ values,
                                                                  const vgr_record* __restrict__ records,
                                                                  const int64_t* __restrict__ counters,
                                                                  int64_t capacity, int64_t visible_capacity) {
  constexpr int64_t kInsertionRun = 32;
  const int64_t n = element_count(counters, capacity);
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x; i + 1 < n;
       i += static_cast<int64_t>(gridDim.x) * kBlockThreads) {
    const uint64_t key = keys[i];
    if (key == kSentinelKey || keys[i + 1] != key || (i > 0 && keys[i - 1] == key)) continue;
    int64_t end = i + 1;
    while (end < n && keys[end] == key) ++end;
    int32_t* run = values + i;
    const int64_t length = end - i;
    if (length <= kInsertionRun) {
      for (int64_t j = 1; j < length; ++j) {
        const int32_t moving = run[j];
        const uint64_t moving_key = tie_key(records, visible_capacity, moving);
        int64_t k = j - 1;
        while (k >= 0 && tie_key(records, visible_capacity, run[k]) > moving_key) {
          run[k + 1] = run[k];
          --k;
        }
        run[k + 1] = moving;
      }
      continue;
    }
    // Max-heap on the keys, then repeated extraction to the back.
    auto sift_down = [&](int64_t root, int64_t size) {
      const int32_t item = run[root];
      const uint64_t item_key = tie_key(records, visible_capacity, item);
      for (;;) {
        int64_t child = 2 * root + 1;
        if (child >= size) break;
        uint64_t child_key = tie_key(records, visible_capacity, run[child]);
        if (child + 1 < size) {
          const uint64_t right_key = tie_key(records, visible_capacity, run[child + 1]);
          if (right_key > child_key) {
            ++child;
            child_key = right_key;
          }
        }
        if (child_key <= item_key) break;
        run[root] = run[child];
        root = child;
      }
      run[root] = item;
    };
    for (int64_t root = length / 2 - 1; root >= 0; --root) sift_down(root, length);
    for (int64_t last = length - 1; last > 0; --last) {
      const int32_t top = run[0];
      run[0] = run[last];
      run[last] = top;
      sift_down(0, last);
    }
  }
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ data, int64_t n, uint32_t value) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    data[i] = value;
  }
}

__global__ void fill_i64_kernel(int64_t* __restrict__ data, int64_t n, int64_t value) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    data[i] = value;
  }
}

__global__ void fill_f32_kernel(float* __restrict__ data, int64_t n, float value) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    data[i] = value;
  }
}

__global__ void copy_i64_kernel(const int64_t* __restrict__ source, int64_t* __restrict__ target) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *target = *source;
}

__global__ void copy_u32_to_i64_kernel(const uint32_t* __restrict__ source, int64_t* __restrict__ target) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *target = static_cast<int64_t>(*source);
}

unsigned fill_grid(int64_t n) {
  int64_t blocks = (n + kBlockThreads - 1) / kBlockThreads;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  return static_cast<unsigned>(blocks);
}

template <int kDigitBits>
void launch_pass(int items, const uint64_t* keys_in, const int32_t* values_in, uint64_t* keys_out,
                 int32_t* values_out, const int64_t* counters, int64_t capacity, SortKeyParams kp, int shift,
                 const uint32_t* digit_base, uint32_t* flags, uint32_t* ticket, int64_t* failures,
                 uint32_t spin_limit, int64_t tiles, hipStream_t stream) {
  if (items == 8) {
    hipLaunchKernelGGL((sort_onesweep_pass_kernel<kDigitBits, 8>), dim3(static_cast<unsigned>(tiles)),
                       dim3(kSortThreads), 0, stream, keys_in, values_in, keys_out, values_out, counters, capacity,
                       kp, shift, digit_base, flags, ticket, failures, spin_limit);
  } else if (items == 12) {
    hipLaunchKernelGGL((sort_onesweep_pass_kernel<kDigitBits, 12>), dim3(static_cast<unsigned>(tiles)),
                       dim3(kSortThreads), 0, stream, keys_in, values_in, keys_out, values_out, counters, capacity,
                       kp, shift, digit_base, flags, ticket, failures, spin_limit);
  } else {
    hipLaunchKernelGGL((sort_onesweep_pass_kernel<kDigitBits, 16>), dim3(static_cast<unsigned>(tiles)),
                       dim3(kSortThreads), 0, stream, keys_in, values_in, keys_out, values_out, counters, capacity,
                       kp, shift, digit_base, flags, ticket, failures, spin_limit);
  }
}

}  // namespace

int launch_fill_u32(uint32_t* data, int64_t n, uint32_t value, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(fill_grid(n)), dim3(kBlockThreads), 0, stream, data, n, value);
  return hipGetLastError() == hipSuccess ? 0 : VGR_E_LAUNCH;
}

int launch_copy_i64(const int64_t* source, int64_t* target, hipStream_t stream) {
  hipLaunchKernelGGL(copy_i64_kernel, dim3(1), dim3(64), 0, stream, source, target);
  return hipGetLastError() == hipSuccess ? 0 : VGR_E_LAUNCH;
}

int launch_copy_u32_to_i64(const uint32_t* source, int64_t* target, hipStream_t stream) {
  hipLaunchKernelGGL(copy_u32_to_i64_kernel, dim3(1), dim3(64), 0, stream, source, target);
  return hipGetLastError() == hipSuccess ? 0 : VGR_E_LAUNCH;
}

int launch_fill_i64(int64_t* data, int64_t n, int64_t value, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fill_i64_kernel, dim3(fill_grid(n)), dim3(kBlockThreads), 0, stream, data, n, value);
  return hipGetLastError() == hipSuccess ? 0 : VGR_E_LAUNCH;
}

int launch_fill_f32(float* data, int64_t n, float value, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(fill_grid(n)), dim3(kBlockThreads), 0, stream, data, n, value);
  return hipGetLastError() == hipSuccess ? 0 : VGR_E_LAUNCH;
}

int launch_sort(const vgr_workspace& ws, const ScratchLayout& layout, char* scratch, uint32_t spin_limit,
                hipStream_t stream, int* result_in_out_buffer, bool restore_id_order) {
  const int digit_bits = layout.sort_digit_bits;
  const int passes = layout.sort_passes;
  const int radix = 1 << digit_bits;
  uint32_t* histogram = reinterpret_cast<uint32_t*>(scratch + layout.sort_histogram);
  uint32_t* digit_base = reinterpret_cast<uint32_t*>(scratch + layout.sort_digit_base);
  uint32_t* tickets = reinterpret_cast<uint32_t*>(scratch + layout.sort_tickets);
  uint32_t* flags = reinterpret_cast<uint32_t*>(scratch + layout.sort_flags);
  SortKeyParams kp;
  kp.depth_bits = layout.depth_bits;
  kp.depth_base = layout.depth_base;
  kp.key_bits = layout.sort_key_bits;
  const int64_t capacity = ws.intersection_capacity;
  const int64_t tiles = layout.sort_tiles;
  int64_t* failures = ws.counters + VGR_COUNTER_SORT_FAILURES;
  const uint32_t limit = spin_limit == 0 ? kDefaultSpinLimit : spin_limit;

  hipLaunchKernelGGL(sort_clear_kernel, dim3(1024), dim3(kBlockThreads), 0, stream, ws.counters, capacity,
                     digit_bits, passes, tiles, layout.sort_tile, histogram, tickets, flags);
  const int items = layout.sort_tile / kSortThreads;
  int64_t histogram_blocks = (capacity + kBlockThreads * 16 - 1) / (kBlockThreads * 16);
  histogram_blocks = histogram_blocks > 2048 ? 2048 : (histogram_blocks < 1 ? 1 : histogram_blocks);
  hipLaunchKernelGGL(sort_histogram_kernel, dim3(static_cast<unsigned>(histogram_blocks)), dim3(kBlockThreads), 0,
                     stream, ws.keys_in, ws.counters, capacity, kp, digit_bits, passes, histogram);
  hipLaunchKernelGGL(sort_digit_scan_kernel, dim3(1), dim3(kSortThreads), 0, stream, histogram, digit_bits, passes,
                     digit_base);
  const uint64_t* source_keys = ws.keys_in;
  const int32_t* source_values = ws.values_in;
  uint64_t* target_keys = ws.keys_out;
  int32_t* target_values = ws.values_out;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * digit_bits;
    uint32_t* pass_flags = flags + static_cast<int64_t>(p) * tiles * radix;
    const uint32_t* pass_base = digit_base + p * radix;
    switch (digit_bits) {
      case 4:
        launch_pass<4>(items, source_keys, source_values, target_keys, target_values, ws.counters, capacity, kp, shift,
                       pass_base, pass_flags, tickets + p, failures, limit, tiles, stream);
        break;
      case 5:
        launch_pass<5>(items, source_keys, source_values, target_keys, target_values, ws.counters, capacity, kp, shift,
                       pass_base, pass_flags, tickets + p, failures, limit, tiles, stream);
        break;
      case 6:
        launch_pass<6>(items, source_keys, source_values, target_keys, target_values, ws.counters, capacity, kp, shift,
                       pass_base, pass_flags, tickets + p, failures, limit, tiles, stream);
        break;
      case 7:
        launch_pass<7>(items, source_keys, source_values, target_keys, target_values, ws.counters, capacity, kp, shift,
                       pass_base, pass_flags, tickets + p, failures, limit, tiles, stream);
        break;
      case 9:
        launch_pass<9>(items, source_keys, source_values, target_keys, target_values, ws.counters, capacity, kp, shift,
                       pass_base, pass_flags, tickets + p, failures, limit, tiles, stream);
        break;
      default:
        launch_pass<8>(items, source_keys, source_values, target_keys, target_values, ws.counters, capacity, kp, shift,
                       pass_base, pass_flags, tickets + p, failures, limit, tiles, stream);
        break;
    }
    const uint64_t* next_keys = target_keys;
    const int32_t* next_values = target_values;
    target_keys = const_cast<uint64_t*>(source_keys);
    target_values = const_cast<int32_t*>(source_values);
    source_keys = next_keys;
    source_values = next_values;
  }
  *result_in_out_buffer = (passes % 2) == 1 ? 1 : 0;
  if (restore_id_order) {
    int64_t tie_blocks = (capacity + kBlockThreads * 8 - 1) / (kBlockThreads * 8);
    tie_blocks = tie_blocks > 4096 ? 4096 : (tie_blocks < 1 ? 1 : tie_blocks);
    hipLaunchKernelGGL(tie_order_kernel, dim3(static_cast<unsigned>(tie_blocks)), dim3(kBlockThreads), 0, stream,
                       *result_in_out_buffer ? ws.keys_out : ws.keys_in,
                       *result_in_out_buffer ? ws.values_out : ws.values_in, ws.records, ws.counters, capacity,
                       ws.visible_capacity);
  }
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("sort launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

int launch_ranges(const vgr_workspace& ws, int64_t global_tiles, bool sorted_in_out, hipStream_t stream) {
  if (global_tiles > 0) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(ws.tile_starts), global_tiles, 0u, stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(ws.tile_ends), global_tiles, 0u, stream);
  }
  const uint64_t* keys = sorted_in_out ? ws.keys_out : ws.keys_in;
  int64_t blocks = (ws.intersection_capacity + kBlockThreads * 16 - 1) / (kBlockThreads * 16);
  blocks = blocks > 4096 ? 4096 : (blocks < 1 ? 1 : blocks);
  hipLaunchKernelGGL(ranges_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlockThreads), 0, stream, keys,
                     ws.counters, ws.intersection_capacity, global_tiles, ws.tile_starts, ws.tile_ends, ws.counters);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("tile range launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
