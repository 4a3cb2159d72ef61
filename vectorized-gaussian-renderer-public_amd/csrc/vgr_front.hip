// Tile-front ordering: the batched path's alternative to the global radix sort.
//
// When every pixel of a tile saturates (transmittance below 1e-4) after the
// nearest part of the tile's depth-sorted list, compositing never reads the
// rest of the list: in a 5M-Gaussian room seen by 1024 cameras of 160x120 the
// compositor consumes about 5 % of the entries (lists of 16k entries on
// average, of which about 700 are read).  The global sort orders all of them.
// This path orders only the front of each tile's list:
//
//   count  : projection of every (camera group x chunk) pair; per workgroup an
//            LDS histogram over the group's tiles, flushed to per-tile counts
//            with one global atomic per non-empty tile (no per-intersection
//            atomics);
//   scan   : per-tile counts -> tile starts (intersection counter, overflow);
//   write  : projection again; per camera, LDS ranks per tile and one global
//            atomic per (workgroup, camera, tile) reserve the workgroup's run
//            inside the tile's bucket; each intersection stores its 27-bit
//            relative depth key and its packed Gaussian index (no 48-byte
//            records: compositing recomputes the projection of the entries it
//            reads);
//   front  : one workgroup per tile selects the entries below a depth
//            threshold chosen by histogram refinement so that at most 4096
//            remain, sorts them in LDS by (depth, Gaussian ID) and writes them
//            as the tile's list;
//   then compositing runs on these lists and reports every tile where a
//   pixel was still open at the end of a truncated list; those tiles are
//   sorted in full and composited again.
// The set of selected entries and their order are functions of the tile's
// multiset of (depth, Gaussian) pairs, so images are deterministic and equal
// to those of the globally sorted path; the order in which the atomics fill
// a bucket only changes what sits in the unsorted tail.
#include "vgr_project_device.h"

namespace vgr {
namespace {

using namespace dev;

constexpr int kFrontThreads = 512;
constexpr int kFrontWaves = kFrontThreads / kWave;
constexpr int kFrontItems = 8;
constexpr int kFrontCapacity = kFrontThreads * kFrontItems;  // entries sorted in LDS per tile (4096)
constexpr int kFrontWaveRun = kFrontCapacity / kFrontWaves;  // entries ranked by one wave per pass
constexpr int kFrontLoads = 4;  // global keys in flight per thread in the bucket scans
constexpr int kFrontDigitBits = 9;
constexpr int kFrontRadix = 1 << kFrontDigitBits;
constexpr uint32_t kFrontInvalidKey = 0xffffffffu;

__device__ __forceinline__ uint64_t front_lanemask_lt() {
  const int lane = threadIdx.x & (kWave - 1);
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

// Lanes of the wave holding the same value in the low `bits` bits of `digit`
// (among the lanes in `live`).
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, int bits, uint64_t live) {
  uint64_t peers = live;
  for (int b = 0; b < bits; ++b) {
    const bool bit = (digit >> b) & 1u;
    const uint64_t vote = __ballot(bit);
    peers &= bit ? vote : ~vote;
  }
  return peers;
}

// ---- count ---------------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void front_count_kernel(ProjectionArgs args, int group_tile_limit,
                                                                    unsigned long long* __restrict__ tile_counts,
                                                                    uint64_t* __restrict__ block_masks,
                                                                    uint64_t* __restrict__ wave_masks,
                                                                    int64_t* __restrict__ block_visible,
                                                                    int64_t* __restrict__ counters) {
  extern __shared__ uint32_t s_hist[];  // [group_tile_limit]
  __shared__ unsigned long long s_mask;
  __shared__ int64_t s_visible[kBlockThreads / kWave];
  int64_t block;
  uint64_t mask;
  if (!resolve_block(args, block, mask)) return;
  const int group = static_cast<int>(block / args.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * args.chunks;
  const int64_t g = chunk * kBlockThreads + threadIdx.x;
  const int first = group * args.cameras_per_block;
  const int last = min(args.active_count, first + args.cameras_per_block);
  const CameraParams& first_camera = args.table[first];
  const CameraParams& last_camera = args.table[last - 1];
  const int64_t group_base = first_camera.tile_base;
  const int64_t group_tiles = static_cast<int64_t>(last_camera.tile_base) +
                              static_cast<int64_t>(last_camera.tiles_x) * last_camera.tiles_y - group_base;
  const int64_t wave_slot = static_cast<int64_t>(blockIdx.x) * (kBlockThreads / kWave) + threadIdx.x / kWave;
  if (group_tiles <= 0 || group_tiles > group_tile_limit) {
    // Layout outside what the host sized the histogram for: fail closed.
    if ((threadIdx.x & (kWave - 1)) == 0) wave_masks[wave_slot] = 0;
    if (threadIdx.x == 0) {
      block_masks[blockIdx.x] = 0;
      block_visible[blockIdx.x] = 0;
      atomicAdd(reinterpret_cast<unsigned long long*>(counters + VGR_COUNTER_SORT_FAILURES), 1ull);
    }
    return;
  }
  for (int64_t i = threadIdx.x; i < group_tiles; i += kBlockThreads) s_hist[i] = 0u;
  if (threadIdx.x == 0) s_mask = 0ull;
  __syncthreads();
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(args, g, loaded, m0, m1, m2, opacity, cov);
  int64_t visible = 0;
  uint64_t wave_mask = 0;
  for (int a = first; a < last; ++a) {
    if (((mask >> (a - first)) & 1ull) == 0) continue;
    const CameraParams& cam = args.table[a];
    Projected p;
    TileRect rect;
    int tiles;
    const bool seen = evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
    if (seen) {
      ++visible;
      const int local = static_cast<int>(static_cast<int64_t>(cam.tile_base) - group_base);
      for (int ty = rect.min_y; ty < rect.max_y; ++ty) {
        for (int tx = rect.min_x; tx < rect.max_x; ++tx) atomicAdd(&s_hist[local + ty * cam.tiles_x + tx], 1u);
      }
    }
    if (__ballot(seen) != 0ull) wave_mask |= 1ull << (a - first);
  }
  // Per-workgroup partial counts (summed by front_totals_kernel): atomics of every
  // wave on one counter word would serialize behind each other.
  visible = wave_sum_i64(visible);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (wave_mask != 0ull) atomicOr(&s_mask, static_cast<unsigned long long>(wave_mask));
    s_visible[threadIdx.x / kWave] = visible;
    wave_masks[wave_slot] = wave_mask;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < group_tiles; i += kBlockThreads) {
    const uint32_t c = s_hist[i];
    if (c != 0u) atomicAdd(&tile_counts[2 * (group_base + i)], static_cast<unsigned long long>(c));
  }
  if (threadIdx.x == 0) {
    block_masks[blockIdx.x] = s_mask;
    int64_t total = 0;
    for (int w = 0; w < kBlockThreads / kWave; ++w) total += s_visible[w];
    block_visible[blockIdx.x] = total;
  }
}

// One workgroup: the visible-pair counter (sum of the per-workgroup counts of the
// kept workgroups) and the active-tile counter (non-empty buckets).
__global__ __launch_bounds__(1024) void front_totals_kernel(const int64_t* __restrict__ block_visible,
                                                            int64_t blocks, const int64_t* __restrict__ kept,
                                                            const int32_t* __restrict__ lengths, int64_t tiles,
                                                            int64_t* __restrict__ counters) {
  __shared__ int64_t s_sum[2][1024 / kWave];
  const int64_t n = kept != nullptr ? (*kept < blocks ? *kept : blocks) : blocks;
  int64_t visible = 0, active = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) visible += block_visible[i];
  for (int64_t t = threadIdx.x; t < tiles; t += 1024) active += lengths[t] > 0 ? 1 : 0;
  visible = wave_sum_i64(visible);
  active = wave_sum_i64(active);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_sum[0][threadIdx.x / kWave] = visible;
    s_sum[1][threadIdx.x / kWave] = active;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t total_visible = 0, total_active = 0;
    for (int w = 0; w < 1024 / kWave; ++w) {
      total_visible += s_sum[0][w];
      total_active += s_sum[1][w];
    }
    counters[VGR_COUNTER_VISIBLE] = total_visible;
    counters[VGR_COUNTER_ACTIVE_TILES] = total_active;
  }
}

// ---- tile starts ---------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void front_starts_kernel(const int64_t* __restrict__ counts,
                                                                     const int64_t* __restrict__ bases,
                                                                     int64_t tiles, const int64_t* __restrict__ totals,
                                                                     int64_t capacity,
                                                                     int32_t* __restrict__ tile_starts,
                                                                     int32_t* __restrict__ tile_ends,
                                                                     int32_t* __restrict__ lengths,
                                                                     unsigned long long* __restrict__ cursors,
                                                                     int64_t* __restrict__ counters) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x;
  if (t == 0) {
    const int64_t total = totals[0];
    counters[VGR_COUNTER_INTERSECTIONS] = total;
    counters[VGR_COUNTER_VISIBLE_OVERFLOW] = 0;
    counters[VGR_COUNTER_INTERSECTION_OVERFLOW] = total > capacity ? total - capacity : 0;
  }
  if (t >= tiles) return;
  const int64_t start = bases[2 * t];
  const int64_t count = counts[2 * t];
  // Starts beyond the capacity are clamped; the writes of such a render are
  // dropped and the overflow counter tells the caller to grow the buffers.
  const int64_t clamped = start < capacity ? start : capacity;
  tile_starts[t] = static_cast<int32_t>(clamped);
  tile_ends[t] = static_cast<int32_t>(clamped);
  const int64_t room = capacity - clamped;
  lengths[t] = static_cast<int32_t>(count < room ? count : room);
  cursors[t] = static_cast<unsigned long long>(start);
}

// ---- write ---------------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void front_write_kernel(ProjectionArgs args, int camera_tile_limit,
                                                                    const uint64_t* __restrict__ block_masks,
                                                                    const uint64_t* __restrict__ wave_masks,
                                                                    unsigned long long* __restrict__ cursors,
                                                                    uint32_t* __restrict__ keys,
                                                                    int32_t* __restrict__ values, int64_t capacity,
                                                                    uint32_t depth_base,
                                                                    int64_t* __restrict__ counters) {
  extern __shared__ uint32_t s_dynamic[];
  uint32_t* s_count = s_dynamic;                          // [camera_tile_limit] ranked entries per tile
  uint32_t* s_base = s_dynamic + camera_tile_limit;       // [camera_tile_limit] first slot of the workgroup's run
  uint32_t* s_large = s_dynamic + 2 * camera_tile_limit;  // [camera_tile_limit] entries of larger splats
  uint32_t* s_cursor = s_dynamic + 3 * camera_tile_limit;  // [camera_tile_limit] next slot for larger splats
  // Splats covering up to four tiles keep their LDS ranks in registers; larger
  // ones are counted per tile and take their slots from an LDS cursor after the
  // ranked entries of the run, so every intersection of the workgroup is placed
  // with one global reservation per (camera, tile) and no per-intersection
  // global atomic.
  constexpr int kRanks = 4;
  int64_t block;
  uint64_t mask;
  if (!resolve_block(args, block, mask)) return;
  const int group = static_cast<int>(block / args.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * args.chunks;
  const int64_t g = chunk * kBlockThreads + threadIdx.x;
  const int first = group * args.cameras_per_block;
  uint64_t cameras = block_masks[blockIdx.x] & mask;
  // Cameras that see none of this wave's Gaussians (per the count pass) are not
  // evaluated by the wave; it still joins the workgroup's barriers.
  const uint64_t wave_cameras = wave_masks[static_cast<int64_t>(blockIdx.x) * (kBlockThreads / kWave) +
                                           threadIdx.x / kWave];
  for (int i = threadIdx.x; i < camera_tile_limit; i += kBlockThreads) {
    s_count[i] = 0u;
    s_large[i] = 0u;
  }
  __syncthreads();
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(args, g, loaded, m0, m1, m2, opacity, cov);
  while (cameras != 0ull) {  // uniform over the workgroup
    const int a = first + __ffsll(static_cast<unsigned long long>(cameras)) - 1;
    cameras &= cameras - 1ull;
    const CameraParams& cam = args.table[a];
    const int camera_tiles = cam.tiles_x * cam.tiles_y;
    if (camera_tiles > camera_tile_limit) {
      if (threadIdx.x == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(counters + VGR_COUNTER_SORT_FAILURES), 1ull);
      }
      continue;
    }
    Projected p;
    TileRect rect;
    int tiles = 0;
    const bool seen = ((wave_cameras >> (a - first)) & 1ull) != 0ull &&
                      evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
    const bool ranked = seen && tiles <= kRanks;
    uint32_t rank[kRanks] = {0u, 0u, 0u, 0u};
    if (ranked) {
      int k = 0;
      for (int ty = rect.min_y; ty < rect.max_y; ++ty) {
        for (int tx = rect.min_x; tx < rect.max_x; ++tx) {
          const uint32_t r = atomicAdd(&s_count[ty * cam.tiles_x + tx], 1u);
#pragma unroll
          for (int q = 0; q < kRanks; ++q) {
            if (q == k) rank[q] = r;
          }
          ++k;
        }
      }
    } else if (seen) {
      for (int ty = rect.min_y; ty < rect.max_y; ++ty) {
        for (int tx = rect.min_x; tx < rect.max_x; ++tx) atomicAdd(&s_large[ty * cam.tiles_x + tx], 1u);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < camera_tiles; i += kBlockThreads) {
      const uint32_t c = s_count[i] + s_large[i];
      if (c != 0u) {
        const unsigned long long base = atomicAdd(&cursors[cam.tile_base + i], static_cast<unsigned long long>(c));
        s_base[i] = base < 0xffffffffull ? static_cast<uint32_t>(base) : 0xffffffffu;
        s_cursor[i] = s_count[i];  // larger splats fill the run after the ranked entries
        s_count[i] = 0u;
        s_large[i] = 0u;
      }
    }
    __syncthreads();
    if (seen) {
      const uint32_t key = vgr_float_bits(p.depth) - depth_base;
      const int32_t value = static_cast<int32_t>(g);
      int k = 0;
      for (int ty = rect.min_y; ty < rect.max_y; ++ty) {
        for (int tx = rect.min_x; tx < rect.max_x; ++tx) {
          const int local = ty * cam.tiles_x + tx;
          uint32_t r = 0u;
          if (ranked) {
#pragma unroll
            for (int q = 0; q < kRanks; ++q) {
              if (q == k) r = rank[q];
            }
          } else {
            r = atomicAdd(&s_cursor[local], 1u);
          }
          const uint64_t slot = static_cast<uint64_t>(s_base[local]) + r;
          if (slot < static_cast<uint64_t>(capacity)) {
            keys[slot] = key;
            values[slot] = value;
          }
          ++k;
        }
      }
    }
  }
}

// ---- front selection and LDS sort ----------------------------------------------

struct FrontShared {
  uint32_t keys[kFrontCapacity];
  int32_t values[kFrontCapacity];
  uint16_t wave_count[kFrontWaves][kFrontRadix];
  uint32_t histogram[kFrontRadix];
  uint32_t digit_start[kFrontRadix];
  uint32_t wave_totals[kFrontWaves];
  uint32_t fill;
  uint32_t pick;
  uint32_t below;
};

__device__ __forceinline__ int32_t front_gaussian_id(const int32_t* gaussian_ids, int32_t value) {
  return gaussian_ids != nullptr ? gaussian_ids[value] : value;
}

// Exclusive scan of the kFrontRadix histogram bins (kFrontBinsPerThread
// consecutive bins per thread) into s.digit_start.  Ends with a barrier.
constexpr int kFrontBinsPerThread = kFrontRadix / kFrontThreads;
static_assert(kFrontBinsPerThread >= 1 && kFrontBinsPerThread * kFrontThreads == kFrontRadix,
              "the bin scan gives every thread the same number of bins");

__device__ __forceinline__ void front_scan_bins(FrontShared& s) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  uint32_t bins[kFrontBinsPerThread];
  uint32_t own = 0u;
#pragma unroll
  for (int b = 0; b < kFrontBinsPerThread; ++b) {
    bins[b] = s.histogram[kFrontBinsPerThread * threadIdx.x + b];
    own += bins[b];
  }
  uint32_t inclusive = own;
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const uint32_t t = __shfl_up(inclusive, offset, kWave);
    if (lane >= offset) inclusive += t;
  }
  if (lane == kWave - 1) s.wave_totals[wave] = inclusive;
  __syncthreads();
  uint32_t running = inclusive - own;
  for (int w = 0; w < wave; ++w) running += s.wave_totals[w];
#pragma unroll
  for (int b = 0; b < kFrontBinsPerThread; ++b) {
    s.digit_start[kFrontBinsPerThread * threadIdx.x + b] = running;
    running += bins[b];
  }
  __syncthreads();
}

// Stable LSD radix sort of the first `count` LDS entries by key bits [0, key_bits).
__device__ __forceinline__ void front_lds_sort(FrontShared& s, int count, int key_bits) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const uint64_t lt = front_lanemask_lt();
  for (int shift = 0; shift < key_bits; shift += kFrontDigitBits) {
    const int bits = key_bits - shift < kFrontDigitBits ? key_bits - shift : kFrontDigitBits;
    const uint32_t digit_mask = (1u << bits) - 1u;
    for (int i = threadIdx.x; i < kFrontWaves * kFrontRadix; i += kFrontThreads) (&s.wave_count[0][0])[i] = 0;
    uint32_t keys[kFrontItems];
    int32_t values[kFrontItems];
    uint32_t digits[kFrontItems];
    uint32_t ranks[kFrontItems];
#pragma unroll
    for (int r = 0; r < kFrontItems; ++r) {
      const int index = wave * kFrontWaveRun + r * kWave + lane;
      keys[r] = index < count ? s.keys[index] : kFrontInvalidKey;
      values[r] = index < count ? s.values[index] : -1;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kFrontItems; ++r) {
      const int index = wave * kFrontWaveRun + r * kWave + lane;
      const bool valid = index < count;
      const uint32_t digit = valid ? (keys[r] >> shift) & digit_mask : 0u;
      const uint64_t peers = match_digit(digit, bits, __ballot(valid));
      digits[r] = digit;
      ranks[r] = 0u;
      if (valid) {
        const uint32_t before = static_cast<uint32_t>(__popcll(peers & lt));
        const uint32_t base = s.wave_count[wave][digit];
        ranks[r] = base + before;
        if (before == 0u) s.wave_count[wave][digit] = static_cast<uint16_t>(base + __popcll(peers));
      }
    }
    __syncthreads();
    // Per digit: totals over the waves (in place: per-wave offsets) -> histogram.
    for (int d = threadIdx.x; d < kFrontRadix; d += kFrontThreads) {
      uint32_t running = 0;
#pragma unroll
      for (int w = 0; w < kFrontWaves; ++w) {
        const uint32_t c = s.wave_count[w][d];
        s.wave_count[w][d] = static_cast<uint16_t>(running);
        running += c;
      }
      s.histogram[d] = running;
    }
    __syncthreads();
    front_scan_bins(s);
#pragma unroll
    for (int r = 0; r < kFrontItems; ++r) {
      const int index = wave * kFrontWaveRun + r * kWave + lane;
      if (index < count) {
        const uint32_t position = s.digit_start[digits[r]] + s.wave_count[wave][digits[r]] + ranks[r];
        s.keys[position] = keys[r];
        s.values[position] = values[r];
      }
    }
    __syncthreads();
  }
}

// Histogram of digit (key >> shift) & mask over the keys whose bits above
// `prefix_shift` equal `prefix`, for keys in global memory.  Ends with a barrier.
__device__ __forceinline__ void front_histogram(FrontShared& s, const uint32_t* __restrict__ keys, int64_t start,
                                                int64_t count, uint32_t prefix, int prefix_shift, int shift,
                                                int bits) {
  for (int i = threadIdx.x; i < kFrontRadix; i += kFrontThreads) s.histogram[i] = 0u;
  __syncthreads();
  const uint32_t digit_mask = (1u << bits) - 1u;
  const uint64_t lt = front_lanemask_lt();
  // kFrontLoads keys per thread per step, all loads issued before any is used.
  for (int64_t base = 0; base < count; base += kFrontThreads * kFrontLoads) {
    uint32_t loaded[kFrontLoads];
#pragma unroll
    for (int r = 0; r < kFrontLoads; ++r) {
      const int64_t i = base + r * kFrontThreads + threadIdx.x;
      loaded[r] = i < count ? keys[start + i] : kFrontInvalidKey;
    }
#pragma unroll
    for (int r = 0; r < kFrontLoads; ++r) {
      const int64_t i = base + r * kFrontThreads + threadIdx.x;
      const uint32_t key = loaded[r];
      const bool in_bucket = i < count && (prefix_shift >= 32 || (key >> prefix_shift) == prefix);
      const uint32_t digit = in_bucket ? (key >> shift) & digit_mask : 0u;
      const uint64_t live = __ballot(in_bucket);
      if (live == 0ull) continue;
      const uint64_t peers = match_digit(digit, bits, live);
      if (in_bucket && (peers & lt) == 0ull) atomicAdd(&s.histogram[digit], static_cast<uint32_t>(__popcll(peers)));
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kFrontThreads) void front_sort_kernel(
    const uint32_t* __restrict__ keys, const int32_t* __restrict__ values, const int32_t* __restrict__ tile_starts,
    const int32_t* __restrict__ lengths, int64_t tiles, int key_bits, const int32_t* __restrict__ gaussian_ids,
    int32_t* __restrict__ values_out, int32_t* __restrict__ tile_ends, int64_t* __restrict__ counters) {
  __shared__ FrontShared s;
  const int64_t t = blockIdx.x;
  if (t >= tiles) return;
  const int64_t start = tile_starts[t];
  const int64_t count = lengths[t];
  if (count <= 0) return;
  if (threadIdx.x == 0) s.fill = 0u;
  __syncthreads();
  // Threshold: every entry with key < threshold is taken (at most kFrontCapacity of them).
  uint32_t threshold = kFrontInvalidKey;
  bool all = count <= kFrontCapacity;
  if (!all) {
    // Radix refinement: at each level the bucket [prefix, prefix + 2^prefix_shift)
    // holds more entries than the room left; its digits split it and the largest
    // run of whole digits that still fits is taken.
    uint32_t prefix = 0u;
    int prefix_shift = key_bits;
    uint32_t taken = 0u;
    while (prefix_shift > 0) {
      const int bits = prefix_shift < kFrontDigitBits ? prefix_shift : kFrontDigitBits;
      const int shift = prefix_shift - bits;
      front_histogram(s, keys, start, count, prefix_shift >= key_bits ? 0u : (prefix >> prefix_shift),
                      prefix_shift >= key_bits ? 32 : prefix_shift, shift, bits);
      front_scan_bins(s);
      // Digit d whose bucket overflows: taken + start(d) <= capacity < taken + start(d) + hist(d).
      if (threadIdx.x == 0) {
        s.pick = 0u;
        s.below = 0u;
      }
      __syncthreads();
      for (int d = threadIdx.x; d < (1 << bits); d += kFrontThreads) {
        const uint32_t before = s.digit_start[d];
        if (taken + before <= static_cast<uint32_t>(kFrontCapacity) &&
            taken + before + s.histogram[d] > static_cast<uint32_t>(kFrontCapacity)) {
          s.pick = static_cast<uint32_t>(d);
          s.below = before;
        }
      }
      __syncthreads();
      const uint32_t digit = s.pick;
      taken += s.below;
      prefix |= digit << shift;
      prefix_shift = shift;
      __syncthreads();
      if (taken >= static_cast<uint32_t>(kFrontCapacity / 2)) break;
    }
    threshold = prefix;  // keys below it: `taken` entries
  }
  // Gather the selected entries into LDS (order irrelevant: sorted next).
  const uint64_t lt = front_lanemask_lt();
  for (int64_t base = 0; base < count; base += kFrontThreads * kFrontLoads) {
    uint32_t loaded[kFrontLoads];
#pragma unroll
    for (int r = 0; r < kFrontLoads; ++r) {
      const int64_t i = base + r * kFrontThreads + threadIdx.x;
      loaded[r] = i < count ? keys[start + i] : kFrontInvalidKey;
    }
#pragma unroll
    for (int r = 0; r < kFrontLoads; ++r) {
      const int64_t i = base + r * kFrontThreads + threadIdx.x;
      const uint32_t key = loaded[r];
      const bool take = i < count && (all || key < threshold);
      const uint64_t vote = __ballot(take);
      if (vote == 0ull) continue;
      uint32_t wave_base = 0u;
      if ((threadIdx.x & (kWave - 1)) == 0) wave_base = atomicAdd(&s.fill, static_cast<uint32_t>(__popcll(vote)));
      wave_base = __shfl(wave_base, 0, kWave);
      if (take) {
        const uint32_t position = wave_base + static_cast<uint32_t>(__popcll(vote & lt));
        if (position < static_cast<uint32_t>(kFrontCapacity)) {
          s.keys[position] = key;
          s.values[position] = values[start + i];
        }
      }
    }
  }
  __syncthreads();
  const int selected = static_cast<int>(s.fill < static_cast<uint32_t>(kFrontCapacity) ? s.fill : kFrontCapacity);
  __syncthreads();
  front_lds_sort(s, selected, key_bits);
  // Equal keys (same depth bits) are ordered by Gaussian ID: each member of a
  // run of equal keys counts the members with a smaller ID.
  for (int i = threadIdx.x; i < selected; i += kFrontThreads) {
    const uint32_t key = s.keys[i];
    const int32_t value = s.values[i];
    const bool tied = (i > 0 && s.keys[i - 1] == key) || (i + 1 < selected && s.keys[i + 1] == key);
    int position = i;
    if (tied) {
      int run_start = i, run_end = i + 1;
      while (run_start > 0 && s.keys[run_start - 1] == key) --run_start;
      while (run_end < selected && s.keys[run_end] == key) ++run_end;
      const int32_t id = front_gaussian_id(gaussian_ids, value);
      int rank = 0;
      for (int j = run_start; j < run_end; ++j) {
        if (front_gaussian_id(gaussian_ids, s.values[j]) < id) ++rank;
      }
      position = run_start + rank;
    }
    values_out[start + position] = value;
  }
  if (threadIdx.x == 0) tile_ends[t] = static_cast<int32_t>(start + selected);
}

// ---- fallback: full sort of the tiles compositing found open ------------------

// Stable LSD sort of a whole bucket through global memory, one workgroup per
// tile: each pass histograms the bucket, then ranks it chunk by chunk (4096
// entries) in LDS and scatters every chunk behind the previous ones.
__global__ __launch_bounds__(kFrontThreads) void front_full_sort_kernel(
    uint32_t* keys_a, int32_t* values_a, uint32_t* keys_b, int32_t* values_b, const int32_t* __restrict__ tile_starts, const int32_t* __restrict__ lengths,
    const int32_t* __restrict__ open_tiles, const uint32_t* __restrict__ open_count, uint32_t* __restrict__ work,
    int key_bits, const int32_t* __restrict__ gaussian_ids, int32_t* __restrict__ result_values,
    int32_t* __restrict__ tile_ends) {
  __shared__ FrontShared s;
  __shared__ uint32_t s_item;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const uint64_t lt = front_lanemask_lt();
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(work, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    __syncthreads();
    if (item >= *open_count) break;
    const int64_t t = open_tiles[item];
    const int64_t start = tile_starts[t];
    const int64_t count = lengths[t];
    // Passes alternate a -> b -> a ...; an odd pass count ends in b.
    uint32_t* src_keys = keys_a;
    int32_t* src_values = values_a;
    uint32_t* dst_keys = keys_b;
    int32_t* dst_values = values_b;
    for (int shift = 0; shift < key_bits; shift += kFrontDigitBits) {
      const int bits = key_bits - shift < kFrontDigitBits ? key_bits - shift : kFrontDigitBits;
      const uint32_t digit_mask = (1u << bits) - 1u;
      front_histogram(s, src_keys, start, count, 0u, 32, shift, bits);
      front_scan_bins(s);
      for (int64_t chunk = 0; chunk < count; chunk += kFrontCapacity) {
        const int chunk_count =
            static_cast<int>(count - chunk < kFrontCapacity ? count - chunk : kFrontCapacity);
        for (int i = threadIdx.x; i < kFrontWaves * kFrontRadix; i += kFrontThreads) (&s.wave_count[0][0])[i] = 0;
        __syncthreads();
        uint32_t keys[kFrontItems];
        int32_t values[kFrontItems];
        uint32_t digits[kFrontItems];
        uint32_t ranks[kFrontItems];
#pragma unroll
        for (int r = 0; r < kFrontItems; ++r) {
          const int index = wave * kFrontWaveRun + r * kWave + lane;
          const bool valid = index < chunk_count;
          keys[r] = valid ? src_keys[start + chunk + index] : kFrontInvalidKey;
          values[r] = valid ? src_values[start + chunk + index] : -1;
          const uint32_t digit = valid ? (keys[r] >> shift) & digit_mask : 0u;
          const uint64_t peers = match_digit(digit, bits, __ballot(valid));
          digits[r] = digit;
          ranks[r] = 0u;
          if (valid) {
            const uint32_t before = static_cast<uint32_t>(__popcll(peers & lt));
            const uint32_t base = s.wave_count[wave][digit];
            ranks[r] = base + before;
            if (before == 0u) s.wave_count[wave][digit] = static_cast<uint16_t>(base + __popcll(peers));
          }
        }
        __syncthreads();
        for (int d = threadIdx.x; d < kFrontRadix; d += kFrontThreads) {
          uint32_t running = 0;
#pragma unroll
          for (int w = 0; w < kFrontWaves; ++w) {
            const uint32_t c = s.wave_count[w][d];
            s.wave_count[w][d] = static_cast<uint16_t>(running);
            running += c;
          }
          s.histogram[d] = running;  // this chunk's count of digit d
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kFrontItems; ++r) {
          const int index = wave * kFrontWaveRun + r * kWave + lane;
          if (index < chunk_count) {
            const int64_t position = start + static_cast<int64_t>(s.digit_start[digits[r]]) +
                                     s.wave_count[wave][digits[r]] + ranks[r];
            dst_keys[position] = keys[r];
            dst_values[position] = values[r];
          }
        }
        __syncthreads();
        for (int d = threadIdx.x; d < kFrontRadix; d += kFrontThreads) s.digit_start[d] += s.histogram[d];
        __syncthreads();
      }
      uint32_t* k = src_keys;
      int32_t* v = src_values;
      src_keys = dst_keys;
      src_values = dst_values;
      dst_keys = k;
      dst_values = v;
    }
    // src now holds the sorted bucket.  Runs of equal keys (equal depth bits)
    // are put in Gaussian-ID order in place by the thread at each run's start.
    for (int64_t i = threadIdx.x; i + 1 < count; i += kFrontThreads) {
      const uint32_t key = src_keys[start + i];
      if (src_keys[start + i + 1] != key || (i > 0 && src_keys[start + i - 1] == key)) continue;
      int64_t run_end = i + 2;
      while (run_end < count && src_keys[start + run_end] == key) ++run_end;
      for (int64_t j = i + 1; j < run_end; ++j) {
        const int32_t moving = src_values[start + j];
        const int32_t moving_id = front_gaussian_id(gaussian_ids, moving);
        int64_t k = j - 1;
        while (k >= i && front_gaussian_id(gaussian_ids, src_values[start + k]) > moving_id) {
          src_values[start + k + 1] = src_values[start + k];
          --k;
        }
        src_values[start + k + 1] = moving;
      }
    }
    __syncthreads();
    if (src_values != result_values) {
      for (int64_t i = threadIdx.x; i < count; i += kFrontThreads) result_values[start + i] = src_values[start + i];
    }
    if (threadIdx.x == 0) tile_ends[t] = static_cast<int32_t>(start + count);
    __syncthreads();
  }
}

}  // namespace

// Scratch of the front path (offsets into workspace.scratch, filled by the caller's layout).
int launch_front_projection(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                            const vgr_workspace& ws, const ScratchLayout& layout, char* scratch, int64_t global_tiles,
                            hipStream_t stream) {
  ProjectionArgs args;
  make_projection_args(scene, cameras, options, layout, scratch, args);
  const bool culled = args.item_blocks != nullptr;
  const int64_t blocks = layout.projection_blocks;
  unsigned long long* tile_counts = reinterpret_cast<unsigned long long*>(scratch + layout.front_counts);
  int64_t* tile_bases = reinterpret_cast<int64_t*>(scratch + layout.front_bases);
  int64_t* totals = reinterpret_cast<int64_t*>(scratch + layout.front_totals);
  unsigned long long* cursors = reinterpret_cast<unsigned long long*>(scratch + layout.front_cursors);
  int32_t* lengths = reinterpret_cast<int32_t*>(scratch + layout.front_lengths);
  uint64_t* block_masks = reinterpret_cast<uint64_t*>(scratch + layout.front_block_masks);
  uint64_t* wave_masks = reinterpret_cast<uint64_t*>(scratch + layout.front_wave_masks);
  int64_t* block_visible = reinterpret_cast<int64_t*>(scratch + layout.front_block_visible);

  launch_fill_i64(ws.counters, VGR_COUNTER_COUNT, 0, stream);
  launch_fill_i64(reinterpret_cast<int64_t*>(tile_counts), 2 * global_tiles, 0, stream);
  if (culled) {
    // The culling kernels of the global path (camera masks per kept block) run on
    // the full candidate grid.
    ProjectionArgs candidates = args;
    candidates.item_blocks = nullptr;
    candidates.item_masks = nullptr;
    candidates.item_count = nullptr;
    const int status = launch_cull(candidates, scene, layout, scratch, options.statistics, stream);
    if (status != 0) return status;
  }
  if (blocks > 0) {
    const size_t count_lds = sizeof(uint32_t) * static_cast<size_t>(layout.front_group_tiles);
    hipLaunchKernelGGL(front_count_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlockThreads), count_lds,
                       stream, args, static_cast<int>(layout.front_group_tiles), tile_counts, block_masks,
                       wave_masks, block_visible, ws.counters);
  }
  scan_pairs(reinterpret_cast<const int64_t*>(tile_counts), global_tiles, tile_bases, tile_bases + 2 * global_tiles,
             0, 0, nullptr, totals, stream);
  if (global_tiles > 0) {
    hipLaunchKernelGGL(front_starts_kernel, dim3(static_cast<unsigned>((global_tiles + kBlockThreads - 1) / kBlockThreads)),
                       dim3(kBlockThreads), 0, stream, reinterpret_cast<const int64_t*>(tile_counts), tile_bases,
                       global_tiles, totals, ws.intersection_capacity, ws.tile_starts, ws.tile_ends, lengths, cursors,
                       ws.counters);
  }
  hipLaunchKernelGGL(front_totals_kernel, dim3(1), dim3(1024), 0, stream, block_visible, blocks,
                     culled ? args.item_count : nullptr, lengths, global_tiles, ws.counters);
  if (blocks > 0) {
    const size_t write_lds = 4 * sizeof(uint32_t) * static_cast<size_t>(layout.front_camera_tiles);
    hipLaunchKernelGGL(front_write_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlockThreads), write_lds,
                       stream, args, static_cast<int>(layout.front_camera_tiles), block_masks, wave_masks, cursors,
                       reinterpret_cast<uint32_t*>(ws.keys_in), ws.values_in, ws.intersection_capacity,
                       layout.depth_base, ws.counters);
  }
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("front projection launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

int launch_front_sort(const vgr_scene& scene, const vgr_workspace& ws, const ScratchLayout& layout, char* scratch,
                      int64_t global_tiles, hipStream_t stream) {
  if (global_tiles <= 0) return 0;
  const int32_t* lengths = reinterpret_cast<const int32_t*>(scratch + layout.front_lengths);
  hipLaunchKernelGGL(front_sort_kernel, dim3(static_cast<unsigned>(global_tiles)), dim3(kFrontThreads), 0, stream,
                     reinterpret_cast<const uint32_t*>(ws.keys_in), ws.values_in, ws.tile_starts, lengths,
                     global_tiles, layout.depth_bits, scene.gaussian_ids, ws.values_out, ws.tile_ends, ws.counters);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("front sort launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

int launch_front_fallback_sort(const vgr_scene& scene, const vgr_workspace& ws, const ScratchLayout& layout,
                               char* scratch, int workgroups, hipStream_t stream) {
  uint32_t* open_state = reinterpret_cast<uint32_t*>(scratch + layout.front_open_state);
  const int32_t* open_tiles = reinterpret_cast<const int32_t*>(scratch + layout.front_open_tiles);
  const int32_t* lengths = reinterpret_cast<const int32_t*>(scratch + layout.front_lengths);
  // Three 9-bit passes (27-bit depth keys) end in the second buffer pair.
  hipLaunchKernelGGL(front_full_sort_kernel, dim3(static_cast<unsigned>(workgroups)), dim3(kFrontThreads), 0, stream,
                     reinterpret_cast<uint32_t*>(ws.keys_in), ws.values_in, reinterpret_cast<uint32_t*>(ws.keys_out),
                     ws.values_out, ws.tile_starts, lengths, open_tiles, open_state + 0, open_state + 1,
                     layout.depth_bits, scene.gaussian_ids, ws.values_out, ws.tile_ends);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("front fallback sort launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
