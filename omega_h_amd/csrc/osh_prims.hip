// Stream-ordered primitives: offset scans, compaction, fills, reductions and
// the stable tuple sort used when building meshes.
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstring>

#include "osh_runtime.hpp"

// Scratch of the scans that allocate nothing (offset_scans_noalloc): the
// tile status words of up to kNoAllocTiles tiles, then the tile ticket
// counter.  Each GPU holds its own copy (128 KiB) for the life of the
// process; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
static constexpr unsigned kNoAllocTiles = 16384;
static __device__ __align__(256) unsigned long long scratch[kNoAllocTiles + 1];

namespace osh {

std::size_t Pool::round_up(std::size_t bytes) {
  if (bytes <= 256) return 256;
  int lg = 63 - __builtin_clzll(bytes - 1);  // 2^lg < bytes <= 2^(lg+1)
  std::size_t step = std::size_t(1) << (lg - 1);
  return (bytes + step - 1) / step * step;
}

void* Pool::get(std::size_t bytes) {
  std::size_t size = round_up(bytes);
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.find(size);
    if (it != free_.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      cached_ -= size;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, size) != hipSuccess) {
    (void)hipGetLastError();
    release_all();
    OSH_HIP(hipMalloc(&p, size));
  }
  return p;
}

void Pool::put(void* p, std::size_t bytes) {
  std::size_t size = round_up(bytes);
  {
    std::lock_guard<std::mutex> g(mu_);
    if (!closed_ && cached_ + size <= cap_) {
      free_[size].push_back(p);
      cached_ += size;
      return;
    }
  }
  (void)hipFree(p);
}

void Pool::release_all() {
  std::unordered_map<std::size_t, std::vector<void*>> blocks;
  {
    std::lock_guard<std::mutex> g(mu_);
    blocks.swap(free_);
    cached_ = 0;
  }
  for (auto& kv : blocks)
    for (void* p : kv.second) (void)hipFree(p);
}

Ctx::Ctx() : pool(std::make_shared<Pool>()) {
  OSH_HIP(hipGetDevice(&device));
  OSH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  OSH_HIP(hipHostMalloc(&pinned, 8 * kSlots + 64, hipHostMallocDefault));
  std::memset(pinned, 0, 8 * kSlots + 64);
  OSH_HIP(hipMalloc(&partials, 1024 * sizeof(double)));
  OSH_HIP(hipMalloc((void**)&tile_counter, sizeof(unsigned)));
  unsigned zero = 0;
  OSH_HIP(hipMemcpy(tile_counter, &zero, sizeof(zero), hipMemcpyHostToDevice));
  noalloc_setup(device);
  profile.from_environment();
}

Ctx::~Ctx() {
  if (profile.print_at_close) {
    std::string r = profile.report(profile.chop, profile.percent, profile_now() - profile.origin);
    std::fputs(r.c_str(), stdout);
    std::fflush(stdout);
  }
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (scan_status) (void)hipFree(scan_status);
  if (tile_counter) (void)hipFree(tile_counter);
  if (pool) pool->close();
  if (partials) (void)hipFree(partials);
  if (pinned) (void)hipHostFree(pinned);
}

void range_begin(Ctx& c, char const* name) {
  if (c.profile.flags & kProfileMarkers) roctxRangePushA(name);
  if (c.profile.flags & kProfileTree) c.profile.start(name, profile_now());
}

void range_end(Ctx& c) {
  // a failed wait is left for the next sync() to report: a range may end
  // while an error unwinds the stack
  if (c.profile.flags & kProfileSync) (void)hipStreamSynchronize(c.stream);
  if (c.profile.flags & kProfileTree) c.profile.stop(profile_now());
  if (c.profile.flags & kProfileMarkers) roctxRangePop();
}

void Ctx::sync() {
  OSH_HIP(hipStreamSynchronize(stream));
  int* err = error_word();
  if (*err != kNoError) {
    int code = *err;
    *err = kNoError;
    if (code == kReflectDownNoMatch)
      throw Error("reflect_down: entity without a unique match");
    if (code == kScanTimeout) throw Error("scan: a tile waited too long for its predecessor");
    if (code == kScanTicket) throw Error("scan: tile tickets out of step");
    if (code == kKeyWithoutUp) throw Error("modify: a key entity has no adjacent entity above it");
    if (code == kNonManifoldSide) throw Error("a side has more than two elements");
    if (code == kBrokenSwapLoop)
      throw Error("swap: the elements around an edge do not close a loop");
    if (code == kBadPart) throw Error("split: a vertex without a valid part number");
    if (code == kFlipNormals)
      throw Error("prevent_coarsen_flip: too many exposed sides at a vertex, or no most normal "
                  "normal");
    if (code == kZeroSizeCavity)
      throw Error("conservation: an integral error in a cavity of zero size");
    if (code == kPolytopeOverflow)
      throw Error("conservation: a simplex intersection left the polytope buffer");
    if (code == kDegenerateEigen)
      throw Error("eigen: no eigenvector for a simple eigenvalue (degenerate metric)");
    throw Error("device error " + std::to_string(code));
  }
}

namespace {

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096 values per block
constexpr int kWave = 64;

__device__ __forceinline__ LO wave_inclusive_scan(LO v) {
  int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    LO u = __shfl_up(v, d, kWave);
    if (lane >= d) v += u;
  }
  return v;
}

__device__ __forceinline__ LO wave_sum(LO v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ LO block_exclusive_scan(LO v, LO* total) {
  __shared__ LO wave_sums[kBlock / kWave];
  int lane = threadIdx.x & (kWave - 1);
  int wave = threadIdx.x / kWave;
  LO inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) wave_sums[wave] = inc;
  __syncthreads();
  LO before = 0;
  LO sum = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    LO s = wave_sums[w];
    if (w < wave) before += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + inc - v;
}

template <class T>
__global__ void fill_kernel(LO n, T* out, T value) {
  LO i = thread_index();
  if (i < n) out[i] = value;
}

// Single-pass exclusive scan with decoupled look-back.  Each workgroup of a
// multi-tile scan takes the next tile ticket from a counter (so every tile it
// waits on belongs to a workgroup that is already running), scans its 4096
// values and publishes its aggregate; its first wave then reads the words of
// the 64 preceding tiles at once, one per lane, and adds the aggregates back
// to the nearest inclusive prefix, moving the window back 64 tiles at a time
// until it meets one.  A status word is [epoch:24 | flag:8 | value:32]; the
// epoch changes every scan, so words left by earlier scans read as "not yet
// published" and no clearing pass is needed.  The word carries the value
// itself, so relaxed agent-scope atomics suffice.  A wait that outlasts
// kSpinLimit polls (it cannot happen on a correct schedule) stops, sets the
// context's error word and lets every workgroup finish; the next stream sync
// raises it.  Values pass through LDS so that global loads and stores are
// coalesced while each thread scans 16 consecutive values; LDS index
// j + j/16 keeps both access patterns free of bank conflicts.
constexpr unsigned kFlagAggregate = 1u;
constexpr unsigned kFlagInclusive = 2u;
constexpr unsigned kSpinLimit = 1u << 20;

__device__ __forceinline__ unsigned long long status_word(unsigned epoch, unsigned flag, LO v) {
  return (static_cast<unsigned long long>(epoch) << 40) |
         (static_cast<unsigned long long>(flag) << 32) | static_cast<unsigned>(v);
}

// the exclusive prefix of `tile` from its predecessors' status words; run by
// the 64 lanes of one wave, returns the same value in every lane
__device__ __forceinline__ LO look_back(unsigned long long* status, unsigned tile,
    unsigned epoch, int* error) {
  int lane = threadIdx.x & (kWave - 1);
  LO prefix = 0;
  long long window = static_cast<long long>(tile) - 1;  // the tile lane 0 reads
  unsigned spins = 0;
  for (;;) {
    long long j = window - lane;
    unsigned long long w = status_word(epoch, kFlagInclusive, 0);  // before tile 0
    if (j >= 0) w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned flag = unsigned(w >> 32) & 0xffu;
    bool ready = unsigned(w >> 40) == epoch && flag != 0;
    if (__ballot(!ready) != 0ull) {
      if (++spins > kSpinLimit) {
        if (lane == 0) *error = kScanTimeout;
        return prefix;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    LO v = LO(unsigned(w));
    unsigned long long inclusive = __ballot(flag == kFlagInclusive);
    if (inclusive != 0ull) {
      int nearest = __ffsll(static_cast<long long>(inclusive)) - 1;
      if (lane > nearest) v = 0;
      return prefix + wave_sum(v);
    }
    prefix += wave_sum(v);
    window -= kWave;
  }
}

// up to kMaxScanJobs independent scans share one launch: job j owns the
// tiles [first[j], first[j + 1]) of the launch and the same range of status
// words, so its look-back never leaves its own tiles
struct ScanBatch {
  void const* in[kMaxScanJobs];
  LO* out[kMaxScanJobs];
  LO* total[kMaxScanJobs];
  LO n[kMaxScanJobs];
  int kind[kMaxScanJobs];
  unsigned first[kMaxScanJobs + 1];
  int count;
};

__device__ __forceinline__ LO scan_input(void const* in, int kind, LO i) {
  if (kind == kScanLO) return static_cast<LO const*>(in)[i];
  I8 x = static_cast<I8 const*>(in)[i];
  return kind == kScanBytes ? LO(x) : LO(x != 0);
}

// without tickets (every job one tile) a workgroup's tile is its index
__global__ void __launch_bounds__(kBlock) scan_lookback(ScanBatch b,
    unsigned long long* status, unsigned* tile_counter, unsigned ticket_base, int tickets,
    unsigned epoch, int* error) {
  __shared__ LO s_vals[kScanTile + kScanTile / 16];
  __shared__ unsigned s_tile;
  __shared__ LO s_prefix;
  if (tickets) {
    if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u) - ticket_base;
    __syncthreads();
  }
  unsigned t = tickets ? s_tile : blockIdx.x;
  if (t >= gridDim.x) {  // the ticket base is out of step: refuse to write
    if (threadIdx.x == 0) *error = kScanTicket;
    return;
  }
  int job = 0;
  while (job + 1 < b.count && t >= b.first[job + 1]) ++job;
  unsigned tile = t - b.first[job];
  unsigned ntiles = b.first[job + 1] - b.first[job];
  LO n = b.n[job];
  int kind = b.kind[job];
  void const* in = b.in[job];
  LO* out = b.out[job];
  unsigned long long* st = status + b.first[job];
  LO tile_base = LO(tile) * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int j = k * kBlock + int(threadIdx.x);
    LO i = tile_base + j;
    s_vals[j + (j >> 4)] = (i < n) ? scan_input(in, kind, i) : 0;
  }
  __syncthreads();
  int row = int(threadIdx.x) * (kScanItems + 1);  // padded index of value 16 * threadIdx.x
  LO vals[kScanItems];
  LO acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    acc += s_vals[row + k];
    vals[k] = acc;
  }
  LO total;
  LO excl = block_exclusive_scan(acc, &total);
  if (threadIdx.x < kWave) {
    LO prefix = 0;
    if (tile != 0) {
      if (threadIdx.x == 0)
        __hip_atomic_store(&st[tile], status_word(epoch, kFlagAggregate, total),
            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      prefix = look_back(st, tile, epoch, error);
    }
    if (threadIdx.x == 0) {
      if (ntiles > 1)
        __hip_atomic_store(&st[tile], status_word(epoch, kFlagInclusive, prefix + total),
            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_prefix = prefix;
      if (tile == 0) out[0] = 0;
      if (b.total[job] && tile == ntiles - 1) *b.total[job] = prefix + total;
    }
  }
  __syncthreads();
  LO prefix = s_prefix + excl;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s_vals[row + k] = prefix + vals[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int j = k * kBlock + int(threadIdx.x);
    LO i = tile_base + j;
    if (i < n) out[1 + i] = s_vals[j + (j >> 4)];
  }
}

__global__ void zero_u64(LO n, unsigned long long* a) {
  LO i = thread_index();
  if (i < n) a[i] = 0ull;
}

// the context's tile status words, grown (and cleared) when a scan needs
// more tiles than they hold; a new epoch per scan, with a clearing pass
// before the 24-bit epoch field would overflow (epochs run 1 .. 2^24 - 1;
// cleared words carry epoch 0)
unsigned long long* scan_status_words(Ctx& c, LO ntiles, unsigned* epoch) {
  if (ntiles > c.scan_status_cap || c.scan_epoch >= Ctx::kMaxScanEpoch) {
    LO cap = ntiles > c.scan_status_cap ? (ntiles > 2 * c.scan_status_cap ? ntiles
                                                                          : 2 * c.scan_status_cap)
                                        : c.scan_status_cap;
    if (cap != c.scan_status_cap) {
      if (c.scan_status) OSH_HIP(hipStreamSynchronize(c.stream));
      if (c.scan_status) (void)hipFree(c.scan_status);
      c.scan_status = nullptr;
      OSH_HIP(hipMalloc(&c.scan_status, sizeof(unsigned long long) * std::size_t(cap)));
      c.scan_status_cap = cap;
    }
    OSH_LAUNCH(zero_u64, cap, c, static_cast<unsigned long long*>(c.scan_status));
    c.scan_epoch = 0;
  }
  *epoch = ++c.scan_epoch;
  return static_cast<unsigned long long*>(c.scan_status);
}

}  // namespace

void offset_scans(Ctx& c, ScanJob const* jobs, int njobs) {
  OSH_CHECK(njobs >= 1 && njobs <= kMaxScanJobs, "offset_scans: 1 to 12 jobs per launch");
  ScanBatch b{};
  unsigned tiles = 0;
  bool multi = false;
  for (int j = 0; j < njobs; ++j) {
    OSH_CHECK(jobs[j].n >= 0 && jobs[j].out, "offset_scans: bad job");
    b.in[j] = jobs[j].in;
    b.out[j] = jobs[j].out;
    b.total[j] = jobs[j].total;
    b.n[j] = jobs[j].n;
    b.kind[j] = jobs[j].kind;
    b.first[j] = tiles;
    // an empty job still takes one tile, which writes its 0 and its total
    unsigned nt = jobs[j].n == 0 ? 1u : unsigned((std::int64_t(jobs[j].n) + kScanTile - 1) / kScanTile);
    multi = multi || nt > 1;
    tiles += nt;
  }
  b.first[njobs] = tiles;
  b.count = njobs;
  unsigned epoch = 0;
  unsigned long long* status = nullptr;
  unsigned ticket_base = c.tile_tickets;
  if (multi) {
    status = scan_status_words(c, LO(tiles), &epoch);
    c.tile_tickets += tiles;
  }
  hipLaunchKernelGGL(scan_lookback, dim3(tiles), dim3(kBlock), 0, c.stream, b, status,
      c.tile_counter, ticket_base, multi ? 1 : 0, epoch, c.error_word());
  OSH_HIP(hipGetLastError());
}

namespace {

// per device: the address of its `scratch`, and the epoch and ticket count
// of the scans made in it so far
struct NoAllocDevice {
  unsigned long long* words = nullptr;
  unsigned epoch = 0;
  unsigned tickets = 0;
  bool cleared = false;
};
constexpr int kMaxNoAllocDevices = 64;
std::mutex noalloc_mu;
NoAllocDevice noalloc_devices[kMaxNoAllocDevices];

}  // namespace

void noalloc_setup(int device) {
  OSH_CHECK(device >= 0 && device < kMaxNoAllocDevices, "device number out of range");
  std::lock_guard<std::mutex> g(noalloc_mu);
  NoAllocDevice& d = noalloc_devices[device];
  if (d.words) return;
  void* p = nullptr;
  OSH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  OSH_CHECK(p != nullptr, "no device address for the scan scratch");
  d.words = static_cast<unsigned long long*>(p);
}

void offset_scans_noalloc(Ctx& c, ScanJob const* jobs, int njobs) {
  OSH_CHECK(njobs >= 1 && njobs <= kMaxScanJobs, "offset_scans: 1 to 12 jobs per launch");
  ScanBatch b{};
  std::int64_t tiles = 0;
  bool multi = false;
  for (int j = 0; j < njobs; ++j) {
    OSH_CHECK(jobs[j].n >= 0 && jobs[j].out, "offset_scans: bad job");
    b.in[j] = jobs[j].in;
    b.out[j] = jobs[j].out;
    b.total[j] = jobs[j].total;
    b.n[j] = jobs[j].n;
    b.kind[j] = jobs[j].kind;
    b.first[j] = unsigned(tiles);
    std::int64_t nt = jobs[j].n == 0 ? 1 : (std::int64_t(jobs[j].n) + kScanTile - 1) / kScanTile;
    multi = multi || nt > 1;
    tiles += nt;
    OSH_CHECK(tiles <= std::int64_t(kNoAllocTiles),
        "offset_scans_noalloc: more than 16384 tiles of 4096 values in one call");
  }
  b.first[njobs] = unsigned(tiles);
  b.count = njobs;
  unsigned epoch = 0, ticket_base = 0;
  unsigned long long* words = nullptr;
  {
    std::lock_guard<std::mutex> g(noalloc_mu);
    OSH_CHECK(c.device >= 0 && c.device < kMaxNoAllocDevices, "device number out of range");
    NoAllocDevice& d = noalloc_devices[c.device];
    OSH_CHECK(d.words, "the scan scratch of this device is not set up");
    words = d.words;
    if (multi) {
      // status words left by earlier scans read as unpublished under a new
      // epoch; before the 24-bit epoch field wraps (and before the first
      // use) a fill kernel clears them
      if (!d.cleared || d.epoch >= Ctx::kMaxScanEpoch) {
        hipLaunchKernelGGL(zero_u64, dim3(nblocks(LO(kNoAllocTiles))), dim3(kBlock), 0,
            c.stream, LO(kNoAllocTiles), words);
        OSH_HIP(hipGetLastError());
        d.epoch = 0;
        d.cleared = true;
      }
      epoch = ++d.epoch;
      ticket_base = d.tickets;
      d.tickets += unsigned(tiles);
    }
  }
  hipLaunchKernelGGL(scan_lookback, dim3(unsigned(tiles)), dim3(kBlock), 0, c.stream, b,
      multi ? words : nullptr, reinterpret_cast<unsigned*>(words + kNoAllocTiles), ticket_base,
      multi ? 1 : 0, epoch, c.error_word());
  OSH_HIP(hipGetLastError());
}

namespace {

// the grand total lands in *total_out (device-visible) when given
template <class T, bool NONZERO>
LOs offset_scan_t(Ctx& c, DArray<T> const& counts, LO* total_out) {
  LOs out(c, counts.size() + 1);
  ScanJob job{counts.data(), sizeof(T) == 1 ? (NONZERO ? kScanNonzero : kScanBytes) : kScanLO,
      counts.size(), out.data(), total_out};
  offset_scans(c, &job, 1);
  return out;
}

__global__ void scatter_marked(LO n, I8 const* marks, LO const* offsets, LO* out) {
  LO i = thread_index();
  if (i < n && marks[i]) out[offsets[i]] = i;
}

__global__ void iota_kernel(LO n, LO* out, LO start, LO stride) {
  LO i = thread_index();
  if (i < n) out[i] = start + i * stride;
}

enum ReduceOp { kMin, kMax, kSum };

template <class T, int Op>
__device__ __forceinline__ T combine(T a, T b) {
  if (Op == kMin) return (b < a) ? b : a;
  if (Op == kMax) return (a < b) ? b : a;
  return a + b;
}

// grid-stride partial reductions into out[blockIdx.x]; a one-block launch
// over the partials then writes the result into host-visible memory (the
// kernel boundary orders the two, so no device-wide fences are needed)
template <class T, int Op>
__global__ void __launch_bounds__(kBlock) reduce_blocks(LO n, T const* in, T* out, T init) {
  __shared__ T buf[kBlock];
  T acc = init;
  for (LO i = LO(blockIdx.x) * kBlock + LO(threadIdx.x); i < n; i += LO(gridDim.x) * kBlock)
    acc = combine<T, Op>(acc, in[i]);
  buf[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) buf[threadIdx.x] = combine<T, Op>(buf[threadIdx.x], buf[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = buf[0];
}

// the result lands in `result` (device-visible) once the stream reaches it
template <class T, int Op>
void reduce_to(Ctx& c, DArray<T> const& a, T init, T* result) {
  LO n = a.size();
  if (n == 0) {
    hipLaunchKernelGGL((fill_kernel<T>), dim3(1), dim3(kBlock), 0, c.stream, 1, result, init);
    OSH_HIP(hipGetLastError());
    return;
  }
  LO nb = LO(nblocks(n, kBlock * 4));
  if (nb > 1024) nb = 1024;
  if (nb == 1) {
    hipLaunchKernelGGL((reduce_blocks<T, Op>), dim3(1), dim3(kBlock), 0, c.stream, n, a.data(),
        result, init);
  } else {
    T* part = static_cast<T*>(c.partials);
    hipLaunchKernelGGL((reduce_blocks<T, Op>), dim3(nb), dim3(kBlock), 0, c.stream, n, a.data(),
        part, init);
    OSH_HIP(hipGetLastError());
    hipLaunchKernelGGL((reduce_blocks<T, Op>), dim3(1), dim3(kBlock), 0, c.stream, nb, part,
        result, init);
  }
  OSH_HIP(hipGetLastError());
}

template <class T, int Op>
T reduce(Ctx& c, DArray<T> const& a, T init) {
  if (a.size() == 0) return init;
  reduce_to<T, Op>(c, a, init, static_cast<T*>(c.pinned));
  return pinned_value<T>(c);
}

__global__ void gather_key_column(LO n, I64 const* keys, LO const* perm, int width,
    int col, I64* out) {
  LO i = thread_index();
  if (i < n) out[i] = keys[std::int64_t(perm[i]) * width + col];
}

}  // namespace

LOs offset_scan(Ctx& c, LOs const& counts, LO* total) {
  LOs out = offset_scan_t<LO, false>(c, counts, total ? static_cast<LO*>(c.pinned) : nullptr);
  if (total) *total = pinned_value<LO>(c);
  return out;
}
LOs offset_scan_bytes(Ctx& c, Bytes const& counts, LO* total) {
  LOs out = offset_scan_t<I8, false>(c, counts, total ? static_cast<LO*>(c.pinned) : nullptr);
  if (total) *total = pinned_value<LO>(c);
  return out;
}

LOs collect_marked(Ctx& c, Bytes const& marks) { return collect_marked(c, marks, nullptr); }

void scatter_marked_list(Ctx& c, Bytes const& marks, LOs const& offsets, LOs& out) {
  OSH_LAUNCH(scatter_marked, marks.size(), c, marks.data(), offsets.data(), out.data());
}

LOs collect_marked(Ctx& c, Bytes const& marks, LOs* offsets_out) {
  LO n = marks.size();
  // the scan counts nonzero marks directly
  LOs offsets = offset_scan_t<I8, true>(c, marks, static_cast<LO*>(c.pinned));
  LO count = pinned_value<LO>(c);
  LOs out(c, count);
  OSH_LAUNCH(scatter_marked, n, c, marks.data(), offsets.data(), out.data());
  if (offsets_out) *offsets_out = offsets;
  return out;
}

LOs iota(Ctx& c, LO n, LO start, LO stride) {
  LOs out(c, n);
  OSH_LAUNCH(iota_kernel, n, c, out.data(), start, stride);
  return out;
}

LOs fill_lo(Ctx& c, LO n, LO value) {
  LOs out(c, n);
  OSH_LAUNCH(fill_kernel<LO>, n, c, out.data(), value);
  return out;
}

Bytes fill_i8(Ctx& c, LO n, I8 value) {
  Bytes out(c, n);
  OSH_LAUNCH(fill_kernel<I8>, n, c, out.data(), value);
  return out;
}

Reals fill_real(Ctx& c, LO n, Real value) {
  Reals out(c, n);
  OSH_LAUNCH(fill_kernel<Real>, n, c, out.data(), value);
  return out;
}

I8 max_i8(Ctx& c, Bytes const& a) { return reduce<I8, kMax>(c, a, I8(-128)); }
Real min_real(Ctx& c, Reals const& a) { return reduce<Real, kMin>(c, a, 1.79769313486231570815e+308); }
void min_real_async(Ctx& c, Reals const& a, int slot) {
  OSH_CHECK(slot >= 0 && slot < Ctx::kSlots, "pinned slot out of range");
  reduce_to<Real, kMin>(c, a, 1.79769313486231570815e+308, slot_ptr<Real>(c, slot));
}
Real max_real(Ctx& c, Reals const& a) { return reduce<Real, kMax>(c, a, -1.79769313486231570815e+308); }
LO sum_lo(Ctx& c, LOs const& a) { return reduce<LO, kSum>(c, a, 0); }

namespace {

// ---- stable LSD radix sort of (64-bit key, 32-bit value) pairs
//
// Keys are signed; flipping the sign bit orders them as unsigned words.  A
// pass sorts by one 8-bit digit: a count kernel histograms the digits of each
// 4096-item tile into counts[digit * ntiles + tile], one exclusive scan of
// those counts (digit-major) gives every (digit, tile) its first output
// position, and a scatter kernel places each item at that position plus its
// rank among the tile's items with the same digit that precede it, so equal
// digits keep their order.  A tile's items are read striped (item r * 256 +
// t is thread t's r-th), so that loads coalesce and the tile order is
// "round, then thread": the rank is the digit's count in earlier rounds (a
// running LDS counter) plus its count in earlier waves of this round (LDS)
// plus the lanes below in the same wave with the same digit (eight ballots).
// Passes run only for the digits in which the keys differ.
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kSortBins = 256;
constexpr int kSortWaves = kBlock / kWave;

__device__ __forceinline__ unsigned long long sort_word(I64 k) {
  return static_cast<unsigned long long>(k) ^ (1ull << 63);
}

// the OR and the AND of every key word (their XOR shows the bits that
// differ); each wave combines its lanes first, so the two words see one
// atomic per wave
__global__ void sort_bits_kernel(LO n, I64 const* keys, unsigned long long* or_and) {
  LO i = thread_index();
  bool valid = i < n;
  unsigned long long w = valid ? sort_word(keys[i]) : 0ull;
  unsigned long long o = w, a = valid ? w : ~0ull;
  for (int d = kWave / 2; d > 0; d >>= 1) {
    o |= __shfl_xor(o, d, kWave);
    a &= __shfl_xor(a, d, kWave);
  }
  // a wave past the end holds o = 0 and a = ~0, which change nothing
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicOr(or_and, o);
    atomicAnd(or_and + 1, a);
  }
}

__global__ void __launch_bounds__(kBlock) sort_count(LO n, I64 const* keys, int shift,
    LO ntiles, LO* counts) {
  __shared__ LO hist[kSortBins];
  hist[threadIdx.x] = 0;
  __syncthreads();
  LO base = LO(blockIdx.x) * kSortTile;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    LO i = base + r * kBlock + LO(threadIdx.x);
    if (i < n) atomicAdd(&hist[int(sort_word(keys[i]) >> shift) & (kSortBins - 1)], 1);
  }
  __syncthreads();
  counts[LO(threadIdx.x) * ntiles + LO(blockIdx.x)] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) sort_scatter(LO n, I64 const* keys, LO const* vals,
    int shift, LO ntiles, LO const* offsets, I64* out_keys, LO* out_vals) {
  __shared__ LO running[kSortBins];
  __shared__ LO wave_counts[kSortWaves][kSortBins];
  int t = int(threadIdx.x);
  int lane = t & (kWave - 1);
  int wave = t / kWave;
  running[t] = offsets[LO(t) * ntiles + LO(blockIdx.x)];
  for (int w = 0; w < kSortWaves; ++w) wave_counts[w][t] = 0;
  __syncthreads();
  LO base = LO(blockIdx.x) * kSortTile;
  unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  for (int r = 0; r < kSortItems; ++r) {
    LO i = base + r * kBlock + t;
    bool valid = i < n;
    I64 k = valid ? keys[i] : 0;
    int d = valid ? int(sort_word(k) >> shift) & (kSortBins - 1) : 0;
    // the lanes of this wave holding the same digit
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned long long bit = __ballot(valid && ((d >> b) & 1));
      same &= ((d >> b) & 1) ? bit : ~bit;
    }
    int rank_in_wave = __popcll(same & below);
    bool leader = valid && rank_in_wave == 0;
    if (leader) wave_counts[wave][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      LO pos = running[d] + rank_in_wave;
      for (int w = 0; w < wave; ++w) pos += wave_counts[w][d];
      out_keys[pos] = k;
      out_vals[pos] = vals[i];
    }
    __syncthreads();
    // advance the running offsets by this round's counts and clear them
    LO add = 0;
    for (int w = 0; w < kSortWaves; ++w) {
      add += wave_counts[w][t];
      wave_counts[w][t] = 0;
    }
    running[t] += add;
    __syncthreads();
  }
}

// sorts (keys, vals) by key in place (stable), using the second pair of
// buffers; returns with the result in the first pair
void radix_sort_pairs(Ctx& c, DArray<I64>& keys, LOs& vals, DArray<I64>& keys_tmp,
    LOs& vals_tmp) {
  LO n = keys.size();
  if (n <= 1) return;
  std::vector<unsigned long long> init = {0ull, ~0ull};
  DArray<unsigned long long> or_and = to_device(c, init);
  OSH_LAUNCH(sort_bits_kernel, n, c, keys.data(), or_and.data());
  std::vector<unsigned long long> oa = to_host(c, or_and);
  unsigned long long differ = oa[0] ^ oa[1];
  LO ntiles = LO((std::int64_t(n) + kSortTile - 1) / kSortTile);
  OSH_CHECK(std::int64_t(ntiles) * kSortBins < std::int64_t(0x7fffffff), "sort: too many items");
  LOs counts(c, ntiles * kSortBins), offsets(c, ntiles * kSortBins + 1);
  for (int shift = 0; shift < 64; shift += 8) {
    if (((differ >> shift) & 0xffull) == 0ull) continue;
    hipLaunchKernelGGL(sort_count, dim3(unsigned(ntiles)), dim3(kBlock), 0, c.stream, n,
        keys.data(), shift, ntiles, counts.data());
    OSH_HIP(hipGetLastError());
    ScanJob job{counts.data(), kScanLO, ntiles * kSortBins, offsets.data(), nullptr};
    offset_scans(c, &job, 1);
    hipLaunchKernelGGL(sort_scatter, dim3(unsigned(ntiles)), dim3(kBlock), 0, c.stream, n,
        keys.data(), vals.data(), shift, ntiles, offsets.data(), keys_tmp.data(),
        vals_tmp.data());
    OSH_HIP(hipGetLastError());
    std::swap(keys, keys_tmp);
    std::swap(vals, vals_tmp);
  }
}

}  // namespace

// LSD over tuple components, each a stable radix sort of (key, position):
// the last component first, so earlier components take precedence.
LOs sort_by_keys(Ctx& c, DArray<I64> const& keys, int width) {
  LO n = keys.size() / width;
  LOs perm = iota(c, n);
  if (n <= 1) return perm;
  DArray<I64> col(c, n), col_tmp(c, n);
  LOs perm_tmp(c, n);
  for (int k = width - 1; k >= 0; --k) {
    OSH_LAUNCH(gather_key_column, n, c, keys.data(), perm.data(), width, k, col.data());
    radix_sort_pairs(c, col, perm, col_tmp, perm_tmp);
  }
  return perm;
}

}  // namespace osh
