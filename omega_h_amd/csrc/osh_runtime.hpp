// Device runtime for the MI355X adapt() path: one HIP stream per context,
// stream-ordered device arrays, launch helpers and error reporting.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
#include <string>
#include <vector>

#include "osh_profile.hpp"

namespace osh {

using LO = std::int32_t;
using I8 = std::int8_t;
using I64 = std::int64_t;
using Real = double;

struct Error : std::runtime_error {
  explicit Error(std::string const& m) : std::runtime_error(m) {}
};

#define OSH_HIP(call)                                                        \
  do {                                                                       \
    hipError_t osh_e_ = (call);                                              \
    if (osh_e_ != hipSuccess)                                                \
      throw ::osh::Error(std::string("HIP error ") +                         \
                         hipGetErrorString(osh_e_) + " at " __FILE__ ":" +   \
                         std::to_string(__LINE__) + " in " #call);            \
  } while (0)

#define OSH_CHECK(cond, msg)                                                 \
  do {                                                                       \
    if (!(cond)) throw ::osh::Error(std::string("check failed: ") + (msg));  \
  } while (0)

// Caching device allocator of one context.  Every array of a context is
// used only on that context's stream, so a block returned by a dead array can
// be handed to the next request at once: the work that used it is already
// ordered before any work that the new owner enqueues.  Sizes are rounded up
// to classes at most 25% apart; blocks stay cached (up to a cap) until the
// context closes, so a steady adapt loop makes no runtime allocation calls.
class Pool {
 public:
  Pool() = default;
  Pool(Pool const&) = delete;
  Pool& operator=(Pool const&) = delete;
  ~Pool() { release_all(); }
  void* get(std::size_t bytes);
  void put(void* p, std::size_t bytes);
  // return every cached block to the runtime (the caller has drained the
  // stream); later puts free directly
  void release_all();
  void close() {
    release_all();
    std::lock_guard<std::mutex> g(mu_);
    closed_ = true;
  }
  static std::size_t round_up(std::size_t bytes);
  // most bytes kept cached; blocks returned beyond it go back to the runtime
  void set_cap(std::size_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    cap_ = bytes;
  }

 private:
  std::mutex mu_;
  std::unordered_map<std::size_t, std::vector<void*>> free_;
  std::size_t cached_ = 0;
  std::size_t cap_ = std::size_t(16) << 30;
  bool closed_ = false;
};

// Lane counts of the kernels that give one row of work several threads
// (powers of two, 1 to 64).  No setting changes a result, only how the work
// is spread over the device; omega_h_amd/autotune.py picks them for a
// workload by timing adapt() and checking that the mesh stays the same.
struct Tuning {
  int coarsen_lanes = 8;  // threads per collapse side (overshoot and quality sweeps)
  int side_lanes = 16;    // threads per swap product side searched for its match
  int choose_lanes = 8;   // threads per swap candidate choosing its triangulation
  int first_lanes = 8;    // threads per edge in the first independent-set round
};

// the lane count for `rows` rows: `lanes`, halved while rows x lanes would
// not fit a launch's 32-bit thread count (results do not depend on it)
inline int fit_lanes(LO rows, int lanes) {
  while (lanes > 1 && std::int64_t(rows) * lanes > std::int64_t(0x7fffffff)) lanes /= 2;
  return lanes;
}

inline int lanes_shift(int lanes) {
  int s = 0;
  while ((1 << s) < lanes) ++s;
  return s;
}

struct Ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  // pinned host words that kernels write results into: 8-byte slots
  // 0..kSlots-1 (totals, reductions) and one error word that kernels set
  // when they find a broken invariant; sync() raises it
  static constexpr int kSlots = 16;
  void* pinned = nullptr;
  int* error_word() const { return reinterpret_cast<int*>(static_cast<char*>(pinned) + 8 * kSlots); }
  // per-block partials of reductions
  void* partials = nullptr;
  // single-pass scans: tile status words (tagged with a per-scan epoch) and
  // the tile ticket counter, whose running total the host tracks
  void* scan_status = nullptr;
  LO scan_status_cap = 0;
  unsigned scan_epoch = 0;
  // the largest epoch a scan may use: status words keep it in 24 bits
  static constexpr unsigned kMaxScanEpoch = (1u << 24) - 1;
  unsigned* tile_counter = nullptr;
  unsigned tile_tickets = 0;
  Tuning tuning;
  // when set (by the batched driver), each pass marks here the parts of a
  // joined mesh that it changes
  int* part_changed = nullptr;
  // device copy of the edge-swap tables, uploaded on first use
  std::shared_ptr<void> swap_tables;
  std::shared_ptr<Pool> pool;
  // named ranges (off unless profiling flags are set)
  Profile profile;
  Ctx();
  ~Ctx();
  Ctx(Ctx const&) = delete;
  Ctx& operator=(Ctx const&) = delete;
  void sync();
};

void range_begin(Ctx& c, char const* name);
void range_end(Ctx& c);

// A named range from construction to end() or destruction (ScopedTimer,
// src/Omega_h_profile.hpp:217-226); ranges of one context nest, so they end
// in the reverse order of their starts
class Range {
 public:
  Range(Ctx& c, char const* name) : c_(c.profile.flags ? &c : nullptr) {
    if (c_) range_begin(*c_, name);
  }
  ~Range() { end(); }
  void end() {
    if (c_) range_end(*c_);
    c_ = nullptr;
  }
  Range(Range const&) = delete;
  Range& operator=(Range const&) = delete;

 private:
  Ctx* c_;
};

// reference-counted device buffer from the context's pool, used in stream
// order; a default-constructed array does not exist (like an empty Omega_h
// Read<T>)
template <class T>
class DArray {
 public:
  DArray() = default;
  DArray(Ctx& c, LO n) : n_(n), exists_(true) {
    OSH_CHECK(n >= 0, "negative array size");
    T* p = nullptr;
    std::size_t bytes = sizeof(T) * std::size_t(n);
    if (n > 0) p = static_cast<T*>(c.pool->get(bytes));
    std::shared_ptr<Pool> pool = c.pool;
    ptr_ = std::shared_ptr<T>(p, [pool, bytes](T* q) {
      if (q) pool->put(q, bytes);
    });
  }
  T* data() const { return ptr_.get(); }
  LO size() const { return n_; }
  bool exists() const { return exists_; }

 private:
  std::shared_ptr<T> ptr_;
  LO n_ = 0;
  bool exists_ = false;
};

using LOs = DArray<LO>;
using Bytes = DArray<I8>;
using Reals = DArray<Real>;

// `h` may be host memory or any device memory (the runtime tells them apart)
template <class T>
inline void upload(Ctx& c, DArray<T>& d, T const* h, LO n) {
  if (n) OSH_HIP(hipMemcpyAsync(d.data(), h, sizeof(T) * size_t(n), hipMemcpyDefault, c.stream));
}

template <class T>
inline DArray<T> to_device(Ctx& c, std::vector<T> const& h) {
  DArray<T> d(c, LO(h.size()));
  upload(c, d, h.data(), LO(h.size()));
  return d;
}

template <class T>
inline std::vector<T> to_host(Ctx& c, DArray<T> const& d) {
  std::vector<T> h(size_t(d.size()));
  if (d.size())
    OSH_HIP(hipMemcpyAsync(h.data(), d.data(), sizeof(T) * size_t(d.size()), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  return h;
}

// a value that a kernel wrote into a pinned slot; read_slot assumes the
// stream has been drained, pinned_value drains it first
template <class T>
inline T read_slot(Ctx& c, int slot) {
  T v;
  std::memcpy(&v, static_cast<char*>(c.pinned) + 8 * slot, sizeof(T));
  return v;
}
template <class T>
inline T pinned_value(Ctx& c, int slot = 0) {
  c.sync();
  return read_slot<T>(c, slot);
}
template <class T>
inline T* slot_ptr(Ctx& c, int slot) {
  return reinterpret_cast<T*>(static_cast<char*>(c.pinned) + 8 * slot);
}

// error codes kernels leave in Ctx::error_word()
enum DeviceError {
  kNoError = 0,
  kReflectDownNoMatch = 1,
  kScanTimeout = 2,
  kScanTicket = 3,
  kKeyWithoutUp = 4,
  kNonManifoldSide = 5,
  kBrokenSwapLoop = 6,
  kDegenerateEigen = 7,
  kBadPart = 8,
  kFlipNormals = 9,
  kZeroSizeCavity = 10,
  kPolytopeOverflow = 11
};

template <class T>
inline T get_value(Ctx& c, T const* dptr) {
  OSH_HIP(hipMemcpyAsync(c.pinned, dptr, sizeof(T), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  T v;
  std::memcpy(&v, c.pinned, sizeof(T));
  return v;
}

template <class T>
inline T last(Ctx& c, DArray<T> const& a) {
  return get_value(c, a.data() + a.size() - 1);
}

constexpr int kBlock = 256;

inline unsigned nblocks(LO n, int block = kBlock) {
  return unsigned((std::int64_t(n) + block - 1) / block);
}

#define OSH_LAUNCH(kernel, n, c, ...)                                          \
  do {                                                                         \
    ::osh::LO osh_n_ = (n);                                                         \
    if (osh_n_ > 0) {                                                          \
      hipLaunchKernelGGL(kernel, dim3(::osh::nblocks(osh_n_)),                 \
          dim3(::osh::kBlock), 0, (c).stream, osh_n_, __VA_ARGS__);            \
      OSH_HIP(hipGetLastError());                                              \
    }                                                                          \
  } while (0)

__device__ __forceinline__ LO thread_index() {
  return LO(blockIdx.x) * LO(blockDim.x) + LO(threadIdx.x);
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  For gather kernels over locality-ordered entities this
// renumbers the workgroups so that each XCD walks one contiguous range and
// neighbouring elements share that XCD's L2 (a bijection of [0, gridDim.x)).
constexpr unsigned kNumXcds = 8;
__device__ __forceinline__ LO xcd_thread_index() {
  unsigned nb = gridDim.x, b = blockIdx.x;
  unsigned per = nb / kNumXcds, rem = nb % kNumXcds;
  unsigned x = b % kNumXcds, k = b / kNumXcds;
  unsigned start = x * per + (x < rem ? x : rem);
  return LO(start + k) * LO(blockDim.x) + LO(threadIdx.x);
}

// ---- primitives (osh_prims.hip) ----
// exclusive scans (n + 1 outputs, out[0] = 0) of LO values, I8 values or I8
// nonzero flags; several share one launch through offset_scans
enum ScanKind { kScanLO = 0, kScanBytes = 1, kScanNonzero = 2 };
constexpr int kMaxScanJobs = 12;
struct ScanJob {
  void const* in;
  int kind;
  LO n;
  LO* out;    // n + 1 entries
  LO* total;  // device-visible grand total (a pinned slot), or nullptr
};
void offset_scans(Ctx& c, ScanJob const* jobs, int njobs);
// the same scans with no allocation: the tile status words and ticket
// counter live in a fixed per-device scratch array (at most 16384 tiles of
// 4096 values per call, all jobs together); noalloc_setup finds that
// array's address on a device (once, with the device current).  Calls that
// overlap on one GPU race for the scratch.
void noalloc_setup(int device);
void offset_scans_noalloc(Ctx& c, ScanJob const* jobs, int njobs);
// n+1 exclusive offsets; with `total`, also the sum (read back after a sync)
LOs offset_scan(Ctx& c, LOs const& counts, LO* total = nullptr);
LOs offset_scan_bytes(Ctx& c, Bytes const& counts, LO* total = nullptr);
LOs collect_marked(Ctx& c, Bytes const& marks);       // indices of nonzero
// out[offsets[i]] = i for nonzero marks[i], with offsets from a nonzero scan
void scatter_marked_list(Ctx& c, Bytes const& marks, LOs const& offsets, LOs& out);
// the same, also returning the n+1 offsets of the nonzero marks
LOs collect_marked(Ctx& c, Bytes const& marks, LOs* offsets_out);
LOs iota(Ctx& c, LO n, LO start = 0, LO stride = 1);
LOs fill_lo(Ctx& c, LO n, LO value);
Bytes fill_i8(Ctx& c, LO n, I8 value);
Reals fill_real(Ctx& c, LO n, Real value);
I8 max_i8(Ctx& c, Bytes const& a);
Real min_real(Ctx& c, Reals const& a);
// the same without a sync: the minimum lands in pinned slot `slot`
void min_real_async(Ctx& c, Reals const& a, int slot);
Real max_real(Ctx& c, Reals const& a);
LO sum_lo(Ctx& c, LOs const& a);
// stable sort of n tuples of `width` int64 keys; returns the permutation
LOs sort_by_keys(Ctx& c, DArray<I64> const& keys, int width);

}  // namespace osh
