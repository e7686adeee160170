// Stream-ordered primitives: offset scans, compaction, fills, reductions and
// the stable tuple sort used when building meshes.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "osh_runtime.hpp"

namespace osh {

Ctx::Ctx() {
  OSH_HIP(hipGetDevice(&device));
  OSH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  OSH_HIP(hipHostMalloc(&pinned, 64, hipHostMallocDefault));
}

Ctx::~Ctx() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (pinned) (void)hipHostFree(pinned);
}

void Ctx::sync() { OSH_HIP(hipStreamSynchronize(stream)); }

namespace {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048 values per block
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

// out[1 + i] = inclusive prefix within the tile; bsum[block] = tile total
template <class T>
__global__ void __launch_bounds__(kBlock) scan_tiles(LO n, T const* in, LO* out, LO* bsum) {
  LO base = LO(blockIdx.x) * kScanTile + LO(threadIdx.x) * kScanItems;
  LO vals[kScanItems];
  LO acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    LO i = base + k;
    LO v = (i < n) ? LO(in[i]) : 0;
    acc += v;
    vals[k] = acc;
  }
  LO total;
  LO excl = block_exclusive_scan(acc, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    LO i = base + k;
    if (i < n) out[1 + i] = excl + vals[k];
  }
  if (threadIdx.x == 0) {
    bsum[blockIdx.x] = total;
    if (blockIdx.x == 0) out[0] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) add_tile_offsets(LO n, LO* out, LO const* boff) {
  LO i = LO(blockIdx.x) * kScanTile + LO(threadIdx.x);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k, i += kBlock) {
    if (i < n) out[1 + i] += boff[blockIdx.x];
  }
}

template <class T>
LOs offset_scan_t(Ctx& c, DArray<T> const& counts) {
  LO n = counts.size();
  LOs out(c, n + 1);
  if (n == 0) {
    LO zero = 0;
    upload(c, out, &zero, 1);
    return out;
  }
  LO nb = LO((std::int64_t(n) + kScanTile - 1) / kScanTile);
  LOs bsum(c, nb);
  hipLaunchKernelGGL(scan_tiles<T>, dim3(nb), dim3(kBlock), 0, c.stream, n,
      counts.data(), out.data(), bsum.data());
  OSH_HIP(hipGetLastError());
  if (nb > 1) {
    LOs boff = offset_scan_t<LO>(c, bsum);
    hipLaunchKernelGGL(add_tile_offsets, dim3(nb), dim3(kBlock), 0, c.stream,
        n, out.data(), boff.data());
    OSH_HIP(hipGetLastError());
  }
  return out;
}

__global__ void nonzero_flags(LO n, I8 const* marks, I8* flags) {
  LO i = thread_index();
  if (i < n) flags[i] = marks[i] != 0;
}

__global__ void scatter_marked(LO n, I8 const* marks, LO const* offsets, LO* out) {
  LO i = thread_index();
  if (i < n && marks[i]) out[offsets[i]] = i;
}

__global__ void iota_kernel(LO n, LO* out, LO start, LO stride) {
  LO i = thread_index();
  if (i < n) out[i] = start + i * stride;
}

template <class T>
__global__ void fill_kernel(LO n, T* out, T value) {
  LO i = thread_index();
  if (i < n) out[i] = value;
}

enum ReduceOp { kMin, kMax, kSum };

template <class T, int Op>
__device__ __forceinline__ T combine(T a, T b) {
  if (Op == kMin) return (b < a) ? b : a;
  if (Op == kMax) return (a < b) ? b : a;
  return a + b;
}

template <class T, int Op>
__global__ void __launch_bounds__(kBlock) reduce_blocks(LO n, T const* in, T* out, T init) {
  __shared__ T buf[kBlock];
  T acc = init;
  for (LO i = LO(blockIdx.x) * kBlock + LO(threadIdx.x); i < n;
       i += LO(gridDim.x) * kBlock)
    acc = combine<T, Op>(acc, in[i]);
  buf[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) buf[threadIdx.x] = combine<T, Op>(buf[threadIdx.x], buf[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = buf[0];
}

template <class T, int Op>
T reduce(Ctx& c, DArray<T> const& a, T init) {
  LO n = a.size();
  if (n == 0) return init;
  LO nb = LO(nblocks(n));
  if (nb > 1024) nb = 1024;
  DArray<T> part(c, nb);
  hipLaunchKernelGGL((reduce_blocks<T, Op>), dim3(nb), dim3(kBlock), 0, c.stream, n,
      a.data(), part.data(), init);
  OSH_HIP(hipGetLastError());
  DArray<T> fin(c, 1);
  hipLaunchKernelGGL((reduce_blocks<T, Op>), dim3(1), dim3(kBlock), 0, c.stream, nb,
      part.data(), fin.data(), init);
  OSH_HIP(hipGetLastError());
  return get_value(c, fin.data());
}

__global__ void gather_key_column(LO n, I64 const* keys, LO const* perm, int width,
    int col, I64* out) {
  LO i = thread_index();
  if (i < n) out[i] = keys[std::int64_t(perm[i]) * width + col];
}

}  // namespace

LOs offset_scan(Ctx& c, LOs const& counts) { return offset_scan_t<LO>(c, counts); }
LOs offset_scan_bytes(Ctx& c, Bytes const& counts) { return offset_scan_t<I8>(c, counts); }

LOs collect_marked(Ctx& c, Bytes const& marks) {
  LO n = marks.size();
  Bytes flags(c, n);
  OSH_LAUNCH(nonzero_flags, n, c, marks.data(), flags.data());
  LOs offsets = offset_scan_bytes(c, flags);
  LO count = last(c, offsets);
  LOs out(c, count);
  OSH_LAUNCH(scatter_marked, n, c, marks.data(), offsets.data(), out.data());
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
Real max_real(Ctx& c, Reals const& a) { return reduce<Real, kMax>(c, a, -1.79769313486231570815e+308); }
LO sum_lo(Ctx& c, LOs const& a) { return reduce<LO, kSum>(c, a, 0); }

// LSD over tuple components, each a stable radix sort of (key, position):
// the last component first, so earlier components take precedence.
LOs sort_by_keys(Ctx& c, DArray<I64> const& keys, int width) {
  LO n = keys.size() / width;
  LOs perm = iota(c, n);
  if (n <= 1) return perm;
  DArray<I64> col(c, n), col_out(c, n);
  LOs perm_out(c, n);
  for (int k = width - 1; k >= 0; --k) {
    OSH_LAUNCH(gather_key_column, n, c, keys.data(), perm.data(), width, k, col.data());
    std::size_t temp_bytes = 0;
    OSH_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, col.data(), col_out.data(),
        perm.data(), perm_out.data(), unsigned(n), 0, 64, c.stream));
    DArray<unsigned char> temp(c, LO(temp_bytes + 1));
    OSH_HIP(rocprim::radix_sort_pairs(temp.data(), temp_bytes, col.data(), col_out.data(),
        perm.data(), perm_out.data(), unsigned(n), 0, 64, c.stream));
    std::swap(perm, perm_out);
  }
  return perm;
}

}  // namespace osh
