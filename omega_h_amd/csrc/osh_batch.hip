// Many independent meshes adapted as one: the batched form of adapt().
//
// Omega_h's adapt() (src/Omega_h_adapt.cpp:272-288) takes one mesh.  Every
// decision inside a pass is local to a cavity, the independent set is a
// fixed point of a rule over star neighbours (src/Omega_h_indset_inline.hpp),
// and modify_ents numbers new entities in the order of the old ones
// (src/Omega_h_modify.cpp:141-517).  So for a mesh made of disjoint parts,
// with part k's entities numbered after those of parts < k in every
// dimension, one pass does to each part exactly what the same pass does to
// that part alone, and the parts stay contiguous blocks.  Passes that find
// nothing to do in a part leave it as it is, so the extra passes a part sees
// while other parts still change are no-ops, and each part ends where its
// own adapt() ends.  The batched form joins the meshes, runs adapt() on the
// joined mesh (one launch per kernel for all meshes of the batch) and splits
// the result.
//
// Each part's vertices carry their part number in the vertex tag
// kPartTag (a linear tag: refinement midpoints average two equal numbers, so
// the value is exact); an entity of any dimension belongs to the part of its
// first vertex.
#include <algorithm>

#include "osh_passes.hpp"

namespace osh {

char const* const kPartTag = "omega_h_amd_part";

namespace {

constexpr int kWave = 64;

__global__ void copy_offset(LO n, LO const* in, LO offset, LO* out) {
  LO i = thread_index();
  if (i < n) out[i] = in[i] + offset;
}

template <class T>
void copy_into(Ctx& c, DArray<T> const& src, DArray<T>& dst, LO at, LO n) {
  if (n > 0)
    OSH_HIP(hipMemcpyAsync(dst.data() + at, src.data(), sizeof(T) * std::size_t(n),
        hipMemcpyDeviceToDevice, c.stream));
}

// One array of a split: `w` items per entity of dimension `d`, `size` bytes
// each, go to the parts' own arrays (dst: one pointer per part).  With
// shift >= 0 the items are indices of entities of dimension `shift`, rebased
// to each part's first entity of that dimension.
struct SplitJob {
  void const* src;
  void* const* dst;
  LO nitems;
  int w;
  int d;
  int size;
  int shift;
};

// every array of a split in one launch: blockIdx.y picks the array, and each
// item finds its part by a binary search over the parts' first entities
// (starts[d * nparts + k], nondecreasing; an empty part starts where the next
// one does, so the last part starting at or before an entity holds it)
__global__ void split_copy(SplitJob const* jobs, LO const* starts, int nparts) {
  SplitJob const j = jobs[blockIdx.y];
  LO const* st = starts + std::size_t(j.d) * std::size_t(nparts);
  for (std::int64_t i = std::int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < j.nitems;
       i += std::int64_t(gridDim.x) * blockDim.x) {
    LO e = LO(i / j.w);
    int lo = 0, hi = nparts - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (st[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    std::int64_t at = i - std::int64_t(st[lo]) * j.w;
    if (j.size == 8) {
      static_cast<Real*>(j.dst[lo])[at] = static_cast<Real const*>(j.src)[i];
    } else if (j.size == 4) {
      LO v = static_cast<LO const*>(j.src)[i];
      if (j.shift >= 0) v -= starts[std::size_t(j.shift) * std::size_t(nparts) + lo];
      static_cast<LO*>(j.dst[lo])[at] = v;
    } else {
      static_cast<I8*>(j.dst[lo])[at] = static_cast<I8 const*>(j.src)[i];
    }
  }
}

// first and last entity of each part, and its entity count, from the part
// number of each entity's first vertex.  Parts are contiguous blocks, so a
// wave's entities nearly always belong to one part: the wave then combines
// its lanes and makes one atomic update per counter instead of 64 (lanes of
// different parts update one by one), which keeps the few counters of a
// batch from serialising the whole mesh's atomics.
__global__ void part_bounds(LO n, LO const* ev2v, int deg, Real const* vpart, int nparts,
    LO* first, LO* last, LO* count, int* error) {
  LO e = thread_index();
  int lane = int(threadIdx.x) & (kWave - 1);
  bool valid = e < n;
  int p = -1;
  if (valid) {
    LO v = ev2v ? ev2v[std::size_t(e) * deg] : e;
    Real pr = vpart[v];
    p = int(pr);
    if (!(pr >= 0.0) || p >= nparts || Real(p) != pr) {
      *error = kBadPart;
      valid = false;
    }
  }
  unsigned long long vmask = __ballot(valid);
  if (vmask == 0ull) return;
  int leader = __ffsll(static_cast<long long>(vmask)) - 1;
  int p0 = __shfl(p, leader, kWave);
  if (__ballot(valid && p != p0) == 0ull) {
    LO lo = valid ? e : 0x7fffffff, hi = valid ? e : -1;
    for (int d = kWave / 2; d > 0; d >>= 1) {
      LO olo = __shfl_xor(lo, d, kWave), ohi = __shfl_xor(hi, d, kWave);
      lo = olo < lo ? olo : lo;
      hi = ohi > hi ? ohi : hi;
    }
    if (lane == leader) {
      atomicMin(first + p0, lo);
      atomicMax(last + p0, hi);
      atomicAdd(count + p0, LO(__popcll(vmask)));
    }
  } else if (valid) {
    atomicMin(first + p, e);
    atomicMax(last + p, e);
    atomicAdd(count + p, 1);
  }
}

}  // namespace

void join_meshes(Ctx& c, std::vector<DMesh const*> const& parts, DMesh& out) {
  OSH_CHECK(!parts.empty(), "join: no meshes");
  DMesh const& p0 = *parts[0];
  int dim = p0.dim;
  for (DMesh const* p : parts) {
    OSH_CHECK(p->dim == dim, "join: meshes of different dimensions");
    OSH_CHECK(p->metric.exists() == p0.metric.exists() && p->metric_ncomps == p0.metric_ncomps,
        "join: meshes with different metric layouts");
    for (int d = 0; d <= dim; ++d)
      OSH_CHECK(p->class_id[d].exists() && p->class_dim[d].exists(),
          "join: every mesh must be classified");
    OSH_CHECK(p->itags.empty() && p->etags.empty(),
        "join: inheritance and element tags are not carried");
    OSH_CHECK(p->vtags.size() == p0.vtags.size(), "join: meshes with different vertex tags");
    for (std::size_t t = 0; t < p->vtags.size(); ++t)
      OSH_CHECK(p->vtags[t].name == p0.vtags[t].name && p->vtags[t].ncomps == p0.vtags[t].ncomps &&
                    p->vtags[t].kind == p0.vtags[t].kind && p->vtags[t].name != kPartTag,
          "join: meshes with different vertex tags");
  }
  DMesh r;
  r.dim = dim;
  r.metric_ncomps = p0.metric_ncomps;
  std::int64_t tot[4] = {0, 0, 0, 0};
  for (DMesh const* p : parts)
    for (int d = 0; d <= dim; ++d) tot[d] += p->nents[d];
  for (int d = 0; d <= dim; ++d) {
    OSH_CHECK(tot[d] * (d + 1) < (std::int64_t(1) << 31), "join: too many entities");
    r.nents[d] = LO(tot[d]);
  }
  for (int d = 1; d <= dim; ++d) {
    r.down[d] = LOs(c, r.nents[d] * (d + 1));
    if (d >= 2) r.down_codes[d] = Bytes(c, r.nents[d] * (d + 1));
  }
  r.coords = Reals(c, r.nents[0] * dim);
  if (p0.metric.exists()) r.metric = Reals(c, r.nents[0] * r.metric_ncomps);
  for (int d = 0; d <= dim; ++d) {
    r.class_id[d] = LOs(c, r.nents[d]);
    r.class_dim[d] = Bytes(c, r.nents[d]);
  }
  for (auto const& t : p0.vtags) r.vtags.push_back({t.name, t.ncomps, t.kind, Reals(c, r.nents[0] * t.ncomps)});
  Reals part(c, r.nents[0]);
  std::vector<Real> part_h(std::size_t(r.nents[0]));
  LO at[4] = {0, 0, 0, 0};
  for (std::size_t k = 0; k < parts.size(); ++k) {
    DMesh const& p = *parts[k];
    for (int d = 1; d <= dim; ++d) {
      LO n = p.nents[d] * (d + 1);
      if (n > 0)
        hipLaunchKernelGGL(copy_offset, dim3(nblocks(n)), dim3(kBlock), 0, c.stream, n,
            p.down[d].data(), at[d - 1], r.down[d].data() + std::size_t(at[d]) * (d + 1));
      OSH_HIP(hipGetLastError());
      if (d >= 2) copy_into(c, p.down_codes[d], r.down_codes[d], at[d] * (d + 1), n);
    }
    copy_into(c, p.coords, r.coords, at[0] * dim, p.nents[0] * dim);
    if (p.metric.exists())
      copy_into(c, p.metric, r.metric, at[0] * r.metric_ncomps, p.nents[0] * r.metric_ncomps);
    for (int d = 0; d <= dim; ++d) {
      copy_into(c, p.class_id[d], r.class_id[d], at[d], p.nents[d]);
      copy_into(c, p.class_dim[d], r.class_dim[d], at[d], p.nents[d]);
    }
    for (std::size_t t = 0; t < p.vtags.size(); ++t)
      copy_into(c, p.vtags[t].data, r.vtags[t].data, at[0] * p.vtags[t].ncomps,
          p.nents[0] * p.vtags[t].ncomps);
    std::fill(part_h.begin() + at[0], part_h.begin() + at[0] + p.nents[0], Real(k));
    for (int d = 0; d <= dim; ++d) at[d] += p.nents[d];
  }
  upload(c, part, part_h.data(), r.nents[0]);
  r.vtags.push_back({kPartTag, 1, 0, part});
  c.sync();
  out = r;
}

void part_ranges(Ctx& c, DMesh& m, int nparts, std::vector<LO> first[4],
    std::vector<LO> count[4]) {
  Reals const* part = nullptr;
  for (auto const& t : m.vtags)
    if (t.name == kPartTag) part = &t.data;
  OSH_CHECK(part != nullptr, "split: the mesh was not joined");
  OSH_CHECK(nparts >= 1, "split: no parts");
  for (int d = 0; d <= m.dim; ++d) {
    std::vector<LO> f(std::size_t(nparts), 0x7fffffff), l(std::size_t(nparts), -1),
        n(std::size_t(nparts), 0);
    LOs df = to_device(c, f), dl = to_device(c, l), dn = to_device(c, n);
    LO const* ev2v = d == 0 ? nullptr : verts_of(c, m, d).data();
    if (m.nents[d] > 0)
      hipLaunchKernelGGL(part_bounds, dim3(nblocks(m.nents[d])), dim3(kBlock), 0, c.stream,
          m.nents[d], ev2v, d + 1, part->data(), nparts, df.data(), dl.data(), dn.data(),
          c.error_word());
    OSH_HIP(hipGetLastError());
    f = to_host(c, df);
    l = to_host(c, dl);
    n = to_host(c, dn);
    LO next = 0;
    for (int k = 0; k < nparts; ++k) {
      if (n[k] == 0) {
        f[k] = next;
        continue;
      }
      // each part is one block, the blocks in part order
      OSH_CHECK(f[k] == next && l[k] - f[k] + 1 == n[k], "split: parts are not contiguous");
      next += n[k];
    }
    OSH_CHECK(next == m.nents[d], "split: entities outside the parts");
    first[d] = f;
    count[d] = n;
  }
}

void split_mesh(Ctx& c, DMesh& m, int nparts, std::vector<DMesh>& out) {
  std::vector<LO> first[4], count[4];
  part_ranges(c, m, nparts, first, count);
  int dim = m.dim;
  out.assign(std::size_t(nparts), DMesh());
  // the parts' arrays are allocated here and filled by one launch
  std::vector<SplitJob> jobs;
  std::vector<void*> dsts;
  auto add = [&](void const* src, LO nitems, int w, int d, int size, int shift,
                 auto&& dst_of) {
    jobs.push_back({src, nullptr, nitems, w, d, size, shift});
    for (int k = 0; k < nparts; ++k) dsts.push_back(dst_of(out[std::size_t(k)], k));
  };
  for (int k = 0; k < nparts; ++k) {
    DMesh& r = out[std::size_t(k)];
    r.dim = dim;
    r.metric_ncomps = m.metric_ncomps;
    for (int d = 0; d <= dim; ++d) r.nents[d] = count[d][k];
  }
  for (int d = 1; d <= dim; ++d) {
    add(m.down[d].data(), m.nents[d] * (d + 1), d + 1, d, 4, d - 1, [&](DMesh& r, int) {
      r.down[d] = LOs(c, r.nents[d] * (d + 1));
      return static_cast<void*>(r.down[d].data());
    });
    if (d >= 2)
      add(m.down_codes[d].data(), m.nents[d] * (d + 1), d + 1, d, 1, -1, [&](DMesh& r, int) {
        r.down_codes[d] = Bytes(c, r.nents[d] * (d + 1));
        return static_cast<void*>(r.down_codes[d].data());
      });
  }
  add(m.coords.data(), m.nents[0] * dim, dim, 0, 8, -1, [&](DMesh& r, int) {
    r.coords = Reals(c, r.nents[0] * dim);
    return static_cast<void*>(r.coords.data());
  });
  if (m.metric.exists()) {
    int nc = m.metric_ncomps;
    add(m.metric.data(), m.nents[0] * nc, nc, 0, 8, -1, [&](DMesh& r, int) {
      r.metric = Reals(c, r.nents[0] * nc);
      return static_cast<void*>(r.metric.data());
    });
  }
  for (int d = 0; d <= dim; ++d) {
    add(m.class_id[d].data(), m.nents[d], 1, d, 4, -1, [&](DMesh& r, int) {
      r.class_id[d] = LOs(c, r.nents[d]);
      return static_cast<void*>(r.class_id[d].data());
    });
    add(m.class_dim[d].data(), m.nents[d], 1, d, 1, -1, [&](DMesh& r, int) {
      r.class_dim[d] = Bytes(c, r.nents[d]);
      return static_cast<void*>(r.class_dim[d].data());
    });
  }
  for (auto const& t : m.vtags) {
    if (t.name == kPartTag) continue;
    int nc = t.ncomps;
    add(t.data.data(), m.nents[0] * nc, nc, 0, 8, -1, [&](DMesh& r, int) {
      r.vtags.push_back({t.name, t.ncomps, t.kind, Reals(c, r.nents[0] * nc)});
      return static_cast<void*>(r.vtags.back().data.data());
    });
  }
  if (m.length.exists())
    add(m.length.data(), m.nents[1], 1, 1, 8, -1, [&](DMesh& r, int) {
      r.length = Reals(c, r.nents[1]);
      return static_cast<void*>(r.length.data());
    });
  if (m.quality.exists())
    add(m.quality.data(), m.nents[dim], 1, dim, 8, -1, [&](DMesh& r, int) {
      r.quality = Reals(c, r.nents[dim]);
      return static_cast<void*>(r.quality.data());
    });
  std::vector<LO> starts(std::size_t(4) * std::size_t(nparts), 0);
  for (int d = 0; d <= dim; ++d)
    std::copy(first[d].begin(), first[d].end(), starts.begin() + std::ptrdiff_t(d) * nparts);
  DArray<void*> ddsts = to_device(c, dsts);
  for (std::size_t j = 0; j < jobs.size(); ++j)
    jobs[j].dst = ddsts.data() + j * std::size_t(nparts);
  DArray<SplitJob> djobs = to_device(c, jobs);
  LOs dstarts = to_device(c, starts);
  LO most = 0;
  for (auto const& j : jobs) most = std::max(most, j.nitems);
  if (most > 0) {
    unsigned gx = std::min(nblocks(most), 4096u);
    hipLaunchKernelGGL(split_copy, dim3(gx, unsigned(jobs.size())), dim3(kBlock), 0, c.stream,
        djobs.data(), dstarts.data(), nparts);
    OSH_HIP(hipGetLastError());
  }
  c.sync();
}

namespace {

__global__ void note_parts(LO nkeys, LO const* keys2kds, LO const* ev2v, Real const* vpart,
    int* changed) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO kd = keys2kds[k];
  LO v = ev2v ? ev2v[std::size_t(kd) * 2] : kd;
  changed[int(vpart[v])] = 1;
}

}  // namespace

namespace {

// an order-preserving integer key of a double (negative values flipped)
__device__ __forceinline__ unsigned long long order_key(double x) {
  unsigned long long u = __double_as_longlong(x);
  return (u >> 63) ? ~u : (u | (1ull << 63));
}

__device__ __forceinline__ double from_order_key(unsigned long long k) {
  unsigned long long u = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __longlong_as_double(u);
}

// per-part minima with one atomic per wave when the wave lies in one part
// (the parts of a joined mesh are contiguous blocks)
__global__ void part_min_kernel(LO ne, int deg, LO const* ev2v, Real const* vpart,
    Real const* quality, unsigned long long* mins) {
  LO e = thread_index();
  int lane = int(threadIdx.x) & (kWave - 1);
  bool valid = e < ne;
  int p = valid ? int(vpart[ev2v[std::size_t(e) * deg]]) : -1;
  unsigned long long key = valid ? order_key(quality[e]) : ~0ull;
  unsigned long long vmask = __ballot(valid);
  if (vmask == 0ull) return;
  int leader = __ffsll(static_cast<long long>(vmask)) - 1;
  int p0 = __shfl(p, leader, kWave);
  if (__ballot(valid && p != p0) == 0ull) {
    for (int d = kWave / 2; d > 0; d >>= 1) {
      unsigned long long o = __shfl_xor(key, d, kWave);
      key = o < key ? o : key;
    }
    if (lane == leader) atomicMin(mins + p0, key);
  } else if (valid) {
    atomicMin(mins + p, key);
  }
}

__global__ void decode_mins(LO n, unsigned long long const* keys, Real* out) {
  LO i = thread_index();
  if (i < n) out[i] = from_order_key(keys[i]);
}

}  // namespace

std::vector<Real> part_min_quality(Ctx& c, DMesh& m, int nparts) {
  Reals const* part = nullptr;
  for (auto const& t : m.vtags)
    if (t.name == kPartTag) part = &t.data;
  OSH_CHECK(part != nullptr, "part_min_quality: the mesh was not joined");
  // the key of +infinity
  std::vector<unsigned long long> init(std::size_t(nparts), 0xfff0000000000000ull);
  DArray<unsigned long long> keys = to_device(c, init);
  int dim = m.dim;
  OSH_LAUNCH(part_min_kernel, m.nents[dim], c, dim + 1, verts_of(c, m, dim).data(),
      part->data(), ask_qualities(c, m).data(), keys.data());
  Reals out(c, nparts);
  OSH_LAUNCH(decode_mins, nparts, c, keys.data(), out.data());
  return to_host(c, out);
}

void note_changed_parts(Ctx& c, DMesh const& m, int key_dim, LOs const& keys2kds) {
  if (!c.part_changed) return;
  Real const* part = nullptr;
  for (auto const& t : m.vtags)
    if (t.name == kPartTag) part = t.data.data();
  if (!part) return;
  OSH_LAUNCH(note_parts, keys2kds.size(), c, keys2kds.data(),
      key_dim == 1 ? m.down[1].data() : (LO const*)nullptr, part, c.part_changed);
}

}  // namespace osh
