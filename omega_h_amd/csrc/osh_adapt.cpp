// The adapt() driver on one GPU.
//
//   adapt / pre_adapt / satisfy_lengths / satisfy_quality
//   src/Omega_h_adapt.cpp:110-288
// A pass is one refine_by_size, coarsen_by_size, swap_edges or
// coarsen_slivers call that rebuilt the mesh; each is timed on the host
// clock with the stream drained at both ends.
#include <chrono>
#include <cstdio>
#include <functional>
#include <vector>

#include "osh_passes.hpp"

namespace osh {

namespace {

char const* plural_name(int d) {
  return d == 3 ? "tets" : (d == 2 ? "triangles" : (d == 1 ? "edges" : "vertices"));
}

// print_goal_stats, src/Omega_h_histogram.cpp:56-85
void print_goal_stats(Ctx& c, int d, char const* name, Reals const& values, Real lo, Real hi) {
  std::vector<Real> v = to_host(c, values);
  LO nlow = 0, nhigh = 0;
  Real mn = v.empty() ? 0.0 : v[0], mx = mn;
  for (Real x : v) {
    if (x < lo) ++nlow;
    if (x > hi) ++nhigh;
    mn = (x < mn) ? x : mn;
    mx = (x > mx) ? x : mx;
  }
  LO nmid = LO(v.size()) - nlow - nhigh;
  std::printf("%d %s, %s [%.2f,%.2f]", int(v.size()), plural_name(d), name, mn, mx);
  if (nlow) std::printf(", %d <%.2f", nlow, lo);
  if (nmid) std::printf(", %d in [%.2f,%.2f]", nmid, lo, hi);
  if (nhigh) std::printf(", %d >%.2f", nhigh, hi);
  std::printf("\n");
}

// print_adapt_status's report, src/Omega_h_adapt.cpp:101-120
void print_status(Ctx& c, DMesh& m, AdaptOptions const& o) {
  if (o.verbosity <= 0) return;
  print_goal_stats(c, m.dim, "quality", ask_qualities(c, m), o.min_quality_allowed,
      o.min_quality_desired);
  print_goal_stats(c, 1, "length", ask_lengths(c, m), o.min_length_desired,
      o.max_length_desired);
  std::fflush(stdout);
}

// get_histogram + print_histogram, src/Omega_h_histogram.cpp:15-52: bin i
// holds the values in [lo + i w, lo + (i + 1) w), the last bin [.., hi]
void print_histogram(std::vector<Real> const& v, int nbins, Real lo, Real hi,
    char const* name) {
  Real w = (hi - lo) / Real(nbins);
  std::printf("%s histogram:\n", name);
  for (int i = 0; i < nbins; ++i) {
    Real floor = w * i + lo;
    Real ceil = (i == nbins - 1) ? hi : w * (i + 1) + lo;
    long long n = 0;
    for (Real x : v)
      if (x >= floor && (i == nbins - 1 ? x <= ceil : x < ceil)) ++n;
    std::printf("%.2f-%.2f: %lld\n", w * i + lo, w * (i + 1) + lo, n);
  }
}

// print_adapt_histograms, src/Omega_h_adapt.cpp:122-138 (verbosity
// EXTRA_STATS)
void print_histograms(Ctx& c, DMesh& m, AdaptOptions const& o) {
  std::vector<Real> q = to_host(c, ask_qualities(c, m));
  std::vector<Real> l = to_host(c, ask_lengths(c, m));
  print_histogram(q, o.nquality_histogram_bins, 0.0, 1.0, "quality");
  print_histogram(l, o.nlength_histogram_bins, o.length_histogram_min, o.length_histogram_max,
      "length");
  Real sum = 0;
  for (Real x : q) sum += x;
  std::printf("average quality: %g\n", q.empty() ? 0.0 : sum / Real(q.size()));
  std::fflush(stdout);
}

constexpr int kQminSlot = Ctx::kSlots - 1;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// *qmin receives the minimum element quality after the pass, reduced before
// the pass's closing sync so that it costs no sync of its own
// the most passes one adapt() records: a guard on the log, far above real
// runs (the reference's loops have no limit; one 1.3M-tet box of the serving
// workload under a constant anisotropic metric takes 4,639 swap and 283
// sliver passes before it stops, unable to reach the quality goal)
constexpr std::size_t kMaxLoggedPasses = std::size_t(1) << 22;

bool run_pass(Ctx& c, DMesh& m, AdaptOptions const& o, int kind,
    std::vector<PassRecord>* log, Real* qmin, PassHook const* hook,
    SwapCache* cache = nullptr) {
  OSH_CHECK(!log || log->size() < kMaxLoggedPasses, "adapt exceeded its pass limit");
  static char const* const kPassNames[4] = {"refine_by_size", "coarsen_by_size", "swap_edges",
      "coarsen_slivers"};
  LO nin = m.nents[m.dim];
  c.sync();
  Range range(c, kPassNames[kind]);
  double t0 = now();
  bool r;
  if (kind == 0)
    r = refine_by_size(c, m, o);
  else if (kind == 1)
    r = coarsen_by_size(c, m, o);
  else if (kind == 2)
    r = swap_edges(c, m, o, cache);
  else
    r = coarsen_slivers(c, m, o);
  min_real_async(c, ask_qualities(c, m), kQminSlot);
  c.sync();
  double t1 = now();
  range.end();
  *qmin = read_slot<Real>(c, kQminSlot);
  if (r && log) log->push_back({kind, nin, m.nents[m.dim], t1 - t0});
  if (r && hook && *hook) (*hook)(m);
  // post_rebuild, src/Omega_h_adapt.cpp:169-171
  if (r && o.verbosity >= 2) print_status(c, m, o);
  return r;
}

}  // namespace

bool adapt(Ctx& c, DMesh& m, AdaptOptions const& o, std::vector<PassRecord>* log,
    PassHook const* hook) {
  OSH_CHECK(0.0 <= o.min_quality_allowed, "min_quality_allowed < 0");
  OSH_CHECK(o.min_quality_allowed <= o.min_quality_desired, "min_quality_allowed > desired");
  OSH_CHECK(o.min_quality_desired <= 1.0, "min_quality_desired > 1");
  OSH_CHECK(o.nsliver_layers >= 0 && o.nsliver_layers < 100, "nsliver_layers out of range");
  // the reports of pre_adapt/post_adapt (src/Omega_h_adapt.cpp:153-166,
  // 244-270) at the reference's verbosity levels: 1 EACH_ADAPT, 2
  // EACH_REBUILD (with the passes' own lines), 3 EXTRA_STATS (histograms)
  double t0 = now();
  // the reference's ScopedTimer adapt_timer("adapt") and the
  // OMEGA_H_TIME_FUNCTION timers of the phases (src/Omega_h_adapt.cpp:
  // 111, 174, 190, 273)
  Range adapt_range(c, "adapt");
  Range pre_range(c, "pre_adapt");
  if (o.verbosity >= 1) std::printf("before adapting:\n");
  print_status(c, m, o);
  Reals const& l = ask_lengths(c, m);
  Real lmin = min_real(c, l), lmax = max_real(c, l);
  Real qmin = min_real(c, ask_qualities(c, m));
  if (qmin >= o.min_quality_desired && lmin >= o.min_length_desired &&
      lmax <= o.max_length_desired)
    return false;
  if (o.verbosity >= 3) print_histograms(c, m, o);
  conserve_setup(c, m);
  pre_range.end();
  if (o.verbosity >= 2) std::printf("addressing edge lengths\n");
  double t1 = now();
  bool did;
  Real q = qmin;  // the minimum quality of the current mesh
  Range lengths_range(c, "satisfy_lengths");
  do {
    did = false;
    if (o.should_refine && run_pass(c, m, o, 0, log, &q, hook)) did = true;
    if (o.should_coarsen && run_pass(c, m, o, 1, log, &q, hook)) did = true;
  } while (did);
  lengths_range.end();
  double t2 = now();
  Range quality_range(c, "satisfy_quality");
  if (q < o.min_quality_desired) {
    if (o.verbosity >= 2) std::printf("addressing element qualities\n");
    // swap choices carry over between swap passes; a sliver pass rebuilds
    // the mesh and drops them
    SwapCache cache;
    do {
      if (o.should_swap && run_pass(c, m, o, 2, log, &q, hook, &cache)) continue;
      cache.clear();
      if (o.should_coarsen_slivers && run_pass(c, m, o, 3, log, &q, hook)) continue;
      if (o.verbosity >= 1) std::printf("could not satisfy quality\n");
      break;
    } while (q < o.min_quality_desired);
  }
  quality_range.end();
  double t3 = now();
  bool conserving = conserve_any(m);
  conserve_correct(c, m, kMaxDiffusionSweeps);
  double t4 = now();
  adapt_range.end();
  if (o.verbosity == 1) {
    std::printf("after adapting:\n");
    print_status(c, m, o);
  }
  if (o.verbosity >= 3) print_histograms(c, m, o);
  if (o.verbosity >= 1) {
    std::printf("addressing edge lengths took %g seconds\n", t2 - t1);
    std::printf("addressing element qualities took %g seconds\n", t3 - t2);
    if (conserving) std::printf("correcting integral errors took %g seconds\n", t4 - t3);
    std::printf("adapting took %g seconds\n\n", now() - t0);
    std::fflush(stdout);
  }
  return true;
}

namespace {

// The batched form runs each phase of adapt() on joined meshes.  Every pass
// marks the parts it changes (Ctx::part_changed); a pass acts on a part as
// a deterministic function of that part alone, so a pass kind that left a
// part unchanged leaves it unchanged again until another kind changes it.
// Parts are taken out of a joined mesh (split off, the rest joined again)
// once a quarter of them no longer follow the joined mesh's pass sequence:
//
// * lengths (refine, then coarsen, until neither changes anything): a part
//   that one round of both left unchanged is finished;
// * quality (swap while swapping changes something, else coarsen slivers,
//   while the minimum quality is below min_quality_desired): a part that a
//   swap pass left unchanged would do a sliver pass next, while the joined
//   mesh goes on swapping.  Such parts are set aside (those already at
//   min_quality_desired are finished) and form a group of their own later,
//   whose first swap pass changes none of them, so that they all go on to
//   their sliver pass together.
//
// Each part thus sees its own pass sequence, with extra passes that leave it
// unchanged, and ends where adapt() alone leaves it.

struct PartFlags {
  Ctx& c;
  LOs flags;
  int n;
  PartFlags(Ctx& ctx, int nparts) : c(ctx), n(nparts) { reset(); }
  ~PartFlags() { c.part_changed = nullptr; }
  void reset() {
    flags = fill_lo(c, n, 0);
    c.part_changed = flags.data();
  }
  // which parts the passes since the last reset changed
  std::vector<LO> read() { return to_host(c, flags); }
};

void put_back(Ctx& c, DMesh& joined, std::vector<DMesh>& meshes, std::vector<int> const& ids) {
  Range range(c, "split_mesh");
  std::vector<DMesh> parts;
  split_mesh(c, joined, int(ids.size()), parts);
  for (std::size_t k = 0; k < ids.size(); ++k) meshes[std::size_t(ids[k])] = parts[k];
}

DMesh joined_of(Ctx& c, std::vector<DMesh>& meshes, std::vector<int> const& ids) {
  Range range(c, "join_meshes");
  std::vector<DMesh const*> ps;
  for (int i : ids) ps.push_back(&meshes[std::size_t(i)]);
  DMesh joined;
  join_meshes(c, ps, joined);
  return joined;
}

void batch_lengths(Ctx& c, std::vector<DMesh>& meshes, std::vector<int> active,
    AdaptOptions const& o, std::vector<PassRecord>* log) {
  if (!o.should_refine && !o.should_coarsen) return;
  Range range(c, "satisfy_lengths");
  while (!active.empty()) {
    int np = int(active.size());
    DMesh joined = joined_of(c, meshes, active);
    PartFlags pf(c, np);
    Real q = 0.0;
    std::vector<char> frozen;
    bool finished = false;
    for (;;) {
      bool did = false;
      if (o.should_refine)
        did = run_pass(c, joined, o, 0, log, &q, nullptr) || did;
      if (o.should_coarsen)
        did = run_pass(c, joined, o, 1, log, &q, nullptr) || did;
      if (!did) {
        finished = true;
        break;
      }
      std::vector<LO> changed = pf.read();
      pf.reset();
      int nfrozen = 0;
      for (LO x : changed) nfrozen += (x == 0);
      if (nfrozen > 0 && 4 * nfrozen >= np) {
        frozen.assign(std::size_t(np), 0);
        for (int k = 0; k < np; ++k) frozen[std::size_t(k)] = changed[std::size_t(k)] == 0;
        break;
      }
    }
    put_back(c, joined, meshes, active);
    if (finished) break;
    std::vector<int> next;
    for (int k = 0; k < np; ++k)
      if (!frozen[std::size_t(k)]) next.push_back(active[std::size_t(k)]);
    active = next;
  }
}

void batch_quality(Ctx& c, std::vector<DMesh>& meshes, std::vector<int> group,
    AdaptOptions const& o, std::vector<PassRecord>* log) {
  if (!o.should_swap && !o.should_coarsen_slivers) return;
  Range range(c, "satisfy_quality");
  std::vector<std::vector<int>> groups;
  if (!group.empty()) groups.push_back(group);
  while (!groups.empty()) {
    std::vector<int> active = groups.back();
    groups.pop_back();
    std::vector<int> stalled;
    while (!active.empty()) {
      int np = int(active.size());
      DMesh joined = joined_of(c, meshes, active);
      PartFlags pf(c, np);
      SwapCache cache;
      Real q = min_real(c, ask_qualities(c, joined));
      std::vector<char> out;
      bool finished = false;
      while (q < o.min_quality_desired) {
        if (o.should_swap) {
          bool did = run_pass(c, joined, o, 2, log, &q, nullptr, &cache);
          std::vector<LO> changed = pf.read();
          pf.reset();
          if (did) {
            int nstalled = 0;
            for (LO x : changed) nstalled += (x == 0);
            if (nstalled > 0 && 4 * nstalled >= np) {
              out.assign(std::size_t(np), 0);
              for (int k = 0; k < np; ++k) out[std::size_t(k)] = changed[std::size_t(k)] == 0;
              break;
            }
            continue;
          }
        }
        // no part swapped: every part goes on to its sliver pass
        cache.clear();
        if (o.should_coarsen_slivers &&
            run_pass(c, joined, o, 3, log, &q, nullptr)) {
          pf.reset();
          continue;
        }
        break;  // could not satisfy quality
      }
      if (out.empty()) finished = true;
      put_back(c, joined, meshes, active);
      if (finished) break;
      std::vector<int> next;
      for (int k = 0; k < np; ++k) {
        int id = active[std::size_t(k)];
        if (!out[std::size_t(k)]) {
          next.push_back(id);
        } else if (min_real(c, ask_qualities(c, meshes[std::size_t(id)])) <
                   o.min_quality_desired) {
          stalled.push_back(id);  // its next pass is a sliver pass
        }
      }
      active = next;
    }
    if (!stalled.empty()) groups.push_back(stalled);
  }
}

}  // namespace

std::vector<int> adapt_batch(Ctx& c, std::vector<DMesh>& meshes, AdaptOptions const& o,
    std::vector<PassRecord>* log) {
  OSH_CHECK(0.0 <= o.min_quality_allowed, "min_quality_allowed < 0");
  OSH_CHECK(o.min_quality_allowed <= o.min_quality_desired, "min_quality_allowed > desired");
  OSH_CHECK(o.min_quality_desired <= 1.0, "min_quality_desired > 1");
  OSH_CHECK(o.nsliver_layers >= 0 && o.nsliver_layers < 100, "nsliver_layers out of range");
  int n = int(meshes.size());
  Range batch_range(c, "adapt_batch");
  Range pre_range(c, "pre_adapt");
  std::vector<int> did(std::size_t(n), 0);
  std::vector<int> todo;
  for (int i = 0; i < n; ++i) {
    DMesh& m = meshes[std::size_t(i)];
    OSH_CHECK(m.etags.empty() && m.itags.empty(),
        "batched adapt: meshes with element or inheritance tags");
    // pre_adapt (src/Omega_h_adapt.cpp:153-167) per mesh
    Reals const& l = ask_lengths(c, m);
    Real lmin = min_real(c, l), lmax = max_real(c, l);
    Real qmin = min_real(c, ask_qualities(c, m));
    if (qmin >= o.min_quality_desired && lmin >= o.min_length_desired &&
        lmax <= o.max_length_desired)
      continue;
    did[std::size_t(i)] = 1;
    todo.push_back(i);
  }
  pre_range.end();
  AdaptOptions quiet = o;
  quiet.verbosity = 0;
  batch_lengths(c, meshes, todo, quiet, log);
  std::vector<int> low;
  for (int i : todo)
    if (min_real(c, ask_qualities(c, meshes[std::size_t(i)])) < o.min_quality_desired)
      low.push_back(i);
  batch_quality(c, meshes, low, quiet, log);
  return did;
}

}  // namespace osh
