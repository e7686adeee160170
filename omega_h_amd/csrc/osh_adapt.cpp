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
bool run_pass(Ctx& c, DMesh& m, AdaptOptions const& o, int kind,
    std::vector<PassRecord>* log, Real* qmin, PassHook const* hook,
    SwapCache* cache = nullptr) {
  OSH_CHECK(!log || log->size() < 4096, "adapt exceeded 4096 passes");
  LO nin = m.nents[m.dim];
  c.sync();
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
  if (o.verbosity >= 2) std::printf("addressing edge lengths\n");
  double t1 = now();
  bool did;
  Real q = qmin;  // the minimum quality of the current mesh
  do {
    did = false;
    if (o.should_refine && run_pass(c, m, o, 0, log, &q, hook)) did = true;
    if (o.should_coarsen && run_pass(c, m, o, 1, log, &q, hook)) did = true;
  } while (did);
  double t2 = now();
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
  double t3 = now();
  bool conserving = conserve_any(m);
  conserve_correct(c, m, kMaxDiffusionSweeps);
  double t4 = now();
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

}  // namespace osh
