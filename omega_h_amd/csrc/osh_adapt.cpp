// The adapt() driver on one GPU.
//
//   adapt / pre_adapt / satisfy_lengths / satisfy_quality
//   src/Omega_h_adapt.cpp:110-288
// A pass is one refine_by_size, coarsen_by_size, swap_edges or
// coarsen_slivers call that rebuilt the mesh; each is timed on the host
// clock with the stream drained at both ends.
#include <chrono>

#include "osh_passes.hpp"

namespace osh {

namespace {

constexpr int kQminSlot = Ctx::kSlots - 1;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// *qmin receives the minimum element quality after the pass, reduced before
// the pass's closing sync so that it costs no sync of its own
bool run_pass(Ctx& c, DMesh& m, AdaptOptions const& o, int kind,
    std::vector<PassRecord>* log, Real* qmin, SwapCache* cache = nullptr) {
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
  return r;
}

}  // namespace

bool adapt(Ctx& c, DMesh& m, AdaptOptions const& o, std::vector<PassRecord>* log) {
  OSH_CHECK(0.0 <= o.min_quality_allowed, "min_quality_allowed < 0");
  OSH_CHECK(o.min_quality_allowed <= o.min_quality_desired, "min_quality_allowed > desired");
  OSH_CHECK(o.min_quality_desired <= 1.0, "min_quality_desired > 1");
  OSH_CHECK(o.nsliver_layers >= 0 && o.nsliver_layers < 100, "nsliver_layers out of range");
  Reals const& l = ask_lengths(c, m);
  Real lmin = min_real(c, l), lmax = max_real(c, l);
  Real qmin = min_real(c, ask_qualities(c, m));
  if (qmin >= o.min_quality_desired && lmin >= o.min_length_desired &&
      lmax <= o.max_length_desired)
    return false;
  bool did;
  Real q = qmin;  // the minimum quality of the current mesh
  do {
    did = false;
    if (o.should_refine && run_pass(c, m, o, 0, log, &q)) did = true;
    if (o.should_coarsen && run_pass(c, m, o, 1, log, &q)) did = true;
  } while (did);
  if (q >= o.min_quality_desired) return true;
  // swap choices carry over between swap passes; a sliver pass rebuilds
  // the mesh and drops them
  SwapCache cache;
  do {
    if (o.should_swap && run_pass(c, m, o, 2, log, &q, &cache)) continue;
    cache.clear();
    if (o.should_coarsen_slivers && run_pass(c, m, o, 3, log, &q)) continue;
    return true;
  } while (q < o.min_quality_desired);
  return true;
}

}  // namespace osh
