// Edge refinement pass on the device.
//
//   refine_by_size / refine_ghosted / refine_element_based src/Omega_h_refine.cpp:17-100
//   refine_qualities (MetricRefineQualities)  src/Omega_h_refine_qualities.cpp:13-86
//   refine_edges_to_pairs / refine_domains_to_pairs / _cuts / combine
//                                             src/Omega_h_refine_topology.cpp:13-203
//   transfer_refine (coords, metric, class, length, quality)
//                                             src/Omega_h_transfer.cpp:182-428
#include <cstdio>

#include "osh_math.hpp"
#include "osh_passes.hpp"

namespace osh {

namespace {

__global__ void gt_kernel(LO n, Real const* a, Real v, I8* out) {
  LO i = thread_index();
  if (i < n) out[i] = a[i] > v;
}

// 2^lshift lanes per candidate edge (consecutive threads of one wave): the
// lanes share the 2 x (elements around the edge) children, each measured
// with the max-determinant metric of its cell-side vertices and the edge
// midpoint, and combine their minima with shuffles.  The minimum does not
// depend on the order (min2 never takes a NaN, and the sign of a zero
// decides nothing downstream: the candidates' qualities are only compared).
template <int D, int M>
__global__ void __launch_bounds__(kBlock) refine_qualities_kernel(LO nt, int lshift,
    LO ncands, LO const* cands, LO const* ev2v, LO const* cv2v, LO const* e2c_a2ab,
    LO const* e2c_ab2b, I8 const* e2c_codes, Real const* coords, Real const* metric,
    Real const* midpt_metrics, Real* out) {
  LO t = thread_index();
  int lanes = 1 << lshift;
  LO cand = t >> lshift;
  int lane = t & (lanes - 1);
  // every lane of a group stays for the shuffles; groups past the end (whole
  // groups, as nt is a multiple of the lane count) do no work
  bool valid = t < nt && cand < ncands;
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double minqual = 1.0;
  LO e = valid ? cands[cand] : 0;
  LO row = valid ? e2c_a2ab[e] : 0;
  LO nchildren = valid ? 2 * (e2c_a2ab[e + 1] - row) : 0;
  double midp[D];
  double mm[nc];
  if (lane < nchildren) {
    for (int j = 0; j < D; ++j)
      midp[j] = (coords[ev2v[e * 2] * D + j] + coords[ev2v[e * 2 + 1] * D + j]) / 2.;
    for (int k = 0; k < nc; ++k) mm[k] = midpt_metrics[cand * nc + k];
  }
  for (LO ch = lane; ch < nchildren; ch += lanes) {
    LO ec = row + ch / 2;
    int eev = ch % 2;
    LO c = e2c_ab2b[ec];
    I8 code = e2c_codes[ec];
    int cce = code_which_down(code);
    int rot = code_rotation(code);
    {
      int cev = eev ^ rot;
      int ccv = down_template(D, 1, cce, cev);
      int ccs = opposite_template(D, 0, ccv);
      double p[D + 1][D];
      double m[D + 1][nc];
      for (int csv = 0; csv < D; ++csv) {
        LO v = cv2v[c * (D + 1) + down_template(D, D - 1, ccs, csv)];
        for (int j = 0; j < D; ++j) p[csv][j] = coords[v * D + j];
        for (int k = 0; k < nc; ++k) m[csv][k] = metric[v * nc + k];
      }
      for (int j = 0; j < D; ++j) p[D][j] = midp[j];
      for (int k = 0; k < nc; ++k) m[D][k] = mm[k];
      double const* pp[D + 1];
      double const* ms[D + 1];
      for (int k = 0; k < D + 1; ++k) {
        pp[k] = p[k];
        ms[k] = m[k];
      }
      if (D == 3) {
        double const* t = pp[1];
        pp[1] = pp[2];
        pp[2] = t;
        t = ms[1];
        ms[1] = ms[2];
        ms[2] = t;
      }
      double q = element_quality<D, M>(pp, ms);
      minqual = math::min2(minqual, q);
    }
  }
  for (int d = lanes / 2; d > 0; d /= 2) minqual = math::min2(minqual, __shfl_xor(minqual, d, 64));
  if (valid && lane == 0) out[cand] = minqual;
}

__global__ void seed_keys(LO ncands, LO const* cands, Real const* quals, Real qmin,
    I8* initial, Real* edge_quals) {
  LO c = thread_index();
  if (c >= ncands) return;
  initial[cands[c]] = quals[c] >= qmin;
  edge_quals[cands[c]] = quals[c];
}

__global__ void prod_counts(LO nkeys, LO const* keys2edges, LO const* pairs_a2ab,
    LO const* cuts_a2ab, LO* counts) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO e = keys2edges[k];
  LO n = pairs_a2ab ? 2 * (pairs_a2ab[e + 1] - pairs_a2ab[e]) : 2;
  if (cuts_a2ab) n += cuts_a2ab[e + 1] - cuts_a2ab[e];
  counts[k] = n;
}

template <int ED>
__global__ void refine_products_kernel(LO nkeys, LO const* keys2edges, LO const* keys2mid,
    LO const* o2n, LO const* ev2v, LO const* keys2prods, LO const* pairs_a2ab,
    LO const* pairs_ab2b, I8 const* pairs_codes, LO const* pair_dom_verts, LO const* cuts_a2ab,
    LO const* cuts_ab2b, I8 const* cuts_codes, LO const* cut_dom_verts, LO* pv, LO* psrc,
    I8* psrc_dim) {
  LO k = thread_index();
  if (k >= nkeys) return;
  constexpr int nv = ED + 1;
  LO e = keys2edges[k];
  LO mid = keys2mid[k];
  LO prod = keys2prods[k];
  if (ED == 1) {
    pv[prod * 2 + 0] = o2n[ev2v[e * 2 + 0]];
    pv[prod * 2 + 1] = mid;
    psrc[prod] = e;
    psrc_dim[prod] = 1;
    ++prod;
    pv[prod * 2 + 0] = mid;
    pv[prod * 2 + 1] = o2n[ev2v[e * 2 + 1]];
    psrc[prod] = e;
    psrc_dim[prod] = 1;
    ++prod;
  } else {
    for (LO ed = pairs_a2ab[e]; ed < pairs_a2ab[e + 1]; ++ed) {
      LO dom = pairs_ab2b[ed];
      I8 code = pairs_codes[ed];
      int dde = code_which_down(code);
      int rot = code_rotation(code);
      for (int eev = 0; eev < 2; ++eev) {
        int dev = eev ^ rot;
        int ddv = down_template(ED, 1, dde, dev);
        int dds = opposite_template(ED, 0, ddv);
        LO* ppv = pv + prod * nv;
        for (int dsv = 0; dsv < ED; ++dsv)
          ppv[dsv] = o2n[pair_dom_verts[dom * (ED + 1) + down_template(ED, ED - 1, dds, dsv)]];
        ppv[ED] = mid;
        if (ED == 3) {
          LO t = ppv[1];
          ppv[1] = ppv[2];
          ppv[2] = t;
        }
        psrc[prod] = dom;
        psrc_dim[prod] = ED;
        ++prod;
      }
    }
  }
  if (cuts_a2ab) {
    constexpr int CD = ED + 1;
    for (LO ed = cuts_a2ab[e]; ed < cuts_a2ab[e + 1]; ++ed) {
      LO dom = cuts_ab2b[ed];
      int dde = code_which_down(cuts_codes[ed]);
      int ddt = opposite_template(CD, 1, dde);
      LO* ccv = pv + prod * nv;
      for (int dtv = 0; dtv < CD - 1; ++dtv)
        ccv[dtv] = o2n[cut_dom_verts[dom * (CD + 1) + down_template(CD, CD - 2, ddt, dtv)]];
      ccv[CD - 1] = mid;
      psrc[prod] = dom;
      psrc_dim[prod] = CD;
      ++prod;
    }
  }
}

__global__ void midpoint_coords(LO nkeys, int dim, LO const* keys2edges, LO const* ev2v,
    Real const* coords, Real* out) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO e = keys2edges[k];
  for (int j = 0; j < dim; ++j) {
    double comp = 0;
    comp += coords[ev2v[e * 2] * dim + j];
    comp += coords[ev2v[e * 2 + 1] * dim + j];
    comp /= 2;
    out[k * dim + j] = comp;
  }
}

}  // namespace

Reals refine_qualities(Ctx& c, DMesh& m, LOs const& cands) {
  Range range(c, "refine_qualities");
  LO n = cands.size();
  int dim = m.dim;
  int nc = m.metric_ncomps;
  int md = (nc == 1) ? 1 : dim;
  Reals mm(c, n * nc);
  mident_metrics(c, m, n, cands.data(), mm.data());
  Reals out(c, n);
  LOs const& cv2v = verts_of(c, m, dim);
  Adj const& e2c = ask_up(c, m, 1, dim);
  // lanes share a candidate's children, as the coarsen sweeps share a side's
  int lanes = fit_lanes(n, c.tuning.coarsen_lanes);
  int lshift = lanes_shift(lanes);
#define OSH_RQ(D, M)                                                                    \
  OSH_LAUNCH((refine_qualities_kernel<D, M>), n * lanes, c, lshift, n, cands.data(),       \
      m.down[1].data(), cv2v.data(), e2c.a2ab.data(), e2c.ab2b.data(), e2c.codes.data(),    \
      m.coords.data(), m.metric.data(), mm.data(), out.data())
  if (dim == 3 && md == 3)
    OSH_RQ(3, 3);
  else if (dim == 3)
    OSH_RQ(3, 1);
  else if (md == 2)
    OSH_RQ(2, 2);
  else
    OSH_RQ(2, 1);
#undef OSH_RQ
  return out;
}

bool refine_by_size(Ctx& c, DMesh& m, AdaptOptions const& o) {
  int dim = m.dim;
  LO ne = m.nents[1];
  Reals const& lengths = ask_lengths(c, m);
  ask_qualities(c, m);
  Bytes cand(c, ne);
  OSH_LAUNCH(gt_kernel, ne, c, lengths.data(), o.max_length_desired, cand.data());
  LOs cands = collect_marked(c, cand);
  if (cands.size() == 0) return false;
  Reals quals = refine_qualities(c, m, cands);
  Bytes initial = fill_i8(c, ne, 0);
  Reals edge_quals = fill_real(c, ne, 0.0);
  OSH_LAUNCH(seed_keys, cands.size(), c, cands.data(), quals.data(), o.min_quality_allowed,
      initial.data(), edge_quals.data());
  bool any;
  Bytes keys = find_indset(c, m, 1, edge_quals, initial, &any);
  if (!any) return false;
  Range modify_range(c, "modify_ents");
  LOs keys2edges = collect_marked(c, keys);
  LO nkeys = keys2edges.size();
  note_changed_parts(c, m, 1, keys2edges);
  if (o.verbosity >= 2) std::printf("refining %d edges\n", int(nkeys));
  int nc = m.metric_ncomps;
  DMesh nw;
  nw.dim = dim;
  nw.metric_ncomps = nc;
  LOs keys2mid, o2n, old_lows2new_lows;
  // products per key of every dimension, numbered by one scan launch
  LOs k2p[4];
  LO nprods[4] = {nkeys, 0, 0, 0};
  Adj const* pairs[4] = {nullptr, nullptr, nullptr, nullptr};
  Adj const* cuts[4] = {nullptr, nullptr, nullptr, nullptr};
  k2p[0] = iota(c, nkeys + 1);
  {
    LOs counts[4];
    ScanJob jobs[3];
    for (int d = 1; d <= dim; ++d) {
      pairs[d] = (d > 1) ? &ask_up(c, m, 1, d) : nullptr;
      cuts[d] = (d < dim) ? &ask_up(c, m, 1, d + 1) : nullptr;
      counts[d] = LOs(c, nkeys);
      OSH_LAUNCH(prod_counts, nkeys, c, keys2edges.data(),
          pairs[d] ? pairs[d]->a2ab.data() : (LO const*)nullptr,
          cuts[d] ? cuts[d]->a2ab.data() : (LO const*)nullptr, counts[d].data());
      k2p[d] = LOs(c, nkeys + 1);
      jobs[d - 1] = {counts[d].data(), kScanLO, nkeys, k2p[d].data(), slot_ptr<LO>(c, d)};
    }
    offset_scans(c, jobs, dim);
    c.sync();
    for (int d = 1; d <= dim; ++d) nprods[d] = read_slot<LO>(c, d);
  }
  // every dimension's new numbering by one more scan launch
  ModPlan plans[4];
  modify_plans(c, m, 1, 0, dim, keys2edges, k2p, plans);
  modify_scan(c, plans, dim + 1);
  for (int d = 0; d <= dim; ++d) {
    LOs pv, psrc;
    Bytes psrc_dim;
    if (d > 0) {
      Adj const* pr = pairs[d];
      Adj const* ct = cuts[d];
      pv = LOs(c, nprods[d] * (d + 1));
      psrc = LOs(c, nprods[d]);
      psrc_dim = Bytes(c, nprods[d]);
      LO const* pdv = (d > 1) ? verts_of(c, m, d).data() : nullptr;
      LO const* cdv = ct ? verts_of(c, m, d + 1).data() : nullptr;
#define OSH_RP(ED)                                                                        \
  OSH_LAUNCH(refine_products_kernel<ED>, nkeys, c, keys2edges.data(), keys2mid.data(),       \
      o2n.data(), m.down[1].data(), k2p[d].data(), pr ? pr->a2ab.data() : (LO const*)nullptr, \
      pr ? pr->ab2b.data() : (LO const*)nullptr, pr ? pr->codes.data() : (I8 const*)nullptr,  \
      pdv, ct ? ct->a2ab.data() : (LO const*)nullptr, ct ? ct->ab2b.data() : (LO const*)nullptr, \
      ct ? ct->codes.data() : (I8 const*)nullptr, cdv, pv.data(), psrc.data(), psrc_dim.data())
      if (d == 1)
        OSH_RP(1);
      else if (d == 2)
        OSH_RP(2);
      else
        OSH_RP(3);
#undef OSH_RP
    }
    ModOut mo;
    modify_build(c, m, nw, plans[d], keys2edges, k2p[d], nprods[d], pv, old_lows2new_lows, mo,
        d > 0 ? &o2n : nullptr);
    if (d == 0) {
      keys2mid = mo.prods2new;
      o2n = mo.old2new;
      Reals pc(c, nkeys * dim);
      OSH_LAUNCH(midpoint_coords, nkeys, c, dim, keys2edges.data(), m.down[1].data(),
          m.coords.data(), pc.data());
      nw.coords = transfer<Real>(c, mo, dim, m.coords, &pc);
      Reals pm(c, nkeys * nc);
      mident_metrics(c, m, nkeys, keys2edges.data(), pm.data());
      nw.metric = transfer<Real>(c, mo, nc, m.metric, &pm);
      for (auto const& t : m.vtags) {
        Reals pt(c, nkeys * t.ncomps);
        if (t.kind == 1)
          mident_values(c, m, nkeys, keys2edges.data(), t.data.data(), t.ncomps, pt.data());
        else
          OSH_LAUNCH(midpoint_coords, nkeys, c, t.ncomps, keys2edges.data(), m.down[1].data(),
              t.data.data(), pt.data());
        nw.vtags.push_back(t.with_data(transfer<Real>(c, mo, t.ncomps, t.data, &pt)));
      }
      transfer_class(c, m, nw, 0, mo, keys2edges, 1, nullptr);
    } else {
      transfer_class(c, m, nw, d, mo, psrc, d, &psrc_dim);
      if (d == dim) {
        transfer_elem_tags(c, m, nw, mo, &psrc);
        conserve_pass(c, m, nw, 0, 1, keys2edges, k2p[d], mo);
      }
      transfer_measures(c, m, nw, d, mo);
    }
    old_lows2new_lows = mo.old2new;
  }
  m = nw;
  return true;
}

}  // namespace osh
