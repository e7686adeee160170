// Device-resident arrays of the C++ facade: Read/Write live in GPU memory
// (data() is a device pointer), tags cross between the mesh and them by
// device-to-device copies, HostRead/HostWrite mirror them on the host, and
// an array keeps its GPU context alive after the Library that made it
// closes.  Prints "ok" and exits 0 when every check holds.  Used by
// tests/test_facade.py.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "Omega_h_amd.hpp"

using namespace Omega_h;

static int failures = 0;

static void expect(bool ok, char const* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

static bool on_device(void const* p) {
  int d = 0;
  check_status(osh_pointer_is_device(p, &d));
  return d == 1;
}

int main() {
  try {
    Reals kept;
    std::vector<Real> kept_host;
    {
      Library lib;
      Mesh mesh = build_box(lib.world(), OMEGA_H_SIMPLEX, 1.0, 1.0, 1.0, 4, 4, 4);
      LO nv = mesh.nverts();
      // a metric written on the host and sent to the device
      HostWrite<Real> hm(nv);
      for (LO v = 0; v < nv; ++v) hm[v] = 16.0 + v % 3;
      Reals metric(hm.write());
      expect(on_device(metric.data()), "a Read made from HostWrite::write() is on the device");
      mesh.add_tag(VERT, "metric", 1, metric);
      Reals coords = mesh.coords();
      expect(on_device(coords.data()), "Mesh::coords() is on the device");
      // a tag set from a device array and read back
      mesh.add_tag(VERT, "target_metric", 1, metric);
      Reals back = mesh.get_array<Real>(VERT, "target_metric");
      expect(on_device(back.data()) && back.data() != metric.data(),
          "get_array returns a device copy of the tag");
      HostRead<Real> hb(back);
      bool same = hb.size() == nv;
      for (LO v = 0; same && v < nv; ++v) same = hb[v] == hm[v];
      expect(same, "a vertex tag round-trips through device memory unchanged");
      // fills, ranges, single items, shared storage
      Write<Real> w(7, 2.5);
      HostRead<Real> hw(w);
      bool filled = hw.size() == 7;
      for (LO i = 0; filled && i < 7; ++i) filled = hw[i] == 2.5;
      expect(filled, "Write(size, value) fills on the device");
      Read<LO> r(5, 3, 2);
      expect(r.get(0) == 3 && r.last() == 11, "Read(size, offset, stride) is a range");
      Read<Real> shared(w);
      w.set(4, -1.0);
      expect(shared.get(4) == -1.0, "a Read made from a Write shares its storage");
      HostWrite<Real> hw2(w);
      hw2[0] = 9.0;
      Write<Real> w2 = hw2.write();
      expect(w2.data() == w.data() && w.get(0) == 9.0,
          "HostWrite(Write).write() writes back into the same device array");
      // reductions on the device agree with the host mirror
      HostRead<Real> hq(mesh.ask_qualities());
      Real qmin = 1.0;
      for (LO i = 0; i < hq.size(); ++i) qmin = hq[i] < qmin ? hq[i] : qmin;
      expect(mesh.min_quality() == qmin, "min_quality equals the host minimum");
      HostRead<Real> hl(mesh.ask_lengths());
      Real lmax = 0.0;
      for (LO i = 0; i < hl.size(); ++i) lmax = hl[i] > lmax ? hl[i] : lmax;
      expect(mesh.max_length() == lmax, "max_length equals the host maximum");
      // an array taken before adapt() keeps the old values
      HostRead<Real> hc(coords);
      kept_host.assign(hc.data(), hc.data() + hc.size());
      kept = coords;
      AdaptOpts opts(&mesh);
      opts.verbosity = SILENT;
      adapt(&mesh, opts);
      expect(mesh.nverts() != nv, "adapt() changed the mesh");
      HostRead<GO> g(mesh.globals(VERT));
      expect(g.size() == mesh.nverts() && g.last() == GO(mesh.nverts() - 1),
          "globals are the indices");
    }
    // the Library is closed; the array still holds its block and context
    HostRead<Real> hk(kept);
    bool unchanged = std::size_t(hk.size()) == kept_host.size();
    for (LO i = 0; unchanged && i < hk.size(); ++i) unchanged = hk[i] == kept_host[std::size_t(i)];
    expect(unchanged, "an array outlives its Library and the mesh's adapt()");
    bool refused = false;
    try {
      Write<Real> none(3);
    } catch (exception const&) {
      refused = true;
    }
    expect(refused, "a new array without an open Library is refused");
  } catch (exception const& e) {
    std::printf("Omega_h::exception: %s\n", e.what());
    return 1;
  }
  std::printf(failures ? "FAILED\n" : "ok\n");
  return failures ? 1 : 0;
}
