// The .osh binary mesh format (Omega_h's own, version 9), one rank.
//
//   binary::write / read (stream)   src/Omega_h_file.cpp:364-470
//   write_array / read_array        src/Omega_h_file.cpp:117-186 (zlib
//                                   compress2 at Z_BEST_SPEED)
//   write_meta / read_meta          src/Omega_h_file.cpp:200-262
//   write_tag / read_tag            src/Omega_h_file.cpp:264-318
//   write / read (directory)        src/Omega_h_file.cpp:472-600: <dir>/0.osh,
//                                   <dir>/nparts, <dir>/version
// A directory holds the mesh's down adjacencies with their alignment codes
// and its tags per dimension.  Written here: coordinates, metric and the
// vertex tags carried through adapt(), class_id and class_dim, and global
// numbers (the index on one rank).  Read: the same, plus any other float64
// vertex tag as a linearly interpolated vertex tag; other tags are skipped.
#include <sys/stat.h>
#include <zlib.h>

#include <array>
#include <cerrno>
#include <cstring>
#include <fstream>
#include <sstream>

#include "osh_mesh.hpp"

namespace osh {

namespace {

constexpr unsigned char kMagic[2] = {0xa1, 0x1a};
constexpr std::int32_t kVersion = 9;
enum TagType : std::int8_t { kI8 = 0, kI32 = 2, kI64 = 3, kF64 = 5 };

template <class T>
void put(std::ostream& s, T v) {
  s.write(reinterpret_cast<char const*>(&v), sizeof(T));
}

template <class T>
T get(std::istream& s) {
  T v;
  s.read(reinterpret_cast<char*>(&v), sizeof(T));
  OSH_CHECK(bool(s), "truncated .osh file");
  return v;
}

template <class T>
void put_array(std::ostream& s, std::vector<T> const& a, bool compressed) {
  put<std::int32_t>(s, std::int32_t(a.size()));
  uLong nbytes = uLong(a.size() * sizeof(T));
  if (!compressed) {
    s.write(reinterpret_cast<char const*>(a.data()), std::streamsize(nbytes));
    return;
  }
  uLong bound = compressBound(nbytes);
  std::vector<Bytef> z(bound);
  int r = compress2(z.data(), &bound, reinterpret_cast<Bytef const*>(a.data()), nbytes,
      Z_BEST_SPEED);
  OSH_CHECK(r == Z_OK, "zlib compression failed");
  put<std::int64_t>(s, std::int64_t(bound));
  s.write(reinterpret_cast<char const*>(z.data()), std::streamsize(bound));
}

template <class T>
std::vector<T> get_array(std::istream& s, bool compressed) {
  std::int32_t n = get<std::int32_t>(s);
  OSH_CHECK(n >= 0, "negative array size in .osh file");
  std::vector<T> a(static_cast<std::size_t>(n));
  uLong nbytes = uLong(a.size() * sizeof(T));
  if (!compressed) {
    s.read(reinterpret_cast<char*>(a.data()), std::streamsize(nbytes));
    OSH_CHECK(bool(s), "truncated .osh file");
    return a;
  }
  std::int64_t zbytes = get<std::int64_t>(s);
  OSH_CHECK(zbytes >= 0, "negative compressed size in .osh file");
  std::vector<Bytef> z(static_cast<std::size_t>(zbytes));
  s.read(reinterpret_cast<char*>(z.data()), std::streamsize(zbytes));
  OSH_CHECK(bool(s), "truncated .osh file");
  uLong out = nbytes;
  int r = uncompress(reinterpret_cast<Bytef*>(a.data()), &out, z.data(), uLong(zbytes));
  OSH_CHECK(r == Z_OK && out == nbytes, "zlib decompression failed");
  return a;
}

void put_string(std::ostream& s, std::string const& v) {
  put<std::int32_t>(s, std::int32_t(v.size()));
  s.write(v.data(), std::streamsize(v.size()));
}

std::string get_string(std::istream& s) {
  std::int32_t n = get<std::int32_t>(s);
  OSH_CHECK(n >= 0 && n < (1 << 20), "bad string in .osh file");
  std::string v(std::size_t(n), '\0');
  s.read(&v[0], n);
  OSH_CHECK(bool(s), "truncated .osh file");
  return v;
}

template <class T>
void put_tag(std::ostream& s, std::string const& name, int ncomps, std::int8_t type,
    std::vector<T> const& data, bool compressed) {
  put_string(s, name);
  put<std::int8_t>(s, std::int8_t(ncomps));
  put<std::int8_t>(s, type);
  put_array(s, data, compressed);
}

void write_int_file(std::string const& path, int v) {
  std::ofstream f(path);
  OSH_CHECK(f.is_open(), "cannot write " + path);
  f << v << '\n';
}

int read_int_file(std::string const& path) {
  std::ifstream f(path);
  OSH_CHECK(f.is_open(), "cannot open " + path);
  int v;
  f >> v;
  OSH_CHECK(bool(f), "cannot read " + path);
  return v;
}

}  // namespace

void write_osh(Ctx& c, DMesh& m, std::string const& dir) {
  if (::mkdir(dir.c_str(), 0777) != 0)
    OSH_CHECK(errno == EEXIST, "cannot create directory " + dir);
  std::ofstream f(dir + "/0.osh", std::ios::binary);
  OSH_CHECK(f.is_open(), "cannot write " + dir + "/0.osh");
  bool const z = true;
  f.write(reinterpret_cast<char const*>(kMagic), 2);
  put<std::int8_t>(f, z ? 1 : 0);
  // meta: family (simplex), dim, one rank, element-based parting, no ghost
  // layers, no RIB hints
  put<std::int8_t>(f, 0);
  put<std::int8_t>(f, std::int8_t(m.dim));
  put<std::int32_t>(f, 1);
  put<std::int32_t>(f, 0);
  put<std::int8_t>(f, 0);
  put<std::int32_t>(f, 0);
  put<std::int8_t>(f, 0);
  put<std::int32_t>(f, m.nents[0]);
  for (int d = 1; d <= m.dim; ++d) {
    put_array(f, to_host(c, m.down[d]), z);
    if (d > 1) put_array(f, to_host(c, m.down_codes[d]), z);
  }
  for (int d = 0; d <= m.dim; ++d) {
    std::vector<std::int64_t> globals(static_cast<std::size_t>(m.nents[d]));
    for (std::size_t i = 0; i < globals.size(); ++i) globals[i] = std::int64_t(i);
    std::int32_t ntags = 1 + (m.class_id[d].exists() ? 1 : 0) + (m.class_dim[d].exists() ? 1 : 0);
    if (d == 0) ntags += 1 + (m.metric.exists() ? 1 : 0) + std::int32_t(m.vtags.size());
    ntags += std::int32_t(m.itags.size());
    put<std::int32_t>(f, ntags);
    if (d == 0) {
      put_tag(f, "coordinates", m.dim, kF64, to_host(c, m.coords), z);
      if (m.metric.exists()) put_tag(f, "metric", m.metric_ncomps, kF64, to_host(c, m.metric), z);
      for (auto const& t : m.vtags) put_tag(f, t.name, t.ncomps, kF64, to_host(c, t.data), z);
    }
    for (auto const& t : m.itags) put_tag(f, t.name, t.ncomps, kF64, to_host(c, t.data[d]), z);
    if (m.class_id[d].exists()) put_tag(f, "class_id", 1, kI32, to_host(c, m.class_id[d]), z);
    if (m.class_dim[d].exists()) put_tag(f, "class_dim", 1, kI8, to_host(c, m.class_dim[d]), z);
    put_tag(f, "global", 1, kI64, globals, z);
  }
  put<std::int32_t>(f, 0);  // no class sets
  put<std::int8_t>(f, 0);   // no parents
  OSH_CHECK(bool(f), "write failed: " + dir + "/0.osh");
  f.close();
  write_int_file(dir + "/nparts", 1);
  write_int_file(dir + "/version", kVersion);
}

void read_osh(Ctx& c, DMesh& m, std::string const& dir) {
  OSH_CHECK(read_int_file(dir + "/nparts") == 1, "only one-part .osh meshes can be read");
  int version = read_int_file(dir + "/version");
  OSH_CHECK(version >= 1 && version <= kVersion, "unsupported .osh version");
  std::ifstream f(dir + "/0.osh", std::ios::binary);
  OSH_CHECK(f.is_open(), "cannot open " + dir + "/0.osh");
  unsigned char magic[2];
  f.read(reinterpret_cast<char*>(magic), 2);
  OSH_CHECK(bool(f) && magic[0] == kMagic[0] && magic[1] == kMagic[1], "not an .osh file");
  bool z = get<std::int8_t>(f) != 0;
  if (version >= 7) OSH_CHECK(get<std::int8_t>(f) == 0, "only simplex meshes are supported");
  int dim = get<std::int8_t>(f);
  OSH_CHECK(dim == 2 || dim == 3, "only 2D and 3D meshes are supported");
  OSH_CHECK(get<std::int32_t>(f) == 1, "only one-rank meshes can be read");
  get<std::int32_t>(f);  // rank
  get<std::int8_t>(f);   // parting
  if (version >= 3) get<std::int32_t>(f);
  if (get<std::int8_t>(f)) {
    std::int32_t naxes = get<std::int32_t>(f);
    for (std::int32_t i = 0; i < 3 * naxes; ++i) get<double>(f);
  }
  if (version < 6) get<std::int8_t>(f);
  DMesh r;
  r.dim = dim;
  r.nents[0] = get<std::int32_t>(f);
  OSH_CHECK(r.nents[0] >= 0, "negative vertex count");
  for (int d = 1; d <= dim; ++d) {
    auto down = get_array<LO>(f, z);
    OSH_CHECK(down.size() % std::size_t(d + 1) == 0, "bad adjacency size");
    LO nlow = r.nents[d - 1];
    for (LO v : down) OSH_CHECK(v >= 0 && v < nlow, "adjacency index out of range");
    r.nents[d] = LO(down.size() / std::size_t(d + 1));
    r.down[d] = to_device(c, down);
    if (d > 1) {
      auto codes = get_array<I8>(f, z);
      OSH_CHECK(codes.size() == down.size(), "bad alignment code count");
      // a down code is make_code(flip, rotation, 0) with rotation below the
      // side's vertex count (d) and no flip for edges (src/Omega_h_align.hpp:43-54)
      for (I8 code : codes)
        OSH_CHECK(code >= 0 && code_which_down(code) == 0 && code_rotation(code) < d &&
                      (d > 2 || !code_is_flipped(code)),
            "alignment code out of range");
      r.down_codes[d] = to_device(c, codes);
    }
  }
  // float64 tags of each dimension by name: present on every dimension they
  // are inheritance tags, on vertices only they are vertex tags
  std::vector<std::pair<std::string, int>> names;
  std::vector<std::array<Reals, 4>> per_dim;
  for (int d = 0; d <= dim; ++d) {
    std::int32_t ntags = get<std::int32_t>(f);
    for (std::int32_t t = 0; t < ntags; ++t) {
      std::string name = get_string(f);
      int ncomps = get<std::int8_t>(f);
      std::int8_t type = get<std::int8_t>(f);
      if (version < 5) {
        get<std::int8_t>(f);
        if (version >= 2) get<std::int8_t>(f);
      }
      std::size_t want = std::size_t(r.nents[d]) * std::size_t(ncomps);
      if (type == kF64) {
        auto a = get_array<Real>(f, z);
        OSH_CHECK(a.size() == want, "tag " + name + " has the wrong size");
        if (d != 0) {
          for (std::size_t k = 0; k < names.size(); ++k)
            if (names[k].first == name && names[k].second == ncomps) per_dim[k][d] = to_device(c, a);
          continue;
        }
        if (name == "coordinates") {
          OSH_CHECK(ncomps == dim, "coordinates must have dim components");
          r.coords = to_device(c, a);
        } else if (name == "metric") {
          OSH_CHECK(ncomps == 1 || ncomps == dim * (dim + 1) / 2,
              "metric must have 1 or dim*(dim+1)/2 components");
          r.metric = to_device(c, a);
          r.metric_ncomps = ncomps;
        } else {
          names.push_back({name, ncomps});
          per_dim.emplace_back();
          per_dim.back()[0] = to_device(c, a);
        }
      } else if (type == kI32) {
        auto a = get_array<std::int32_t>(f, z);
        OSH_CHECK(a.size() == want, "tag " + name + " has the wrong size");
        if (name == "class_id") r.class_id[d] = to_device(c, a);
      } else if (type == kI8) {
        auto a = get_array<I8>(f, z);
        OSH_CHECK(a.size() == want, "tag " + name + " has the wrong size");
        if (name == "class_dim") r.class_dim[d] = to_device(c, a);
      } else if (type == kI64) {
        get_array<std::int64_t>(f, z);
      } else {
        OSH_CHECK(false, "unknown tag type in .osh file");
      }
    }
  }
  for (std::size_t k = 0; k < names.size(); ++k) {
    bool everywhere = true;
    for (int d = 1; d <= dim; ++d) everywhere = everywhere && per_dim[k][d].exists();
    std::string const& name = names[k].first;
    if (everywhere) {
      DMesh::InheritTag t;
      t.name = name;
      t.ncomps = names[k].second;
      for (int d = 0; d <= dim; ++d) t.data[d] = per_dim[k][d];
      r.itags.push_back(t);
    } else {
      bool is_metric = name.size() >= 6 && name.compare(name.size() - 6, 6, "metric") == 0;
      r.vtags.push_back({name, names[k].second, is_metric ? 1 : 0, per_dim[k][0]});
    }
  }
  OSH_CHECK(r.coords.exists(), "the mesh has no coordinates");
  // classification is all or nothing: every dimension carries both tags, or
  // none does (the passes read both on every dimension)
  bool any_class = false, all_class = true;
  for (int d = 0; d <= dim; ++d) {
    for (bool has : {r.class_id[d].exists(), r.class_dim[d].exists()}) {
      any_class = any_class || has;
      all_class = all_class && has;
    }
  }
  OSH_CHECK(all_class || !any_class, "class_id/class_dim missing on some dimensions");
  c.sync();
  m = r;
}

}  // namespace osh
