// Mesh input, geometry and orderings. Each routine follows the reference algorithm named in
// mesh.hpp exactly (same tie-breaks, same summation order), because the cell numbering and the
// sweep orders decide the LU-SGS iterates bit for bit.

#include "mesh.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <set>

#include "json_lite.hpp"

namespace gsk {

int PolyMesh::find_patch(const std::string& name) const {
  for (std::size_t p = 0; p < patches.size(); ++p)
    if (patches[p].name == name) return static_cast<int>(p);
  return -1;
}

std::vector<int> PolyMesh::boundary_face_patches() const {
  std::vector<int> out(static_cast<std::size_t>(n_boundary_faces()), -1);
  for (std::size_t p = 0; p < patches.size(); ++p)
    for (Index f = patches[p].start; f < patches[p].end(); ++f)
      out[static_cast<std::size_t>(f - n_internal_faces())] = static_cast<int>(p);
  return out;
}

// ---------------------------------------------------------------------------------------------
// .pmesh v1 (docs/pmesh_format.md §1-§3)
// ---------------------------------------------------------------------------------------------

namespace {

bool known_patch_kind(const std::string& k) {
  return k == "wall" || k == "inlet" || k == "outlet" || k == "symmetry" || k == "patch";
}

template <class T>
std::vector<T> read_index_array(const std::string& bytes, const json::Value& desc, std::size_t expect_len,
                                const char* name) {
  const std::string dtype = desc.at("dtype").as_str();
  const auto off = static_cast<std::size_t>(desc.at("offset").as_int());
  const json::Value& shape = desc.at("shape");
  GSK_CHECK(shape.kind == json::Value::Array && shape.arr.size() == 1,
            strf("pmesh: array %s must be one-dimensional", name));
  const auto n = static_cast<std::size_t>(shape.arr[0].as_int());
  GSK_CHECK(expect_len == static_cast<std::size_t>(-1) || n == expect_len,
            strf("pmesh: array %s has length %zu, expected %zu", name, n, expect_len));
  std::vector<T> out(n);
  if (dtype == "int32") {
    GSK_CHECK(off + 4 * n <= bytes.size(), strf("pmesh: array %s runs past the end of the file", name));
    for (std::size_t i = 0; i < n; ++i) {
      std::int32_t v;
      std::memcpy(&v, bytes.data() + off + 4 * i, 4);
      out[i] = static_cast<T>(v);
    }
  } else if (dtype == "int64") {
    GSK_CHECK(off + 8 * n <= bytes.size(), strf("pmesh: array %s runs past the end of the file", name));
    for (std::size_t i = 0; i < n; ++i) {
      std::int64_t v;
      std::memcpy(&v, bytes.data() + off + 8 * i, 8);
      out[i] = static_cast<T>(v);
    }
  } else {
    GSK_CHECK(false, strf("pmesh: array %s has unsupported dtype \"%s\"", name, dtype.c_str()));
  }
  return out;
}

}  // namespace

PolyMesh read_pmesh(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  GSK_CHECK(in.good(), "cannot open mesh file \"" + path + "\"");
  std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  GSK_CHECK(bytes.size() >= 16 && std::memcmp(bytes.data(), "PMESH\0\0\0", 8) == 0,
            "\"" + path + "\" is not a .pmesh file (bad magic)");
  std::uint64_t hlen = 0;
  std::memcpy(&hlen, bytes.data() + 8, 8);
  GSK_CHECK(16 + hlen <= bytes.size(), "pmesh: header runs past the end of the file");
  const json::Value h = json::parse(bytes.substr(16, static_cast<std::size_t>(hlen)));
  GSK_CHECK(h.at("format").as_str() == "pmesh", "pmesh: format is not \"pmesh\"");
  GSK_CHECK(h.at("version").as_int() == 1, "pmesh: unsupported version (only v1 is known)");
  GSK_CHECK(!h.has("length_unit") || h.at("length_unit").as_str() == "m", "pmesh: length_unit must be \"m\"");
  PolyMesh m;
  const auto n_points = static_cast<std::size_t>(h.at("n_points").as_int());
  const auto n_faces = static_cast<std::size_t>(h.at("n_faces").as_int());
  const auto n_int = static_cast<std::size_t>(h.at("n_internal_faces").as_int());
  m.n_cells = static_cast<Index>(h.at("n_cells").as_int());
  const json::Value& arrays = h.at("arrays");
  {
    const json::Value& d = arrays.at("points");
    GSK_CHECK(d.at("dtype").as_str() == "float64", "pmesh: points must be float64");
    const auto off = static_cast<std::size_t>(d.at("offset").as_int());
    const json::Value& shape = d.at("shape");
    GSK_CHECK(shape.arr.size() == 2 && static_cast<std::size_t>(shape.arr[0].as_int()) == n_points &&
                  shape.arr[1].as_int() == 3,
              "pmesh: points must have shape [n_points, 3]");
    GSK_CHECK(off + 24 * n_points <= bytes.size(), "pmesh: points run past the end of the file");
    m.points.resize(n_points);
    for (std::size_t i = 0; i < n_points; ++i) {
      double v[3];
      std::memcpy(v, bytes.data() + off + 24 * i, 24);
      m.points[i] = {v[0], v[1], v[2]};
    }
  }
  m.face_offsets = read_index_array<Offset>(bytes, arrays.at("face_offsets"), n_faces + 1, "face_offsets");
  m.face_nodes = read_index_array<Index>(bytes, arrays.at("face_nodes"), static_cast<std::size_t>(-1), "face_nodes");
  m.owner = read_index_array<Index>(bytes, arrays.at("owner"), n_faces, "owner");
  m.neighbour = read_index_array<Index>(bytes, arrays.at("neighbour"), n_int, "neighbour");
  for (const json::Value& p : h.at("patches").arr) {
    Patch patch;
    patch.name = p.at("name").as_str();
    patch.kind = p.has("kind") ? p.at("kind").as_str() : "patch";
    patch.start = static_cast<Index>(p.at("start").as_int());
    patch.count = static_cast<Index>(p.at("count").as_int());
    m.patches.push_back(patch);
  }
  validate_topology(m);
  return m;
}

namespace {
void put_bytes(std::string& out, std::size_t off, const void* src, std::size_t n) {
  if (out.size() < off + n) out.resize(off + n, '\0');
  std::memcpy(out.data() + off, src, n);
}
}  // namespace

void write_pmesh(const std::string& path, const PolyMesh& m) {
  // compute offsets with 64-byte alignment after the header
  const std::size_t np = m.points.size(), nf = m.owner.size(), ni = m.neighbour.size();
  std::string patches = "[";
  for (std::size_t p = 0; p < m.patches.size(); ++p) {
    patches += strf("%s{\"name\": \"%s\", \"kind\": \"%s\", \"start\": %d, \"count\": %d}", p ? ", " : "",
                    m.patches[p].name.c_str(), m.patches[p].kind.c_str(), m.patches[p].start, m.patches[p].count);
  }
  patches += "]";
  auto header_for = [&](std::size_t base) {
    std::size_t off = (base + 63) / 64 * 64;
    const std::size_t o_pts = off; off = (off + 24 * np + 63) / 64 * 64;
    const std::size_t o_fo = off; off = (off + 8 * (nf + 1) + 63) / 64 * 64;
    const std::size_t o_fn = off; off = (off + 4 * m.face_nodes.size() + 63) / 64 * 64;
    const std::size_t o_ow = off; off = (off + 4 * nf + 63) / 64 * 64;
    const std::size_t o_nb = off;
    return strf("{\"format\": \"pmesh\", \"version\": 1, \"length_unit\": \"m\", \"n_points\": %zu, \"n_faces\": %zu, "
                "\"n_internal_faces\": %zu, \"n_cells\": %d, \"patches\": %s, \"arrays\": {"
                "\"points\": {\"dtype\": \"float64\", \"shape\": [%zu, 3], \"offset\": %zu}, "
                "\"face_offsets\": {\"dtype\": \"int64\", \"shape\": [%zu], \"offset\": %zu}, "
                "\"face_nodes\": {\"dtype\": \"int32\", \"shape\": [%zu], \"offset\": %zu}, "
                "\"owner\": {\"dtype\": \"int32\", \"shape\": [%zu], \"offset\": %zu}, "
                "\"neighbour\": {\"dtype\": \"int32\", \"shape\": [%zu], \"offset\": %zu}}}",
                np, nf, ni, m.n_cells, patches.c_str(), np, o_pts, nf + 1, o_fo, m.face_nodes.size(), o_fn, nf, o_ow, ni,
                o_nb);
  };
  // the header length changes the offsets; iterate until stable (at most a few rounds)
  std::string hdr = header_for(16);
  for (int it = 0; it < 8; ++it) {
    const std::string next = header_for(16 + hdr.size());
    if (next.size() == hdr.size()) { hdr = next; break; }
    hdr = next;
  }
  const json::Value h = json::parse(hdr);
  std::string out("PMESH\0\0\0", 8);
  const std::uint64_t hl = hdr.size();
  put_bytes(out, 8, &hl, 8);
  put_bytes(out, 16, hdr.data(), hdr.size());
  const json::Value& a = h.at("arrays");
  for (std::size_t i = 0; i < np; ++i) {
    const double v[3] = {m.points[i].x, m.points[i].y, m.points[i].z};
    put_bytes(out, static_cast<std::size_t>(a.at("points").at("offset").as_int()) + 24 * i, v, 24);
  }
  put_bytes(out, static_cast<std::size_t>(a.at("face_offsets").at("offset").as_int()), m.face_offsets.data(), 8 * (nf + 1));
  put_bytes(out, static_cast<std::size_t>(a.at("face_nodes").at("offset").as_int()), m.face_nodes.data(),
            4 * m.face_nodes.size());
  put_bytes(out, static_cast<std::size_t>(a.at("owner").at("offset").as_int()), m.owner.data(), 4 * nf);
  if (ni > 0) put_bytes(out, static_cast<std::size_t>(a.at("neighbour").at("offset").as_int()), m.neighbour.data(), 4 * ni);
  std::ofstream o(path, std::ios::binary);
  GSK_CHECK(o.good(), "cannot write mesh file \"" + path + "\"");
  o.write(out.data(), static_cast<std::streamsize>(out.size()));
}

void validate_topology(const PolyMesh& m) {
  const Index np = m.n_points(), nf = m.n_faces(), ni = m.n_internal_faces(), nc = m.n_cells;
  GSK_CHECK(nc > 0, "mesh has no cells");
  GSK_CHECK(np > 0, "mesh has no points");
  GSK_CHECK(ni <= nf, "more internal faces than faces");
  GSK_CHECK(m.face_offsets.size() == static_cast<std::size_t>(nf) + 1 && m.face_offsets.front() == 0 &&
                m.face_offsets.back() == static_cast<Offset>(m.face_nodes.size()),
            "face_offsets is inconsistent with face_nodes");
  for (Index f = 0; f < nf; ++f) GSK_CHECK(m.face_size(f) >= 3, strf("face %d has fewer than 3 nodes", f));
  for (Index f = 0; f < nf; ++f) {
    const Index* nodes = m.face(f);
    const Index n = m.face_size(f);
    for (Index i = 0; i < n; ++i) {
      GSK_CHECK(nodes[i] >= 0 && nodes[i] < np, strf("face %d node %d out of range", f, nodes[i]));
      for (Index j = 0; j < i; ++j) GSK_CHECK(nodes[i] != nodes[j], strf("face %d repeats node %d", f, nodes[i]));
    }
  }
  std::vector<Index> per_cell(static_cast<std::size_t>(nc), 0);
  for (Index f = 0; f < nf; ++f) {
    const Index o = m.owner[static_cast<std::size_t>(f)];
    GSK_CHECK(o >= 0 && o < nc, strf("face %d owner %d out of range", f, o));
    ++per_cell[static_cast<std::size_t>(o)];
  }
  for (Index f = 0; f < ni; ++f) {
    const Index o = m.owner[static_cast<std::size_t>(f)], n = m.neighbour[static_cast<std::size_t>(f)];
    GSK_CHECK(n >= 0 && n < nc, strf("face %d neighbour %d out of range", f, n));
    GSK_CHECK(o < n, strf("internal face %d: owner %d is not below neighbour %d", f, o, n));
    ++per_cell[static_cast<std::size_t>(n)];
    if (f > 0) {
      const Index po = m.owner[static_cast<std::size_t>(f) - 1], pn = m.neighbour[static_cast<std::size_t>(f) - 1];
      GSK_CHECK(po < o || (po == o && pn <= n), strf("internal faces not in (owner, neighbour) order at face %d", f));
    }
  }
  for (Index c = 0; c < nc; ++c)
    GSK_CHECK(per_cell[static_cast<std::size_t>(c)] >= 4, strf("cell %d has fewer than 4 faces", c));
  std::set<std::string> names;
  Index expect = ni;
  for (const Patch& p : m.patches) {
    GSK_CHECK(!p.name.empty(), "patch with an empty name");
    GSK_CHECK(names.insert(p.name).second, "duplicate patch name \"" + p.name + "\"");
    GSK_CHECK(known_patch_kind(p.kind), "patch \"" + p.name + "\" has unknown kind \"" + p.kind + "\"");
    GSK_CHECK(p.count >= 0 && p.start == expect, "patch \"" + p.name + "\" does not follow the previous one");
    expect = p.start + p.count;
  }
  GSK_CHECK(expect == nf, "patches do not cover the boundary faces exactly");
}

// ---------------------------------------------------------------------------------------------
// Geometry (docs/pmesh_format.md §4)
// ---------------------------------------------------------------------------------------------

namespace {
void face_geometry(const PolyMesh& m, Index f, Vec3& area_vector, Vec3& center) {
  const Index n = m.face_size(f);
  const Index* nodes = m.face(f);
  const auto P = [&](Index k) -> const Vec3& { return m.points[static_cast<std::size_t>(nodes[k])]; };
  if (n == 3) {
    area_vector = 0.5 * cross(P(1) - P(0), P(2) - P(0));
    center = (P(0) + P(1) + P(2)) / 3.0;
    return;
  }
  Vec3 xbar{};
  for (Index k = 0; k < n; ++k) xbar += P(k);
  xbar /= static_cast<double>(n);
  Vec3 sum_n{};
  for (Index k = 0; k < n; ++k) sum_n += cross(P(k) - xbar, P((k + 1) % n) - xbar);
  area_vector = 0.5 * sum_n;
  const double mag = norm(sum_n);
  if (!(mag > 0.0)) {
    center = xbar;
    return;
  }
  const Vec3 nhat = sum_n / mag;
  double sum_a = 0.0;
  Vec3 sum_ac{};
  for (Index k = 0; k < n; ++k) {
    const Vec3& pk = P(k);
    const Vec3& pk1 = P((k + 1) % n);
    const double a = dot(cross(pk - xbar, pk1 - xbar), nhat);
    const Vec3 c = (pk + pk1 + xbar) / 3.0;
    sum_a += a;
    sum_ac += a * c;
  }
  center = std::abs(sum_a) > 0.0 ? sum_ac / sum_a : xbar;
}
}  // namespace

MeshGeometry compute_geometry(const PolyMesh& m) {
  const Index nf = m.n_faces(), ni = m.n_internal_faces();
  const auto nc = static_cast<std::size_t>(m.n_cells);
  MeshGeometry g;
  g.face_area_vector.resize(static_cast<std::size_t>(nf));
  g.face_area.resize(static_cast<std::size_t>(nf));
  g.face_normal.resize(static_cast<std::size_t>(nf));
  g.face_center.resize(static_cast<std::size_t>(nf));
  for (Index f = 0; f < nf; ++f) {
    const auto fi = static_cast<std::size_t>(f);
    Vec3 S, c;
    face_geometry(m, f, S, c);
    const double a = norm(S);
    GSK_CHECK(a > 0.0, strf("face %d has zero area", f));
    g.face_area_vector[fi] = S;
    g.face_area[fi] = a;
    g.face_normal[fi] = S / a;
    g.face_center[fi] = c;
  }
  std::vector<Vec3> cbar(nc, Vec3{});
  std::vector<int> nfc(nc, 0);
  for (Index f = 0; f < nf; ++f) {
    const auto fi = static_cast<std::size_t>(f);
    const auto o = static_cast<std::size_t>(m.owner[fi]);
    cbar[o] += g.face_center[fi];
    ++nfc[o];
    if (f < ni) {
      const auto n = static_cast<std::size_t>(m.neighbour[fi]);
      cbar[n] += g.face_center[fi];
      ++nfc[n];
    }
  }
  for (std::size_t c = 0; c < nc; ++c) {
    GSK_CHECK(nfc[c] > 0, strf("cell %zu has no faces", c));
    cbar[c] /= static_cast<double>(nfc[c]);
  }
  g.cell_volume.assign(nc, 0.0);
  std::vector<Vec3> mom(nc, Vec3{});
  auto pyramid = [&](std::size_t c, const Vec3& s_out, const Vec3& xf) {
    const double v = dot(s_out, xf - cbar[c]) / 3.0;
    g.cell_volume[c] += v;
    mom[c] += v * (0.75 * xf + 0.25 * cbar[c]);
  };
  for (Index f = 0; f < nf; ++f) {
    const auto fi = static_cast<std::size_t>(f);
    pyramid(static_cast<std::size_t>(m.owner[fi]), g.face_area_vector[fi], g.face_center[fi]);
    if (f < ni) pyramid(static_cast<std::size_t>(m.neighbour[fi]), -g.face_area_vector[fi], g.face_center[fi]);
  }
  g.cell_center.resize(nc);
  for (std::size_t c = 0; c < nc; ++c) {
    const double v = g.cell_volume[c];
    g.cell_center[c] = std::abs(v) > 0.0 ? mom[c] / v : cbar[c];
  }
  return g;
}

// ---------------------------------------------------------------------------------------------
// CSR and least squares
// ---------------------------------------------------------------------------------------------

CellFaceCsr build_cell_faces(const PolyMesh& m) {
  const Index nf = m.n_faces(), ni = m.n_internal_faces();
  const auto nc = static_cast<std::size_t>(m.n_cells);
  CellFaceCsr csr;
  csr.offsets.assign(nc + 1, 0);
  for (Index f = 0; f < nf; ++f) {
    ++csr.offsets[static_cast<std::size_t>(m.owner[static_cast<std::size_t>(f)]) + 1];
    if (f < ni) ++csr.offsets[static_cast<std::size_t>(m.neighbour[static_cast<std::size_t>(f)]) + 1];
  }
  for (std::size_t c = 0; c < nc; ++c) csr.offsets[c + 1] += csr.offsets[c];
  const auto total = static_cast<std::size_t>(csr.offsets.back());
  csr.faces.resize(total);
  csr.other.resize(total);
  csr.sign.resize(total);
  std::vector<Offset> fill(csr.offsets.begin(), csr.offsets.end() - 1);
  for (Index f = 0; f < nf; ++f) {
    const auto fi = static_cast<std::size_t>(f);
    const Index o = m.owner[fi];
    const Index nb = f < ni ? m.neighbour[fi] : kInvalid;
    const auto k = static_cast<std::size_t>(fill[static_cast<std::size_t>(o)]++);
    csr.faces[k] = f;
    csr.other[k] = nb;
    csr.sign[k] = 1;
    if (nb != kInvalid) {
      const auto k2 = static_cast<std::size_t>(fill[static_cast<std::size_t>(nb)]++);
      csr.faces[k2] = f;
      csr.other[k2] = o;
      csr.sign[k2] = -1;
    }
  }
  return csr;
}

std::vector<Vec3> build_lsq_coefficients(const PolyMesh& m, const MeshGeometry& g, const CellFaceCsr& csr,
                                         double weight_power) {
  GSK_CHECK(weight_power >= 0.0, "lsq_weight_power must be >= 0");
  std::vector<Vec3> coeff(csr.faces.size());
  auto disp = [&](Index c, Offset k) {
    const auto ki = static_cast<std::size_t>(k);
    const Vec3& xc = g.cell_center[static_cast<std::size_t>(c)];
    if (csr.other[ki] != kInvalid) return g.cell_center[static_cast<std::size_t>(csr.other[ki])] - xc;
    return g.face_center[static_cast<std::size_t>(csr.faces[ki])] - xc;
  };
  auto weight = [&](const Vec3& r) { return weight_power == 0.0 ? 1.0 : std::pow(norm(r), -weight_power); };
  for (Index c = 0; c < m.n_cells; ++c) {
    // symmetric G = sum w r r^T, entries (00, 01, 02, 11, 12, 22)
    double a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0;
    for (Offset k = csr.begin(c); k < csr.end(c); ++k) {
      const Vec3 r = disp(c, k);
      const double w = weight(r);
      a00 += w * r.x * r.x;
      a01 += w * r.x * r.y;
      a02 += w * r.x * r.z;
      a11 += w * r.y * r.y;
      a12 += w * r.y * r.z;
      a22 += w * r.z * r.z;
    }
    const double det = a00 * (a11 * a22 - a12 * a12) - a01 * (a01 * a22 - a12 * a02) + a02 * (a01 * a12 - a11 * a02);
    const double tr3 = (a00 + a11 + a22) / 3.0;
    const double cond = tr3 > 0.0 ? det / (tr3 * tr3 * tr3) : 0.0;
    GSK_CHECK(cond > 1e-12, strf("cell %d: least-squares matrix is near singular (cond %.3e)", c, cond));
    const double i00 = (a11 * a22 - a12 * a12) / det;
    const double i01 = (a02 * a12 - a01 * a22) / det;
    const double i02 = (a01 * a12 - a02 * a11) / det;
    const double i11 = (a00 * a22 - a02 * a02) / det;
    const double i12 = (a01 * a02 - a00 * a12) / det;
    const double i22 = (a00 * a11 - a01 * a01) / det;
    for (Offset k = csr.begin(c); k < csr.end(c); ++k) {
      const Vec3 r = disp(c, k);
      const Vec3 gr{i00 * r.x + i01 * r.y + i02 * r.z, i01 * r.x + i11 * r.y + i12 * r.z,
                    i02 * r.x + i12 * r.y + i22 * r.z};
      coeff[static_cast<std::size_t>(k)] = weight(r) * gr;
    }
  }
  return coeff;
}

// ---------------------------------------------------------------------------------------------
// Orderings
// ---------------------------------------------------------------------------------------------

CellAdjacency build_cell_adjacency(const PolyMesh& m) {
  const auto nc = static_cast<std::size_t>(m.n_cells);
  const Index ni = m.n_internal_faces();
  std::vector<Offset> off(nc + 1, 0);
  for (Index f = 0; f < ni; ++f) {
    ++off[static_cast<std::size_t>(m.owner[static_cast<std::size_t>(f)]) + 1];
    ++off[static_cast<std::size_t>(m.neighbour[static_cast<std::size_t>(f)]) + 1];
  }
  for (std::size_t c = 0; c < nc; ++c) off[c + 1] += off[c];
  std::vector<Index> raw(static_cast<std::size_t>(off.back()));
  std::vector<Offset> fill(off.begin(), off.end() - 1);
  for (Index f = 0; f < ni; ++f) {
    const Index o = m.owner[static_cast<std::size_t>(f)], n = m.neighbour[static_cast<std::size_t>(f)];
    raw[static_cast<std::size_t>(fill[static_cast<std::size_t>(o)]++)] = n;
    raw[static_cast<std::size_t>(fill[static_cast<std::size_t>(n)]++)] = o;
  }
  CellAdjacency a;
  a.offsets.assign(nc + 1, 0);
  a.cells.reserve(raw.size());
  for (std::size_t c = 0; c < nc; ++c) {
    auto b = raw.begin() + off[c], e = raw.begin() + off[c + 1];
    std::sort(b, e);
    e = std::unique(b, e);
    a.cells.insert(a.cells.end(), b, e);
    a.offsets[c + 1] = static_cast<Offset>(a.cells.size());
  }
  return a;
}

// rcm_ordering and rcb_ordering: ordering.cpp

std::vector<Index> invert_permutation(const std::vector<Index>& perm) {
  std::vector<Index> inv(perm.size(), kInvalid);
  for (std::size_t i = 0; i < perm.size(); ++i) {
    const Index p = perm[i];
    GSK_CHECK(p >= 0 && static_cast<std::size_t>(p) < perm.size() && inv[static_cast<std::size_t>(p)] == kInvalid,
              strf("not a permutation (position %zu, value %d)", i, p));
    inv[static_cast<std::size_t>(p)] = static_cast<Index>(i);
  }
  return inv;
}

PolyMesh renumber_cells(const PolyMesh& m, const std::vector<Index>& new_to_old) {
  GSK_CHECK(new_to_old.size() == static_cast<std::size_t>(m.n_cells), "permutation length differs from cell count");
  const std::vector<Index> o2n = invert_permutation(new_to_old);
  const Index ni = m.n_internal_faces();
  struct Ref {
    Index face, owner, neighbour;
    bool flip;
  };
  std::vector<Ref> internal(static_cast<std::size_t>(ni));
  for (Index f = 0; f < ni; ++f) {
    Index o = o2n[static_cast<std::size_t>(m.owner[static_cast<std::size_t>(f)])];
    Index n = o2n[static_cast<std::size_t>(m.neighbour[static_cast<std::size_t>(f)])];
    const bool flip = o > n;
    if (flip) std::swap(o, n);
    internal[static_cast<std::size_t>(f)] = {f, o, n, flip};
  }
  std::stable_sort(internal.begin(), internal.end(), [](const Ref& a, const Ref& b) {
    return a.owner != b.owner ? a.owner < b.owner : a.neighbour < b.neighbour;
  });
  PolyMesh out;
  out.points = m.points;
  out.patches = m.patches;
  out.n_cells = m.n_cells;
  out.face_offsets.assign(1, 0);
  out.face_nodes.reserve(m.face_nodes.size());
  auto append = [&](Index f, bool flip) {
    const Index* nodes = m.face(f);
    const Index n = m.face_size(f);
    if (!flip) {
      out.face_nodes.insert(out.face_nodes.end(), nodes, nodes + n);
    } else {
      out.face_nodes.push_back(nodes[0]);
      for (Index k = n - 1; k >= 1; --k) out.face_nodes.push_back(nodes[k]);
    }
    out.face_offsets.push_back(static_cast<Offset>(out.face_nodes.size()));
  };
  for (const Ref& r : internal) {
    append(r.face, r.flip);
    out.owner.push_back(r.owner);
    out.neighbour.push_back(r.neighbour);
  }
  for (const Patch& p : m.patches) {
    std::vector<std::pair<Index, Index>> faces;
    for (Index f = p.start; f < p.end(); ++f)
      faces.emplace_back(o2n[static_cast<std::size_t>(m.owner[static_cast<std::size_t>(f)])], f);
    std::stable_sort(faces.begin(), faces.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (const auto& [o, f] : faces) {
      append(f, false);
      out.owner.push_back(o);
    }
  }
  return out;
}

std::vector<Index> match_boundary_faces(const PolyMesh& orig, const PolyMesh& ren) {
  const Index ni_o = orig.n_internal_faces(), ni_r = ren.n_internal_faces();
  GSK_CHECK(ni_o == ni_r && orig.patches.size() == ren.patches.size(), "renumbered mesh differs in shape");
  std::vector<Index> new_to_old(static_cast<std::size_t>(orig.n_boundary_faces()), kInvalid);
  for (std::size_t pi = 0; pi < orig.patches.size(); ++pi) {
    const Patch& po = orig.patches[pi];
    const Patch& pr = ren.patches[pi];
    std::map<std::vector<Index>, Index> idx;
    for (Index f = po.start; f < po.end(); ++f)
      idx.emplace(std::vector<Index>(orig.face(f), orig.face(f) + orig.face_size(f)), f);
    for (Index f = pr.start; f < pr.end(); ++f) {
      const auto it = idx.find(std::vector<Index>(ren.face(f), ren.face(f) + ren.face_size(f)));
      GSK_CHECK(it != idx.end(), strf("renumbered boundary face %d not found in the original mesh", f));
      new_to_old[static_cast<std::size_t>(f - ni_r)] = it->second - ni_o;
    }
  }
  return new_to_old;
}

std::vector<Index> partition_contiguous(Index n_cells, int n_blocks) {
  GSK_CHECK(n_blocks >= 1 && n_cells >= n_blocks, "invalid block count");
  std::vector<Index> off(static_cast<std::size_t>(n_blocks) + 1, 0);
  const Index base = n_cells / n_blocks, rem = n_cells % n_blocks;
  for (int b = 0; b < n_blocks; ++b)
    off[static_cast<std::size_t>(b) + 1] = off[static_cast<std::size_t>(b)] + base + (b < rem ? 1 : 0);
  return off;
}

}  // namespace gsk
