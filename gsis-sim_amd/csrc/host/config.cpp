// TOML subset reader and case/kinetic parameter parsing (see config.hpp for the reference
// locations each part follows).

#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <limits>
#include <sstream>

namespace gsk {

double TomlValue::number() const {
  if (kind == Int) return static_cast<double>(i);
  GSK_CHECK(kind == Float, "TOML value is not a number");
  return d;
}

// ---------------------------------------------------------------------------------------------
// TOML subset
// ---------------------------------------------------------------------------------------------
namespace {

struct Cursor {
  const std::string& s;
  std::size_t i = 0;
  int line = 1;
  const std::string& src;
  [[noreturn]] void error(const std::string& msg) const { fail_at(__FILE__, __LINE__, strf("%s:%d: ", src.c_str(), line) + msg); }
  bool eof() const { return i >= s.size(); }
  char peek() const { return eof() ? '\0' : s[i]; }
  void skip_ws_comments(bool newlines) {
    while (!eof()) {
      const char c = s[i];
      if (c == ' ' || c == '\t' || c == '\r') {
        ++i;
      } else if (c == '#') {
        while (!eof() && s[i] != '\n') ++i;
      } else if (c == '\n' && newlines) {
        ++i;
        ++line;
      } else {
        break;
      }
    }
  }
};

std::string parse_key(Cursor& c) {
  c.skip_ws_comments(false);
  std::string k;
  if (c.peek() == '"') {
    ++c.i;
    while (!c.eof() && c.peek() != '"') k += c.s[c.i++];
    if (c.eof()) c.error("unterminated quoted key");
    ++c.i;
    return k;
  }
  while (!c.eof() && (std::isalnum(static_cast<unsigned char>(c.peek())) || c.peek() == '_' || c.peek() == '-'))
    k += c.s[c.i++];
  if (k.empty()) c.error("expected a key");
  return k;
}

TomlValue parse_value(Cursor& c) {
  c.skip_ws_comments(false);
  TomlValue v;
  const char ch = c.peek();
  if (ch == '"') {
    ++c.i;
    v.kind = TomlValue::String;
    while (true) {
      if (c.eof() || c.peek() == '\n') c.error("unterminated string");
      const char x = c.s[c.i++];
      if (x == '"') break;
      if (x == '\\') {
        const char e = c.s[c.i++];
        v.s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
      } else {
        v.s += x;
      }
    }
    return v;
  }
  if (ch == '\'') {
    ++c.i;
    v.kind = TomlValue::String;
    while (!c.eof() && c.peek() != '\'') v.s += c.s[c.i++];
    if (c.eof()) c.error("unterminated literal string");
    ++c.i;
    return v;
  }
  if (ch == '[') {
    ++c.i;
    v.kind = TomlValue::Array;
    while (true) {
      c.skip_ws_comments(true);
      if (c.peek() == ']') {
        ++c.i;
        break;
      }
      v.arr.push_back(parse_value(c));
      c.skip_ws_comments(true);
      if (c.peek() == ',') {
        ++c.i;
        continue;
      }
      if (c.peek() == ']') {
        ++c.i;
        break;
      }
      c.error("expected ',' or ']' in array");
    }
    return v;
  }
  if (c.s.compare(c.i, 4, "true") == 0) {
    c.i += 4;
    v.kind = TomlValue::Bool;
    v.b = true;
    return v;
  }
  if (c.s.compare(c.i, 5, "false") == 0) {
    c.i += 5;
    v.kind = TomlValue::Bool;
    v.b = false;
    return v;
  }
  std::string tok;
  while (!c.eof() && (std::isalnum(static_cast<unsigned char>(c.peek())) || c.peek() == '+' || c.peek() == '-' ||
                      c.peek() == '.' || c.peek() == '_'))
    tok += c.s[c.i++];
  tok.erase(std::remove(tok.begin(), tok.end(), '_'), tok.end());
  if (tok.empty()) c.error("expected a value");
  const bool is_float = tok.find_first_of(".eE") != std::string::npos || tok == "inf" || tok == "nan";
  try {
    if (is_float) {
      v.kind = TomlValue::Float;
      v.d = std::stod(tok);
    } else {
      v.kind = TomlValue::Int;
      v.i = std::stoll(tok);
    }
  } catch (...) {
    c.error("bad number \"" + tok + "\"");
  }
  return v;
}

}  // namespace

TomlDoc parse_toml(const std::string& text, const std::string& source) {
  TomlDoc doc;
  Cursor c{text, 0, 1, source};
  TomlTable* cur = nullptr;
  while (true) {
    c.skip_ws_comments(true);
    if (c.eof()) break;
    if (c.peek() == '[') {
      const bool arr = c.s.compare(c.i, 2, "[[") == 0;
      c.i += arr ? 2 : 1;
      const std::string name = parse_key(c);
      c.skip_ws_comments(false);
      if (arr) {
        if (c.s.compare(c.i, 2, "]]") != 0) c.error("expected ']]'");
        c.i += 2;
        auto& list = doc.table_arrays[name];
        if (list.empty()) doc.order.push_back(name);
        list.emplace_back();
        cur = &list.back();
      } else {
        if (c.peek() != ']') c.error("expected ']'");
        ++c.i;
        if (doc.tables.count(name) || doc.table_arrays.count(name)) c.error("section [" + name + "] defined twice");
        doc.order.push_back(name);
        cur = &doc.tables[name];
      }
      continue;
    }
    const std::string key = parse_key(c);
    c.skip_ws_comments(false);
    if (c.peek() != '=') c.error("expected '=' after key \"" + key + "\"");
    ++c.i;
    TomlValue v = parse_value(c);
    if (cur == nullptr) c.error("key \"" + key + "\" outside of a section");
    if (cur->count(key)) c.error("duplicate key \"" + key + "\"");
    (*cur)[key] = std::move(v);
    c.skip_ws_comments(false);
    if (!c.eof() && c.peek() != '\n') c.error("unexpected text after value of \"" + key + "\"");
  }
  return doc;
}

// ---------------------------------------------------------------------------------------------
// Case sections
// ---------------------------------------------------------------------------------------------
namespace {

struct Reader {
  Reader(const TomlTable& table, std::string n) : t(table), name(std::move(n)) {}
  const TomlTable& t;
  std::string name;
  std::set<std::string> used;
  const TomlValue* find(const std::string& k) {
    auto it = t.find(k);
    if (it == t.end()) return nullptr;
    used.insert(k);
    return &it->second;
  }
  double num(const std::string& k) {
    const TomlValue* v = find(k);
    GSK_CHECK(v != nullptr, name + "." + k + " is required");
    GSK_CHECK(v->kind == TomlValue::Int || v->kind == TomlValue::Float, name + "." + k + " must be a number");
    return v->number();
  }
  std::optional<double> opt_num(const std::string& k) {
    if (!t.count(k)) return std::nullopt;
    return num(k);
  }
  std::string str(const std::string& k) {
    const TomlValue* v = find(k);
    GSK_CHECK(v != nullptr && v->kind == TomlValue::String, name + "." + k + " must be a string");
    return v->s;
  }
  std::optional<std::string> opt_str(const std::string& k) {
    if (!t.count(k)) return std::nullopt;
    return str(k);
  }
  std::optional<Vec3> opt_vec3(const std::string& k) {
    const TomlValue* v = find(k);
    if (v == nullptr) return std::nullopt;
    GSK_CHECK(v->kind == TomlValue::Array && v->arr.size() == 3, name + "." + k + " must be an array of 3 numbers");
    return Vec3{v->arr[0].number(), v->arr[1].number(), v->arr[2].number()};
  }
  void finish() {
    for (const auto& kv : t)
      GSK_CHECK(used.count(kv.first), "unknown key " + name + "." + kv.first);
  }
};

void require_positive(double v, const std::string& what) { GSK_CHECK(v > 0.0 && std::isfinite(v), what + " must be positive"); }

BoundaryType boundary_type(const std::string& s, const std::string& where) {
  if (s == "maxwell_wall") return BoundaryType::MaxwellWall;
  if (s == "mass_flow_inlet") return BoundaryType::MassFlowInlet;
  if (s == "diffuse_wall") return BoundaryType::DiffuseWall;
  if (s == "pressure_inlet") return BoundaryType::PressureInlet;
  if (s == "pressure_outlet") return BoundaryType::PressureOutlet;
  if (s == "vacuum") return BoundaryType::Vacuum;
  if (s == "symmetry") return BoundaryType::Symmetry;
  if (s == "periodic_slab") return BoundaryType::PeriodicSlab;
  GSK_CHECK(false, where + ".type = \"" + s + "\" is not a known boundary condition");
  return BoundaryType::DiffuseWall;
}

}  // namespace

const BoundarySpec* CaseConfig::find_boundary(const std::string& patch) const {
  for (const BoundarySpec& b : boundaries)
    if (b.patch == patch) return &b;
  return nullptr;
}

std::string CaseConfig::mesh_path() const {
  if (mesh.empty() || mesh[0] == '/' || source_dir.empty()) return mesh;
  return source_dir + "/" + mesh;
}

CaseConfig parse_case_toml(const std::string& text, const std::string& source) {
  const TomlDoc doc = parse_toml(text, source);
  CaseConfig cfg;
  const auto slash = source.find_last_of('/');
  cfg.source_dir = slash == std::string::npos ? std::string() : source.substr(0, slash);
  static const char* known[] = {"case", "gas", "conditions", "boundary", "velocity_grid", "gsis", "kinetic", "macro", "output"};
  for (const std::string& s : doc.order) {
    bool ok = false;
    for (const char* k : known) ok = ok || s == k;
    if (!ok) std::fprintf(stderr, "%s: ignoring unknown section [%s]\n", source.c_str(), s.c_str());
  }
  for (const char* req : {"case", "gas", "conditions", "velocity_grid"})
    GSK_CHECK(doc.tables.count(req), source + ": required section [" + req + "] is missing");
  GSK_CHECK(doc.table_arrays.count("boundary") && !doc.table_arrays.at("boundary").empty(),
            source + ": no [[boundary]] entries");
  {
    Reader r{doc.tables.at("case"), "case"};
    cfg.name = r.str("name");
    cfg.mesh = r.str("mesh");
    const std::string mode = r.opt_str("mode").value_or("gsis");
    GSK_CHECK(mode == "gsis" || mode == "cis" || mode == "nsf", "case.mode must be gsis, cis or nsf");
    cfg.mode = mode == "cis" ? SolverMode::Cis : mode == "nsf" ? SolverMode::Nsf : SolverMode::Gsis;
    r.finish();
  }
  {
    Reader r{doc.tables.at("gas"), "gas"};
    cfg.molecular_mass = r.num("molecular_mass");
    cfg.mu_ref = r.num("mu_ref");
    cfg.T_ref = r.num("T_ref");
    cfg.omega = r.num("omega");
    r.finish();
    require_positive(cfg.molecular_mass, "gas.molecular_mass");
    require_positive(cfg.mu_ref, "gas.mu_ref");
    require_positive(cfg.T_ref, "gas.T_ref");
    require_positive(cfg.omega, "gas.omega");
    GSK_CHECK(cfg.molecular_mass < 1e-20, "gas.molecular_mass is too large for a molecular mass in kg");
  }
  {
    Reader r{doc.tables.at("conditions"), "conditions"};
    cfg.L_ref = r.num("L_ref");
    cfg.T0 = r.num("T0");
    cfg.delta_rp = r.opt_num("delta_rp");
    cfg.Kn = r.opt_num("Kn");
    cfg.p_ref = r.opt_num("p_ref");
    r.finish();
    require_positive(cfg.L_ref, "conditions.L_ref");
    require_positive(cfg.T0, "conditions.T0");
    const int given = int(cfg.delta_rp.has_value()) + int(cfg.Kn.has_value()) + int(cfg.p_ref.has_value());
    GSK_CHECK(given == 1, "conditions needs exactly one of delta_rp, Kn, p_ref");
    if (cfg.delta_rp) require_positive(*cfg.delta_rp, "conditions.delta_rp");
    if (cfg.Kn) require_positive(*cfg.Kn, "conditions.Kn");
    if (cfg.p_ref) require_positive(*cfg.p_ref, "conditions.p_ref");
  }
  const auto& blist = doc.table_arrays.at("boundary");
  for (std::size_t i = 0; i < blist.size(); ++i) {
    const std::string nm = strf("boundary[%zu]", i);
    Reader r{blist[i], nm};
    BoundarySpec b;
    b.patch = r.str("patch");
    GSK_CHECK(!b.patch.empty(), nm + ".patch is empty");
    b.type = boundary_type(r.str("type"), nm);
    switch (b.type) {
      case BoundaryType::MaxwellWall:
        b.accommodation = r.num("accommodation");
        GSK_CHECK(b.accommodation > 0.0 && b.accommodation <= 1.0, nm + ".accommodation must be in (0, 1]");
        [[fallthrough]];
      case BoundaryType::DiffuseWall:
        b.temperature = r.num("temperature");
        b.velocity = r.opt_vec3("velocity");
        break;
      case BoundaryType::MassFlowInlet:
      case BoundaryType::PressureInlet: {
        b.temperature = r.num("temperature");
        if (b.type == BoundaryType::MassFlowInlet) {
          b.mass_flow = r.num("mass_flow");
          require_positive(*b.mass_flow, nm + ".mass_flow");
        } else {
          b.pressure = r.opt_num("pressure");
        }
        const std::string tv = r.opt_str("tangential_velocity").value_or("extrapolate");
        GSK_CHECK(tv == "extrapolate" || tv == "zero", nm + ".tangential_velocity must be \"extrapolate\" or \"zero\"");
        b.inlet_zero_tangential = tv == "zero";
        break;
      }
      case BoundaryType::PressureOutlet:
        b.pressure = r.num("pressure");
        break;
      default:
        break;
    }
    r.finish();
    if (b.temperature) require_positive(*b.temperature, nm + ".temperature");
    if (b.pressure) {
      if (b.type == BoundaryType::PressureOutlet)
        GSK_CHECK(*b.pressure >= 0.0, nm + ".pressure must be >= 0");
      else
        require_positive(*b.pressure, nm + ".pressure");
    }
    GSK_CHECK(cfg.find_boundary(b.patch) == nullptr, "duplicate boundary for patch \"" + b.patch + "\"");
    GSK_CHECK(b.type != BoundaryType::PressureInlet || b.pressure || cfg.delta_rp,
              nm + ": pressure_inlet may omit pressure only when conditions.delta_rp is given");
    cfg.boundaries.push_back(b);
  }
  {
    Reader r{doc.tables.at("velocity_grid"), "velocity_grid"};
    const TomlValue* aut = r.find("auto");
    GSK_CHECK(aut == nullptr || aut->kind == TomlValue::Bool, "velocity_grid.auto must be true or false");
    const TomlValue* n = r.find("n");
    const TomlValue* range = r.find("range");
    long long total = 1;
    if (aut != nullptr && aut->b) {
      // velocity_grid.auto (case_config.cpp:589-610): a cube [-L, L]^3 with L = u_max + 5 sqrt(R T_max)
      // (the fastest wall plus five thermal speeds of the hottest boundary) and an even point count
      // per axis, 2 ceil(L / (0.7 sqrt(R T_min))) (at least 2), so the coldest boundary Maxwellian
      // has about 1.4 points per thermal speed; T_min / T_max over the boundary temperatures (T0
      // when none is given), u_max over the wall velocities. In units of sqrt(R T0).
      GSK_CHECK(n == nullptr && range == nullptr, "velocity_grid.auto = true cannot be combined with n or range");
      double t_min = std::numeric_limits<double>::infinity(), t_max = 0.0, u_max = 0.0;
      for (const BoundarySpec& b : cfg.boundaries) {
        if (b.temperature) {
          t_min = std::min(t_min, *b.temperature);
          t_max = std::max(t_max, *b.temperature);
        }
        if (b.velocity) u_max = std::max(u_max, norm(*b.velocity));
      }
      if (t_max == 0.0) t_min = t_max = cfg.T0;
      const double R = cfg.gas_constant();
      const double limit = u_max + 5.0 * std::sqrt(R * t_max);
      const double points = std::max(2.0, 2.0 * std::ceil(limit / (0.7 * std::sqrt(R * t_min))));
      GSK_CHECK(std::isfinite(points) && points * points * points <= 100000000.0,
                "velocity_grid.auto gives too many velocity points (at most 1e8)");
      const double half = limit / std::sqrt(R * cfg.T0);
      GSK_CHECK(std::isfinite(half), "velocity_grid.auto gives a range that is not finite");
      for (std::size_t a = 0; a < 3; ++a) {
        cfg.vn[a] = static_cast<int>(points);
        cfg.vlo[a] = -half;
        cfg.vhi[a] = half;
        total *= cfg.vn[a];
      }
    } else {
      GSK_CHECK(n && n->kind == TomlValue::Array && n->arr.size() == 3, "velocity_grid.n must be [nx, ny, nz]");
      GSK_CHECK(range && range->kind == TomlValue::Array && range->arr.size() == 3,
                "velocity_grid.range must be [[lo, hi] x 3]");
      for (int a = 0; a < 3; ++a) {
        GSK_CHECK(n->arr[static_cast<std::size_t>(a)].kind == TomlValue::Int, "velocity_grid.n must hold integers");
        cfg.vn[static_cast<std::size_t>(a)] = static_cast<int>(n->arr[static_cast<std::size_t>(a)].i);
        const TomlValue& ra = range->arr[static_cast<std::size_t>(a)];
        GSK_CHECK(ra.kind == TomlValue::Array && ra.arr.size() == 2, "velocity_grid.range entries must be [lo, hi]");
        cfg.vlo[static_cast<std::size_t>(a)] = ra.arr[0].number();
        cfg.vhi[static_cast<std::size_t>(a)] = ra.arr[1].number();
        GSK_CHECK(cfg.vlo[static_cast<std::size_t>(a)] < cfg.vhi[static_cast<std::size_t>(a)], "velocity_grid.range needs lo < hi");
        GSK_CHECK(cfg.vn[static_cast<std::size_t>(a)] >= 1, "velocity_grid.n must be >= 1");
        total *= cfg.vn[static_cast<std::size_t>(a)];
      }
    }
    GSK_CHECK(total <= 100000000, "too many velocity points");
    // extension: tanh-stretched grid (no reference counterpart; 0 keeps the uniform midpoint grid)
    if (const TomlValue* st = r.find("stretch")) {
      cfg.vstretch = st->number();
      GSK_CHECK(std::isfinite(cfg.vstretch) && cfg.vstretch >= 0.0 && cfg.vstretch <= 10.0,
                "velocity_grid.stretch must be in [0, 10]");
    }
    r.finish();
  }
  if (doc.tables.count("gsis")) {
    const TomlTable& g = doc.tables.at("gsis");
    auto it = g.find("tolerance");
    if (it != g.end()) cfg.gsis_tolerance = it->second.number();
    it = g.find("mass_balance_tolerance");
    if (it != g.end()) cfg.gsis_mass_balance_tolerance = it->second.number();
    require_positive(cfg.gsis_mass_balance_tolerance, "gsis.mass_balance_tolerance");
    it = g.find("mass_balance_required");
    if (it != g.end()) {
      GSK_CHECK(it->second.kind == TomlValue::Bool, "gsis.mass_balance_required must be true or false");
      cfg.gsis_mass_balance_required = it->second.b;
    }
  }
  if (doc.tables.count("output")) {  // section positions for the mass-balance record (section_cut.cpp:11-38)
    const TomlTable& o = doc.tables.at("output");
    auto ax = o.find("mass_flow_section_axis");
    if (ax != o.end()) {
      GSK_CHECK(ax->second.kind == TomlValue::String, "output.mass_flow_section_axis must be a string");
      const std::string& a = ax->second.s;
      GSK_CHECK(a == "x" || a == "X" || a == "y" || a == "Y" || a == "z" || a == "Z",
                "output.mass_flow_section_axis must be \"x\", \"y\" or \"z\" (value: \"" + a + "\")");
      cfg.section_axis = (a == "x" || a == "X") ? 0 : (a == "y" || a == "Y") ? 1 : 2;
    }
    auto ps = o.find("mass_flow_sections");
    if (ps != o.end()) {
      GSK_CHECK(ps->second.kind == TomlValue::String, "output.mass_flow_sections must be a string of positions");
      const std::string& t = ps->second.s;
      auto sep = [](char c) { return c == ',' || c == ';' || c == ' ' || c == '\t'; };
      std::size_t i = 0;
      while (i < t.size()) {
        while (i < t.size() && sep(t[i])) ++i;
        if (i >= t.size()) break;
        std::size_t j = i;
        while (j < t.size() && !sep(t[j])) ++j;
        const std::string tok = t.substr(i, j - i);
        char* end = nullptr;
        const double v = std::strtod(tok.c_str(), &end);
        GSK_CHECK(end == tok.c_str() + tok.size() && std::isfinite(v),
                  "section position \"" + tok + "\" is not a number (list: \"" + t + "\")");
        cfg.section_positions.push_back(v);
        i = j;
      }
    }
  }
  if (doc.tables.count("kinetic")) cfg.kinetic = doc.tables.at("kinetic");
  return cfg;
}

CaseConfig load_case(const std::string& path) {
  std::ifstream in(path);
  GSK_CHECK(in.good(), "cannot open case file \"" + path + "\"");
  std::stringstream ss;
  ss << in.rdbuf();
  return parse_case_toml(ss.str(), path);
}

DerivedQuantities derive_quantities(const CaseConfig& c) {
  DerivedQuantities d;
  d.R = c.gas_constant();
  d.mu0 = c.viscosity(c.T0);
  d.most_probable_speed = std::sqrt(2.0 * d.R * c.T0);
  const double kn_factor = d.mu0 / c.L_ref * std::sqrt(kPi * d.R * c.T0 / 2.0);
  if (c.delta_rp) {
    d.p_ref = *c.delta_rp * d.mu0 * d.most_probable_speed / c.L_ref;
    d.has_p_in = true;
  } else if (c.Kn) {
    d.p_ref = kn_factor / *c.Kn;
  } else {
    d.p_ref = *c.p_ref;
  }
  d.p_in = d.p_ref;
  d.rho_in = d.p_in / (d.R * c.T0);
  d.Kn = kn_factor / d.p_ref;
  d.delta_from_Kn = kSqrtPi / (2.0 * d.Kn);
  d.ref_length = c.L_ref;
  d.ref_temperature = c.T0;
  d.ref_pressure = d.p_ref;
  d.ref_density = d.ref_pressure / (d.R * d.ref_temperature);
  d.ref_velocity = std::sqrt(d.R * d.ref_temperature);
  d.ref_mass_flux = d.ref_density * d.ref_velocity;
  d.ref_heat_flux = d.ref_pressure * d.ref_velocity;
  const double ref_viscosity = d.ref_pressure * d.ref_length / d.ref_velocity;
  d.mu_hat0 = d.mu0 / ref_viscosity;
  d.mass_flow_scale = kSqrtPi * c.L_ref * c.L_ref * d.p_in / d.most_probable_speed;
  return d;
}

// ---------------------------------------------------------------------------------------------
// [kinetic]
// ---------------------------------------------------------------------------------------------

KineticParams parse_kinetic_params(const TomlTable& t) {
  KineticParams p;
  std::set<std::string> used;
  auto get = [&](const char* k) -> const TomlValue* {
    auto it = t.find(k);
    if (it == t.end()) return nullptr;
    used.insert(k);
    return &it->second;
  };
  auto num = [&](const char* k, double& dst) {
    if (const TomlValue* v = get(k)) {
      GSK_CHECK(v->kind == TomlValue::Int || v->kind == TomlValue::Float, strf("kinetic.%s must be a number", k));
      dst = v->number();
    }
  };
  auto integer = [&](const char* k, long long& dst) {
    if (const TomlValue* v = get(k)) {
      GSK_CHECK(v->kind == TomlValue::Int, strf("kinetic.%s must be an integer", k));
      dst = v->i;
    }
  };
  auto str = [&](const char* k, std::string& dst) {
    if (const TomlValue* v = get(k)) {
      GSK_CHECK(v->kind == TomlValue::String, strf("kinetic.%s must be a string", k));
      dst = v->s;
    }
  };
  auto boolean = [&](const char* k, bool& dst) {
    if (const TomlValue* v = get(k)) {
      GSK_CHECK(v->kind == TomlValue::Bool, strf("kinetic.%s must be a boolean", k));
      dst = v->b;
    }
  };
  num("pseudo_cfl", p.pseudo_cfl);
  str("pseudo_time", p.pseudo_time);
  num("limiter_k", p.limiter_k);
  str("limiter", p.limiter);
  str("limiter_eps", p.limiter_eps);
  num("lsq_weight_power", p.lsq_weight_power);
  boolean("positivity_limiter", p.positivity_limiter);
  integer("velocity_chunk", p.velocity_chunk);
  integer("sweep_blocks", p.sweep_blocks);
  str("cell_ordering", p.cell_ordering);
  str("precision", p.precision);
  str("moment_sums", p.moment_sums);
  str("backend", p.backend);
  integer("max_iterations", p.max_iterations);
  num("initial_pressure", p.initial_pressure);
  num("initial_temperature", p.initial_temperature);
  num("convergence_u_floor", p.convergence_u_floor);
  num("wall_relaxation", p.wall_relaxation);
  integer("boundary_iterations", p.boundary_iterations);
  num("interface_damping", p.interface_damping);
  str("gsis_negative_check", p.gsis_negative_check);
  integer("launch_cells", p.launch_cells);
  integer("sweep_cells_per_launch", p.sweep_cells_per_launch);
  integer("inner_sweeps", p.inner_sweeps);
  boolean("inner_wall_update", p.inner_wall_update);
  integer("block_passes", p.block_passes);
  str("sweep_order", p.sweep_order);
  str("equilibrium", p.equilibrium);
  str("deq_fallback", p.deq_fallback);
  str("pressure_boundary_reconstruction", p.pressure_boundary_reconstruction);
  num("cold_bgk_temperature", p.cold_bgk_temperature);
  for (const auto& kv : t) GSK_CHECK(used.count(kv.first), "unknown key kinetic." + kv.first);
  validate_kinetic_params(p);
  return p;
}

void validate_kinetic_params(const KineticParams& p) {
  auto one_of = [](const std::string& v, std::initializer_list<const char*> allowed) {
    for (const char* a : allowed)
      if (v == a) return true;
    return false;
  };
  GSK_CHECK(p.pseudo_cfl >= 0.0, "kinetic.pseudo_cfl must be >= 0");
  GSK_CHECK(one_of(p.pseudo_time, {"global", "velocity"}), "kinetic.pseudo_time must be \"global\" or \"velocity\"");
  GSK_CHECK(p.limiter_k >= 0.0, "kinetic.limiter_k must be >= 0");
  GSK_CHECK(one_of(p.limiter, {"venkatakrishnan", "none"}), "kinetic.limiter must be \"venkatakrishnan\" or \"none\"");
  GSK_CHECK(one_of(p.limiter_eps, {"absolute", "relative"}), "kinetic.limiter_eps must be \"absolute\" or \"relative\"");
  GSK_CHECK(p.lsq_weight_power >= 0.0, "kinetic.lsq_weight_power must be >= 0");
  GSK_CHECK(p.velocity_chunk >= 0, "kinetic.velocity_chunk must be >= 0");
  GSK_CHECK(p.sweep_blocks >= 0, "kinetic.sweep_blocks must be >= 0");
  GSK_CHECK(one_of(p.cell_ordering, {"rcb", "rcm", "none"}), "kinetic.cell_ordering must be rcb, rcm or none");
  GSK_CHECK(one_of(p.precision, {"fp32", "fp64"}), "kinetic.precision must be \"fp32\" or \"fp64\"");
  GSK_CHECK(one_of(p.moment_sums, {"fp64", "mfma"}), "kinetic.moment_sums must be \"fp64\" or \"mfma\"");
  GSK_CHECK(p.moment_sums == "fp64" || p.precision == "fp32",
            "kinetic.moment_sums = \"mfma\" needs precision = \"fp32\" (the FP64 form sums in FP64)");
  GSK_CHECK(one_of(p.backend, {"cuda", "hip", "cpu"}), "kinetic.backend must be \"cuda\", \"hip\" or \"cpu\"");
  GSK_CHECK(p.max_iterations >= 1, "kinetic.max_iterations must be >= 1");
  GSK_CHECK(p.initial_pressure > 0.0, "kinetic.initial_pressure must be positive");
  GSK_CHECK(p.initial_temperature > 0.0, "kinetic.initial_temperature must be positive");
  GSK_CHECK(p.convergence_u_floor >= 0.0, "kinetic.convergence_u_floor must be >= 0");
  GSK_CHECK(p.wall_relaxation > 0.0 && p.wall_relaxation <= 1.0, "kinetic.wall_relaxation must be in (0, 1]");
  GSK_CHECK(p.boundary_iterations >= 1 && p.boundary_iterations <= 1000, "kinetic.boundary_iterations must be 1..1000");
  GSK_CHECK(p.interface_damping >= 0.0, "kinetic.interface_damping must be >= 0");
  GSK_CHECK(one_of(p.gsis_negative_check, {"all", "created", "limit", "blend"}),
            "kinetic.gsis_negative_check must be all, created, limit or blend");
  GSK_CHECK(p.launch_cells >= 1, "kinetic.launch_cells must be >= 1");
  GSK_CHECK(p.sweep_cells_per_launch >= 1, "kinetic.sweep_cells_per_launch must be >= 1");
  GSK_CHECK(p.inner_sweeps >= 1 && p.inner_sweeps <= 1000, "kinetic.inner_sweeps must be 1..1000");
  GSK_CHECK(p.block_passes >= 1 && p.block_passes <= 100, "kinetic.block_passes must be 1..100");
  GSK_CHECK(one_of(p.sweep_order, {"block", "octant", "dir13"}), "kinetic.sweep_order must be block, octant or dir13");
  GSK_CHECK(one_of(p.equilibrium, {"continuous", "discrete"}), "kinetic.equilibrium must be continuous or discrete");
  GSK_CHECK(one_of(p.deq_fallback, {"collisionless", "continuous"}), "kinetic.deq_fallback must be collisionless or continuous");
  GSK_CHECK(one_of(p.pressure_boundary_reconstruction, {"limited", "cell"}),
            "kinetic.pressure_boundary_reconstruction must be limited or cell");
  GSK_CHECK(p.cold_bgk_temperature >= 0.0, "kinetic.cold_bgk_temperature must be >= 0");
}

}  // namespace gsk
