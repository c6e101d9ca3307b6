#pragma once
// Case input: a TOML subset reader, the case sections the kinetic path reads, the derived
// nondimensional scales and the [kinetic] parameters with their validation.
//
// Reference behaviour followed:
//   sections / keys / boundary types      src/config/case_config.cpp:312-460, :540-556
//   derived scales (mu_hat0, Kn, Mdot)    src/config/case_config.cpp:633-663
//   [kinetic] keys, defaults, checks      src/kinetic/kinetic_params.hpp:11-114, kinetic_params.cpp:9-92
//   case-vs-mesh checks                   src/config/case_mesh_check.cpp:10-70

#include <array>
#include <map>
#include <optional>
#include <set>
#include <string>
#include <variant>
#include <vector>

#include "core.hpp"

namespace gsk {

// ---- TOML subset: tables, arrays of tables, scalars and (nested) arrays ----
struct TomlValue {
  enum Kind { Bool, Int, Float, String, Array } kind = Int;
  bool b = false;
  long long i = 0;
  double d = 0.0;
  std::string s;
  std::vector<TomlValue> arr;
  double number() const;
};
using TomlTable = std::map<std::string, TomlValue>;
struct TomlDoc {
  std::map<std::string, TomlTable> tables;                    // [name]
  std::map<std::string, std::vector<TomlTable>> table_arrays;  // [[name]]
  std::vector<std::string> order;                             // section names in file order
};
TomlDoc parse_toml(const std::string& text, const std::string& source);

enum class BoundaryType { MaxwellWall, MassFlowInlet, DiffuseWall, PressureInlet, PressureOutlet, Vacuum, Symmetry, PeriodicSlab };
enum class SolverMode { Gsis, Cis, Nsf };

struct BoundarySpec {
  std::string patch;
  BoundaryType type = BoundaryType::DiffuseWall;
  double accommodation = 1.0;
  std::optional<double> mass_flow, temperature, pressure;
  std::optional<Vec3> velocity;
  bool inlet_zero_tangential = false;
};

struct CaseConfig {
  std::string source_dir;  // directory of the TOML file (mesh paths are relative to it)
  std::string name, mesh;
  SolverMode mode = SolverMode::Gsis;
  // [gas]
  double molecular_mass = 0, mu_ref = 0, T_ref = 0, omega = 0, Pr = 2.0 / 3.0, gamma = 5.0 / 3.0;
  // [conditions]
  double L_ref = 0, T0 = 0;
  std::optional<double> delta_rp, Kn, p_ref;
  std::vector<BoundarySpec> boundaries;
  // [velocity_grid]
  std::array<int, 3> vn{};
  std::array<double, 3> vlo{}, vhi{};
  double vstretch = 0.0;  // velocity_grid.stretch: tanh stretching strength (0 = the reference's uniform grid)
  // [kinetic] raw values (validated by parse_kinetic_params)
  TomlTable kinetic;
  // [gsis] values used by CIS-only runs
  double gsis_tolerance = 1e-5;
  double gsis_mass_balance_tolerance = 2e-3;  // [gsis].mass_balance_tolerance (cis-run's balance check)
  bool gsis_mass_balance_required = false;    // [gsis].mass_balance_required
  // [output].mass_flow_sections ("0.2, 0.5 0.8": positions in L_ref units) and mass_flow_section_axis
  std::vector<double> section_positions;
  int section_axis = 0;

  double gas_constant() const { return kBoltzmann / molecular_mass; }
  double viscosity(double T) const { return mu_ref * std::pow(T / T_ref, omega); }
  const BoundarySpec* find_boundary(const std::string& patch) const;
  std::string mesh_path() const;
};

CaseConfig parse_case_toml(const std::string& text, const std::string& source);
CaseConfig load_case(const std::string& path);

struct DerivedQuantities {
  double R = 0, mu0 = 0, most_probable_speed = 0, p_ref = 0, p_in = 0, rho_in = 0, Kn = 0, delta_from_Kn = 0;
  double mu_hat0 = 0, mass_flow_scale = 0;
  bool has_p_in = false;
  // reference scales
  double ref_length = 0, ref_temperature = 0, ref_pressure = 0, ref_density = 0, ref_velocity = 0;
  double ref_mass_flux = 0, ref_heat_flux = 0;
};
DerivedQuantities derive_quantities(const CaseConfig& c);

struct KineticParams {
  double pseudo_cfl = 0.0;
  std::string pseudo_time = "global";
  double limiter_k = 3.0;
  std::string limiter = "venkatakrishnan";
  std::string limiter_eps = "relative";
  double lsq_weight_power = 0.0;
  bool positivity_limiter = false;
  long long velocity_chunk = 0;
  long long sweep_blocks = 0;
  std::string cell_ordering = "rcb";
  std::string precision = "fp32";
  std::string moment_sums = "fp64";  // extension: "mfma" = matrix-core moment contraction (FP32 form)
  std::string backend = "cuda";
  long long max_iterations = 100000;
  double initial_pressure = 1.0;
  double initial_temperature = 1.0;
  double convergence_u_floor = 1.0e-2;
  double wall_relaxation = 1.0;
  long long boundary_iterations = 1;
  double interface_damping = 0.5;
  std::string gsis_negative_check = "blend";
  long long launch_cells = 16384;
  long long sweep_cells_per_launch = 4096;
  long long inner_sweeps = 4;
  bool inner_wall_update = true;
  long long block_passes = 2;
  std::string sweep_order = "octant";
  std::string equilibrium = "discrete";
  std::string deq_fallback = "collisionless";
  std::string pressure_boundary_reconstruction = "limited";
  double cold_bgk_temperature = 0.0;
};
// Reads [kinetic] with the reference defaults; unknown keys and out-of-range values are errors.
// backend accepts "cuda" (the GPU backend, here HIP), "hip" and "cpu".
KineticParams parse_kinetic_params(const TomlTable& t);
void validate_kinetic_params(const KineticParams& p);

}  // namespace gsk
