#pragma once
// Boundary state of the CIS iteration, host FP64: one inflow state (scale x Maxwellian) per boundary
// face, and what each round of boundary sums makes of it (face fluxes F+ / F-, face states, patch
// mass flows, wall densities). Reference behaviour (src/kinetic/kinetic_solver.cpp):
//   states at rest, discrete-equilibrium inflow        :103-142
//   wall densities from the cells (first prepare)      :156-164
//   pressure-face states (reconstruction +             :172-260
//     characteristic condition)
//   Maxwell-wall unit-density sums                     :263-277
//   boundary sums -> fluxes, states, wall densities    :280-335, :456-480

#include <array>
#include <vector>

#include "setup.hpp"

namespace gsk {

// The discrete conservative Maxwellian of an inflow state (rho, u, T) in FP64: replaces the state by
// the Maxwellian whose discrete moments on the velocity grid are (rho, rho u, rho E), when the grid
// does not resolve it and the Newton solve succeeds; otherwise leaves it.
void discrete_inflow_state(const KineticSetup& s, double tolerance, gsk_inflow_state& state);

class BoundaryModel {
 public:
  explicit BoundaryModel(const KineticSetup& s);

  std::vector<PatchCondition>& patches() { return patches_; }
  const std::vector<PatchCondition>& patches() const { return patches_; }
  const PatchCondition& patch_of(Index b) const {
    return patches_[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(b)])];
  }

  // States before the first prepare: wall faces carry the unit-density wall Maxwellian (discrete when
  // the case uses the discrete equilibrium), pressure faces (p / T, 0, T), the others nothing. Maxwell
  // walls also get their unit-density half-range sums.
  void reset(double inflow_tolerance);
  // First prepare only: wall densities from the owner cells, alpha rho_c sqrt(T_c / T_wall).
  void seed_wall_densities(const std::vector<double>& moments);
  // New states of the pressure faces from the owner cells' moments; returns the faces set.
  std::vector<Index> update_pressure_faces(const std::vector<double>& moments);
  // One round of boundary sums (20 per face): fluxes, face states, patch mass flows, and the wall
  // densities they imply (diffuse walls relaxed by wall_relaxation, Maxwell walls by mass balance).
  void absorb(const std::vector<double>& sums);
  // Wall densities only, from the sums of the wall faces (the refresh between inner sweeps).
  void refresh_walls(const std::vector<double>& sums, const std::vector<int>& wall_faces);
  // Branches of the pressure-face reconstruction's limiter (see reconstruct_face_state): Off, Record
  // (compute normally and keep every choice) or Follow (take the recorded choices), for
  // sensitivity solves that differentiate one smooth branch.
  enum class Branches { Off, Record, Follow };
  void set_branches(Branches b) { branches_ = b; }
  void forget_face_states() { has_face_state_ = false; }
  bool has_face_states() const { return has_face_state_; }
  // the face states were set elsewhere (the device prepare, which runs the same rounds)
  void mark_face_states() { has_face_state_ = true; }

  std::vector<gsk_inflow_state> inflow;
  std::vector<std::array<double, 5>> flux_out, flux_in, face_state;  // F+, F-, (rho, u, T) per face
  std::vector<double> patch_flow;
  std::vector<std::array<double, 10>> wall_unit;  // Maxwell walls: sum over incoming xi of w M (xi.n) psi, w M psi

 private:
  // The limited linear reconstruction of (p, u, T) at pressure face b from its owner cell.
  std::array<double, 5> reconstruct_face_state(Index b, const std::vector<double>& moments) const;
  std::array<double, 10> unit_wall_sums(Index b) const;
  double diffuse_wall_density(double out_flux, double unit_in_flux) const;

  const KineticSetup& s_;
  std::vector<PatchCondition> patches_;  // mutable copy: mass-flow inlets adjust p, Maxwell walls alpha
  bool has_face_state_ = false;
  // per boundary face and variable: neighbour (CSR slot) whose limiter value set phi (-1: phi = 1),
  // its clamp and the sign of its projection, the slots of the largest and smallest neighbour
  // value (-1: the owner's), and whether the positivity fallback kept the owner's state
  struct FaceBranch {
    int phi_slot[5], hi_slot[5], lo_slot[5];
    unsigned clamp[5];
    bool positive_side[5];
    bool owner_state;
  };
  Branches branches_ = Branches::Off;
  mutable std::vector<FaceBranch> branch_;
  FaceBranch& face_branch(Index b) const;
};

}  // namespace gsk
