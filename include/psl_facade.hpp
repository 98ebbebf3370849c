// psl_facade.hpp -- header-only C++ facade with the reference's spellings over
// the C ABI of include/psl.h.
//
// A host driver written against the reference's include/ API
// (flowAroundSphere/flowAroundSphereMainSingle.cu) keeps its call sequence:
//
//   GridStruct grid(info);                 // include/types.h:77
//   grid.setMarkers(getMarkers);           // getMarkers callback, evaluated on the host
//   grid.setGivenRhoUxUyUz(getGiven);      // getGivenRhoUxUyUz callback
//   fillEquilibriumFromFunction(grid, getInitialRhoUxUyUz);
//   for (...) { applyStreaming(grid); applyLocalCellUpdate(grid); }
//
// The per-cell callbacks of the reference are __cuda_callable__ functions
// compiled into its kernels.  Here they are ordinary host functions evaluated
// once per cell into the flag bytes / value table that the HIP kernels read;
// call setMarkers again whenever their inputs change.
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "psl.h"

namespace psl_facade {

inline void check(int rc) {
  if (rc != PSL_OK) throw std::runtime_error(std::string("psl: ") + psl_last_error());
}

// MarkerStruct (include/types.h:79-80)
struct MarkerStruct {
  bool fluid = false, bounceback = false, movingBounceback = false, forcedVelocity = false;
  bool periodicX = false, periodicY = false, periodicZ = false, givenRho = false, givenUxUyUz = false;
  bool nonReflectiveOutlet = false, refinement = false, ghost = false;
};

using InfoStruct = psl_info;

// Flag byte of a resolved marker (the order of tests in
// include/applyLocalCellUpdate.h:17-67).
inline uint8_t flag_of(const MarkerStruct& m, int slot = 0) {
  int kind;
  if (m.ghost) kind = PSL_GHOST;
  else if (m.bounceback) kind = PSL_BOUNCEBACK;
  else if (m.fluid) kind = PSL_FLUID;
  else if (m.givenRho && !m.givenUxUyUz) kind = PSL_GIVEN_RHO;
  else if (!m.givenRho && m.givenUxUyUz) kind = PSL_GIVEN_U;
  else if (!m.givenRho && !m.givenUxUyUz) kind = PSL_GIVEN_NONE;
  else kind = PSL_GIVEN_RHOU;
  int f = kind | (slot << 6);
  if (m.periodicX) f |= PSL_PERIODIC_X;
  if (m.periodicY) f |= PSL_PERIODIC_Y;
  if (m.periodicZ) f |= PSL_PERIODIC_Z;
  return (uint8_t)f;
}

class GridStruct {
 public:
  InfoStruct Info{};

  explicit GridStruct(const InfoStruct& info, int storage = PSL_F32) {
    check(psl_grid_create(&info, storage, &h_));
    check(psl_grid_get_info(h_, &Info));
  }
  GridStruct(const GridStruct&) = delete;
  GridStruct& operator=(const GridStruct&) = delete;
  ~GridStruct() { psl_grid_destroy(h_); }

  psl_grid* handle() const { return h_; }
  int64_t cellCount() const { return (int64_t)Info.cellCountX * Info.cellCountY * Info.cellCountZ; }

  // push Info changes (sub-grid bounds, nu) to the library
  void pushInfo() { check(psl_grid_set_info(h_, &Info)); }

  // getMarkers(i, j, k, Marker, Info); the bounceback array (if any) is applied first,
  // as Grid.bouncebackMarkerArray is read before the callback in the reference.
  template <class GetMarkers>
  void setMarkers(GetMarkers getMarkers, const std::vector<uint8_t>* bounceback = nullptr,
                  std::function<int(int, int, int)> slot = nullptr) {
    std::vector<uint8_t> flags((size_t)cellCount());
    for (int k = 0; k < Info.cellCountZ; ++k)
      for (int j = 0; j < Info.cellCountY; ++j)
        for (int i = 0; i < Info.cellCountX; ++i) {
          const size_t c = ((size_t)k * Info.cellCountY + j) * Info.cellCountX + i;
          MarkerStruct m;
          if (bounceback) m.bounceback = (*bounceback)[c] != 0;
          getMarkers(i, j, k, m, Info);
          flags[c] = flag_of(m, slot ? slot(i, j, k) : 0);
        }
    check(psl_grid_set_flags(h_, flags.data()));
  }

  // getGivenRhoUxUyUz for the (up to 4) value slots used by the markers
  void setGivenValues(const float (&v)[16]) { check(psl_grid_set_bc_values(h_, v)); }
  void setSmagorinsky(float cs) { check(psl_grid_set_smagorinsky(h_, cs, nullptr)); }
  void setSmagorinsky(const std::vector<float>& perCell) {
    check(psl_grid_set_smagorinsky(h_, 0.f, perCell.data()));
  }
  void setPeriodicCopies(bool x, bool y) { check(psl_grid_set_periodic_copies(h_, (x ? 1 : 0) | (y ? 2 : 0))); }

 private:
  psl_grid* h_ = nullptr;
};

// fillEquilibriumFromFunction (include/fillEquilibrium.h:1-22) with
// getInitialRhoUxUyUz(i, j, k, rho, ux, uy, uz, Marker, Info).
template <class GetInitial>
inline void fillEquilibriumFromFunction(GridStruct& g, GetInitial getInitial) {
  const size_t n = (size_t)g.cellCount();
  std::vector<float> r(n, 1.f), a(n, 0.f), b(n, 0.f), c(n, 0.f);
  const InfoStruct& I = g.Info;
  for (int k = 0; k < I.cellCountZ; ++k)
    for (int j = 0; j < I.cellCountY; ++j)
      for (int i = 0; i < I.cellCountX; ++i) {
        const size_t id = ((size_t)k * I.cellCountY + j) * I.cellCountX + i;
        MarkerStruct m;
        getInitial(i, j, k, r[id], a[id], b[id], c[id], m, I);
      }
  check(psl_fill_equilibrium(g.handle(), r.data(), a.data(), b.data(), c.data()));
}

inline void applyStreaming(GridStruct& g) { check(psl_apply_streaming(g.handle())); }
inline void applyLocalCellUpdate(GridStruct& g) { check(psl_apply_local_cell_update(g.handle())); }
inline void applyBCPeriodicX(GridStruct& g) { check(psl_apply_bc_periodic_x(g.handle())); }
inline void applyBCPeriodicY(GridStruct& g) { check(psl_apply_bc_periodic_y(g.handle())); }
inline void applyCoarseFineGridCommunication(GridStruct& c, GridStruct& f) {
  check(psl_coarse_fine_communication(c.handle(), f.handle()));
}
inline void fillCoarseGridFromFine(GridStruct& c, GridStruct& f) { check(psl_fill_coarse_from_fine(c.handle(), f.handle())); }

// updateGrid recursion (flowAroundSphereMainSingle.cu:158-167)
inline void updateGrid(std::vector<GridStruct*>& grids, int steps = 1) {
  std::vector<psl_grid*> h;
  for (auto* g : grids) h.push_back(g->handle());
  check(psl_update_grid(h.data(), (int)h.size(), steps));
}

// getSphereDrag-style momentum exchange over bounceback cells of a box (FP64)
inline void momentumExchange(GridStruct& g, const int (&lo)[3], const int (&hi)[3], double (&out)[3]) {
  check(psl_reduce_momentum_exchange(g.handle(), lo, hi, out));
}

// Restart state (no reference counterpart): the populations as stored on the
// device with the shifter; a grid with the same settings continues from it bit
// for bit.
inline std::vector<uint8_t> saveState(GridStruct& g) {
  int64_t n = 0;
  check(psl_grid_state_bytes(g.handle(), &n));
  std::vector<uint8_t> s((size_t)n);
  check(psl_grid_save_state(g.handle(), s.data(), n));
  return s;
}
inline void loadState(GridStruct& g, const std::vector<uint8_t>& s) {
  check(psl_grid_load_state(g.handle(), s.data(), (int64_t)s.size()));
}

// Cells with a non-finite population or density <= 0, and the density range
// of the others (no reference counterpart: the reference has no guard).
struct Health {
  int64_t badCells = 0;
  double rhoMin = 0.0, rhoMax = 0.0;
  bool ok() const { return badCells == 0; }
};
inline Health checkHealth(GridStruct& g) {
  Health h;
  check(psl_grid_check_health(g.handle(), &h.badCells, &h.rhoMin, &h.rhoMax));
  return h;
}

// ------------------------------------------------------------- geometry
class STLStruct {
 public:
  explicit STLStruct(const std::string& path) { check(psl_stl_read(path.c_str(), &h_)); }
  STLStruct(const float* tri9, int n) { check(psl_stl_from_triangles(tri9, n, &h_)); }
  STLStruct(const STLStruct&) = delete;
  STLStruct& operator=(const STLStruct&) = delete;
  ~STLStruct() { psl_stl_destroy(h_); }
  psl_stl* handle() const { return h_; }
  int triangleCount() const {
    int c = 0;
    check(psl_stl_info(h_, &c, nullptr));
    return c;
  }
  void setOrigin(float ox, float oy, float oz) { check(psl_stl_set_origin(h_, ox, oy, oz)); }

 private:
  psl_stl* h_ = nullptr;
};

class BoolArrayType {
 public:
  BoolArrayType(int64_t n, bool value) { check(psl_mask_create(n, value ? 1 : 0, &h_)); n_ = n; }
  BoolArrayType(const BoolArrayType&) = delete;
  BoolArrayType& operator=(const BoolArrayType&) = delete;
  ~BoolArrayType() { psl_mask_destroy(h_); }
  psl_mask* handle() const { return h_; }
  std::vector<uint8_t> host() const {
    std::vector<uint8_t> v((size_t)n_);
    check(psl_mask_get(h_, v.data()));
    return v;
  }

 private:
  psl_mask* h_ = nullptr;
  int64_t n_ = 0;
};

inline int checkSTLEdges(const STLStruct& s) {
  int bad = 0;
  check(psl_stl_check_edges(s.handle(), &bad, nullptr));
  return bad;
}
inline void applyMarkersInsideSTL(BoolArrayType& m, const STLStruct& s, bool inside, const InfoStruct& info) {
  check(psl_markers_inside_stl(m.handle(), s.handle(), inside ? 1 : 0, &info));
}
inline void multiplyBoolArrays(BoolArrayType& a, BoolArrayType& b, BoolArrayType& r) {
  check(psl_mask_and(a.handle(), b.handle(), r.handle()));
}
inline void sumBoolArrays(BoolArrayType& a, BoolArrayType& b, BoolArrayType& r) {
  check(psl_mask_or(a.handle(), b.handle(), r.handle()));
}
inline void invertBoolArray(BoolArrayType& a) { check(psl_mask_not(a.handle())); }

// ------------------------------------------------------- DIAD (esotwist) levels
// buildDIADGrids (include/DIADFunctions.h:640-1046) on the device.  The drivers'
// getMarkers decisions for bounceback, refinement and outlet cells are given as
// an ordered psl_marker_rule list (see psl.h).
class DIADLevels {
 public:
  DIADLevels(const InfoStruct& base, int levels, int wallRefinementSpan, float nuPhys,
             const std::vector<const STLStruct*>& stls, const std::vector<psl_marker_rule>& rules) {
    std::vector<const psl_stl*> h;
    for (auto* s : stls) h.push_back(s->handle());
    check(psl_diad_build(&base, levels, wallRefinementSpan, nuPhys, h.data(), (int)h.size(), rules.data(),
                         (int)rules.size(), &h_));
    levels_ = levels;
  }
  DIADLevels(const DIADLevels&) = delete;
  DIADLevels& operator=(const DIADLevels&) = delete;
  ~DIADLevels() { psl_diad_levels_destroy(h_); }
  psl_diad_levels* handle() const { return h_; }
  int levels() const { return levels_; }
  InfoStruct info(int level) const {
    InfoStruct i{};
    check(psl_diad_levels_info(h_, level, &i, nullptr));
    return i;
  }
  // one of the PSL_DIAD_* arrays of a level (int or uint8 elements, host copy)
  template <class T>
  std::vector<T> array(int level, int which) const {
    int64_t counts[PSL_DIAD_ARRAYS];
    check(psl_diad_levels_info(h_, level, nullptr, counts));
    std::vector<T> v((size_t)counts[which]);
    if (!v.empty()) check(psl_diad_levels_get(h_, level, which, v.data()));
    return v;
  }

 private:
  psl_diad_levels* h_ = nullptr;
  int levels_ = 0;
};

// DIADGridStruct (include/types.h:166-171) of one built level
class DIADGridStruct {
 public:
  DIADGridStruct(const DIADLevels& lv, int level, int storage = PSL_F32) : info_(lv.info(level)) {
    check(psl_diad_create_level(lv.handle(), level, storage, &h_));
    n_ = info_.cellCount;
  }
  DIADGridStruct(const DIADGridStruct&) = delete;
  DIADGridStruct& operator=(const DIADGridStruct&) = delete;
  ~DIADGridStruct() { psl_diad_destroy(h_); }
  psl_diad* handle() const { return h_; }
  const InfoStruct& info() const { return info_; }
  int64_t cellCount() const { return n_; }
  // flag bytes of the level's cells: getMarkers resolved on the host from the
  // cell table (ijk) and the builder's bounceback marker
  void setFlags(const std::vector<uint8_t>& flags) { check(psl_diad_set_flags(h_, flags.data())); }
  void setGivenValues(const std::vector<float>& values16) { check(psl_diad_set_bc_values(h_, values16.data())); }
  void setSmagorinsky(float cs) { check(psl_diad_set_smagorinsky(h_, cs)); }
  void fillEquilibrium() { check(psl_diad_fill_equilibrium(h_, nullptr, nullptr, nullptr, nullptr)); }

 private:
  psl_diad* h_ = nullptr;
  InfoStruct info_{};
  int64_t n_ = 0;
};

inline void applyStreaming(DIADGridStruct& g) { check(psl_diad_apply_streaming(g.handle())); }
inline void applyLocalCellUpdate(DIADGridStruct& g) { check(psl_diad_apply_local_cell_update(g.handle())); }
inline void applyCoarseFineGridCommunication(DIADGridStruct& c, DIADGridStruct& f) {
  check(psl_diad_coarse_fine_communication(c.handle(), f.handle()));
}
inline void applyNonReflectiveOutletZ(DIADGridStruct& g) { check(psl_diad_apply_nonreflective_outlet_z(g.handle())); }
// updateGrid over DIAD levels (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:200-209)
inline void updateGrid(std::vector<DIADGridStruct*>& grids, int steps = 1) {
  std::vector<psl_diad*> h;
  for (auto* g : grids) h.push_back(g->handle());
  check(psl_diad_update_grid(h.data(), (int)h.size(), steps));
}
// getSphereDrag(DIADGridStruct&), all components, FP64
inline void momentumExchange(DIADGridStruct& g, double (&out)[3]) {
  check(psl_diad_reduce_momentum_exchange(g.handle(), out));
}
// restart state and health of an esotwist grid (the parity in place of the shifter)
inline std::vector<uint8_t> saveState(DIADGridStruct& g) {
  int64_t n = 0;
  check(psl_diad_state_bytes(g.handle(), &n));
  std::vector<uint8_t> s((size_t)n);
  check(psl_diad_save_state(g.handle(), s.data(), n));
  return s;
}
inline void loadState(DIADGridStruct& g, const std::vector<uint8_t>& s) {
  check(psl_diad_load_state(g.handle(), s.data(), (int64_t)s.size()));
}
inline Health checkHealth(DIADGridStruct& g) {
  Health h;
  check(psl_diad_check_health(g.handle(), &h.badCells, &h.rhoMin, &h.rhoMax));
  return h;
}
// movingBounceback cells with their wall velocities (cells flagged PSL_BOUNCEBACK)
inline void setMovingBounceback(DIADGridStruct& g, const std::vector<int>& cells, const std::vector<float>& u3) {
  check(psl_diad_set_moving_bounceback(g.handle(), cells.data(), (int64_t)cells.size(), u3.data()));
}
// forcedVelocity cells with their target velocities (cells flagged PSL_BOUNCEBACK)
inline void setForcedVelocity(DIADGridStruct& g, const std::vector<int>& cells, const std::vector<float>& ut3) {
  check(psl_diad_set_forced_velocity(g.handle(), cells.data(), (int64_t)cells.size(), ut3.data()));
}
// rotateSTLAlongZ (include/STLFunctions.h:774-809)
inline void rotateSTLAlongZ(STLStruct& s, float radians) { check(psl_stl_rotate(s.handle(), 2, radians)); }
// ApplyMarkersInsideSTL on a DIAD grid's cells (include/STLFunctions.h:482-696): 1 inside, 0 outside
inline std::vector<uint8_t> ApplyMarkersInsideSTL(DIADGridStruct& g, STLStruct& s) {
  std::vector<uint8_t> m((size_t)g.cellCount());
  check(psl_diad_markers_inside_stl(g.handle(), s.handle(), m.data()));
  return m;
}
// getTorque (M35Full/m35Main.cu:176-260): torque about z in N m of the forced cells
inline double getTorque(DIADGridStruct& g, float rhoNominalPhys) {
  double out[6];
  check(psl_diad_reduce_forcing(g.handle(), rhoNominalPhys, out));
  return out[5];
}

// ------------------------------------------------------------------ output
// exportSectionCutPlotXY/ZY/ZX (include/plotter/exportSectionCutPlot.h): the cut
// written to path in the plotters' sim_data.bin format
enum PlaneEnum { XY = 0, ZY = 1, ZX = 2 };
inline void exportSectionCutPlot(GridStruct& g, PlaneEnum plane, int cut, int plotNumber, float rhoNominalPhys,
                                 float soundspeedPhys, const char* path = "/dev/shm/sim_data.bin") {
  int dims[2];
  check(psl_section_cut(g.handle(), plane, cut, rhoNominalPhys, soundspeedPhys, nullptr, dims));
  std::vector<float> data((size_t)dims[0] * dims[1] * 5);
  check(psl_section_cut(g.handle(), plane, cut, rhoNominalPhys, soundspeedPhys, data.data(), dims));
  check(psl_write_sim_data(path, plotNumber, dims[0], dims[1], 5, data.data()));
}
inline void exportSectionCutPlot(std::vector<DIADGridStruct*>& grids, PlaneEnum plane, int cut, int plotNumber,
                                 float rhoNominalPhys, float soundspeedPhys,
                                 const char* path = "/dev/shm/sim_data.bin") {
  std::vector<psl_diad*> h;
  for (auto* g : grids) h.push_back(g->handle());
  int dims[3];
  check(psl_diad_section_cut(h.data(), (int)h.size(), plane, cut, rhoNominalPhys, soundspeedPhys, 20000000, nullptr,
                             dims));
  std::vector<float> data((size_t)dims[0] * dims[1] * 6);
  check(psl_diad_section_cut(h.data(), (int)h.size(), plane, cut, rhoNominalPhys, soundspeedPhys, 20000000,
                             data.data(), dims));
  check(psl_write_sim_data(path, plotNumber, dims[0], dims[1], 6, data.data()));
}
// exportHistoryData (exportSectionCutPlot.h:994-1001)
inline void exportHistoryData(const std::vector<float>& history, int currentIteration,
                              const char* path = "/dev/shm/historyData.bin") {
  check(psl_write_history(path, history.data(), currentIteration + 1));
}

}  // namespace psl_facade
