// Host driver in the shape of the reference's flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:
// wall-refined esotwist (DIAD) levels around a sphere, built on the device and
// advanced by the level recursion, through include/psl_facade.hpp.
//
// The driver's getMarkers splits into the builder's rules (sphere wall on every
// level, refinement regions per level) and the per-level flag bytes resolved on
// the host from each level's cell table.  Prints the drag coefficient and the
// throughput every `chunk` coarse steps.  Build (host compiler only):
//   g++ -O2 -std=c++17 -I include periodicshiftlbm_amd/examples/flow_around_sphere_diad.cpp \
//       -L periodicshiftlbm_amd -lpsl_amd -Wl,-rpath,$PWD/periodicshiftlbm_amd -o flow_around_sphere_diad
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "psl_facade.hpp"

using namespace psl_facade;

constexpr float sphereDiameterPhys = 2000.f;  // mm
constexpr float resGlobal = 100.f;            // mm
constexpr float uxInlet = 0.07f;
constexpr float reynoldsNumber = 1000000.f;
constexpr float SmagorinskyConstantGlobal = 0.1f;
constexpr float rhoNominalPhys = 1.225f;

static psl_marker_rule rule(int what, int shape, std::vector<float> p, unsigned level_mask, int value = 1) {
  psl_marker_rule r{};
  r.what = what;
  r.value = value;
  r.level_mask = level_mask;
  r.shape = shape;
  for (size_t a = 0; a < p.size() && a < 8; ++a) r.p[a] = p[a];
  return r;
}

int main(int argc, char** argv) {
  const int levels = argc > 1 ? std::atoi(argv[1]) : 6;
  const int iterationCount = argc > 2 ? std::atoi(argv[2]) : 100;
  const int iterationChunk = argc > 3 ? std::atoi(argv[3]) : 20;
  const float D = sphereDiameterPhys, R = 0.5f * D;
  const float nuPhys = uxInlet * (D / 1000.f) / reynoldsNumber;
  const float dt = resGlobal / 1000.f;
  const int n = (int)std::ceil(11.f * D / resGlobal);

  InfoStruct base{};
  base.res = resGlobal;
  base.dtPhys = dt;
  base.nu = dt * nuPhys / ((resGlobal / 1000.f) * (resGlobal / 1000.f));
  base.cellCountX = base.cellCountY = base.cellCountZ = n;
  base.cellCount = n * n * n;
  base.ox = -2.f * D;
  base.oy = base.oz = -5.5f * D;

  // getMarkers, builder part (flowAroundSphereMainSingleDIAD.cu:47-80)
  std::vector<psl_marker_rule> rules = {rule(PSL_RULE_BOUNCEBACK, PSL_SHAPE_SPHERE_LE, {0, 0, 0, R}, 0xffffffffu)};
  const float widths[4][2] = {{1.5f, 4.f}, {1.f, 2.5f}, {0.7f, 1.f}, {0.6f, 0.6f}};
  for (int l = 0; l < 4 && l < levels - 1; ++l) {
    rules.push_back(rule(PSL_RULE_REFINEMENT, PSL_SHAPE_XCYLINDER, {0, 0, 0, widths[l][0] * D, widths[l][1] * D}, 1u << l));
    rules.push_back(rule(PSL_RULE_REFINEMENT, PSL_SHAPE_SPHERE_LT, {0, 0, 0, widths[l][0] * D}, 1u << l));
  }
  DIADLevels lv(base, levels, 3, nuPhys, {}, rules);
  std::printf("build %.1f ms\n", [&] { double ms = 0; psl_diad_levels_build_ms(lv.handle(), &ms); return ms; }());

  std::vector<std::unique_ptr<DIADGridStruct>> owned;
  std::vector<DIADGridStruct*> grids;
  long long updates = 0;
  for (int l = 0; l < levels; ++l) {
    owned.emplace_back(new DIADGridStruct(lv, l, PSL_F16DEV));
    DIADGridStruct& g = *owned.back();
    const InfoStruct I = g.info();
    const std::vector<int> ijk = lv.array<int>(l, PSL_DIAD_IJK);
    const std::vector<uint8_t> bb = lv.array<uint8_t>(l, PSL_DIAD_BOUNCEBACK);
    const int64_t cells = g.cellCount();
    // getMarkers, per-cell part: inlet / walls / outlet on level 0, ghost box faces above
    std::vector<uint8_t> flags((size_t)cells);
    for (int64_t c = 0; c < cells; ++c) {
      const int i = ijk[c], j = ijk[cells + c], k = ijk[2 * cells + c];
      MarkerStruct m;
      m.bounceback = bb[c];
      if (!m.bounceback) {
        const bool face_i = i == 0 || i == I.cellCountX - 1;
        const bool face_j = j == 0 || j == I.cellCountY - 1, face_k = k == 0 || k == I.cellCountZ - 1;
        if (l == 0) {
          if (face_j || face_k || i == 0) m.givenUxUyUz = true;
          else if (i == I.cellCountX - 1) m.givenRho = true;
          else m.fluid = true;
        } else if (face_i || face_j || face_k) {
          m.ghost = true;
        } else {
          m.fluid = true;
        }
      }
      flags[c] = flag_of(m, m.givenRho ? 1 : 0);
    }
    g.setFlags(flags);
    g.setGivenValues({1.f, uxInlet, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f});
    g.setSmagorinsky(SmagorinskyConstantGlobal);
    g.fillEquilibrium();
    grids.push_back(&g);
    updates += (long long)cells << l;
    std::printf("level %d: %lld cells\n", l, (long long)cells);
  }
  const float Dfine = D / lv.info(levels - 1).res;
  auto t0 = std::chrono::steady_clock::now();
  for (int it = 1; it <= iterationCount; ++it) {
    updateGrid(grids, 1);
    if (it % iterationChunk == 0) {
      double f[3];
      momentumExchange(*grids.back(), f);
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      const double cd = -8.0 * f[0] / (uxInlet * uxInlet * M_PI * Dfine * Dfine);
      std::printf("iteration %d: C_D %.4f, %.2f GLUPS\n", it, cd, updates * (double)iterationChunk / s / 1e9);
      // a diverged run stops here instead of writing NaN drag (the reference has no guard)
      for (size_t l = 0; l < grids.size(); ++l) {
        const Health h = checkHealth(*grids[l]);
        if (!h.ok()) {
          std::printf("level %zu: %lld cells with non-finite populations or rho <= 0; stopping\n", l,
                      (long long)h.badCells);
          return 2;
        }
      }
      t0 = std::chrono::steady_clock::now();
    }
  }
  (void)rhoNominalPhys;
  return 0;
}
