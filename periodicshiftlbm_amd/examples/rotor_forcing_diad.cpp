// Host driver in the shape of the reference's M35Full/m35Main.cu: a rotor inside a
// z channel on wall-refined esotwist (DIAD) levels, through include/psl_facade.hpp.
//
// Every two iterations the rotor's STL is turned by angularVelocity * dtPhys *
// iteration about z (rotateSTLAlongZ on a fresh copy, m35Main.cu:333-344), each
// level's cells inside it become forced-velocity cells driven towards the rigid
// rotation (getForcing, m35Main.cu:100-118), and the torque summed over the
// levels (getTorque, m35Main.cu:176-260) is printed every `chunk` iterations.
// The rotor is four blades given as triangles; the driver's STL files and rotor
// profile are replaced by that rigid rotation.  Build (host compiler only):
//   g++ -O2 -std=c++17 -I include periodicshiftlbm_amd/examples/rotor_forcing_diad.cpp \
//       -L periodicshiftlbm_amd -lpsl_amd -Wl,-rpath,$PWD/periodicshiftlbm_amd -o rotor_forcing_diad
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "psl_facade.hpp"

using namespace psl_facade;

constexpr float resGlobal = 10.f;           // mm
constexpr float dtGlobal = 1e-3f;           // s
constexpr float angularVelocity = 3.3f;     // rad/s
constexpr float uzInletPhys = 0.2f;         // m/s
constexpr float nuPhys = 1e-6f;             // m2/s, water
constexpr float rhoNominalPhys = 1000.f;    // kg/m3
constexpr float SmagorinskyConstantGlobal = 0.1f;

// 12 outward-wound triangles of an axis-aligned box
static void addBox(std::vector<float>& tri, const float lo[3], const float hi[3]) {
  const float v[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {hi[0], hi[1], lo[2]}, {lo[0], hi[1], lo[2]},
                         {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]}, {hi[0], hi[1], hi[2]}, {lo[0], hi[1], hi[2]}};
  const int f[12][3] = {{0, 2, 1}, {0, 3, 2}, {4, 5, 6}, {4, 6, 7}, {0, 1, 5}, {0, 5, 4},
                        {2, 3, 7}, {2, 7, 6}, {1, 2, 6}, {1, 6, 5}, {3, 0, 4}, {3, 4, 7}};
  for (const auto& t : f)
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) tri.push_back(v[t[a]][c]);
}

static psl_marker_rule rule(int what, int shape, std::vector<float> p, unsigned level_mask) {
  psl_marker_rule r{};
  r.what = what;
  r.value = 1;
  r.level_mask = level_mask;
  r.shape = shape;
  for (size_t a = 0; a < p.size() && a < 8; ++a) r.p[a] = p[a];
  return r;
}

int main(int argc, char** argv) {
  const int levels = argc > 1 ? std::atoi(argv[1]) : 2;
  const int iterationCount = argc > 2 ? std::atoi(argv[2]) : 200;
  const int iterationChunk = argc > 3 ? std::atoi(argv[3]) : 50;
  const int nxy = 48, nz = 64;

  InfoStruct base{};
  base.res = resGlobal;
  base.dtPhys = dtGlobal;
  base.nu = dtGlobal * nuPhys / ((resGlobal / 1000.f) * (resGlobal / 1000.f));
  base.cellCountX = base.cellCountY = nxy;
  base.cellCountZ = nz;
  base.cellCount = nxy * nxy * nz;
  base.ox = base.oy = -0.5f * (nxy - 1) * resGlobal;
  base.oz = -0.5f * (nz - 1) * resGlobal;
  const float uzInlet = uzInletPhys * dtGlobal / (resGlobal / 1000.f);

  // the rotor: four blades around the z axis
  std::vector<float> tri;
  const float bl[4][2][3] = {{{20, -15, -40}, {150, 15, 40}}, {{-150, -15, -40}, {-20, 15, 40}},
                             {{-15, 20, -40}, {15, 150, 40}}, {{-15, -150, -40}, {15, -20, 40}}};
  for (const auto& b : bl) addBox(tri, b[0], b[1]);
  const int triCount = (int)tri.size() / 9;

  // refinement around the rotor on every level but the finest
  std::vector<psl_marker_rule> rules;
  for (int l = 0; l + 1 < levels; ++l) {
    const float w = 180.f - 20.f * l, h = 70.f - 10.f * l;
    rules.push_back(rule(PSL_RULE_REFINEMENT, PSL_SHAPE_BOX, {-w, -w, -h, w, w, h}, 1u << l));
  }
  DIADLevels lv(base, levels, 2, nuPhys, {}, rules);

  std::vector<std::unique_ptr<DIADGridStruct>> owned;
  std::vector<DIADGridStruct*> grids;
  std::vector<std::vector<uint8_t>> baseFlags;
  std::vector<std::vector<int>> ijk;
  for (int l = 0; l < levels; ++l) {
    owned.emplace_back(new DIADGridStruct(lv, l, PSL_F16DEV));
    DIADGridStruct& g = *owned.back();
    const InfoStruct I = g.info();
    ijk.push_back(lv.array<int>(l, PSL_DIAD_IJK));
    const int64_t cells = g.cellCount();
    std::vector<uint8_t> flags((size_t)cells);
    for (int64_t c = 0; c < cells; ++c) {
      const int i = ijk[l][c], j = ijk[l][cells + c], k = ijk[l][2 * cells + c];
      MarkerStruct m;
      const bool face = i == 0 || j == 0 || k == 0 || i == I.cellCountX - 1 || j == I.cellCountY - 1 ||
                        k == I.cellCountZ - 1;
      if (l == 0 && face) {
        if (k == I.cellCountZ - 1) m.givenRho = true;
        else m.givenUxUyUz = true;
      } else if (face) {
        m.ghost = true;
      } else {
        m.fluid = true;
      }
      flags[c] = flag_of(m, m.givenRho ? 1 : 0);
    }
    baseFlags.push_back(flags);
    g.setFlags(flags);
    g.setGivenValues({1.f, 0.f, 0.f, uzInlet, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f});
    g.setSmagorinsky(SmagorinskyConstantGlobal);
    g.fillEquilibrium();
    grids.push_back(&g);
    std::printf("level %d: %lld cells\n", l, (long long)cells);
  }

  auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it <= iterationCount; ++it) {
    if (it % 2 == 0) {
      STLStruct rotated(tri.data(), triCount);
      rotateSTLAlongZ(rotated, angularVelocity * dtGlobal * it);
      for (int l = 0; l < levels; ++l) {
        DIADGridStruct& g = *grids[l];
        const InfoStruct I = g.info();
        const int64_t cells = g.cellCount();
        const std::vector<uint8_t> inside = ApplyMarkersInsideSTL(g, rotated);
        std::vector<uint8_t> flags = baseFlags[l];
        std::vector<int> forced;
        std::vector<float> target;
        // lattice velocity per mm of radius: omega * dt_l / res_l, the same on every level
        const float w = angularVelocity * I.dtPhys / (I.res / 1000.f) / 1000.f;
        for (int64_t c = 0; c < cells; ++c) {
          if (!inside[c] || baseFlags[l][c] != PSL_FLUID) continue;  // forced cells replace fluid cells only
          const float x = ijk[l][c] * I.res + I.ox, y = ijk[l][cells + c] * I.res + I.oy;
          forced.push_back((int)c);
          target.insert(target.end(), {-w * y, w * x, 0.f});
          flags[c] = PSL_BOUNCEBACK;  // skipped by the bulk update, updated by the forced-cell kernel
        }
        g.setFlags(flags);
        setForcedVelocity(g, forced, target);
      }
    }
    updateGrid(grids, 1);
    if (it > 0 && it % iterationChunk == 0) {
      double torque = 0.0;
      for (auto* g : grids) torque += getTorque(*g, rhoNominalPhys);
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("iteration %d: torque %.6e N m, %.1f ms per iteration\n", it, torque, 1e3 * s / iterationChunk);
      t0 = std::chrono::steady_clock::now();
    }
  }
  return 0;
}
