// Host driver in the shape of the reference's flowAroundSphere/flowAroundSphereMainSingle.cu,
// calling the MI355X library through include/psl_facade.hpp (C ABI underneath).
//
// Sphere of diameter D in a channel: velocity inlet, pressure outlet, periodic
// y (copies) and z (shift), Smagorinsky LES; prints the drag coefficient and
// the throughput every `chunk` steps.  Build (host compiler only):
//   g++ -O2 -std=c++17 -I include periodicshiftlbm_amd/examples/flow_around_sphere.cpp \
//       -L periodicshiftlbm_amd -lpsl_amd -Wl,-rpath,$PWD/periodicshiftlbm_amd -o flow_around_sphere
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "psl_facade.hpp"

using namespace psl_facade;

constexpr float uxInlet = 0.04f;
constexpr float reynoldsNumber = 1000.f;
constexpr float SmagorinskyConstantGlobal = 0.17f;

int main(int argc, char** argv) {
  const int nx = argc > 1 ? std::atoi(argv[1]) : 128;
  const int ny = nx / 2, nz = nx / 2;
  const int iterationCount = argc > 2 ? std::atoi(argv[2]) : 200;
  const int iterationChunk = argc > 3 ? std::atoi(argv[3]) : 100;
  const float D = ny / 4.f;
  const float cx = nx / 4.f, cy = ny / 2.f, cz = nz / 2.f;

  InfoStruct info{};
  info.res = 1.f;
  info.cellCountX = nx;
  info.cellCountY = ny;
  info.cellCountZ = nz;
  info.nu = uxInlet * D / reynoldsNumber;
  GridStruct grid(info);

  // getMarkers (flowAroundSphereMainSingle.cu:41-67), with periodic y/z faces
  auto getMarkers = [&](int i, int j, int k, MarkerStruct& m, const InfoStruct& I) {
    const float r2 = (i - cx) * (i - cx) + (j - cy) * (j - cy) + (k - cz) * (k - cz);
    if (r2 <= 0.25f * D * D) m.bounceback = true;
    else if (i == 0) { m.givenUxUyUz = true; m.periodicY = m.periodicZ = true; }
    else if (i == I.cellCountX - 1) { m.givenRho = true; m.periodicY = m.periodicZ = true; }
    else m.fluid = true;
  };
  grid.setMarkers(getMarkers);
  // getGivenRhoUxUyUz (lines 69-77): one slot
  const float given[16] = {1.f, uxInlet, 0.f, 0.f};
  grid.setGivenValues(given);
  grid.setSmagorinsky(SmagorinskyConstantGlobal);
  grid.setPeriodicCopies(false, true);
  fillEquilibriumFromFunction(grid, [](int, int, int, float& rho, float& ux, float& uy, float& uz,
                                       const MarkerStruct&, const InfoStruct&) {
    rho = 1.f;
    ux = uy = uz = 0.f;
  });

  const int lo[3] = {(int)(cx - D / 2), (int)(cy - D / 2), (int)(cz - D / 2)};
  const int hi[3] = {(int)(cx + D / 2) + 2, (int)(cy + D / 2) + 2, (int)(cz + D / 2) + 2};
  std::vector<GridStruct*> grids{&grid};
  auto t0 = std::chrono::steady_clock::now();
  for (int it = 1; it <= iterationCount; ++it) {
    updateGrid(grids);
    if (it % iterationChunk == 0) {
      double f[3];
      momentumExchange(grid, lo, hi, f);  // synchronises
      const auto t1 = std::chrono::steady_clock::now();
      const double s = std::chrono::duration<double>(t1 - t0).count();
      const double cd = -8.0 * f[0] / (uxInlet * uxInlet * M_PI * D * D);
      std::printf("iteration %d drag %.6e Cd %.5f MLUPS %.1f\n", it, f[0], cd,
                  (double)nx * ny * nz * iterationChunk / s / 1e6);
      t0 = std::chrono::steady_clock::now();
    }
  }
  std::printf("Finished successfully\n");
  return 0;
}
