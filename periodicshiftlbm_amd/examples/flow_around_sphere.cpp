// Host driver in the shape of the reference's flowAroundSphere/flowAroundSphereMainSingle.cu,
// calling the MI355X library through include/psl_facade.hpp (C ABI underneath).
//
// A sphere of 2000 mm in a cube of 11 diameters, Re = 1e6, Smagorinsky Cs = 0.1,
// velocity inlet and lateral walls (given u), pressure outlet; four nested grids,
// each 2x finer, placed by the driver's rule.  Every iteration runs the level
// recursion and records the finest grid's drag coefficient; every `chunk`
// iterations it prints drag, C_D and GLUPS, fills coarse grids from fine ones,
// writes the mid-plane XY cut of every level (sim_data.bin) and the C_D history
// (historyData.bin), as the reference driver does before calling its plotters.
//
//   flow_around_sphere [cells=220] [iterations=2000] [chunk=500] [output dir=/dev/shm]
//
// `cells` is the coarse edge (220: the driver's 100 mm cells; the sphere is then
// 160 cells across on the finest level).  Build (host compiler only):
//   g++ -O2 -std=c++17 -ffp-contract=off -I include periodicshiftlbm_amd/examples/flow_around_sphere.cpp \
//       -L periodicshiftlbm_amd -lpsl_amd -Wl,-rpath,$PWD/periodicshiftlbm_amd -o flow_around_sphere
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "psl_facade.hpp"

using namespace psl_facade;

constexpr float sphereDiameterPhys = 2000.f;  // mm
constexpr float uxInlet = 0.07f;              // also the nominal lattice Mach number
constexpr float reynoldsNumber = 1000000.f;
constexpr float SmagorinskyConstantGlobal = 0.1f;
constexpr float sphereRadiusPhys = 0.5 * sphereDiameterPhys;
constexpr float uxInletPhys = uxInlet;  // m/s: physical velocity equals the lattice velocity
constexpr float nuPhys = uxInletPhys * (sphereDiameterPhys / 1000.f) / reynoldsNumber;
constexpr float rhoNominalPhys = 1.225f;
constexpr float invSqrt3 = 0.577350269f;
constexpr float domainSizePhys = 11.f * sphereDiameterPhys;
constexpr float sphereXPhys = 2.f * sphereDiameterPhys;
constexpr float sphereYPhys = 5.5f * sphereDiameterPhys;
constexpr float sphereZPhys = 5.5f * sphereDiameterPhys;
constexpr int gridLevelCount = 4;

int main(int argc, char** argv) {
  const int cells = argc > 1 ? std::atoi(argv[1]) : 220;
  const int iterationCount = argc > 2 ? std::atoi(argv[2]) : 2000;
  const int iterationChunk = argc > 3 ? std::atoi(argv[3]) : 500;
  const std::string outDir = argc > 4 ? argv[4] : "/dev/shm";
  const float resGlobal = 22000.f / cells;  // mm
  const float dtPhysGlobal = (uxInlet / uxInletPhys) * (resGlobal / 1000);
  const float soundspeedPhys = invSqrt3 * (resGlobal / 1000) / dtPhysGlobal;

  // level Infos (flowAroundSphereMainSingle.cu:171-221)
  std::vector<InfoStruct> info(gridLevelCount);
  info[0].res = resGlobal;
  info[0].dtPhys = dtPhysGlobal;
  info[0].nu = (info[0].dtPhys * nuPhys) / ((info[0].res / 1000) * (info[0].res / 1000));
  info[0].cellCountX = info[0].cellCountY = info[0].cellCountZ = (int)std::ceil(domainSizePhys / resGlobal);
  for (int level = 1; level < gridLevelCount; level++) {
    InfoStruct& c = info[level - 1];
    InfoStruct& f = info[level];
    f.gridID = c.gridID + 1;
    f.res = c.res * 0.5f;
    f.dtPhys = c.dtPhys * 0.5f;
    f.nu = (f.dtPhys * nuPhys) / ((f.res / 1000.f) * (f.res / 1000.f));
    float progress = (float)level / (float)(gridLevelCount - 1);
    progress = std::pow(progress, 0.2f);
    const float xStart = progress * 1.4f * sphereDiameterPhys;
    const float xEnd = (1 - progress) * domainSizePhys + progress * 2.6f * sphereDiameterPhys;
    const float yStart = progress * 4.9f * sphereDiameterPhys;
    c.iSubgridStart = (int)((xStart - c.ox) / c.res + 0.5f);
    c.iSubgridEnd = (int)((xEnd - c.ox) / c.res + 0.5f);
    c.jSubgridStart = (int)((yStart - c.oy) / c.res + 0.5f);
    c.jSubgridEnd = c.jSubgridStart + (c.iSubgridEnd - c.iSubgridStart);
    c.kSubgridStart = c.jSubgridStart;
    c.kSubgridEnd = c.jSubgridEnd;
    f.ox = c.ox + c.iSubgridStart * c.res - f.res * 0.5f;
    f.oy = c.oy + c.jSubgridStart * c.res - f.res * 0.5f;
    f.oz = c.oz + c.kSubgridStart * c.res - f.res * 0.5f;
    f.cellCountX = (c.iSubgridEnd - c.iSubgridStart) * 2;
    f.cellCountY = (c.jSubgridEnd - c.jSubgridStart) * 2;
    f.cellCountZ = (c.kSubgridEnd - c.kSubgridStart) * 2;
  }

  // getMarkers (lines 41-67)
  auto getMarkers = [](int iCell, int jCell, int kCell, MarkerStruct& Marker, const InfoStruct& Info) {
    const float xPhys = iCell * Info.res + Info.ox;
    const float yPhys = jCell * Info.res + Info.oy;
    const float zPhys = kCell * Info.res + Info.oz;
    const float r2 = (xPhys - sphereXPhys) * (xPhys - sphereXPhys) + (yPhys - sphereYPhys) * (yPhys - sphereYPhys) +
                     (zPhys - sphereZPhys) * (zPhys - sphereZPhys);
    if (Info.gridID != 0) {
      if (r2 <= sphereRadiusPhys * sphereRadiusPhys) Marker.bounceback = true;
      else if (iCell == 0 || iCell == Info.cellCountX - 1) Marker.ghost = true;
      else if (jCell == 0 || jCell == Info.cellCountY - 1) Marker.ghost = true;
      else if (kCell == 0 || kCell == Info.cellCountZ - 1) Marker.ghost = true;
      else Marker.fluid = true;
    } else {
      if (r2 <= sphereRadiusPhys * sphereRadiusPhys) Marker.bounceback = true;
      else if (jCell == 0 || jCell == Info.cellCountY - 1) Marker.givenUxUyUz = true;
      else if (kCell == 0 || kCell == Info.cellCountZ - 1) Marker.givenUxUyUz = true;
      else if (iCell == 0) Marker.givenUxUyUz = true;
      else if (iCell == Info.cellCountX - 1) Marker.givenRho = true;
      else Marker.fluid = true;
    }
  };
  // getGivenRhoUxUyUz (lines 69-77): one value slot
  const float given[16] = {1.f, uxInlet, 0.f, 0.f};

  std::vector<std::unique_ptr<GridStruct>> grids;
  std::vector<GridStruct*> levels;
  int64_t cellCountTotal = 0, cellUpdatesPerIteration = 0;
  for (int level = 0; level < gridLevelCount; level++) {
    grids.emplace_back(new GridStruct(info[level]));
    GridStruct& g = *grids.back();
    g.setMarkers(getMarkers);
    g.setGivenValues(given);
    g.setSmagorinsky(SmagorinskyConstantGlobal);
    // getInitialRhoUxUyUz (lines 84-92): rest
    fillEquilibriumFromFunction(g, [](int, int, int, float& rho, float& ux, float& uy, float& uz, const MarkerStruct&,
                                      const InfoStruct&) {
      rho = 1.f;
      ux = uy = uz = 0.f;
    });
    levels.push_back(&g);
    std::printf("Cell count on grid %d: %lld\n", level, (long long)g.cellCount());
    // updates per iteration (lines 223-236); the cells a finer grid covers are skipped,
    // which are those at least two cells inside its box on the coarser grid
    cellCountTotal += g.cellCount();
    const InfoStruct& I = g.Info;
    int64_t covered = 0;
    if (level < gridLevelCount - 1)
      covered = (int64_t)(I.iSubgridEnd - I.iSubgridStart - 4) * (I.jSubgridEnd - I.jSubgridStart - 4) *
                (I.kSubgridEnd - I.kSubgridStart - 4);
    cellUpdatesPerIteration += (g.cellCount() - covered) << level;
  }
  std::printf("Cell count total: %lld\n", (long long)cellCountTotal);
  std::printf("Cell updates per iteration: %lld\n", (long long)cellUpdatesPerIteration);

  // getSphereDrag's box on the finest grid (lines 99-112)
  GridStruct& finest = *levels.back();
  const InfoStruct& F = finest.Info;
  const int lo[3] = {(int)((sphereXPhys - sphereRadiusPhys - F.ox) / F.res),
                     (int)((sphereYPhys - sphereRadiusPhys - F.oy) / F.res),
                     (int)((sphereZPhys - sphereRadiusPhys - F.oz) / F.res)};
  const int hi[3] = {(int)((sphereXPhys + sphereRadiusPhys - F.ox) / F.res) + 2,
                     (int)((sphereYPhys + sphereRadiusPhys - F.oy) / F.res) + 2,
                     (int)((sphereZPhys + sphereRadiusPhys - F.oz) / F.res) + 2};
  // C_D = -8 F / (rho U^2 pi D^2) with F converted to newtons (convertToPhysicalForce,
  // include/cellFunctions.h:221-226, on the finest grid's res and dtPhys)
  const double toNewton = (double)rhoNominalPhys * std::pow(F.res / 1000.0, 4) / ((double)F.dtPhys * F.dtPhys);
  const double cdScale = -8.0 / (rhoNominalPhys * uxInletPhys * uxInletPhys * M_PI * (sphereDiameterPhys / 1000.0) *
                                 (sphereDiameterPhys / 1000.0));

  std::printf("Starting simulation\n");
  std::vector<float> historyDragCoefficient(iterationCount + 1, 0.f);
  auto t0 = std::chrono::steady_clock::now();
  for (int iteration = 0; iteration <= iterationCount; iteration++) {
    updateGrid(levels);
    double force[3];
    momentumExchange(finest, lo, hi, force);
    const double drag = force[0] * toNewton;
    historyDragCoefficient[iteration] = (float)(cdScale * drag);
    if (iteration % iterationChunk == 0 && iteration != 0) {
      const auto t1 = std::chrono::steady_clock::now();
      const double lapTime = std::chrono::duration<double>(t1 - t0).count();
      std::printf("Finished iteration %d\nDrag %.6e\nDrag coefficient %.6f\nGLUPS: %.3f\n", iteration, drag,
                  historyDragCoefficient[iteration], (double)cellUpdatesPerIteration * iterationChunk / lapTime / 1e9);
      // a diverged run stops here instead of plotting NaN fields (the reference has no guard)
      for (int level = 0; level < gridLevelCount; level++) {
        const Health h = checkHealth(*levels[level]);
        if (!h.ok()) {
          std::printf("Grid %d: %lld cells with non-finite populations or rho <= 0; stopping\n", level,
                      (long long)h.badCells);
          return 2;
        }
      }
      for (int level = gridLevelCount - 2; level >= 0; level--) fillCoarseGridFromFine(*levels[level], *levels[level + 1]);
      for (int level = 0; level < gridLevelCount; level++)
        exportSectionCutPlot(*levels[level], XY, levels[level]->Info.cellCountZ / 2, iteration + level, rhoNominalPhys,
                             soundspeedPhys, (outDir + "/sim_data.bin").c_str());
      exportHistoryData(historyDragCoefficient, iteration, (outDir + "/historyData.bin").c_str());
      std::fflush(stdout);
      t0 = std::chrono::steady_clock::now();
    }
  }
  std::printf("Finished successfully\n");
  return 0;
}
