// Golden-vector harness around the reference's own boundary-condition headers
// (include/boundaryConditions/{applyBounceback,restoreRho,restoreUxUyUz,
// restoreRhoUxUyUz,applyMBBC}.h), compiled where they lie under the reference
// tree.  Test infrastructure only: used by tests/golden/make_bc_golden.py to
// write tests/golden/bc_vectors.npz.
#include <hip/hip_runtime.h>

#include "applyBounceback.h"
#include "restoreRho.h"
#include "restoreUxUyUz.h"
#include "restoreRhoUxUyUz.h"
#include "applyMBBC.h"

extern "C" {
// which: 0 applyMBBC, 1 restoreRhoUxUyUz, 2 restoreUxUyUz, 3 restoreRho, 4 applyBounceback
int ref_bc_f32(int which, int nx, int ny, int nz, float* f, float* rho_u) {
  float (&ff)[27] = *reinterpret_cast<float (*)[27]>(f);
  float rho = rho_u[0], ux = rho_u[1], uy = rho_u[2], uz = rho_u[3];
  switch (which) {
    case 0: applyMBBC(nx, ny, nz, rho, ux, uy, uz, ff); break;
    case 1: restoreRhoUxUyUz(nx, ny, nz, rho, ux, uy, uz, ff); break;
    case 2: restoreUxUyUz(nx, ny, nz, rho, ux, uy, uz, ff); break;
    case 3: restoreRho(nx, ny, nz, rho, ux, uy, uz, ff); break;
    case 4: applyBounceback(ff); break;
    default: return -1;
  }
  rho_u[0] = rho; rho_u[1] = ux; rho_u[2] = uy; rho_u[3] = uz;
  return 0;
}
}
