// Host build of the library's per-cell numerics (periodicshiftlbm_amd/csrc/psl_math.h)
// so CPU tests can compare the exact device formulation with the oracle
// without a GPU.  Test infrastructure only.
#include "../../periodicshiftlbm_amd/csrc/psl_math.h"

extern "C" {
void dm_collide(float* f, long long n, float nu, float cs, int w_off) {
  const float tau = 3.f * nu + 0.5f;
  psl::Relax rx{1.f / tau, tau, cs};
  for (long long c = 0; c < n; ++c) {
    float v[27];
    for (int q = 0; q < 27; ++q) v[q] = f[27 * c + q] - (w_off ? psl::weight(q) : 0.f);
    if (w_off) psl::collide<true>(v, rx, cs);
    else psl::collide<false>(v, rx, cs);
    for (int q = 0; q < 27; ++q) f[27 * c + q] = v[q] + (w_off ? psl::weight(q) : 0.f);
  }
}
void dm_rescale(float* f, long long n, int c2f) {
  for (long long c = 0; c < n; ++c) {
    float v[27];
    for (int q = 0; q < 27; ++q) v[q] = f[27 * c + q];
    psl::rescale(v, c2f != 0);
    for (int q = 0; q < 27; ++q) f[27 * c + q] = v[q];
  }
}
int dm_bc(int which, int code, float* f, float* ru) {
  float v[27];
  for (int q = 0; q < 27; ++q) v[q] = f[q];
  float rho = ru[0], ux = ru[1], uy = ru[2], uz = ru[3];
  if (which == 0) psl::bc_mbbc<float>(code, v, rho, ux, uy, uz);
  else if (which == 1) psl::bc_restore_rhou<float>(code, v, rho, ux, uy, uz);
  else if (which == 2) psl::bc_restore_u<float>(code, v, rho, ux, uy, uz);
  else psl::bc_restore_rho<float>(code, v, rho, ux, uy, uz);
  for (int q = 0; q < 27; ++q) f[q] = v[q];
  ru[0] = rho; ru[1] = ux; ru[2] = uy; ru[3] = uz;
  return 0;
}
float dm_feq(int q, float rho, float ux, float uy, float uz) { return psl::feq(q, rho, ux, uy, uz); }
}
