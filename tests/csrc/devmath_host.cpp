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

extern "C" {
// Host emulation of one fused update (bulk kernel over fluid/bounceback cells,
// then the boundary-list kernel), same per-cell code as the device kernels.
void dm_ps_update(float* fa, const int* shift, int nx, int ny, int nz, const unsigned char* flags, const float* bc,
                  float nu, float cs) {
  const long long n = (long long)nx * ny * nz;
  const float tau = 3.f * nu + 0.5f;
  psl::Relax rx{1.f / tau, tau, cs};
  for (int pass = 0; pass < 2; ++pass)
    for (long long cell = 0; cell < n; ++cell) {
      const int fl = flags ? flags[cell] : 0;
      const int kind = fl & 7;
      const bool bulk = kind <= psl::KIND_BOUNCEBACK;
      const bool bnd = kind >= psl::KIND_GIVEN_RHO && kind <= psl::KIND_GIVEN_RHOU;
      if ((pass == 0 && !bulk) || (pass == 1 && !bnd)) continue;
      float v[27];
      long long sl[27];
      for (int q = 0; q < 27; ++q) {
        long long s = cell + shift[q];
        if (s >= n) s -= n;
        sl[q] = (long long)q * n + s;
        v[q] = fa[sl[q]];
      }
      if (pass == 0) {
        if (kind == psl::KIND_BOUNCEBACK) psl::bounceback(v);
        else psl::collide<false>(v, rx, cs);
      } else {
        const int k = (int)(cell / ((long long)nx * ny));
        const long long rem = cell - (long long)k * nx * ny;
        const int j = (int)(rem / nx), i = (int)(rem % nx);
        const int code = psl::outer_normal_code(i, j, k, nx, ny, nz, fl);
        const int slot = (fl >> 6) & 3;
        psl::boundary_cell(v, kind, code, bc[4 * slot], bc[4 * slot + 1], bc[4 * slot + 2], bc[4 * slot + 3]);
        psl::collide<false>(v, rx, cs);
      }
      for (int q = 0; q < 27; ++q) fa[sl[q]] = v[q];
    }
}
}
