// Scalar transport on periodic-shift grids.
//
// include/scalarTransportFunctions.h:10-175 (Zhang, Fan, Chen 2011): a second
// set of 27 populations T_q per cell, stored like f (FP32, slot
// (cell + shift[q]) mod N) so that the same shifter streams both.  The update
// is applyLocalCellUpdate(GridStruct&, ScalarTransportStruct&): T bounces back
// with f, takes the prescribed value where one is given, and after the fluid
// collision relaxes with the post-collision f:
//   Tsum = sum f_q T_q / rho,  S = sum c_q f_q (T_q - Tsum),
//   T_q = Tsum + (1 - 1 / tauT) (3 / rho) (c_q - u) . S.
// The reference's mirror marker has no definition in its sources and is not
// provided.
#include "psl_internal.h"

#include <cmath>
#include <vector>

namespace psl {
namespace {

struct ScalarParams {
  float* f;
  float* T;
  Shift27 sh;
  uint32_t n;
  int nx, ny, nz;
  const uint8_t* flags;
  const float4* bc;
  const float* given;
  const float* cs_cell;
  float omega, tau, cs, tau_t;
};

__device__ __forceinline__ void scalar_collide(const float (&f)[27], float (&T)[27], float tau_t) {
  float rho, ux, uy, uz;
  rho_u(f, rho, ux, uy, uz);
  float tsum = 0.f;
#pragma unroll
  for (int q = 0; q < 27; ++q) tsum += f[q] * T[q];
  tsum = tsum / rho;
  float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
  for (int q = 1; q < 27; ++q) {
    const float d = f[q] * (T[q] - tsum);
    if (CX[q]) sx += CX[q] > 0 ? d : -d;
    if (CY[q]) sy += CY[q] > 0 ? d : -d;
    if (CZ[q]) sz += CZ[q] > 0 ? d : -d;
  }
  const float rhodiv = 3.f / rho;
  const float bracket = 1.f - 1.f / tau_t;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const float fi = rhodiv * (((float)CX[q] - ux) * sx + ((float)CY[q] - uy) * sy + ((float)CZ[q] - uz) * sz);
    T[q] = tsum + bracket * fi;
  }
}

__global__ __launch_bounds__(256) void k_scalar_update(ScalarParams p) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= p.n) return;
  const int fl = p.flags ? p.flags[cell] : 0;
  const int kind = fl & 7;
  if (kind == KIND_GHOST || kind == KIND_COVERED) return;
  uint32_t slot[27];
  float f[27], T[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = cell + (uint32_t)p.sh.s[q];
    if (s >= p.n) s -= p.n;
    slot[q] = s;
    f[q] = p.f[(size_t)q * p.n + s];
    T[q] = p.T[(size_t)q * p.n + s];
  }
  if (kind == KIND_BOUNCEBACK) {
    bounceback(f);
    bounceback(T);
  } else {
    if (kind != KIND_FLUID) {
      const uint32_t nxy = (uint32_t)p.nx * p.ny;
      const int k = (int)(cell / nxy), rem = (int)(cell % nxy);
      const int j = rem / p.nx, i = rem % p.nx;
      const int code = outer_normal_code(i, j, k, p.nx, p.ny, p.nz, fl);
      const float4 g = p.bc[(fl >> 6) & 3];
      boundary_cell(f, kind, code, g.x, g.y, g.z, g.w);
    }
    if (p.given) {
      const float gv = p.given[cell];
      if (gv == gv) {
#pragma unroll
        for (int q = 0; q < 27; ++q) T[q] = gv;
      }
    }
    const float cs = p.cs_cell ? p.cs_cell[cell] : p.cs;
    collide<false>(f, Relax{p.omega, p.tau, cs}, cs);
    scalar_collide(f, T, p.tau_t);
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    p.f[(size_t)q * p.n + slot[q]] = f[q];
    p.T[(size_t)q * p.n + slot[q]] = T[q];
  }
}

// getScalarTransport (include/cellFunctions.h:228-235): sum f T / sum f per cell
__global__ __launch_bounds__(256) void k_scalar_field(const float* __restrict__ f, const float* __restrict__ T,
                                                      Shift27 sh, uint32_t n, float* __restrict__ out) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
  float fsum = 0.f, ft = 0.f;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = cell + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    const float a = f[(size_t)q * n + s];
    fsum += a;
    ft += a * T[(size_t)q * n + s];
  }
  out[cell] = ft / fsum;
}

// applyZeroGradientOutletScalarTransportBC (scalarTransport/scalarTransportMain.cu:95-135):
// a non-bounceback cell of the last x layer takes the 27 scalar populations of
// its -x neighbour at their shifted slots.  One thread per (q, j, k): the layer
// is strided by nx in memory, so every copy is its own line and the launch
// needs the 27x more waves to keep enough of them in flight.  The written and
// read cells are disjoint layers, so no slot is both read and written.
__global__ __launch_bounds__(256) void k_scalar_outlet_x(float* __restrict__ T, Shift27 sh, uint32_t n, int nx,
                                                         int ny, int nz, const uint8_t* __restrict__ flags) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t m = (uint32_t)ny * nz;
  if (t >= 27u * m) return;
  const uint32_t q = t / m, r = t - q * m;
  const uint32_t j = r % ny, k = r / ny;
  const uint32_t cw = (k * ny + j) * nx + (nx - 1), cr = cw - 1;
  if (flags && (flags[cw] & 7) == KIND_BOUNCEBACK) return;
  uint32_t sw = cw + (uint32_t)sh.s[q], sr = cr + (uint32_t)sh.s[q];
  if (sw >= n) sw -= n;
  if (sr >= n) sr -= n;
  T[(size_t)q * n + sw] = T[(size_t)q * n + sr];
}

}  // namespace

// the scalar value of every cell (n floats, device) on the grid's stream
int scalar_field_device(psl_grid* g, float* out) {
  PSL_CHECK_ARG(g && out && g->scalar_T, "scalar transport not enabled");
  PSL_CHECK_ARG(g->storage == PSL_F32, "scalar transport needs FP32 distributions");
  Shift27 sh;
  for (int q = 0; q < 27; ++q) sh.s[q] = g->shift[q];
  const uint32_t n = (uint32_t)g->info.cellCount;
  k_scalar_field<<<(n + 255) / 256, 256, 0, g->stream>>>((const float*)g->f, g->scalar_T, sh, n, out);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}
}  // namespace psl

using namespace psl;

extern "C" {

int psl_scalar_enable(psl_grid* g, float tau_t) {
  PSL_CHECK_ARG(g, "null grid");
  PSL_CHECK_ARG(g->storage == PSL_F32, "scalar transport needs FP32 distributions");
  PSL_CHECK_ARG(tau_t > 0.5f, "tauT must exceed 1/2");
  const size_t bytes = 27ull * (size_t)g->info.cellCount * sizeof(float);
  if (!g->scalar_T) PSL_HIP_TRY(hipMalloc(&g->scalar_T, bytes));
  PSL_HIP_TRY(hipMemset(g->scalar_T, 0, bytes));  // TArray.setValue(0)
  g->tau_t = tau_t;
  return PSL_OK;
}

int psl_scalar_set_given(psl_grid* g, const float* given) {
  PSL_CHECK_ARG(g, "null grid");
  if (!given) {
    if (g->scalar_given) PSL_HIP_TRY(hipFree(g->scalar_given));
    g->scalar_given = nullptr;
    return PSL_OK;
  }
  const size_t bytes = (size_t)g->info.cellCount * sizeof(float);
  if (!g->scalar_given) PSL_HIP_TRY(hipMalloc(&g->scalar_given, bytes));
  PSL_HIP_TRY(hipMemcpy(g->scalar_given, given, bytes, hipMemcpyHostToDevice));
  return PSL_OK;
}

// T_q(cell) in logical order, host[q * N + cell]
int psl_scalar_get(psl_grid* g, float* host) {
  PSL_CHECK_ARG(g && host && g->scalar_T, "scalar transport not enabled");
  const size_t n = (size_t)g->info.cellCount;
  std::vector<float> raw(27 * n);
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  PSL_HIP_TRY(hipMemcpy(raw.data(), g->scalar_T, raw.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int q = 0; q < 27; ++q)
    for (size_t c = 0; c < n; ++c) host[q * n + c] = raw[q * n + (c + (size_t)g->shift[q]) % n];
  return PSL_OK;
}

int psl_scalar_set(psl_grid* g, const float* host) {
  PSL_CHECK_ARG(g && host && g->scalar_T, "scalar transport not enabled");
  const size_t n = (size_t)g->info.cellCount;
  std::vector<float> raw(27 * n);
  for (int q = 0; q < 27; ++q)
    for (size_t c = 0; c < n; ++c) raw[q * n + (c + (size_t)g->shift[q]) % n] = host[q * n + c];
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  PSL_HIP_TRY(hipMemcpy(g->scalar_T, raw.data(), raw.size() * sizeof(float), hipMemcpyHostToDevice));
  return PSL_OK;
}

int psl_scalar_apply_zero_gradient_outlet_x(psl_grid* g) {
  PSL_CHECK_ARG(g && g->scalar_T, "scalar transport not enabled");
  PSL_CHECK_ARG(g->info.cellCountX >= 2, "the outlet needs two x layers");
  Shift27 sh;
  for (int q = 0; q < 27; ++q) sh.s[q] = g->shift[q];
  const psl_info& I = g->info;
  const uint32_t m = 27u * (uint32_t)I.cellCountY * I.cellCountZ;
  k_scalar_outlet_x<<<(m + 255) / 256, 256, 0, g->stream>>>(g->scalar_T, sh, (uint32_t)I.cellCount, I.cellCountX,
                                                            I.cellCountY, I.cellCountZ, g->flags);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

int psl_scalar_apply_local_cell_update(psl_grid* g) {
  PSL_CHECK_ARG(g && g->scalar_T, "scalar transport not enabled");
  PSL_CHECK_ARG(g->storage == PSL_F32, "scalar transport needs FP32 distributions");
  ScalarParams p;
  p.f = (float*)g->f;
  p.T = g->scalar_T;
  for (int q = 0; q < 27; ++q) p.sh.s[q] = g->shift[q];
  p.n = (uint32_t)g->info.cellCount;
  p.nx = g->info.cellCountX;
  p.ny = g->info.cellCountY;
  p.nz = g->info.cellCountZ;
  p.flags = g->flags;
  p.bc = g->bc_dev;
  p.given = g->scalar_given;
  p.cs_cell = g->cs_cell;
  p.tau = 3.f * g->info.nu + 0.5f;
  p.omega = 1.f / p.tau;
  p.cs = g->cs;
  p.tau_t = g->tau_t;
  k_scalar_update<<<(p.n + 255) / 256, 256, 0, g->stream>>>(p);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

}  // extern "C"
