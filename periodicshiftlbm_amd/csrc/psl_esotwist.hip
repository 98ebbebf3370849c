// Esotwist streaming on directly addressed (DIAD) grids.
//
// include/applyStreaming.h:44-192 and include/applyLocalCellUpdate.h:77-171.
// A DIAD grid stores only existing cells; cell c carries its (i, j, k) and the
// indices of its +i, +j, +k, +ij, +ik, +jk, +ijk neighbours.  Population q of
// a cell is read from and written to the cell itself or one of those
// neighbours (tables ESO_READ / ESO_WRITE), and a parity bit flips between
// steps: with the flipper set, reads use q and writes the opposite direction,
// otherwise the reverse.  Streaming is therefore only the parity flip; the
// update is in place.  Bounceback cells are skipped and reflect implicitly.
//
// Layout: fArray(q, cell) at f[q * n + cell]; ijk[3][n]; nbr[7][n].
#include "psl_internal.h"

#include <cstring>

namespace psl {

// neighbour slot per population: 0 = cell, 1 = i, 2 = j, 3 = k, 4 = ij, 5 = ik, 6 = jk, 7 = ijk
constexpr int ESO_READ[27] = {0, 0, 1, 3, 0, 2, 0, 3, 1, 0, 5, 4, 0, 3, 2, 1, 2, 0, 6, 5, 2, 4, 3, 6, 1, 7, 0};
constexpr int ESO_WRITE[27] = {0, 1, 0, 0, 3, 0, 2, 1, 3, 5, 0, 0, 4, 2, 3, 2, 1, 6, 0, 2, 5, 3, 4, 1, 6, 0, 7};
__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

struct EsoParams {
  uint32_t n;
  int nx, ny, nz;
  float omega, tau, cs;
  const uint8_t* flags;
  const int* ijk;  // 3 x n
  const int* nbr;  // 7 x n
  const float4* bc;  // 4 given-value slots (device)
};

template <class St>
__device__ __forceinline__ void eso_slots(const EsoParams& p, uint32_t cell, uint32_t (&slot)[8]) {
  slot[0] = cell;
#pragma unroll
  for (int t = 0; t < 7; ++t) slot[t + 1] = (uint32_t)p.nbr[(size_t)t * p.n + cell];
}

template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_update(typename St::T* __restrict__ f, EsoParams p) {
  typedef typename St::T T;
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= p.n) return;
  const int fl = p.flags ? p.flags[cell] : 0;
  int kind = fl & 7;
  if (kind == KIND_BOUNCEBACK) return;  // reflected implicitly by the maps
  if (kind == KIND_GHOST || kind == KIND_COVERED) kind = KIND_GIVEN_NONE;
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
  float v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int fi = FLIP ? q : opp(q);
    v[q] = St::load(f[(size_t)fi * p.n + slot[ESO_READ[q]]]);
  }
  Relax rx{p.omega, p.tau, p.cs};
  if (kind == KIND_FLUID) {
    collide<St::W_OFF>(v, rx, p.cs);
  } else {
    if (St::W_OFF) {
#pragma unroll
      for (int q = 0; q < 27; ++q) v[q] += weight(q);
    }
    const int i = p.ijk[cell], j = p.ijk[(size_t)p.n + cell], k = p.ijk[2 * (size_t)p.n + cell];
    const int code = outer_normal_code(i, j, k, p.nx, p.ny, p.nz, fl);
    const float4 g = p.bc[(fl >> 6) & 3];  // indexed load, no select chain
    boundary_cell(v, kind, code, g.x, g.y, g.z, g.w);
    collide<false>(v, rx, p.cs);
    if (St::W_OFF) {
#pragma unroll
      for (int q = 0; q < 27; ++q) v[q] -= weight(q);
    }
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int fi = FLIP ? opp(q) : q;
    f[(size_t)fi * p.n + slot[ESO_WRITE[q]]] = St::store(v[q]);
  }
}

// fillEquilibriumFromFunction(DIADGridStruct&) (include/fillEquilibrium.h:25-70):
// feq written through the write maps with the current parity.
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_fill(typename St::T* __restrict__ f, EsoParams p, const float* rho,
                                                  const float* ux, const float* uy, const float* uz) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= p.n) return;
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
  const float r = rho ? rho[cell] : 1.f, a = ux ? ux[cell] : 0.f, b = uy ? uy[cell] : 0.f, c = uz ? uz[cell] : 0.f;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    float e = feq(q, r, a, b, c);
    if (St::W_OFF) e -= weight(q);
    const int fi = FLIP ? opp(q) : q;
    f[(size_t)fi * p.n + slot[ESO_WRITE[q]]] = St::store(e);
  }
}

// Post-collision populations of every cell, read back through the write maps
// of the current parity (the state a following streaming step would move).
template <class St, bool FLIP, bool SET>
__global__ __launch_bounds__(256) void k_eso_io(typename St::T* __restrict__ f, EsoParams p, float* buf) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= p.n) return;
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int fi = FLIP ? opp(q) : q;
    const size_t at = (size_t)fi * p.n + slot[ESO_WRITE[q]];
    if (SET) {
      float v = buf[(size_t)q * p.n + cell];
      if (St::W_OFF) v -= weight(q);
      f[at] = St::store(v);
    } else {
      float v = St::load(f[at]);
      if (St::W_OFF) v += weight(q);
      buf[(size_t)q * p.n + cell] = v;
    }
  }
}

// Populations of a cell as its last update left them: the write maps of the
// current parity (getEsotwistWriteIndex, include/applyStreaming.h:138-192),
// full values (the lattice weights added back for 16-bit storage).
template <class St, bool FLIP>
__device__ __forceinline__ void eso_get(const typename St::T* __restrict__ f, const EsoParams& p, uint32_t cell,
                                        float (&v)[27]) {
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int fi = FLIP ? opp(q) : q;
    float x = St::load(f[(size_t)fi * p.n + slot[ESO_WRITE[q]]]);
    if (St::W_OFF) x += weight(q);
    v[q] = x;
  }
}

template <class St, bool FLIP>
__device__ __forceinline__ void eso_put(typename St::T* __restrict__ f, const EsoParams& p, uint32_t cell,
                                        const float (&v)[27]) {
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int fi = FLIP ? opp(q) : q;
    float x = v[q];
    if (St::W_OFF) x -= weight(q);
    f[(size_t)fi * p.n + slot[ESO_WRITE[q]]] = St::store(x);
  }
}

// writeToCoarseGridInterface(DIAD): include/gridRefinementFunctions.h:540-622.
// Eight lanes per interface cell, one per fine child (the read cell and its
// +i +j +k +ij +ik +jk +ijk neighbours): each lane gathers its child, the
// group sums the eight in child order through lane shuffles (the reference's
// sequential order, so every lane holds the same sum), and lane 0 rescales
// fine -> coarse and writes the coarse cell.  Eight independent gathers per
// cell keep enough loads in flight on the thin interface lists.
template <class St, bool FC, bool FF>
__global__ __launch_bounds__(256) void k_eso_to_coarse(typename St::T* __restrict__ fc, EsoParams pc,
                                                       const typename St::T* __restrict__ ff, EsoParams pf,
                                                       const int* __restrict__ write, const int* __restrict__ read,
                                                       uint32_t m) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t g = t >> 3, w = t & 7;
  const bool valid = g < m;  // the eight lanes of a group agree
  float v[27];
  if (valid) {
    const uint32_t rc = (uint32_t)read[g];
    const uint32_t child = w == 0 ? rc : (uint32_t)pf.nbr[(size_t)(w - 1) * pf.n + rc];
    eso_get<St, FF>(ff, pf, child, v);
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = 0.f;
  }
  float acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    float a = 0.f;
#pragma unroll
    for (int src = 0; src < 8; ++src) a += __shfl(v[q], src, 8);
    acc[q] = a * 0.125f;
  }
  if (!valid || w != 0) return;
  rescale(acc, false);
  eso_put<St, FC>(fc, pc, (uint32_t)write[g], acc);
}

// writeToFineGridInterface(DIAD): include/gridRefinementFunctions.h:624-702.
// Eight lanes per interface cell: each reads the coarse cell (the same lines),
// rescales coarse -> fine (identical in every lane) and writes one child: the
// write cell or its +i +j +k +ij +ik +jk +ijk neighbour.
template <class St, bool FC, bool FF>
__global__ __launch_bounds__(256) void k_eso_to_fine(const typename St::T* __restrict__ fc, EsoParams pc,
                                                     typename St::T* __restrict__ ff, EsoParams pf,
                                                     const int* __restrict__ write, const int* __restrict__ read,
                                                     uint32_t m) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t g = t >> 3, w = t & 7;
  if (g >= m) return;
  float v[27];
  eso_get<St, FC>(fc, pc, (uint32_t)read[g], v);
  rescale(v, true);
  const uint32_t wc = (uint32_t)write[g];
  const uint32_t child = w == 0 ? wc : (uint32_t)pf.nbr[(size_t)(w - 1) * pf.n + wc];
  eso_put<St, FF>(ff, pf, child, v);
}

// applyNonReflectiveOutletZ (include/boundaryConditions/applyNonReflectiveOutlet.h:9-73):
// for the nine populations with c_z < 0 of every listed outlet cell,
// f = invSqrt3 f(-k neighbour) + (1 - invSqrt3) f(cell), both as the last update
// wrote them, stored where the next step reads them (the read maps of the
// opposite parity).
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_outlet_z(typename St::T* __restrict__ f, EsoParams p,
                                                      const int* __restrict__ outlet, const int* __restrict__ kmin,
                                                      uint32_t m) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= m) return;
  constexpr int QZ[9] = {3, 7, 10, 13, 18, 19, 22, 23, 25};
  constexpr float INV_SQRT3 = 0.577350269f;  // the drivers' invSqrt3
  const uint32_t cell = (uint32_t)outlet[t];
  const uint32_t km = (uint32_t)kmin[cell];
  uint32_t sc[8], sk[8];
  eso_slots<St>(p, cell, sc);
  eso_slots<St>(p, km, sk);
  float v[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    const int q = QZ[a];
    const int fi = FLIP ? opp(q) : q;
    float x = St::load(f[(size_t)fi * p.n + sk[ESO_WRITE[q]]]);
    float y = St::load(f[(size_t)fi * p.n + sc[ESO_WRITE[q]]]);
    if (St::W_OFF) {
      x += weight(q);
      y += weight(q);
    }
    v[a] = INV_SQRT3 * x + (1.f - INV_SQRT3) * y;
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    const int q = QZ[a];
    const int fi = FLIP ? opp(q) : q;  // the read maps of parity !FLIP
    float x = v[a];
    if (St::W_OFF) x -= weight(q);
    f[(size_t)fi * p.n + sc[ESO_READ[q]]] = St::store(x);
  }
}

// applyMovingBounceback (include/boundaryConditions/applyMovingBounceback.h:8-62)
// on the listed cells, which the local update skips as bounceback cells
// (M35Full/applyLocalCellUpdateCustomized.h:69-77): populations read through
// the read maps, each opposite pair swapped with f_q = f_opp(q) + 6 w_q (c_q . u) and
// f_opp(q) = f_q - 6 w_q (c_q . u) for the odd member q of the pair,
// written through the write maps.  w_q = w_opp(q), so the swap acts on the
// stored values (f or f - w) directly, and a wall at rest leaves them exactly
// as they were, which is the implicit bounceback of a skipped cell.
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_moving_bb(typename St::T* __restrict__ f, EsoParams p,
                                                       const int* __restrict__ cells, const float* __restrict__ u3,
                                                       uint32_t m) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= m) return;
  const uint32_t cell = (uint32_t)cells[t];
  const float ux = u3[3 * (size_t)t], uy = u3[3 * (size_t)t + 1], uz = u3[3 * (size_t)t + 2];
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
  float v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] = St::load(f[(size_t)(FLIP ? q : opp(q)) * p.n + slot[ESO_READ[q]]]);
#pragma unroll
  for (int q = 1; q < 27; q += 2) {
    const float cu = (float)CX[q] * ux + (float)CY[q] * uy + (float)CZ[q] * uz;
    // 2 w_q / c_s^2 as the reference writes it: 4/9 faces, 1/9 edges, 1/36 corners
    const float a = (q < 7 ? 4.0f / 9.0f : (q < 19 ? 1.0f / 9.0f : 1.0f / 36.0f)) * cu;
    const float lo = v[q];
    v[q] = v[q + 1] + a;
    v[q + 1] = lo - a;
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) f[(size_t)(FLIP ? opp(q) : q) * p.n + slot[ESO_WRITE[q]]] = St::store(v[q]);
}

// Forced-velocity cells (M35Full/applyLocalCellUpdateCustomized.h:80-86,114-121,
// getForcing in M35Full/m35Main.cu:100-118) on the listed cells, which the
// local update skips as bounceback cells: populations read through the read
// maps, force g = rho (u_target - u) from the cell's own rho and u, the
// forcing collision, written through the write maps.  No box-face treatment.
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_forced(typename St::T* __restrict__ f, EsoParams p,
                                                    const int* __restrict__ cells, const float* __restrict__ ut3,
                                                    uint32_t m) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= m) return;
  const uint32_t cell = (uint32_t)cells[t];
  uint32_t slot[8];
  eso_slots<St>(p, cell, slot);
  float v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    float x = St::load(f[(size_t)(FLIP ? q : opp(q)) * p.n + slot[ESO_READ[q]]]);
    if (St::W_OFF) x += weight(q);
    v[q] = x;
  }
  float r, a, b, c;
  rho_u(v, r, a, b, c);
  const float gx = r * (ut3[3 * (size_t)t] - a), gy = r * (ut3[3 * (size_t)t + 1] - b),
              gz = r * (ut3[3 * (size_t)t + 2] - c);
  const Relax rx{p.omega, p.tau, p.cs};
  collide_forced(v, rx, p.cs, gx, gy, gz);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    float x = v[q];
    if (St::W_OFF) x -= weight(q);
    f[(size_t)(FLIP ? opp(q) : q) * p.n + slot[ESO_WRITE[q]]] = St::store(x);
  }
}

// getTorque (M35Full/m35Main.cu:176-260) over the listed cells, all six
// components in FP64 per-block partial sums: per cell the force
// g = rho (u_target - u) from the populations the next update will read
// (read maps of the parity after FLIP, as getTorque reads them), times
// `scale` (convertToPhysicalForce, include/cellFunctions.h:221-226: newtons),
// and r x g with r = (i res + ox, j res + oy, k res + oz) in millimetres.
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_forcing_sum(const typename St::T* __restrict__ f, EsoParams p,
                                                         double* __restrict__ partials, const int* __restrict__ cells,
                                                         const float* __restrict__ ut3, uint32_t m, float scale,
                                                         float res, float ox, float oy, float oz) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (t < m) {
    const uint32_t cell = (uint32_t)cells[t];
    uint32_t slot[8];
    eso_slots<St>(p, cell, slot);
    float v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      float x = St::load(f[(size_t)(FLIP ? opp(q) : q) * p.n + slot[ESO_READ[q]]]);
      if (St::W_OFF) x += weight(q);
      v[q] = x;
    }
    float r, a, b, c;
    rho_u(v, r, a, b, c);
    const float gx = r * (ut3[3 * (size_t)t] - a) * scale, gy = r * (ut3[3 * (size_t)t + 1] - b) * scale,
                gz = r * (ut3[3 * (size_t)t + 2] - c) * scale;
    const float x = (float)p.ijk[cell] * res + ox, y = (float)p.ijk[(size_t)p.n + cell] * res + oy,
                z = (float)p.ijk[2 * (size_t)p.n + cell] * res + oz;
    acc[0] = (double)gx;
    acc[1] = (double)gy;
    acc[2] = (double)gz;
    acc[3] = (double)(y * gz - z * gy);
    acc[4] = (double)(z * gx - x * gz);
    acc[5] = (double)(-gx * y + gy * x);
  }
  __shared__ double sh[6][256];
  for (int a = 0; a < 6; ++a) sh[a][threadIdx.x] = acc[a];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < (unsigned)s)
      for (int a = 0; a < 6; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) partials[6 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// rho, ux, uy, uz of every cell (getRhoUxUyUz, include/cellFunctions.h:135-155)
// from the populations the last update wrote: out[a * n + cell]
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_macro(const typename St::T* __restrict__ f, EsoParams p,
                                                   float* __restrict__ out) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= p.n) return;
  float v[27];
  eso_get<St, FLIP>(f, p, cell, v);
  float r, a, b, c;
  rho_u(v, r, a, b, c);
  out[cell] = r;
  out[(size_t)p.n + cell] = a;
  out[2 * (size_t)p.n + cell] = b;
  out[3 * (size_t)p.n + cell] = c;
}

// getSphereDrag(DIADGridStruct&) (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:132-198),
// all three components: per bounceback cell rho (u - u after bounceback), in
// FP64 per-block partial sums.
// With `list`, the cells list[0, m) instead (the moving-wall cells, whose
// share is taken out of the sum: the reference does not count them as
// bounceback cells).
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_momentum(const typename St::T* __restrict__ f, EsoParams p,
                                                      double* __restrict__ partials, const int* __restrict__ list,
                                                      uint32_t m) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const bool in = list ? t < m : (t < p.n && p.flags && (p.flags[t] & 7) == KIND_BOUNCEBACK);
  const uint32_t cell = list ? (in ? (uint32_t)list[t] : 0u) : t;
  double acc[3] = {0.0, 0.0, 0.0};
  if (in) {
    float v[27];
    eso_get<St, FLIP>(f, p, cell, v);
    float r, a, b, c;
    rho_u(v, r, a, b, c);
    bounceback(v);
    float r2, a2, b2, c2;
    rho_u(v, r2, a2, b2, c2);
    acc[0] = (double)(r * (a - a2));
    acc[1] = (double)(r * (b - b2));
    acc[2] = (double)(r * (c - c2));
  }
  __shared__ double sh[3][256];
  for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = acc[a];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < (unsigned)s)
      for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) partials[3 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace psl

using namespace psl;

static EsoParams eso_params(const psl_diad* d) {
  EsoParams p;
  p.n = (uint32_t)d->n;
  p.nx = d->info.cellCountX;
  p.ny = d->info.cellCountY;
  p.nz = d->info.cellCountZ;
  const float tau = 3.f * d->info.nu + 0.5f;
  p.tau = tau;
  p.omega = 1.f / tau;
  p.cs = d->cs;
  p.flags = d->flags;
  p.ijk = d->ijk;
  p.nbr = d->nbr;
  p.bc = d->bc_dev;
  return p;
}

template <class St>
static int eso_update(psl_diad* d) {
  const EsoParams p = eso_params(d);
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  if (d->flipper) k_eso_update<St, true><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p);
  else k_eso_update<St, false><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_fill(psl_diad* d, const float* r, const float* a, const float* b, const float* c) {
  const EsoParams p = eso_params(d);
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  if (d->flipper) k_eso_fill<St, true><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p, r, a, b, c);
  else k_eso_fill<St, false><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p, r, a, b, c);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St, bool SET>
static int eso_io(psl_diad* d, float* buf) {
  const EsoParams p = eso_params(d);
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  if (d->flipper) k_eso_io<St, true, SET><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p, buf);
  else k_eso_io<St, false, SET><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p, buf);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St, bool FC, bool FF>
static int eso_comm_t(psl_diad* c, psl_diad* f) {
  typedef typename St::T T;
  const EsoParams pc = eso_params(c), pf = eso_params(f);
  if (c->n_f2c) {
    k_eso_to_coarse<St, FC, FF><<<(unsigned)((8 * c->n_f2c + 255) / 256), 256, 0, c->stream>>>(
        (T*)c->f, pc, (const T*)f->f, pf, c->f2c_write, c->f2c_read, (uint32_t)c->n_f2c);
    PSL_LAUNCH_CHECK();
  }
  if (c->n_c2f) {
    k_eso_to_fine<St, FC, FF><<<(unsigned)((8 * c->n_c2f + 255) / 256), 256, 0, c->stream>>>(
        (const T*)c->f, pc, (T*)f->f, pf, c->c2f_write, c->c2f_read, (uint32_t)c->n_c2f);
    PSL_LAUNCH_CHECK();
  }
  return PSL_OK;
}

template <class St>
static int eso_comm(psl_diad* c, psl_diad* f) {
  if (c->flipper) return f->flipper ? eso_comm_t<St, true, true>(c, f) : eso_comm_t<St, true, false>(c, f);
  return f->flipper ? eso_comm_t<St, false, true>(c, f) : eso_comm_t<St, false, false>(c, f);
}

template <class St>
static int eso_momentum(psl_diad* d, uint32_t nb, const int* list, uint32_t m, double* partials) {
  const EsoParams p = eso_params(d);
  if (d->flipper)
    k_eso_momentum<St, true><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, partials, list, m);
  else
    k_eso_momentum<St, false><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, partials, list, m);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_outlet(psl_diad* d) {
  if (!d->n_outlet) return PSL_OK;
  const EsoParams p = eso_params(d);
  const unsigned nb = (unsigned)((d->n_outlet + 255) / 256);
  typedef typename St::T T;
  if (d->flipper) k_eso_outlet_z<St, true><<<nb, 256, 0, d->stream>>>((T*)d->f, p, d->outlet, d->kmin, (uint32_t)d->n_outlet);
  else k_eso_outlet_z<St, false><<<nb, 256, 0, d->stream>>>((T*)d->f, p, d->outlet, d->kmin, (uint32_t)d->n_outlet);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_moving(psl_diad* d) {
  if (!d->n_moving) return PSL_OK;
  const EsoParams p = eso_params(d);
  const unsigned nb = (unsigned)((d->n_moving + 255) / 256);
  typedef typename St::T T;
  if (d->flipper)
    k_eso_moving_bb<St, true><<<nb, 256, 0, d->stream>>>((T*)d->f, p, d->moving, d->moving_u, (uint32_t)d->n_moving);
  else
    k_eso_moving_bb<St, false><<<nb, 256, 0, d->stream>>>((T*)d->f, p, d->moving, d->moving_u, (uint32_t)d->n_moving);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_forced(psl_diad* d) {
  if (!d->n_forced) return PSL_OK;
  const EsoParams p = eso_params(d);
  const unsigned nb = (unsigned)((d->n_forced + 255) / 256);
  typedef typename St::T T;
  if (d->flipper)
    k_eso_forced<St, true><<<nb, 256, 0, d->stream>>>((T*)d->f, p, d->forced, d->forced_u, (uint32_t)d->n_forced);
  else
    k_eso_forced<St, false><<<nb, 256, 0, d->stream>>>((T*)d->f, p, d->forced, d->forced_u, (uint32_t)d->n_forced);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_forcing_sum(psl_diad* d, uint32_t nb, double* partials, float scale) {
  const EsoParams p = eso_params(d);
  const psl_info& I = d->info;
  typedef typename St::T T;
  if (d->flipper)
    k_eso_forcing_sum<St, true><<<nb, 256, 0, d->stream>>>((const T*)d->f, p, partials, d->forced, d->forced_u,
                                                           (uint32_t)d->n_forced, scale, I.res, I.ox, I.oy, I.oz);
  else
    k_eso_forcing_sum<St, false><<<nb, 256, 0, d->stream>>>((const T*)d->f, p, partials, d->forced, d->forced_u,
                                                            (uint32_t)d->n_forced, scale, I.res, I.ox, I.oy, I.oz);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_macro(psl_diad* d, float* out) {
  const EsoParams p = eso_params(d);
  const unsigned nb = (unsigned)((d->n + 255) / 256);
  if (d->flipper) k_eso_macro<St, true><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, out);
  else k_eso_macro<St, false><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, out);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

#define PSL_ESO_DISPATCH(fn, ...)                                         \
  (d->storage == PSL_F32 ? fn<StF32>(__VA_ARGS__)                         \
                         : d->storage == PSL_F16DEV ? fn<StF16>(__VA_ARGS__) : fn<StBF16>(__VA_ARGS__))

extern "C" {

int psl_diad_create(const psl_info* info, int storage, int64_t n, const int* ijk3, const int* nbr7, psl_diad** out) {
  PSL_CHECK_ARG(info && out && ijk3 && nbr7 && n > 0, "bad DIAD grid arguments");
  PSL_CHECK_ARG(n < (1ll << 31), "cell count must be below 2^31");
  PSL_CHECK_ARG(storage == PSL_F32 || storage == PSL_F16DEV || storage == PSL_BF16DEV, "unknown storage");
  for (int64_t t = 0; t < 7 * n; ++t)
    PSL_CHECK_ARG(nbr7[t] >= 0 && nbr7[t] < n, "neighbour index out of range");
  psl_diad* d = new psl_diad();
  d->info = *info;
  d->storage = storage;
  d->n = n;
  const size_t fb = 27ull * n * storage_bytes(storage);
  if (hipMalloc(&d->f, fb) != hipSuccess || hipMalloc(&d->bc_dev, 4 * sizeof(float4)) != hipSuccess ||
      hipMemset(d->bc_dev, 0, 4 * sizeof(float4)) != hipSuccess || hipMalloc(&d->ijk, 3ull * n * sizeof(int)) != hipSuccess ||
      hipMalloc(&d->nbr, 7ull * n * sizeof(int)) != hipSuccess ||
      hipMemcpy(d->ijk, ijk3, 3ull * n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->nbr, nbr7, 7ull * n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("DIAD allocation / upload failed");
    if (d->f) (void)hipFree(d->f);
    if (d->bc_dev) (void)hipFree(d->bc_dev);
    if (d->ijk) (void)hipFree(d->ijk);
    if (d->nbr) (void)hipFree(d->nbr);
    delete d;
    return PSL_E_NOMEM;
  }
  *out = d;
  return PSL_OK;
}

int psl_diad_destroy(psl_diad* d) {
  if (!d) return PSL_OK;
  (void)hipDeviceSynchronize();
  if (d->f) (void)hipFree(d->f);
  if (d->ijk) (void)hipFree(d->ijk);
  if (d->nbr) (void)hipFree(d->nbr);
  if (d->flags) (void)hipFree(d->flags);
  if (d->bc_dev) (void)hipFree(d->bc_dev);
  for (int* a : {d->c2f_read, d->c2f_write, d->f2c_write, d->f2c_read, d->outlet, d->kmin, d->moving, d->forced})
    if (a) (void)hipFree(a);
  if (d->moving_u) (void)hipFree(d->moving_u);
  if (d->forced_u) (void)hipFree(d->forced_u);
  if (d->partials) (void)hipFree(d->partials);
  delete d;
  return PSL_OK;
}

// Moving-wall and forced cells are updated by their own list kernels and must
// be skipped by the local update, so they must be flagged bounceback.
static int lists_are_bounceback(const psl_diad* d, const uint8_t* flags) {
  for (const std::vector<int>* list : {&d->moving_host, &d->forced_host})
    for (int c : *list)
      if (!flags || (flags[c] & 7) != KIND_BOUNCEBACK) {
        set_error("moving-wall and forced-velocity cells must be flagged PSL_BOUNCEBACK (cell " + std::to_string(c) +
                  ")");
        return PSL_E_ARG;
      }
  return PSL_OK;
}

int psl_diad_set_flags(psl_diad* d, const uint8_t* host) {
  PSL_CHECK_ARG(d, "null grid");
  const int rc = lists_are_bounceback(d, host);
  if (rc) return rc;
  if (!host) {
    if (d->flags) PSL_HIP_TRY(hipFree(d->flags));
    d->flags = nullptr;
    return PSL_OK;
  }
  if (!d->flags) PSL_HIP_TRY(hipMalloc(&d->flags, (size_t)d->n));
  PSL_HIP_TRY(hipMemcpy(d->flags, host, (size_t)d->n, hipMemcpyHostToDevice));
  return PSL_OK;
}

int psl_diad_set_bc_values(psl_diad* d, const float* v16) {
  PSL_CHECK_ARG(d && v16, "null argument");
  PSL_HIP_TRY(hipMemcpy(d->bc_dev, v16, 4 * sizeof(float4), hipMemcpyHostToDevice));
  return PSL_OK;
}

int psl_diad_set_smagorinsky(psl_diad* d, float cs) {
  PSL_CHECK_ARG(d, "null grid");
  d->cs = cs;
  return PSL_OK;
}

int psl_diad_set_stream(psl_diad* d, void* s) {
  PSL_CHECK_ARG(d, "null grid");
  d->stream = (hipStream_t)s;
  return PSL_OK;
}

int psl_diad_fill_equilibrium(psl_diad* d, const float* rho, const float* ux, const float* uy, const float* uz) {
  PSL_CHECK_ARG(d, "null grid");
  const size_t n = (size_t)d->n;
  float* dv[4] = {nullptr, nullptr, nullptr, nullptr};
  const float* hv[4] = {rho, ux, uy, uz};
  int rc = PSL_OK;
  for (int a = 0; a < 4 && rc == PSL_OK; ++a)
    if (hv[a] && (hipMalloc(&dv[a], n * sizeof(float)) != hipSuccess ||
                  hipMemcpy(dv[a], hv[a], n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))
      rc = PSL_E_HIP;
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_fill, d, dv[0], dv[1], dv[2], dv[3]);
  if (hipStreamSynchronize(d->stream) != hipSuccess && rc == PSL_OK) rc = PSL_E_HIP;
  for (int a = 0; a < 4; ++a)
    if (dv[a]) (void)hipFree(dv[a]);
  if (rc == PSL_E_HIP) set_error("DIAD fill_equilibrium failed");
  return rc;
}

int psl_diad_apply_streaming(psl_diad* d) {
  PSL_CHECK_ARG(d, "null grid");
  d->flipper = !d->flipper;
  return PSL_OK;
}

int psl_diad_apply_local_cell_update(psl_diad* d) {
  PSL_CHECK_ARG(d, "null grid");
  int rc = PSL_ESO_DISPATCH(eso_update, d);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_moving, d);
  if (rc != PSL_OK) return rc;
  return PSL_ESO_DISPATCH(eso_forced, d);
}

int psl_diad_get_flipper(const psl_diad* d, int* out) {
  PSL_CHECK_ARG(d && out, "null argument");
  *out = d->flipper ? 1 : 0;
  return PSL_OK;
}

int psl_diad_get_f(psl_diad* d, float* host) {
  PSL_CHECK_ARG(d && host, "null argument");
  const size_t bytes = 27ull * d->n * sizeof(float);
  float* buf = nullptr;
  PSL_HIP_TRY(hipMalloc(&buf, bytes));
  int rc;
  switch (d->storage) {
    case PSL_F32: rc = eso_io<StF32, false>(d, buf); break;
    case PSL_F16DEV: rc = eso_io<StF16, false>(d, buf); break;
    default: rc = eso_io<StBF16, false>(d, buf); break;
  }
  if (rc == PSL_OK && hipMemcpyAsync(host, buf, bytes, hipMemcpyDeviceToHost, d->stream) != hipSuccess) rc = PSL_E_HIP;
  if (hipStreamSynchronize(d->stream) != hipSuccess && rc == PSL_OK) rc = PSL_E_HIP;
  (void)hipFree(buf);
  if (rc == PSL_E_HIP) set_error("DIAD get_f failed");
  return rc;
}

int psl_diad_set_f(psl_diad* d, const float* host) {
  PSL_CHECK_ARG(d && host, "null argument");
  const size_t bytes = 27ull * d->n * sizeof(float);
  float* buf = nullptr;
  PSL_HIP_TRY(hipMalloc(&buf, bytes));
  int rc = hipMemcpy(buf, host, bytes, hipMemcpyHostToDevice) == hipSuccess ? PSL_OK : PSL_E_HIP;
  if (rc == PSL_OK) {
    switch (d->storage) {
      case PSL_F32: rc = eso_io<StF32, true>(d, buf); break;
      case PSL_F16DEV: rc = eso_io<StF16, true>(d, buf); break;
      default: rc = eso_io<StBF16, true>(d, buf); break;
    }
  }
  if (hipStreamSynchronize(d->stream) != hipSuccess && rc == PSL_OK) rc = PSL_E_HIP;
  (void)hipFree(buf);
  if (rc == PSL_E_HIP) set_error("DIAD set_f failed");
  return rc;
}

static int upload_ints(int** dst, const int* src, int64_t n) {
  if (*dst) PSL_HIP_TRY(hipFree(*dst));
  *dst = nullptr;
  if (n == 0) return PSL_OK;
  PSL_HIP_TRY(hipMalloc(dst, (size_t)n * sizeof(int)));
  PSL_HIP_TRY(hipMemcpy(*dst, src, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  return PSL_OK;
}

int psl_diad_set_interface(psl_diad* coarse, const psl_diad* fine, const int* c2f_read, const int* c2f_write,
                           int64_t n_c2f, const int* f2c_write, const int* f2c_read, int64_t n_f2c) {
  PSL_CHECK_ARG(coarse && fine && n_c2f >= 0 && n_f2c >= 0, "bad interface arguments");
  PSL_CHECK_ARG((n_c2f == 0 || (c2f_read && c2f_write)) && (n_f2c == 0 || (f2c_write && f2c_read)), "null table");
  for (int64_t t = 0; t < n_c2f; ++t)
    PSL_CHECK_ARG(c2f_read[t] >= 0 && c2f_read[t] < coarse->n && c2f_write[t] >= 0 && c2f_write[t] < fine->n,
                  "coarse-to-fine index out of range");
  for (int64_t t = 0; t < n_f2c; ++t)
    PSL_CHECK_ARG(f2c_write[t] >= 0 && f2c_write[t] < coarse->n && f2c_read[t] >= 0 && f2c_read[t] < fine->n,
                  "fine-to-coarse index out of range");
  int rc = upload_ints(&coarse->c2f_read, c2f_read, n_c2f);
  if (rc == PSL_OK) rc = upload_ints(&coarse->c2f_write, c2f_write, n_c2f);
  if (rc == PSL_OK) rc = upload_ints(&coarse->f2c_write, f2c_write, n_f2c);
  if (rc == PSL_OK) rc = upload_ints(&coarse->f2c_read, f2c_read, n_f2c);
  coarse->n_c2f = rc == PSL_OK ? n_c2f : 0;
  coarse->n_f2c = rc == PSL_OK ? n_f2c : 0;
  return rc;
}

int psl_diad_set_outlet(psl_diad* d, const int* outlet_cells, int64_t n_outlet, const int* kmin) {
  PSL_CHECK_ARG(d && n_outlet >= 0 && (n_outlet == 0 || (outlet_cells && kmin)), "bad outlet arguments");
  for (int64_t t = 0; t < n_outlet; ++t) PSL_CHECK_ARG(outlet_cells[t] >= 0 && outlet_cells[t] < d->n, "outlet cell out of range");
  if (n_outlet)
    for (int64_t c = 0; c < d->n; ++c) PSL_CHECK_ARG(kmin[c] >= 0 && kmin[c] < d->n, "-k neighbour out of range");
  int rc = upload_ints(&d->outlet, outlet_cells, n_outlet);
  if (rc == PSL_OK) rc = upload_ints(&d->kmin, kmin, n_outlet ? d->n : 0);
  d->n_outlet = rc == PSL_OK ? n_outlet : 0;
  return rc;
}

// Upload of a cell list with three floats per cell (moving walls, forced
// cells).  Cells must be in range, listed once, and not in `other`.
static int set_cell_list(psl_diad* d, const int* cells, int64_t n, const float* u3, const std::vector<int>& other,
                         const char* what, int** dev, float** dev_u, int64_t* dev_n, std::vector<int>* host) {
  PSL_CHECK_ARG(n >= 0 && (n == 0 || (cells && u3)), std::string("bad ") + what + " arguments");
  for (int64_t t = 0; t < n; ++t) PSL_CHECK_ARG(cells[t] >= 0 && cells[t] < d->n, std::string(what) + " cell out of range");
  if (n) {
    std::vector<uint8_t> seen((size_t)d->n, 0);
    for (int c : other) seen[c] = 2;
    for (int64_t t = 0; t < n; ++t) {
      PSL_CHECK_ARG(seen[cells[t]] != 1, std::string(what) + " cell listed twice");
      PSL_CHECK_ARG(seen[cells[t]] != 2, std::string(what) + " cell is also a moving-wall or forced cell");
      seen[cells[t]] = 1;
    }
  }
  if (n) {
    // the listed cells must be flagged bounceback (flags as the device holds them)
    PSL_CHECK_ARG(d->flags, std::string(what) + " cells need flags (psl_diad_set_flags) first");
    std::vector<uint8_t> fl((size_t)d->n);
    PSL_HIP_TRY(hipMemcpy(fl.data(), d->flags, (size_t)d->n, hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < n; ++t)
      PSL_CHECK_ARG((fl[cells[t]] & 7) == KIND_BOUNCEBACK,
                    std::string(what) + " cell " + std::to_string(cells[t]) + " is not flagged PSL_BOUNCEBACK");
  }
  *dev_n = 0;
  host->clear();
  int rc = upload_ints(dev, cells, n);
  if (rc == PSL_OK) {
    if (*dev_u) (void)hipFree(*dev_u);
    *dev_u = nullptr;
    if (n) {
      if (hipMalloc(dev_u, 3 * (size_t)n * sizeof(float)) != hipSuccess ||
          hipMemcpy(*dev_u, u3, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        set_error(std::string(what) + " upload failed");
        rc = PSL_E_HIP;
      }
    }
  }
  if (rc == PSL_OK) {
    *dev_n = n;
    host->assign(cells, cells + n);
  }
  return rc;
}

int psl_diad_set_moving_bounceback(psl_diad* d, const int* cells, int64_t n, const float* u3) {
  PSL_CHECK_ARG(d, "null grid");
  return set_cell_list(d, cells, n, u3, d->forced_host, "moving-bounceback", &d->moving, &d->moving_u, &d->n_moving,
                       &d->moving_host);
}

int psl_diad_set_forced_velocity(psl_diad* d, const int* cells, int64_t n, const float* ut3) {
  PSL_CHECK_ARG(d, "null grid");
  return set_cell_list(d, cells, n, ut3, d->moving_host, "forced-velocity", &d->forced, &d->forced_u, &d->n_forced,
                       &d->forced_host);
}

int psl_diad_apply_nonreflective_outlet_z(psl_diad* d) {
  PSL_CHECK_ARG(d, "null grid");
  return PSL_ESO_DISPATCH(eso_outlet, d);
}

int psl_diad_coarse_fine_communication(psl_diad* coarse, psl_diad* fine) {
  PSL_CHECK_ARG(coarse && fine, "null grid");
  PSL_CHECK_ARG(coarse->storage == fine->storage, "levels must share the storage type");
  PSL_CHECK_ARG(coarse->stream == fine->stream, "levels must share the stream");
  psl_diad* d = coarse;
  return PSL_ESO_DISPATCH(eso_comm, coarse, fine);
}

// updateGrid(std::vector<DIADGridStruct>&, level)
// (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:200-209)
static int diad_recurse(psl_diad* const* lv, int nl, int l) {
  psl_diad* d = lv[l];
  d->flipper = !d->flipper;
  int rc = PSL_ESO_DISPATCH(eso_update, d);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_moving, d);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_forced, d);
  if (rc != PSL_OK || l == nl - 1) return rc;
  for (int s = 0; s < 2 && rc == PSL_OK; ++s) rc = diad_recurse(lv, nl, l + 1);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_comm, d, lv[l + 1]);
  return rc;
}

int psl_diad_update_grid(psl_diad* const* levels, int nlevels, int nsteps) {
  PSL_CHECK_ARG(levels && nlevels >= 1 && nsteps >= 0, "bad level list");
  for (int l = 0; l < nlevels; ++l) {
    PSL_CHECK_ARG(levels[l], "null level");
    PSL_CHECK_ARG(levels[l]->storage == levels[0]->storage, "levels must share the storage type");
    PSL_CHECK_ARG(levels[l]->stream == levels[0]->stream, "levels must share the stream");
  }
  for (int s = 0; s < nsteps; ++s) {
    const int rc = diad_recurse(levels, nlevels, 0);
    if (rc != PSL_OK) return rc;
  }
  return PSL_OK;
}

int psl_diad_reduce_momentum_exchange(psl_diad* d, double out[3]) {
  PSL_CHECK_ARG(d && out, "null argument");
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  const uint32_t nbm = (uint32_t)((d->n_moving + 255) / 256);
  const uint32_t nbf = (uint32_t)((d->n_forced + 255) / 256);
  const size_t need = 3 * ((size_t)nb + nbm + nbf);
  if (d->n_partials < need) {
    if (d->partials) PSL_HIP_TRY(hipFree(d->partials));
    d->partials = nullptr;
    d->n_partials = 0;
    PSL_HIP_TRY(hipMalloc(&d->partials, need * sizeof(double)));
    d->n_partials = need;
  }
  int rc = PSL_ESO_DISPATCH(eso_momentum, d, nb, nullptr, 0u, d->partials);
  if (rc == PSL_OK && nbm)
    rc = PSL_ESO_DISPATCH(eso_momentum, d, nbm, d->moving, (uint32_t)d->n_moving, d->partials + 3 * (size_t)nb);
  if (rc == PSL_OK && nbf)
    rc = PSL_ESO_DISPATCH(eso_momentum, d, nbf, d->forced, (uint32_t)d->n_forced,
                          d->partials + 3 * ((size_t)nb + nbm));
  if (rc != PSL_OK) return rc;
  std::vector<double> h(need);
  PSL_HIP_TRY(hipMemcpyAsync(h.data(), d->partials, need * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  PSL_HIP_TRY(hipStreamSynchronize(d->stream));
  double acc[3] = {0.0, 0.0, 0.0}, mov[3] = {0.0, 0.0, 0.0};
  for (uint32_t b = 0; b < nb; ++b)
    for (int a = 0; a < 3; ++a) acc[a] += h[3 * (size_t)b + a];
  for (uint32_t b = 0; b < nbm + nbf; ++b)
    for (int a = 0; a < 3; ++a) mov[a] += h[3 * ((size_t)nb + b) + a];
  for (int a = 0; a < 3; ++a) out[a] = acc[a] - mov[a];
  return PSL_OK;
}

int psl_diad_reduce_forcing(psl_diad* d, float rho_nominal_phys, double out[6]) {
  PSL_CHECK_ARG(d && out, "null argument");
  const uint32_t nb = (uint32_t)((d->n_forced + 255) / 256);
  for (int a = 0; a < 6; ++a) out[a] = 0.0;
  if (!nb) return PSL_OK;
  const size_t need = 6 * (size_t)nb;
  if (d->n_partials < need) {
    if (d->partials) PSL_HIP_TRY(hipFree(d->partials));
    d->partials = nullptr;
    d->n_partials = 0;
    PSL_HIP_TRY(hipMalloc(&d->partials, need * sizeof(double)));
    d->n_partials = need;
  }
  const float h = d->info.res / 1000.f;
  const float scale = rho_nominal_phys * h * h * h * h / (d->info.dtPhys * d->info.dtPhys);
  const int rc = PSL_ESO_DISPATCH(eso_forcing_sum, d, nb, d->partials, scale);
  if (rc != PSL_OK) return rc;
  std::vector<double> hp(need);
  PSL_HIP_TRY(hipMemcpyAsync(hp.data(), d->partials, need * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  PSL_HIP_TRY(hipStreamSynchronize(d->stream));
  for (uint32_t b = 0; b < nb; ++b)
    for (int a = 0; a < 6; ++a) out[a] += hp[6 * (size_t)b + a];
  for (int a = 3; a < 6; ++a) out[a] /= 1000.0;  // N mm -> N m
  return PSL_OK;
}

// Neighbour maps of a dense nx*ny*nz box (cell = (k*ny + j)*nx + i); axes with
// the periodic bit wrap, the others clamp to the last layer.
int psl_diad_dense_box(int nx, int ny, int nz, int periodic_mask, int* ijk3, int* nbr7) {
  PSL_CHECK_ARG(nx > 0 && ny > 0 && nz > 0 && ijk3 && nbr7, "bad box");
  const int64_t n = (int64_t)nx * ny * nz;
  auto wrap = [](int a, int m, bool per) { return per ? (a % m) : (a < m ? a : m - 1); };
  const bool px = periodic_mask & 1, py = periodic_mask & 2, pz = periodic_mask & 4;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t c = ((int64_t)k * ny + j) * nx + i;
        ijk3[c] = i;
        ijk3[n + c] = j;
        ijk3[2 * n + c] = k;
        const int i1 = wrap(i + 1, nx, px), j1 = wrap(j + 1, ny, py), k1 = wrap(k + 1, nz, pz);
        auto id = [&](int a, int b, int cc) { return (int)(((int64_t)cc * ny + b) * nx + a); };
        nbr7[0 * n + c] = id(i1, j, k);
        nbr7[1 * n + c] = id(i, j1, k);
        nbr7[2 * n + c] = id(i, j, k1);
        nbr7[3 * n + c] = id(i1, j1, k);
        nbr7[4 * n + c] = id(i1, j, k1);
        nbr7[5 * n + c] = id(i, j1, k1);
        nbr7[6 * n + c] = id(i1, j1, k1);
      }
  return PSL_OK;
}

}  // extern "C"

namespace psl {
int diad_macro_device(psl_diad* d, float* out4n) { return PSL_ESO_DISPATCH(eso_macro, d, out4n); }
}  // namespace psl
