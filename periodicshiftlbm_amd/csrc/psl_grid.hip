// Grid lifecycle, the fused periodic-shift step kernels, periodic copies,
// equilibrium fill, field access and FP64 reductions.
//
// Data layout in HBM (one allocation per grid): fArray(q, slot) at
// f[q * N + slot], N = cellCount, cell = k*Nx*Ny + j*Nx + i, and population q
// of cell c lives at slot (c + shift[q]) mod N (include/types.h:42-45,
// include/applyStreaming.h:33-41).  Streaming only changes shift[q]
// (include/applyStreaming.h:11-31), so one step is a read-modify-write of
// 27 slots per cell with no cross-cell dependency: every slot is owned by
// exactly one cell (paper section 1.1.2, eqs 1.5-1.6).
//
// Kernels of one step:
//   k_ps_bulk      fluid + bounceback cells, one thread per cell, the hot path;
//                  it also stages the box-face cells' loaded f (see Stage);
//   k_ps_boundary_staged  box-face boundary cells from a compact list
//                  (restore + MBBC + collision) on the staged f; the cell
//                  sets are disjoint, so the two launches together equal the
//                  reference's single update (k_ps_boundary: the same from
//                  the slots, next to the learned-closure bulk kernel);
//   k_ps_fused     both in one launch (a tuning variant);
//   k_periodic     applyBCPeriodicX/Y surface copies (before the update).
#include "psl_internal.h"
#include "psl_stream.h"
#include "psl_closure.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace psl {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------- the step
// The local cell update (include/applyLocalCellUpdate.h:1-74) splits the
// cells into two disjoint sets: the listed box-face cells (the branch at
// lines 46-67: restore + MBBC + collision) and the bulk (fluid and
// bounceback cells, in cell order).  After applyStreaming every cell owns its
// 27 slots, so the two sets can be updated in any order or concurrently.
//   k_ps_bulk + k_ps_boundary  two launches, one after the other;
//   k_ps_fused                 one launch in which each boundary-list
//                              workgroup sits right before the bulk blocks
//                              whose cells it covers (fused_order), so the
//                              short, latency-bound list overlaps the
//                              bandwidth-bound bulk and its strided lines are
//                              shared with those blocks through L2.  The
//                              boundary path needs ~100 VGPRs against ~70 for
//                              the bulk, so the kernel is held at 7 waves per
//                              SIMD and spills only inside the boundary path.
// The per-cell paths (bulk_update, boundary_update) are shared with the
// batched kernels, so a batched grid computes bit-identical results.
template <class St>
__global__ __launch_bounds__(256) void k_ps_boundary(StepParams p, BoundaryParams b) {
  boundary_update<St>(p, b, blockIdx.x * 256u + threadIdx.x);
}

// The listed cells from the records k_ps_bulk<..., STAGE = true> wrote.
template <class St>
__global__ __launch_bounds__(64) void k_ps_boundary_staged(StagedParams p) {
  typedef typename St::T T;
  __shared__ float4 given[4];
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  const bool live = t < p.nlist;
  // the record (issued first) and the given-value table are in flight together
  float v[27];
  uint32_t cell;
  int fl;
  load_staged<St>(p.stage, live ? t : p.nlist - 1, v, cell, fl);
  if (threadIdx.x < 4) given[threadIdx.x] = p.bc[threadIdx.x];
  __syncthreads();
  if (!live) return;
  const float cs = p.cs_cell ? ld<float, false>(p.cs_cell + cell) : p.cs;
  T* f = (T*)p.f;
  boundary_core<St>(cell, fl, v, given[(fl >> 6) & 3], p.nx, p.ny, p.nz, Relax{p.omega, p.tau, p.cs}, cs,
                    [&](int q, T w) {
                      uint32_t slot = cell + (uint32_t)p.s[q];
                      if (slot >= p.n) slot -= p.n;
                      st<T, false>(f + (size_t)q * p.n + slot, w);
                    });
}

// (held at the plain kernel's 7 waves per SIMD: without the bound the staging
// branch raises the register count to 5 waves)
template <class St, bool FLAGS, bool LES, int TUNE, bool STAGE, bool HELD = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_ps_bulk(StepParams p, Stage sg) {
  bulk_update<St, FLAGS, LES, TUNE, STAGE, HELD>(p, p.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x,
                                                 sg);
}

// Bulk cells with the learned closure in place of the Smagorinsky rate.
template <class St, bool FLAGS, int TUNE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_ps_bulk_closure(
    StepParams p, const MlpNet* __restrict__ net, float nu, uint16_t* __restrict__ y_out) {
  __shared__ TanhTable tt;
  bulk_update_closure<St, FLAGS, TUNE>(p, p.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x, net, nu, tt,
                                       y_out);
}

// One workgroup per entry of the fused order (fused_order): boundary-list
// ranges interleaved with the bulk blocks they border.
template <class St, bool LES, int TUNE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_ps_fused(
    StepParams p, BoundaryParams b, const FusedEntry* __restrict__ order) {
  const uint32_t w = p.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const FusedEntry e = order[w];
  if (e.b != FUSED_BULK) {
    b.n = e.b;
    boundary_update<St>(p, b, e.a + threadIdx.x);
    return;
  }
  bulk_update<St, true, LES, TUNE>(p, e.a);
}

// ------------------------------------------------------------ periodic copy
// include/boundaryConditions/periodicXY.h:38-128: +c populations right->left,
// -c populations left->right, raw storage words.
__constant__ int c_pos_x[9] = {1, 7, 9, 12, 16, 20, 22, 23, 26};
__constant__ int c_neg_x[9] = {2, 8, 10, 11, 15, 19, 21, 24, 25};
__constant__ int c_pos_y[9] = {6, 12, 13, 15, 17, 19, 22, 24, 26};
__constant__ int c_neg_y[9] = {5, 11, 14, 16, 18, 20, 21, 23, 25};

template <class T>
__global__ __launch_bounds__(256) void k_periodic(T* __restrict__ f, Shift27 sh, uint32_t n, int nx, int ny,
                                                  uint32_t row_blocks, int axis) {
  // one thread per (surface cell, copied population): blockIdx.y is the
  // copied population (18), blockIdx.x the 256-cell run of a surface row and
  // the row's k; the surface coordinate runs fastest, so for the y copies
  // (rows along x) a wave moves one contiguous run of each population
  const uint32_t na = axis == 0 ? (uint32_t)ny : (uint32_t)nx;
  const uint32_t k = blockIdx.x / row_blocks;  // workgroup-uniform
  const uint32_t a = (blockIdx.x - k * row_blocks) * 256u + threadIdx.x;
  if (a >= na) return;
  const int c = (int)blockIdx.y;
  const uint32_t plane = k * (uint32_t)nx * (uint32_t)ny;
  uint32_t left, right;
  if (axis == 0) {
    left = plane + a * (uint32_t)nx;
    right = left + (uint32_t)nx - 1u;
  } else {
    left = plane + a;
    right = plane + (uint32_t)(ny - 1) * (uint32_t)nx + a;
  }
  const bool pos = c < 9;
  const int q = pos ? (axis == 0 ? c_pos_x[c] : c_pos_y[c]) : (axis == 0 ? c_neg_x[c - 9] : c_neg_y[c - 9]);
  auto slot = [&](uint32_t cell) -> int64_t {
    uint32_t s = cell + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    return (int64_t)q * n + s;
  };
  // +c populations right -> left, -c populations left -> right
  if (pos) f[slot(left)] = f[slot(right)];
  else f[slot(right)] = f[slot(left)];
}

// --------------------------------------------------------- equilibrium fill
// include/fillEquilibrium.h:1-22.  Written at the shifted slot, which is the
// reference's unshifted index whenever the shifter is zero (its only use).
template <class St>
__global__ __launch_bounds__(256) void k_fill_eq(typename St::T* __restrict__ f, Shift27 sh, uint32_t n,
                                                 const float* rho, const float* ux, const float* uy, const float* uz) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
  const float r = rho ? rho[cell] : 1.f;
  const float a = ux ? ux[cell] : 0.f, b = uy ? uy[cell] : 0.f, c = uz ? uz[cell] : 0.f;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = cell + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    float e = feq(q, r, a, b, c);
    if (St::W_OFF) e -= weight(q);
    f[(int64_t)q * n + s] = St::store(e);
  }
}

// --------------------------------------------------------------- field I/O
template <class St>
__global__ __launch_bounds__(256) void k_get_f(const typename St::T* __restrict__ f, Shift27 sh, uint32_t n,
                                               float* __restrict__ out) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = cell + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    float v = St::load(f[(int64_t)q * n + s]);
    if (St::W_OFF) v += weight(q);
    out[(int64_t)q * n + cell] = v;
  }
}

template <class St>
__global__ __launch_bounds__(256) void k_set_f(typename St::T* __restrict__ f, Shift27 sh, uint32_t n,
                                               const float* __restrict__ in) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = cell + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    float v = in[(int64_t)q * n + cell];
    if (St::W_OFF) v -= weight(q);
    f[(int64_t)q * n + s] = St::store(v);
  }
}

template <class St>
__device__ __forceinline__ void load_full(const typename St::T* f, const Shift27& sh, uint32_t n, uint32_t cell,
                                          float (&v)[27]) {
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = cell + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    v[q] = St::load(f[(int64_t)q * n + s]);
    if (St::W_OFF) v[q] += weight(q);
  }
}

template <class St>
__global__ __launch_bounds__(256) void k_macro(const typename St::T* __restrict__ f, Shift27 sh, uint32_t n,
                                               float* rho, float* ux, float* uy, float* uz) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
  float v[27];
  load_full<St>(f, sh, n, cell, v);
  float r, a, b, c;
  rho_u(v, r, a, b, c);
  if (rho) rho[cell] = r;
  if (ux) ux[cell] = a;
  if (uy) uy[cell] = b;
  if (uz) uz[cell] = c;
}

// --------------------------------------------------------------- reductions
// Deterministic FP64 reductions: per-block partials in a fixed order, then
// one block sums the partials.
__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  return x;
}

template <int K>
__device__ __forceinline__ void block_sum_store(double (&x)[K], double* out) {
  __shared__ double sm[4][K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    const double s = wave_sum(x[c]);
    if (lane == 0) sm[wave][c] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) out[threadIdx.x] = ((sm[0][threadIdx.x] + sm[1][threadIdx.x]) + sm[2][threadIdx.x]) + sm[3][threadIdx.x];
}

// getSphereDrag / getLift momentum exchange (flowAroundSphereMainSingle.cu:115-146).
template <class St>
__global__ __launch_bounds__(256) void k_momentum_exchange(const typename St::T* __restrict__ f, Shift27 sh, uint32_t n,
                                                           const uint8_t* __restrict__ flags, int nx, int ny, int lo0,
                                                           int lo1, int lo2, int sx, int sy, int sz,
                                                           double* partials) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t total = (uint32_t)sx * sy * sz;
  double acc[3] = {0.0, 0.0, 0.0};
  if (t < total && flags) {
    const int kr = t / (sx * sy);
    const int rem = t - kr * sx * sy;
    const int jr = rem / sx, ir = rem - jr * sx;
    const uint32_t cell = (uint32_t)(kr + lo2) * nx * ny + (uint32_t)(jr + lo1) * nx + (ir + lo0);
    if ((flags[cell] & 7) == KIND_BOUNCEBACK) {
      float v[27];
      load_full<St>(f, sh, n, cell, v);
      float r, a, b, c;
      rho_u(v, r, a, b, c);
      bounceback(v);
      float r2, a2, b2, c2;
      rho_u(v, r2, a2, b2, c2);
      acc[0] = (double)(r * (a - a2));
      acc[1] = (double)(r * (b - b2));
      acc[2] = (double)(r * (c - c2));
    }
  }
  block_sum_store<3>(acc, partials + 3 * (size_t)blockIdx.x);
}

template <class St>
__global__ __launch_bounds__(256) void k_mass(const typename St::T* __restrict__ f, Shift27 sh, uint32_t n,
                                              const uint8_t* __restrict__ flags, double* partials) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  double acc[1] = {0.0};
  if (cell < n) {
    const int kind = flags ? (flags[cell] & 7) : KIND_FLUID;
    if (kind != KIND_GHOST && kind != KIND_COVERED && kind != KIND_BOUNCEBACK) {
      double m = 0.0;
#pragma unroll
      for (int q = 0; q < 27; ++q) {
        uint32_t s = cell + (uint32_t)sh.s[q];
        if (s >= n) s -= n;
        m += (double)St::load(f[(int64_t)q * n + s]);
      }
      if (St::W_OFF) m += 1.0;
      acc[0] = m;
    }
  }
  block_sum_store<1>(acc, partials + blockIdx.x);
}

template <int K>
__global__ __launch_bounds__(256) void k_sum_partials(const double* partials, uint32_t nblocks, double* out) {
  double acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.0;
  for (uint32_t b = threadIdx.x; b < nblocks; b += 256)
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] += partials[(size_t)K * b + c];
  block_sum_store<K>(acc, out);
}

// ------------------------------------------------------------- plain copy
// One 16-byte element per thread: many short-lived waves keep more requests
// in flight than a grid-stride loop over a capped grid (6.1 vs 4.6 TB/s for a
// 2 GiB copy on MI355X).
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t < n16) dst[t] = src[t];
}

// The data movement of one periodic-shift local update without the
// collision: each cell reads its 27 shifted slots, then writes them back
// (+1), in the bulk kernel's order.  Measures the HBM rate the access pattern
// itself allows on a grid of n cells.
template <class T>
__global__ __launch_bounds__(256) void k_ps_pattern(T* __restrict__ f, Shift27 sh, uint32_t n) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= n) return;
  T v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = c + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    v[q] = f[(size_t)q * n + s];
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    uint32_t s = c + (uint32_t)sh.s[q];
    if (s >= n) s -= n;
    f[(size_t)q * n + s] = (T)(v[q] + (T)1);
  }
}
__global__ void k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

// ------------------------------------------------------------ flag helpers
// Covered region of a grid that carries a finer sub-box
// (include/applyLocalCellUpdate.h:25-30).
__global__ void k_flags_overlay(uint8_t* flags, uint32_t n, int nx, int ny, int is, int ie, int js, int je, int ks,
                                int ke) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
  const uint32_t nxy = (uint32_t)nx * ny;
  const int k = cell / nxy;
  const uint32_t rem = cell - (uint32_t)k * nxy;
  const int j = rem / nx, i = rem - j * nx;
  if ((i > is + 1 && i < ie - 2) && (j > js + 1 && j < je - 2) && (k > ks + 1 && k < ke - 2))
    flags[cell] = (uint8_t)((flags[cell] & ~7) | KIND_COVERED);
}

__global__ void k_apply_bb_mask(uint8_t* flags, const uint8_t* mask, uint32_t n) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= n) return;
  if (mask[cell] && (flags[cell] & 7) != KIND_COVERED) flags[cell] = (uint8_t)((flags[cell] & ~7) | KIND_BOUNCEBACK);
}

// ------------------------------------------------------------ host helpers
static inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

static inline Shift27 shift_of(const psl_grid* g) {
  Shift27 s;
  for (int q = 0; q < 27; ++q) s.s[q] = g->shift[q];
  return s;
}

static inline uint32_t ncells(const psl_grid* g) { return (uint32_t)g->info.cellCount; }

int ensure_partials(psl_grid* g, size_t n) {
  if (g->n_partials >= n && g->partials) return PSL_OK;
  if (g->partials) PSL_HIP_TRY(hipFree(g->partials));
  g->partials = nullptr;
  PSL_HIP_TRY(hipMalloc(&g->partials, n * sizeof(double)));
  g->n_partials = n;
  return PSL_OK;
}

static bool has_subgrid(const psl_info& I) {
  return (I.iSubgridEnd - I.iSubgridStart) > 4 && (I.jSubgridEnd - I.jSubgridStart) > 4 &&
         (I.kSubgridEnd - I.kSubgridStart) > 4;
}

// Upload the user flags (or all-fluid), overlay the covered sub-box, and
// rebuild the compact list of box-face boundary cells.
std::vector<FusedEntry> fused_order(const std::vector<uint32_t>& list, uint32_t n) {
  std::vector<FusedEntry> order;
  if (list.empty()) return order;
  const uint32_t nb = (n + 255u) / 256u;
  size_t pos = 0;  // first list entry not yet assigned
  uint32_t b = 0;
  while (b < nb) {
    // grow a span of bulk blocks [b, e) while its list entries fit one workgroup
    size_t end = pos;
    uint32_t e = b;
    while (e < nb && e - b < FUSED_SPAN) {
      size_t next = end;
      while (next < list.size() && list[next] < (e + 1) * 256u) ++next;
      if (next - pos > 256) break;
      end = next;
      ++e;
    }
    if (end > pos) order.push_back({(uint32_t)pos, (uint32_t)end});
    for (uint32_t k = b; k < e; ++k) order.push_back({k, FUSED_BULK});
    pos = end;
    b = e;
  }
  return order;
}

int grid_rebuild_flags(psl_grid* g) {
  const uint32_t n = ncells(g);
  const bool need = !g->host_flags.empty() || has_subgrid(g->info);
  if (!need) {
    g->census[0] = n;
    g->census[1] = g->census[2] = 0;
    if (g->flags) PSL_HIP_TRY(hipFree(g->flags));
    if (g->bc_list) PSL_HIP_TRY(hipFree(g->bc_list));
    if (g->fused_dev) PSL_HIP_TRY(hipFree(g->fused_dev));
    if (g->stage_buf) PSL_HIP_TRY(hipFree(g->stage_buf));
    if (g->stage_base) PSL_HIP_TRY(hipFree(g->stage_base));
    g->flags = nullptr;
    g->bc_list = nullptr;
    g->fused_dev = nullptr;
    g->stage_buf = nullptr;
    g->stage_base = nullptr;
    g->fused_host.clear();
    g->n_bc = 0;
    return PSL_OK;
  }
  if (!g->flags) PSL_HIP_TRY(hipMalloc(&g->flags, n));
  if (g->host_flags.empty()) PSL_HIP_TRY(hipMemsetAsync(g->flags, 0, n, g->stream));
  else PSL_HIP_TRY(hipMemcpyAsync(g->flags, g->host_flags.data(), n, hipMemcpyHostToDevice, g->stream));
  const psl_info& I = g->info;
  if (has_subgrid(I)) {
    k_flags_overlay<<<blocks_for(n), 256, 0, g->stream>>>(g->flags, n, I.cellCountX, I.cellCountY, I.iSubgridStart,
                                                          I.iSubgridEnd, I.jSubgridStart, I.jSubgridEnd,
                                                          I.kSubgridStart, I.kSubgridEnd);
    PSL_LAUNCH_CHECK();
  }
  // boundary list from the final flags
  std::vector<uint8_t> fin(n);
  PSL_HIP_TRY(hipMemcpyAsync(fin.data(), g->flags, n, hipMemcpyDeviceToHost, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  std::vector<uint32_t> list;
  int64_t skipped = 0;
  for (uint32_t c = 0; c < n; ++c) {
    const int kind = fin[c] & 7;
    if (kind >= KIND_GIVEN_RHO && kind <= KIND_GIVEN_RHOU) list.push_back(c);
    else if (kind >= KIND_GHOST) ++skipped;
  }
  g->census[1] = (int64_t)list.size();
  g->census[2] = skipped;
  g->census[0] = (int64_t)n - g->census[1] - skipped;
  if (g->bc_list) PSL_HIP_TRY(hipFree(g->bc_list));
  g->bc_list = nullptr;
  g->n_bc = (uint32_t)list.size();
  if (!list.empty()) {
    PSL_HIP_TRY(hipMalloc(&g->bc_list, list.size() * sizeof(uint32_t)));
    PSL_HIP_TRY(hipMemcpy(g->bc_list, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  if (g->fused_dev) PSL_HIP_TRY(hipFree(g->fused_dev));
  g->fused_dev = nullptr;
  if (g->stage_buf) PSL_HIP_TRY(hipFree(g->stage_buf));
  if (g->stage_base) PSL_HIP_TRY(hipFree(g->stage_base));
  g->stage_buf = nullptr;
  g->stage_base = nullptr;
  if (!list.empty()) {
    // staging records and the per-wave count of listed cells below each wave
    const uint32_t nw = (n + 63) / 64;
    std::vector<uint32_t> base(nw);
    size_t at = 0;
    for (uint32_t w = 0; w < nw; ++w) {
      while (at < list.size() && list[at] < 64u * w) ++at;
      base[w] = (uint32_t)at;
    }
    PSL_HIP_TRY(hipMalloc(&g->stage_buf, list.size() * STAGE_STRIDE * sizeof(float)));
    PSL_HIP_TRY(hipMalloc(&g->stage_base, nw * sizeof(uint32_t)));
    PSL_HIP_TRY(hipMemcpy(g->stage_base, base.data(), nw * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  g->fused_host = fused_order(list, n);
  if (!g->fused_host.empty()) {
    const size_t bytes = g->fused_host.size() * sizeof(FusedEntry);
    PSL_HIP_TRY(hipMalloc(&g->fused_dev, bytes));
    PSL_HIP_TRY(hipMemcpy(g->fused_dev, g->fused_host.data(), bytes, hipMemcpyHostToDevice));
  }
  return PSL_OK;
}

static void drop_profile_events(psl_grid* g) {
  for (auto& e : g->prof_events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  g->prof_events.clear();
}

template <class St, int TUNE>
static void launch_bulk(bool flags, bool les, bool held, uint32_t nb, size_t lds, hipStream_t stream,
                        const StepParams& p, const Stage& sg) {
  if (held) {
    // between learned-closure evaluations (never staged)
    if (flags) k_ps_bulk<St, true, true, TUNE, false, true><<<nb, 256, lds, stream>>>(p, sg);
    else k_ps_bulk<St, false, true, TUNE, false, true><<<nb, 256, lds, stream>>>(p, sg);
  } else if (sg.buf) {
    if (les) k_ps_bulk<St, true, true, TUNE, true><<<nb, 256, lds, stream>>>(p, sg);
    else k_ps_bulk<St, true, false, TUNE, true><<<nb, 256, lds, stream>>>(p, sg);
  } else if (flags) {
    if (les) k_ps_bulk<St, true, true, TUNE, false><<<nb, 256, lds, stream>>>(p, sg);
    else k_ps_bulk<St, true, false, TUNE, false><<<nb, 256, lds, stream>>>(p, sg);
  } else {
    if (les) k_ps_bulk<St, false, true, TUNE, false><<<nb, 256, lds, stream>>>(p, sg);
    else k_ps_bulk<St, false, false, TUNE, false><<<nb, 256, lds, stream>>>(p, sg);
  }
}

template <class St, int TUNE>
static void launch_closure(bool flags, uint32_t nb, hipStream_t stream, const StepParams& p, const MlpNet* net,
                           float nu, uint16_t* y_out) {
  if (flags) k_ps_bulk_closure<St, true, TUNE><<<nb, 256, 0, stream>>>(p, net, nu, y_out);
  else k_ps_bulk_closure<St, false, TUNE><<<nb, 256, 0, stream>>>(p, net, nu, y_out);
}

template <class St, int TUNE>
static void launch_fused(bool les, uint32_t nb, hipStream_t stream, const StepParams& p, const BoundaryParams& b,
                         const FusedEntry* order) {
  if (les) k_ps_fused<St, true, TUNE><<<nb, 256, 0, stream>>>(p, b, order);
  else k_ps_fused<St, false, TUNE><<<nb, 256, 0, stream>>>(p, b, order);
}

template <class St>
static int launch_local_update(psl_grid* g, bool dominant_only = false) {
  typedef typename St::T T;
  StepParams p;
  p.n = ncells(g);
  T* f = (T*)g->f;
  for (int q = 0; q < 27; ++q) {
    T* row = f + (size_t)q * p.n + g->shift[q];
    p.lo[q] = row;
    p.hi[q] = row - p.n;
    p.thr[q] = p.n - (uint32_t)g->shift[q];
  }
  const float tau = 3.f * g->info.nu + 0.5f;
  p.tau = tau;
  p.omega = 1.f / tau;
  p.cs = g->cs;
  p.cs_cell = g->cs_cell;
  p.flags = g->flags;
  p.xcd_remap = (g->tuning & PSL_TUNE_XCD_REMAP) ? 1u : 0u;
  p.nut_y = nullptr;
  p.nu = g->info.nu;
  p.nut_scale = MLP_NU_SCALE;
  bool les = g->cs != 0.f || g->cs_cell != nullptr;
  BoundaryParams b;
  b.list = g->bc_list;
  b.n = g->flags ? g->n_bc : 0u;
  b.nx = g->info.cellCountX;
  b.ny = g->info.cellCountY;
  b.nz = g->info.cellCountZ;
  b.bc = g->bc_dev;
  const MlpNet* net = psl_mlp_device_net(g->closure);
  // a closure evaluated every closure_every steps: the evaluating step stores
  // the network output per cell, the steps between relax with the stored one
  uint16_t* y_out = nullptr;
  if (net && g->closure_every > 1) {
    if (g->closure_count % (uint64_t)g->closure_every == 0) {
      y_out = g->closure_y;
    } else {
      p.nut_y = g->closure_y;
      les = true;
      net = nullptr;
    }
    ++g->closure_count;
  }
  const bool held = p.nut_y != nullptr;
  const bool fused = b.n && (g->tuning & PSL_TUNE_FUSED_BOUNDARY) && !net && !held;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (g->profile) {
    PSL_HIP_TRY(hipEventCreate(&ev0));
    PSL_HIP_TRY(hipEventCreate(&ev1));
    PSL_HIP_TRY(hipEventRecord(ev0, g->stream));
  }
  const int nt = g->tuning & (TUNE_NT_LOAD | TUNE_NT_STORE);
  const uint32_t nb = blocks_for(p.n);
  if (net) {
    const bool fl = g->flags != nullptr;
    const float nu = g->info.nu;
    switch (nt) {
      case 0: launch_closure<St, 0>(fl, nb, g->stream, p, net, nu, y_out); break;
      case 1: launch_closure<St, 1>(fl, nb, g->stream, p, net, nu, y_out); break;
      case 2: launch_closure<St, 2>(fl, nb, g->stream, p, net, nu, y_out); break;
      default: launch_closure<St, 3>(fl, nb, g->stream, p, net, nu, y_out); break;
    }
  } else if (fused) {
    const uint32_t nf = (uint32_t)g->fused_host.size();
    switch (nt) {
      case 0: launch_fused<St, 0>(les, nf, g->stream, p, b, g->fused_dev); break;
      case 1: launch_fused<St, 1>(les, nf, g->stream, p, b, g->fused_dev); break;
      case 2: launch_fused<St, 2>(les, nf, g->stream, p, b, g->fused_dev); break;
      default: launch_fused<St, 3>(les, nf, g->stream, p, b, g->fused_dev); break;
    }
  } else {
    const bool fl = g->flags != nullptr;
    // the listed cells are staged by the bulk kernel whenever there are any
    // (not between closure evaluations: the closure path updates them from
    // their slots)
    const Stage sg{b.n && !held ? g->stage_buf : nullptr, g->stage_base};
    switch (nt) {
      case 0: launch_bulk<St, 0>(fl, les, held, nb, g->bulk_lds, g->stream, p, sg); break;
      case 1: launch_bulk<St, 1>(fl, les, held, nb, g->bulk_lds, g->stream, p, sg); break;
      case 2: launch_bulk<St, 2>(fl, les, held, nb, g->bulk_lds, g->stream, p, sg); break;
      default: launch_bulk<St, 3>(fl, les, held, nb, g->bulk_lds, g->stream, p, sg); break;
    }
  }
  PSL_LAUNCH_CHECK();
  if (g->profile) {
    PSL_HIP_TRY(hipEventRecord(ev1, g->stream));
    g->prof_events.emplace_back(ev0, ev1);
  }
  if (b.n && !fused && !dominant_only) {
    if (net || held) k_ps_boundary<St><<<blocks_for(b.n), 256, 0, g->stream>>>(p, b);
    else {
      StagedParams sp;
      sp.f = g->f;
      sp.stage = g->stage_buf;
      sp.cs_cell = g->cs_cell;
      sp.bc = g->bc_dev;
      sp.n = p.n;
      sp.nlist = b.n;
      for (int q = 0; q < 27; ++q) sp.s[q] = g->shift[q];
      sp.omega = p.omega;
      sp.tau = p.tau;
      sp.cs = p.cs;
      sp.nx = b.nx;
      sp.ny = b.ny;
      sp.nz = b.nz;
      k_ps_boundary_staged<St><<<(b.n + 63) / 64, 64, 0, g->stream>>>(sp);
    }
    PSL_LAUNCH_CHECK();
  }
  return PSL_OK;
}

template <template <class> class Fn, class... A>
static int dispatch_storage(int storage, A... args) {
  switch (storage) {
    case PSL_F32: return Fn<StF32>::run(args...);
    case PSL_F16DEV: return Fn<StF16>::run(args...);
    case PSL_BF16DEV: return Fn<StBF16>::run(args...);
  }
  set_error("unknown storage");
  return PSL_E_ARG;
}

template <class St> struct LocalUpdateFn { static int run(psl_grid* g) { return launch_local_update<St>(g); } };
template <class St> struct DominantFn { static int run(psl_grid* g) { return launch_local_update<St>(g, true); } };

template <class St> struct FillEqFn {
  static int run(psl_grid* g, const float* r, const float* a, const float* b, const float* c) {
    const uint32_t n = ncells(g);
    k_fill_eq<St><<<blocks_for(n), 256, 0, g->stream>>>((typename St::T*)g->f, shift_of(g), n, r, a, b, c);
    PSL_LAUNCH_CHECK();
    return PSL_OK;
  }
};

template <class St> struct GetFFn {
  static int run(psl_grid* g, float* d) {
    const uint32_t n = ncells(g);
    k_get_f<St><<<blocks_for(n), 256, 0, g->stream>>>((const typename St::T*)g->f, shift_of(g), n, d);
    PSL_LAUNCH_CHECK();
    return PSL_OK;
  }
};
template <class St> struct SetFFn {
  static int run(psl_grid* g, const float* d) {
    const uint32_t n = ncells(g);
    k_set_f<St><<<blocks_for(n), 256, 0, g->stream>>>((typename St::T*)g->f, shift_of(g), n, d);
    PSL_LAUNCH_CHECK();
    return PSL_OK;
  }
};
template <class St> struct MacroFn {
  static int run(psl_grid* g, float* r, float* a, float* b, float* c) {
    const uint32_t n = ncells(g);
    k_macro<St><<<blocks_for(n), 256, 0, g->stream>>>((const typename St::T*)g->f, shift_of(g), n, r, a, b, c);
    PSL_LAUNCH_CHECK();
    return PSL_OK;
  }
};
template <class St> struct MassFn {
  static int run(psl_grid* g, uint32_t nb) {
    k_mass<St><<<nb, 256, 0, g->stream>>>((const typename St::T*)g->f, shift_of(g), ncells(g), g->flags, g->partials);
    PSL_LAUNCH_CHECK();
    return PSL_OK;
  }
};

static int periodic_copy(psl_grid* g, int axis) {
  const psl_info& I = g->info;
  const int na = axis == 0 ? I.cellCountY : I.cellCountX;
  const uint32_t row_blocks = blocks_for((uint64_t)na);
  // row_blocks * nz <= cellCount / 256 + nz < 2^31 (cells are below 2^31)
  const dim3 grid(row_blocks * (uint32_t)I.cellCountZ, 18);
  if (g->storage == PSL_F32)
    k_periodic<float><<<grid, 256, 0, g->stream>>>((float*)g->f, shift_of(g), ncells(g), I.cellCountX, I.cellCountY,
                                                   row_blocks, axis);
  else
    k_periodic<uint16_t><<<grid, 256, 0, g->stream>>>((uint16_t*)g->f, shift_of(g), ncells(g), I.cellCountX,
                                                      I.cellCountY, row_blocks, axis);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

}  // namespace psl

using namespace psl;

// ======================================================================= C ABI
extern "C" {

const char* psl_last_error(void) { return g_last_error.c_str(); }
int psl_abi_version(void) { return PSL_ABI_VERSION; }

int psl_device_count(int* count) {
  PSL_CHECK_ARG(count, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = e == hipSuccess ? c : 0;
  return PSL_OK;
}

int psl_grid_create(const psl_info* info, int storage, psl_grid** out) {
  PSL_CHECK_ARG(info && out, "null argument");
  PSL_CHECK_ARG(storage == PSL_F32 || storage == PSL_F16DEV || storage == PSL_BF16DEV, "unknown storage");
  PSL_CHECK_ARG(info->cellCountX > 0 && info->cellCountY > 0 && info->cellCountZ > 0, "cell counts must be positive");
  const uint64_t n = (uint64_t)info->cellCountX * info->cellCountY * info->cellCountZ;
  PSL_CHECK_ARG(n < (1ull << 31), "cellCount must be below 2^31");
  psl_grid* g = new psl_grid();
  g->info = *info;
  g->info.cellCount = (int)n;
  g->storage = storage;
  const size_t bytes = 27ull * n * storage_bytes(storage);
  hipError_t e = hipMalloc(&g->f, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc fArray: ") + hipGetErrorString(e));
    delete g;
    return PSL_E_NOMEM;
  }
  e = hipMalloc(&g->result, 4 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&g->bc_dev, 4 * sizeof(float4));
  if (e == hipSuccess) e = hipMemset(g->bc_dev, 0, 4 * sizeof(float4));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc grid tables: ") + hipGetErrorString(e));
    (void)hipFree(g->f);
    if (g->result) (void)hipFree(g->result);
    if (g->bc_dev) (void)hipFree(g->bc_dev);
    delete g;
    return PSL_E_NOMEM;
  }
  int rc = grid_rebuild_flags(g);
  if (rc) {
    psl_grid_destroy(g);
    return rc;
  }
  *out = g;
  return PSL_OK;
}

int psl_grid_destroy(psl_grid* g) {
  if (!g) return PSL_OK;
  (void)hipDeviceSynchronize();
  drop_profile_events(g);
  if (g->f) (void)hipFree(g->f);
  if (g->flags) (void)hipFree(g->flags);
  if (g->bc_list) (void)hipFree(g->bc_list);
  if (g->fused_dev) (void)hipFree(g->fused_dev);
  if (g->stage_buf) (void)hipFree(g->stage_buf);
  if (g->stage_base) (void)hipFree(g->stage_base);
  if (g->cs_cell) (void)hipFree(g->cs_cell);
  if (g->partials) (void)hipFree(g->partials);
  if (g->closure_y) (void)hipFree(g->closure_y);
  if (g->result) (void)hipFree(g->result);
  if (g->bc_dev) (void)hipFree(g->bc_dev);
  if (g->scalar_T) (void)hipFree(g->scalar_T);
  if (g->scalar_given) (void)hipFree(g->scalar_given);
  delete g;
  return PSL_OK;
}

int psl_grid_get_info(const psl_grid* g, psl_info* out) {
  PSL_CHECK_ARG(g && out, "null argument");
  *out = g->info;
  return PSL_OK;
}

int psl_grid_set_info(psl_grid* g, const psl_info* info) {
  PSL_CHECK_ARG(g && info, "null argument");
  PSL_CHECK_ARG(info->cellCountX == g->info.cellCountX && info->cellCountY == g->info.cellCountY &&
                    info->cellCountZ == g->info.cellCountZ,
                "cell counts of an existing grid cannot change");
  const bool sub_changed = memcmp(&info->iSubgridStart, &g->info.iSubgridStart, 6 * sizeof(int)) != 0;
  const int n = g->info.cellCount;
  g->info = *info;
  g->info.cellCount = n;
  if (sub_changed) return grid_rebuild_flags(g);
  return PSL_OK;
}

int psl_grid_set_stream(psl_grid* g, void* s) {
  PSL_CHECK_ARG(g, "null grid");
  g->stream = (hipStream_t)s;
  return PSL_OK;
}

int psl_grid_set_flags(psl_grid* g, const uint8_t* host_flags) {
  PSL_CHECK_ARG(g, "null grid");
  if (host_flags) g->host_flags.assign(host_flags, host_flags + g->info.cellCount);
  else g->host_flags.clear();
  return grid_rebuild_flags(g);
}

int psl_grid_get_flags(const psl_grid* g, uint8_t* host) {
  PSL_CHECK_ARG(g && host, "null argument");
  if (!g->flags) {
    memset(host, 0, (size_t)g->info.cellCount);
    return PSL_OK;
  }
  PSL_HIP_TRY(hipMemcpyAsync(host, g->flags, (size_t)g->info.cellCount, hipMemcpyDeviceToHost, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  return PSL_OK;
}

int psl_grid_set_bc_values(psl_grid* g, const float* v) {
  PSL_CHECK_ARG(g && v, "null argument");
  memcpy(g->bcvals.v, v, sizeof(g->bcvals.v));
  PSL_HIP_TRY(hipMemcpyAsync(g->bc_dev, g->bcvals.v, sizeof(g->bcvals.v), hipMemcpyHostToDevice, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  return PSL_OK;
}

int psl_grid_set_smagorinsky(psl_grid* g, float cs, const float* per_cell) {
  PSL_CHECK_ARG(g, "null grid");
  g->cs = cs;
  if (per_cell) {
    if (!g->cs_cell) PSL_HIP_TRY(hipMalloc(&g->cs_cell, (size_t)g->info.cellCount * sizeof(float)));
    PSL_HIP_TRY(hipMemcpy(g->cs_cell, per_cell, (size_t)g->info.cellCount * sizeof(float), hipMemcpyHostToDevice));
  } else if (g->cs_cell) {
    PSL_HIP_TRY(hipFree(g->cs_cell));
    g->cs_cell = nullptr;
  }
  return PSL_OK;
}

int psl_grid_set_periodic_copies(psl_grid* g, int mask) {
  PSL_CHECK_ARG(g, "null grid");
  g->periodic_copies = mask & 3;
  return PSL_OK;
}

int psl_fill_equilibrium(psl_grid* g, const float* rho, const float* ux, const float* uy, const float* uz) {
  PSL_CHECK_ARG(g, "null grid");
  const size_t n = (size_t)g->info.cellCount;
  float* d[4] = {nullptr, nullptr, nullptr, nullptr};
  const float* h[4] = {rho, ux, uy, uz};
  int rc = PSL_OK;
  for (int a = 0; a < 4 && rc == PSL_OK; ++a) {
    if (!h[a]) continue;
    if (hipMalloc(&d[a], n * sizeof(float)) != hipSuccess ||
        hipMemcpy(d[a], h[a], n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("fill_equilibrium: upload failed");
      rc = PSL_E_HIP;
    }
  }
  if (rc == PSL_OK) rc = dispatch_storage<FillEqFn>(g->storage, g, (const float*)d[0], (const float*)d[1],
                                                    (const float*)d[2], (const float*)d[3]);
  if (hipStreamSynchronize(g->stream) != hipSuccess && rc == PSL_OK) {
    set_error("fill_equilibrium: sync failed");
    rc = PSL_E_HIP;
  }
  for (int a = 0; a < 4; ++a)
    if (d[a]) (void)hipFree(d[a]);
  return rc;
}

int psl_apply_streaming(psl_grid* g) {
  PSL_CHECK_ARG(g, "null grid");
  const int64_t N = g->info.cellCount;
  const int64_t nx = g->info.cellCountX, nxy = nx * g->info.cellCountY;
  for (int q = 0; q < 27; ++q) {
    int64_t s = g->shift[q];
    s -= CX[q];
    s -= nx * CY[q];
    s -= nxy * CZ[q];
    // reduced modulo N: the reference's two wraps (include/applyStreaming.h:24-27)
    // are the same on every grid with cellCountY * cellCountZ >= 2, and leave
    // the shift outside [0, N) on thinner ones
    s %= N;
    if (s < 0) s += N;
    g->shift[q] = (int)s;
  }
  return PSL_OK;
}

int psl_apply_local_cell_update(psl_grid* g) {
  PSL_CHECK_ARG(g, "null grid");
  return dispatch_storage<LocalUpdateFn>(g->storage, g);
}

int psl_apply_bc_periodic_x(psl_grid* g) {
  PSL_CHECK_ARG(g, "null grid");
  return periodic_copy(g, 0);
}

int psl_apply_bc_periodic_y(psl_grid* g) {
  PSL_CHECK_ARG(g, "null grid");
  return periodic_copy(g, 1);
}

int psl_get_shifter(const psl_grid* g, int* out) {
  PSL_CHECK_ARG(g && out, "null argument");
  memcpy(out, g->shift, sizeof(g->shift));
  return PSL_OK;
}

int psl_set_shifter(psl_grid* g, const int* in) {
  PSL_CHECK_ARG(g && in, "null argument");
  for (int q = 0; q < 27; ++q) PSL_CHECK_ARG(in[q] >= 0 && in[q] < g->info.cellCount, "shift out of range");
  memcpy(g->shift, in, sizeof(g->shift));
  return PSL_OK;
}

int psl_get_f(psl_grid* g, float* host) {
  PSL_CHECK_ARG(g && host, "null argument");
  const size_t bytes = 27ull * g->info.cellCount * sizeof(float);
  float* d = nullptr;
  PSL_HIP_TRY(hipMalloc(&d, bytes));
  int rc = dispatch_storage<GetFFn>(g->storage, g, d);
  if (rc == PSL_OK && hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, g->stream) != hipSuccess) rc = PSL_E_HIP;
  if (hipStreamSynchronize(g->stream) != hipSuccess && rc == PSL_OK) rc = PSL_E_HIP;
  (void)hipFree(d);
  if (rc == PSL_E_HIP) set_error("psl_get_f: copy failed");
  return rc;
}

int psl_set_f(psl_grid* g, const float* host) {
  PSL_CHECK_ARG(g && host, "null argument");
  const size_t bytes = 27ull * g->info.cellCount * sizeof(float);
  float* d = nullptr;
  PSL_HIP_TRY(hipMalloc(&d, bytes));
  int rc = PSL_OK;
  if (hipMemcpy(d, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("psl_set_f: upload failed");
    rc = PSL_E_HIP;
  }
  if (rc == PSL_OK) rc = dispatch_storage<SetFFn>(g->storage, g, (const float*)d);
  if (hipStreamSynchronize(g->stream) != hipSuccess && rc == PSL_OK) rc = PSL_E_HIP;
  (void)hipFree(d);
  return rc;
}

int psl_get_macroscopic_device(psl_grid* g, float* out) {
  PSL_CHECK_ARG(g && out, "null argument");
  const size_t n = (size_t)g->info.cellCount;
  return dispatch_storage<MacroFn>(g->storage, g, out, out + n, out + 2 * n, out + 3 * n);
}

int psl_get_macroscopic(psl_grid* g, float* rho, float* ux, float* uy, float* uz) {
  PSL_CHECK_ARG(g, "null grid");
  const size_t n = (size_t)g->info.cellCount;
  float* d = nullptr;
  PSL_HIP_TRY(hipMalloc(&d, 4 * n * sizeof(float)));
  int rc = dispatch_storage<MacroFn>(g->storage, g, d, d + n, d + 2 * n, d + 3 * n);
  float* h[4] = {rho, ux, uy, uz};
  for (int a = 0; a < 4 && rc == PSL_OK; ++a)
    if (h[a] && hipMemcpyAsync(h[a], d + a * n, n * sizeof(float), hipMemcpyDeviceToHost, g->stream) != hipSuccess)
      rc = PSL_E_HIP;
  if (hipStreamSynchronize(g->stream) != hipSuccess && rc == PSL_OK) rc = PSL_E_HIP;
  (void)hipFree(d);
  if (rc == PSL_E_HIP) set_error("psl_get_macroscopic: copy failed");
  return rc;
}

int psl_grid_device_f(psl_grid* g, void** dev_ptr, size_t* bytes) {
  PSL_CHECK_ARG(g && dev_ptr, "null argument");
  *dev_ptr = g->f;
  if (bytes) *bytes = 27ull * g->info.cellCount * storage_bytes(g->storage);
  return PSL_OK;
}

// The box [lo, hi) of a momentum-exchange call clipped to the grid; false when
// it is empty (or the grid has no flags), with the result then zero.
static bool momentum_box(const psl_grid* g, const int* lo3, const int* hi3, int (&lo)[3], int (&hi)[3],
                         uint32_t& nb) {
  const psl_info& I = g->info;
  lo[0] = lo[1] = lo[2] = 0;
  hi[0] = I.cellCountX;
  hi[1] = I.cellCountY;
  hi[2] = I.cellCountZ;
  if (lo3) for (int a = 0; a < 3; ++a) lo[a] = std::max(0, lo3[a]);
  if (hi3) {
    hi[0] = std::min(I.cellCountX, hi3[0]);
    hi[1] = std::min(I.cellCountY, hi3[1]);
    hi[2] = std::min(I.cellCountZ, hi3[2]);
  }
  if (!g->flags || hi[0] <= lo[0] || hi[1] <= lo[1] || hi[2] <= lo[2]) return false;
  nb = blocks_for((uint64_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]));
  return true;
}

// Per-block partials into `partials` (3 nb doubles), their fixed-order sum
// into `result` (3 doubles), and the copy to the host: the kernels and block
// count of both momentum-exchange entry points, so they return the same bits.
static int momentum_exchange_into(psl_grid* g, const int (&lo)[3], const int (&hi)[3], uint32_t nb, double* partials,
                                  double* result, double* out3) {
  const psl_info& I = g->info;
#define PSL_MOMENTUM_LAUNCH(ST)                                                                                   \
  k_momentum_exchange<ST><<<nb, 256, 0, g->stream>>>((const typename ST::T*)g->f, shift_of(g), ncells(g), g->flags, \
                                                     I.cellCountX, I.cellCountY, lo[0], lo[1], lo[2], hi[0] - lo[0],  \
                                                     hi[1] - lo[1], hi[2] - lo[2], partials)
  if (g->storage == PSL_F32) PSL_MOMENTUM_LAUNCH(StF32);
  else if (g->storage == PSL_F16DEV) PSL_MOMENTUM_LAUNCH(StF16);
  else if (g->storage == PSL_BF16DEV) PSL_MOMENTUM_LAUNCH(StBF16);
  else {
    set_error("unknown storage");
    return PSL_E_ARG;
  }
#undef PSL_MOMENTUM_LAUNCH
  PSL_LAUNCH_CHECK();
  k_sum_partials<3><<<1, 256, 0, g->stream>>>(partials, nb, result);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipMemcpyAsync(out3, result, 3 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  return PSL_OK;
}

int psl_reduce_momentum_exchange(psl_grid* g, const int* lo3, const int* hi3, double* out3) {
  PSL_CHECK_ARG(g && out3, "null argument");
  int lo[3], hi[3];
  uint32_t nb = 0;
  out3[0] = out3[1] = out3[2] = 0.0;
  if (!momentum_box(g, lo3, hi3, lo, hi, nb)) return PSL_OK;
  int rc = ensure_partials(g, 3ull * nb);
  if (rc) return rc;
  return momentum_exchange_into(g, lo, hi, nb, g->partials, g->result, out3);
}

// ------------------------------------------- momentum exchange, no allocation
// Scratch of psl_reduce_momentum_exchange_noalloc: 3 FP64 partials per block
// of 256 box cells, then the 3 sums.  Every device the library runs on holds
// its own copy until the process ends.  In host code `scratch` is only a
// symbol: its device address comes from hipGetSymbolAddress, once per device.
constexpr uint32_t NOALLOC_MAX_BLOCKS = PSL_NOALLOC_MAX_BOX_CELLS / 256;
static __device__ __align__(256) double scratch[3 * NOALLOC_MAX_BLOCKS + 4];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[PSL_NOALLOC_MAX_DEVICES] = {};

static int scratch_of_current_device(double** out) {
  int dev = -1;
  PSL_HIP_TRY(hipGetDevice(&dev));
  PSL_CHECK_ARG(dev >= 0 && dev < PSL_NOALLOC_MAX_DEVICES, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return PSL_E_HIP;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return PSL_OK;
}

int psl_reduce_momentum_exchange_noalloc(psl_grid* g, const int* lo3, const int* hi3, double* out3) {
  PSL_CHECK_ARG(g && out3, "null argument");
  int lo[3], hi[3];
  uint32_t nb = 0;
  out3[0] = out3[1] = out3[2] = 0.0;
  if (!momentum_box(g, lo3, hi3, lo, hi, nb)) return PSL_OK;
  if (nb > NOALLOC_MAX_BLOCKS) {
    set_error("psl_reduce_momentum_exchange_noalloc: box of " + std::to_string((uint64_t)nb * 256) +
              " cells (rounded up to 256) exceeds the scratch limit of " +
              std::to_string((uint64_t)PSL_NOALLOC_MAX_BOX_CELLS) + " cells");
    return PSL_E_ARG;
  }
  double* sp = nullptr;
  const int rc = scratch_of_current_device(&sp);
  if (rc) return rc;
  // every block writes its partials and the sum kernel overwrites the result,
  // so nothing left by an earlier call is read
  return momentum_exchange_into(g, lo, hi, nb, sp, sp + 3ull * NOALLOC_MAX_BLOCKS, out3);
}

int psl_reduce_mass(psl_grid* g, double* out) {
  PSL_CHECK_ARG(g && out, "null argument");
  const uint32_t nb = blocks_for(ncells(g));
  int rc = ensure_partials(g, nb);
  if (rc) return rc;
  rc = dispatch_storage<MassFn>(g->storage, g, nb);
  if (rc) return rc;
  k_sum_partials<1><<<1, 256, 0, g->stream>>>(g->partials, nb, g->result);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipMemcpyAsync(out, g->result, sizeof(double), hipMemcpyDeviceToHost, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  return PSL_OK;
}

int psl_device_copy(void* dst, const void* src, size_t bytes) {
  PSL_CHECK_ARG(dst && src, "null pointer");
  const size_t n16 = bytes / 16;
  PSL_CHECK_ARG(n16 / 256 < (1ull << 31), "copy too large");
  if (((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0 && n16) {
    const unsigned nb = (unsigned)((n16 + 255) / 256);
    k_copy16<<<nb, 256>>>((const uint4*)src, (uint4*)dst, n16);
    PSL_LAUNCH_CHECK();
    const size_t done = n16 * 16;
    if (bytes > done) {
      k_copy1<<<1, 256>>>((const uint8_t*)src + done, (uint8_t*)dst + done, bytes - done);
      PSL_LAUNCH_CHECK();
    }
  } else if (bytes) {
    k_copy1<<<(unsigned)std::min<size_t>((bytes + 255) / 256, 8192), 256>>>((const uint8_t*)src, (uint8_t*)dst, bytes);
    PSL_LAUNCH_CHECK();
  }
  return PSL_OK;
}

// Dynamic LDS per workgroup that leaves room for exactly `workgroups_per_cu`
// workgroups on a CU of the current device (0: no cap, no LDS); the kernels
// that take it do not use it.
static int occupancy_lds(int workgroups_per_cu, size_t* lds_out) {
  PSL_CHECK_ARG(workgroups_per_cu == 0 || (workgroups_per_cu >= 2 && workgroups_per_cu <= 6),
                "workgroups per CU must be 0 (no cap) or 2-6");
  size_t lds = 0;
  if (workgroups_per_cu) {
    int dev = 0, per_cu = 0, per_block = 0;
    PSL_HIP_TRY(hipGetDevice(&dev));
    PSL_HIP_TRY(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    PSL_HIP_TRY(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // whole KiB below the CU's share: the allocation's rounding up then still
    // leaves room for exactly that many workgroups
    lds = ((size_t)per_cu / (size_t)workgroups_per_cu) & ~(size_t)1023;
    PSL_CHECK_ARG(lds > 0 && lds <= (size_t)per_block && (size_t)per_cu / lds == (size_t)workgroups_per_cu,
                  "this device's LDS cannot express that cap");
  }
  *lds_out = lds;
  return PSL_OK;
}

int psl_probe_ps_pattern(int nx, int ny, int nz, int storage, int launches, double* gbs) {
  return psl_probe_ps_pattern_capped(nx, ny, nz, storage, 0, launches, gbs);
}

int psl_probe_ps_pattern_capped(int nx, int ny, int nz, int storage, int workgroups_per_cu, int launches,
                                double* gbs) {
  PSL_CHECK_ARG(nx > 0 && ny > 0 && nz > 0 && launches > 0 && gbs, "bad arguments");
  PSL_CHECK_ARG(storage == PSL_F32 || storage == PSL_F16DEV || storage == PSL_BF16DEV, "unknown storage");
  const uint64_t n64 = (uint64_t)nx * ny * nz;
  PSL_CHECK_ARG(n64 < (1ull << 31), "cellCount must be below 2^31");
  const uint32_t n = (uint32_t)n64;
  const size_t esz = storage_bytes(storage);
  size_t lds = 0;
  const int lrc = occupancy_lds(workgroups_per_cu, &lds);
  if (lrc) return lrc;
  void* f = nullptr;
  PSL_HIP_TRY(hipMalloc(&f, 27ull * n * esz));
  // shifts after 1000 streaming steps of an nx x ny x nz grid
  Shift27 sh;
  for (int q = 0; q < 27; ++q) {
    const int64_t d = (int64_t)CX[q] + (int64_t)nx * CY[q] + (int64_t)nx * ny * CZ[q];
    int64_t v = (-(d * 1000)) % (int64_t)n;
    if (v < 0) v += n;
    sh.s[q] = (int)v;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PSL_OK;
  auto launch = [&]() {
    if (esz == 4) k_ps_pattern<float><<<blocks_for(n), 256, lds>>>((float*)f, sh, n);
    else k_ps_pattern<uint16_t><<<blocks_for(n), 256, lds>>>((uint16_t*)f, sh, n);
  };
  if (hipMemset(f, 0, 27ull * n * esz) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = PSL_E_HIP;
  float ms = 0.f;
  if (rc == PSL_OK) {
    launch();  // warm
    (void)hipEventRecord(e0, nullptr);
    for (int k = 0; k < launches; ++k) launch();
    (void)hipEventRecord(e1, nullptr);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = PSL_E_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(f);
  if (rc) {
    set_error("pattern probe failed");
    return rc;
  }
  *gbs = 2.0 * 27.0 * n * esz * launches / (ms * 1e-3) / 1e9;
  return PSL_OK;
}

int psl_device_alloc(size_t bytes, void** out) {
  PSL_CHECK_ARG(out, "null argument");
  PSL_HIP_TRY(hipMalloc(out, bytes));
  return PSL_OK;
}

int psl_device_free(void* ptr) {
  if (ptr) PSL_HIP_TRY(hipFree(ptr));
  return PSL_OK;
}

int psl_device_upload(void* dst, const void* host, size_t bytes) {
  PSL_CHECK_ARG(dst && host, "null argument");
  PSL_HIP_TRY(hipMemcpy(dst, host, bytes, hipMemcpyHostToDevice));
  return PSL_OK;
}

int psl_device_download(void* host, const void* src, size_t bytes) {
  PSL_CHECK_ARG(host && src, "null argument");
  PSL_HIP_TRY(hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost));
  return PSL_OK;
}

int psl_device_synchronize(void) {
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

int psl_grid_profile_enable(psl_grid* g, int enable) {
  PSL_CHECK_ARG(g, "null grid");
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  drop_profile_events(g);
  g->profile = enable != 0;
  return PSL_OK;
}

int psl_grid_profile_read(psl_grid* g, double* total_ms, int* launches) {
  PSL_CHECK_ARG(g && total_ms, "null argument");
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  double t = 0.0;
  for (auto& e : g->prof_events) {
    float ms = 0.f;
    PSL_HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
    t += ms;
  }
  *total_ms = t;
  if (launches) *launches = (int)g->prof_events.size();
  drop_profile_events(g);
  return PSL_OK;
}

int psl_grid_set_tuning(psl_grid* g, int tuning) {
  PSL_CHECK_ARG(g, "null grid");
  PSL_CHECK_ARG((tuning & ~(PSL_TUNE_NT_LOAD | PSL_TUNE_NT_STORE | PSL_TUNE_XCD_REMAP | PSL_TUNE_FUSED_BOUNDARY)) == 0,
                "unknown tuning bits");
  g->tuning = tuning;
  return PSL_OK;
}

int psl_grid_get_tuning(const psl_grid* g, int* tuning) {
  PSL_CHECK_ARG(g && tuning, "null argument");
  *tuning = g->tuning;
  return PSL_OK;
}

// Time every tuning variant of the bulk kernel on this grid (each for `steps`
// updates with the grid's own data; the distributions advance by
// 8 * steps updates) and keep the fastest.  times_ms (optional, 8 entries)
// receives the mean bulk-kernel time of each variant.
int psl_grid_set_occupancy(psl_grid* g, int workgroups_per_cu) {
  PSL_CHECK_ARG(g, "null grid");
  size_t lds = 0;
  const int rc = occupancy_lds(workgroups_per_cu, &lds);
  if (rc) return rc;
  g->occupancy = workgroups_per_cu;
  g->bulk_lds = lds;
  return PSL_OK;
}

int psl_grid_get_occupancy(const psl_grid* g, int* workgroups_per_cu) {
  PSL_CHECK_ARG(g && workgroups_per_cu, "null argument");
  *workgroups_per_cu = g->occupancy;
  return PSL_OK;
}

int psl_grid_autotune(psl_grid* g, int steps, int* best, double* times_ms) {
  PSL_CHECK_ARG(g && steps > 0, "bad autotune arguments");
  // whole steps are timed, so that the fused and the two-launch forms compare
  // on equal terms
  const int saved = g->tuning, saved_occ = g->occupancy;
  double best_t = 1e300;
  int best_v = saved;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PSL_HIP_TRY(hipEventCreate(&e0));
  PSL_HIP_TRY(hipEventCreate(&e1));
  int rc = psl_grid_set_occupancy(g, 0);
  // one whole step (streaming, periodic copies, local update): timed with the
  // streaming, the rows are misaligned as in a run (repeated updates at one
  // shift keep them where they are, at zero shift all line-aligned, where
  // the non-temporal variants win and lose by a third once the rows move)
  auto step = [&]() {
    int r = psl_apply_streaming(g);
    if (r == PSL_OK && (g->periodic_copies & 1)) r = psl_apply_bc_periodic_x(g);
    if (r == PSL_OK && (g->periodic_copies & 2)) r = psl_apply_bc_periodic_y(g);
    if (r == PSL_OK) r = psl_apply_local_cell_update(g);
    return r;
  };
  // mean ms of `n` steps after one warm-up, at the current settings
  auto time_updates = [&](double& mean, int n) {
    int r = step();
    if (r == PSL_OK && hipEventRecord(e0, g->stream) != hipSuccess) r = PSL_E_HIP;
    for (int s = 0; s < n && r == PSL_OK; ++s) r = step();
    if (r == PSL_OK && hipEventRecord(e1, g->stream) != hipSuccess) r = PSL_E_HIP;
    float ms = 0.f;
    if (r == PSL_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      r = PSL_E_HIP;
    mean = (double)ms / n;
    return r;
  };
  double first[PSL_TUNE_VARIANTS];
  for (int v = 0; v < PSL_TUNE_VARIANTS && rc == PSL_OK; ++v) {
    g->tuning = v;
    double mean = 0.0;
    rc = time_updates(mean, steps);
    if (rc != PSL_OK) break;
    first[v] = mean;
    if (times_ms) times_ms[v] = mean;
    if (mean < best_t) {
      best_t = mean;
      best_v = v;
    }
  }
  // The variants within 2 % of the fastest are timed again over at least 256
  // steps. Each step moves every row by a fixed number of elements across
  // cache lines and memory channels, so the time per step runs in phases tens
  // of steps long (on the FP32 256x128x128 sphere one variant measures 211 to
  // 235 us per step over 16- to 128-step windows, and within 1 % over 256 or
  // 512); a short window sees a different phase for each variant. The caps
  // below are timed over the same count.
  const int long_steps = steps > 256 ? steps : 256;
  if (rc == PSL_OK) {
    const double limit = 1.02 * best_t;
    best_t = 1e300;
    for (int v = 0; v < PSL_TUNE_VARIANTS && rc == PSL_OK; ++v) {
      if (first[v] > limit) continue;
      g->tuning = v;
      double mean = 0.0;
      rc = time_updates(mean, long_steps);
      if (rc == PSL_OK && mean < best_t) {
        best_t = mean;
        best_v = v;
      }
    }
  }
  // occupancy caps for the chosen variant when its bulk kernel has a launch of
  // its own (always, unless the boundary cells are fused into it: a grid
  // without listed cells runs the bulk launch under every variant); a cap is
  // kept only if it wins by more than the timing noise
  int best_occ = 0;
  const bool fused_launch = (best_v & PSL_TUNE_FUSED_BOUNDARY) && g->flags && g->n_bc && !g->closure;
  if (rc == PSL_OK && !fused_launch) {
    g->tuning = best_v;
    const double base_t = best_t;
    for (const int occ : {5, 4, 3}) {
      if (psl_grid_set_occupancy(g, occ) != PSL_OK) continue;  // not expressible on this device
      double mean = 0.0;
      rc = time_updates(mean, long_steps);
      if (rc != PSL_OK) break;
      if (mean < best_t && mean < 0.995 * base_t) {
        best_t = mean;
        best_occ = occ;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PSL_OK) {
    g->tuning = saved;
    (void)psl_grid_set_occupancy(g, saved_occ);
    if (rc == PSL_E_HIP) set_error("autotune timing failed");
    return rc;
  }
  g->tuning = best_v;
  rc = psl_grid_set_occupancy(g, best_occ);
  if (rc) return rc;
  if (best) *best = best_v;
  return PSL_OK;
}

int psl_grid_set_closure(psl_grid* g, const psl_mlp* m) {
  PSL_CHECK_ARG(g, "null grid");
  g->closure = m;
  g->closure_count = 0;  // the next step evaluates the network
  return PSL_OK;
}

int psl_grid_set_closure_every(psl_grid* g, int steps) {
  PSL_CHECK_ARG(g && steps >= 1, "closure_every must be at least 1");
  if (steps > 1 && !g->closure_y) PSL_HIP_TRY(hipMalloc(&g->closure_y, (size_t)g->info.cellCount * sizeof(uint16_t)));
  g->closure_every = steps;
  g->closure_count = 0;
  return PSL_OK;
}

int psl_grid_get_closure_y(psl_grid* g, float* host) {
  PSL_CHECK_ARG(g && host, "null argument");
  PSL_CHECK_ARG(g->closure_y, "the grid holds no network output (psl_grid_set_closure_every with steps > 1)");
  const size_t n = (size_t)g->info.cellCount;
  std::vector<uint16_t> h(n);
  PSL_HIP_TRY(hipMemcpyAsync(h.data(), g->closure_y, n * sizeof(uint16_t), hipMemcpyDeviceToHost, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  for (size_t c = 0; c < n; ++c) {
    _Float16 v;
    memcpy(&v, &h[c], sizeof(v));
    host[c] = (float)v;
  }
  return PSL_OK;
}

// The dominant update kernel launched back to back between two events; with
// `advance`, the shifter moves on before every launch (streaming is a host-side
// shift update), so the launches see the shifts of consecutive steps.
static int time_dominant(psl_grid* g, int launches, bool advance, double* ms_per_launch, int* kernel) {
  PSL_CHECK_ARG(g && launches > 0 && ms_per_launch, "bad arguments");
  const bool fused = g->flags && g->n_bc && (g->tuning & PSL_TUNE_FUSED_BOUNDARY) && !g->closure;
  const bool was = g->profile;
  g->profile = false;
  hipEvent_t e0, e1;
  PSL_HIP_TRY(hipEventCreate(&e0));
  PSL_HIP_TRY(hipEventCreate(&e1));
  int rc = dispatch_storage<DominantFn>(g->storage, g);  // warm
  if (rc == PSL_OK && hipEventRecord(e0, g->stream) != hipSuccess) rc = PSL_E_HIP;
  for (int k = 0; k < launches && rc == PSL_OK; ++k) {
    if (advance) rc = psl_apply_streaming(g);
    if (rc == PSL_OK) rc = dispatch_storage<DominantFn>(g->storage, g);
  }
  if (rc == PSL_OK && hipEventRecord(e1, g->stream) != hipSuccess) rc = PSL_E_HIP;
  float ms = 0.f;
  if (rc == PSL_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = PSL_E_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g->profile = was;
  if (rc == PSL_E_HIP) set_error("kernel timing failed");
  if (rc) return rc;
  *ms_per_launch = (double)ms / launches;
  if (kernel) *kernel = g->closure ? 2 : fused ? 1 : 0;
  return PSL_OK;
}

int psl_grid_time_update_kernel(psl_grid* g, int launches, double* ms_per_launch, int* kernel) {
  return time_dominant(g, launches, false, ms_per_launch, kernel);
}

int psl_grid_time_update_kernel_streaming(psl_grid* g, int launches, double* ms_per_launch, int* kernel) {
  return time_dominant(g, launches, true, ms_per_launch, kernel);
}

int psl_grid_cell_census(const psl_grid* g, int64_t* bulk, int64_t* boundary, int64_t* skipped) {
  PSL_CHECK_ARG(g, "null grid");
  if (bulk) *bulk = g->census[0];
  if (boundary) *boundary = g->census[1];
  if (skipped) *skipped = g->census[2];
  return PSL_OK;
}

int psl_grid_apply_bounceback_mask(psl_grid* g, const psl_mask* m) {
  PSL_CHECK_ARG(g && m, "null argument");
  PSL_CHECK_ARG(m->n == g->info.cellCount, "mask size does not match the grid");
  const uint32_t n = ncells(g);
  // materialise host flags so that later set_info rebuilds keep the cells
  if (g->host_flags.empty()) g->host_flags.assign(n, 0);
  int rc = grid_rebuild_flags(g);
  if (rc) return rc;
  k_apply_bb_mask<<<blocks_for(n), 256, 0, g->stream>>>(g->flags, m->d, n);
  PSL_LAUNCH_CHECK();
  std::vector<uint8_t> hm(n);
  PSL_HIP_TRY(hipMemcpyAsync(hm.data(), m->d, n, hipMemcpyDeviceToHost, g->stream));
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  for (uint32_t c = 0; c < n; ++c)
    if (hm[c]) g->host_flags[c] = (uint8_t)((g->host_flags[c] & ~7) | KIND_BOUNCEBACK);
  return grid_rebuild_flags(g);
}

}  // extern "C"
