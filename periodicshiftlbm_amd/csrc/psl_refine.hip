// Block refinement on periodic-shift grids: fine->coarse averaging,
// coarse->fine gradient interpolation, both with central-moment rescaling,
// and the updateGrid level recursion.
//
// include/gridRefinementFunctions.h:283-536 and paper section 3 (eqs 3.21-3.26).
// The six faces of each direction are processed by one launch each
// (blockIdx.y = face): faces overlap only on edges/corners, where every
// writer stores identical values, so the concurrent faces equal the
// reference's sequential ones.  Fine->coarse runs before coarse->fine because
// the interpolation stencil reads the coarse ring that the averaging writes.
#include "psl_internal.h"

#include <algorithm>

namespace psl {

struct Box {
  int s[3];
  int e[3];
};
struct Boxes6 {
  Box b[6];
  int count;
  int dirs[6][9];  // populations written to ghost fine cells, per face
};

struct GridView {
  Shift27 sh;
  uint32_t n;
  int nx, ny, nz;
};

__device__ __forceinline__ int64_t slot_of(const GridView& g, uint32_t cell, int q) {
  uint32_t s = cell + (uint32_t)g.sh.s[q];
  if (s >= g.n) s -= g.n;
  return (int64_t)q * g.n + s;
}

template <class St>
__device__ __forceinline__ void load_cell(const typename St::T* f, const GridView& g, uint32_t cell, float (&v)[27]) {
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    v[q] = St::load(f[slot_of(g, cell, q)]);
    if (St::W_OFF) v[q] += weight(q);
  }
}

// writeToCoarseGridInterface (lines 283-336)
template <class St>
__global__ __launch_bounds__(256) void k_to_coarse(typename St::T* __restrict__ fc, GridView gc, int iss, int jss,
                                                   int kss, const typename St::T* __restrict__ ff, GridView gf,
                                                   Boxes6 boxes) {
  const Box& b = boxes.b[blockIdx.y];
  const int sx = b.e[0] - b.s[0], sy = b.e[1] - b.s[1], sz = b.e[2] - b.s[2];
  if (sx <= 0 || sy <= 0 || sz <= 0) return;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)sx * sy * sz) return;
  const int kr = t / (sx * sy);
  const int rem = t - kr * sx * sy;
  const int jr = rem / sx, ir = rem - jr * sx;
  const int ic = b.s[0] + ir, jc = b.s[1] + jr, kc = b.s[2] + kr;
  const int i0 = (ic - iss) * 2, j0 = (jc - jss) * 2, k0 = (kc - kss) * 2;
  float acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q] = 0.f;
  for (int ka = 0; ka < 2; ++ka)
    for (int ja = 0; ja < 2; ++ja)
      for (int ia = 0; ia < 2; ++ia) {
        const uint32_t cf = (uint32_t)(k0 + ka) * gf.nx * gf.ny + (uint32_t)(j0 + ja) * gf.nx + (i0 + ia);
        float v[27];
        load_cell<St>(ff, gf, cf, v);
#pragma unroll
        for (int q = 0; q < 27; ++q) acc[q] += v[q];
      }
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q] *= 0.125f;
  rescale(acc, false);
  const uint32_t cc = (uint32_t)kc * gc.nx * gc.ny + (uint32_t)jc * gc.nx + ic;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const float o = St::W_OFF ? acc[q] - weight(q) : acc[q];
    fc[slot_of(gc, cc, q)] = St::store(o);
  }
}

// writeToFineGridInterface (lines 338-431)
template <class St>
__global__ __launch_bounds__(256) void k_to_fine(const typename St::T* __restrict__ fc, GridView gc, int iss, int jss,
                                                 int kss, typename St::T* __restrict__ ff, GridView gf,
                                                 const uint8_t* __restrict__ fine_flags, Boxes6 boxes) {
  const Box& b = boxes.b[blockIdx.y];
  const int sx = b.e[0] - b.s[0], sy = b.e[1] - b.s[1], sz = b.e[2] - b.s[2];
  if (sx <= 0 || sy <= 0 || sz <= 0) return;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)sx * sy * sz) return;
  const int kr = t / (sx * sy);
  const int rem = t - kr * sx * sy;
  const int jr = rem / sx, ir = rem - jr * sx;
  const int i_f = b.s[0] + ir, jf = b.s[1] + jr, kf = b.s[2] + kr;
  const uint32_t cellf = (uint32_t)kf * gf.nx * gf.ny + (uint32_t)jf * gf.nx + i_f;
  const int ic = i_f / 2 + iss, jc = jf / 2 + jss, kc = kf / 2 + kss;
  const float dx = (i_f & 1) ? 0.25f : -0.25f;
  const float dy = (jf & 1) ? 0.25f : -0.25f;
  const float dz = (kf & 1) ? 0.25f : -0.25f;
  float f[27];
  load_cell<St>(fc, gc, (uint32_t)kc * gc.nx * gc.ny + (uint32_t)jc * gc.nx + ic, f);
  int ip = ic + 1, im = ic - 1, jp = jc + 1, jm = jc - 1, kp = kc + 1, km = kc - 1;
  float dix = 0.5f, diy = 0.5f, diz = 0.5f;
  if (ip >= gc.nx) { ip = ic; dix = 1.f; }
  if (im < 0) { im = 0; dix = 1.f; }
  if (jp >= gc.ny) { jp = jc; diy = 1.f; }
  if (jm < 0) { jm = 0; diy = 1.f; }
  if (kp >= gc.nz) { kp = kc; diz = 1.f; }
  if (km < 0) { km = 0; diz = 1.f; }
  const int ni[6] = {ip, im, ic, ic, ic, ic};
  const int nj[6] = {jc, jc, jp, jm, jc, jc};
  const int nk[6] = {kc, kc, kc, kc, kp, km};
  const float w6[6] = {dix * dx, -dix * dx, diy * dy, -diy * dy, diz * dz, -diz * dz};
  for (int nb = 0; nb < 6; ++nb) {
    const uint32_t cn = (uint32_t)nk[nb] * gc.nx * gc.ny + (uint32_t)nj[nb] * gc.nx + ni[nb];
    float v[27];
    load_cell<St>(fc, gc, cn, v);
#pragma unroll
    for (int q = 0; q < 27; ++q) f[q] += v[q] * w6[nb];
  }
  rescale(f, true);
  const bool ghost = fine_flags && ((fine_flags[cellf] & 7) == KIND_GHOST);
  if (ghost) {
    for (int t9 = 0; t9 < 9; ++t9) {
      const int q = boxes.dirs[blockIdx.y][t9];
      float o = f[0];
#pragma unroll
      for (int r = 0; r < 27; ++r)
        if (r == q) o = f[r];
      if (St::W_OFF) o -= weight(q);
      ff[slot_of(gf, cellf, q)] = St::store(o);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const float o = St::W_OFF ? f[q] - weight(q) : f[q];
      ff[slot_of(gf, cellf, q)] = St::store(o);
    }
  }
}

static GridView view_of(const psl_grid* g) {
  GridView v;
  for (int q = 0; q < 27; ++q) v.sh.s[q] = g->shift[q];
  v.n = (uint32_t)g->info.cellCount;
  v.nx = g->info.cellCountX;
  v.ny = g->info.cellCountY;
  v.nz = g->info.cellCountZ;
  return v;
}

static uint32_t box_blocks(const Boxes6& B) {
  uint64_t m = 0;
  for (int i = 0; i < B.count; ++i) {
    const Box& b = B.b[i];
    const int64_t v = (int64_t)std::max(0, b.e[0] - b.s[0]) * std::max(0, b.e[1] - b.s[1]) * std::max(0, b.e[2] - b.s[2]);
    m = std::max<uint64_t>(m, (uint64_t)v);
  }
  return (uint32_t)((m + 255) / 256);
}

static void set_box(Box& b, int s0, int s1, int s2, int e0, int e1, int e2) {
  b.s[0] = s0; b.s[1] = s1; b.s[2] = s2;
  b.e[0] = e0; b.e[1] = e1; b.e[2] = e2;
}

static const int POS_X[9] = {1, 7, 9, 12, 16, 20, 22, 23, 26};
static const int NEG_X[9] = {2, 8, 10, 11, 15, 19, 21, 24, 25};
static const int POS_Y[9] = {6, 12, 13, 15, 17, 19, 22, 24, 26};
static const int NEG_Y[9] = {5, 11, 14, 16, 18, 20, 21, 23, 25};
static const int POS_Z[9] = {4, 8, 9, 14, 17, 20, 21, 24, 26};
static const int NEG_Z[9] = {3, 7, 10, 13, 18, 19, 22, 23, 25};

template <class St>
static int to_coarse(psl_grid* c, psl_grid* f, const Boxes6& B) {
  if (B.count == 0) return PSL_OK;
  const uint32_t nb = box_blocks(B);
  if (nb == 0) return PSL_OK;
  dim3 grid(nb, B.count);
  k_to_coarse<St><<<grid, 256, 0, c->stream>>>((typename St::T*)c->f, view_of(c), c->info.iSubgridStart,
                                               c->info.jSubgridStart, c->info.kSubgridStart,
                                               (const typename St::T*)f->f, view_of(f), B);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int to_fine(psl_grid* c, psl_grid* f, const Boxes6& B) {
  if (B.count == 0) return PSL_OK;
  const uint32_t nb = box_blocks(B);
  if (nb == 0) return PSL_OK;
  dim3 grid(nb, B.count);
  k_to_fine<St><<<grid, 256, 0, c->stream>>>((const typename St::T*)c->f, view_of(c), c->info.iSubgridStart,
                                             c->info.jSubgridStart, c->info.kSubgridStart, (typename St::T*)f->f,
                                             view_of(f), f->flags, B);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int communicate(psl_grid* c, psl_grid* f, bool full_fill) {
  const psl_info& C = c->info;
  const psl_info& F = f->info;
  Boxes6 B{};
  if (full_fill) {
    set_box(B.b[0], C.iSubgridStart + 1, C.jSubgridStart + 1, C.kSubgridStart + 1, C.iSubgridEnd - 1,
            C.jSubgridEnd - 1, C.kSubgridEnd - 1);
    B.count = 1;
    return to_coarse<St>(c, f, B);
  }
  // fine -> coarse ring (lines 447-486)
  B.count = 0;
  if (C.cellCountX > 1) {
    set_box(B.b[B.count++], C.iSubgridStart + 1, C.jSubgridStart + 1, C.kSubgridStart + 1, C.iSubgridStart + 2,
            C.jSubgridEnd - 1, C.kSubgridEnd - 1);
    set_box(B.b[B.count++], C.iSubgridEnd - 2, C.jSubgridStart + 1, C.kSubgridStart + 1, C.iSubgridEnd - 1,
            C.jSubgridEnd - 1, C.kSubgridEnd - 1);
  }
  if (C.cellCountY > 1) {
    set_box(B.b[B.count++], C.iSubgridStart + 1, C.jSubgridStart + 1, C.kSubgridStart + 1, C.iSubgridEnd - 1,
            C.jSubgridStart + 2, C.kSubgridEnd - 1);
    set_box(B.b[B.count++], C.iSubgridStart + 1, C.jSubgridEnd - 2, C.kSubgridStart + 1, C.iSubgridEnd - 1,
            C.jSubgridEnd - 1, C.kSubgridEnd - 1);
  }
  if (C.cellCountZ > 1) {
    set_box(B.b[B.count++], C.iSubgridStart + 1, C.jSubgridStart + 1, C.kSubgridStart + 1, C.iSubgridEnd - 1,
            C.jSubgridEnd - 1, C.kSubgridStart + 2);
    set_box(B.b[B.count++], C.iSubgridStart + 1, C.jSubgridStart + 1, C.kSubgridEnd - 2, C.iSubgridEnd - 1,
            C.jSubgridEnd - 1, C.kSubgridEnd - 1);
  }
  int rc = to_coarse<St>(c, f, B);
  if (rc) return rc;
  // coarse -> fine layers (lines 488-525)
  Boxes6 Fb{};
  Fb.count = 0;
  auto add = [&](int s0, int s1, int s2, int e0, int e1, int e2, const int* dirs) {
    set_box(Fb.b[Fb.count], s0, s1, s2, e0, e1, e2);
    for (int t = 0; t < 9; ++t) Fb.dirs[Fb.count][t] = dirs[t];
    Fb.count++;
  };
  if (C.iSubgridStart > 0) add(0, 0, 0, 2, F.cellCountY, F.cellCountZ, POS_X);
  if (C.iSubgridEnd < C.cellCountX - 1) add(F.cellCountX - 2, 0, 0, F.cellCountX, F.cellCountY, F.cellCountZ, NEG_X);
  if (C.jSubgridStart > 0) add(0, 0, 0, F.cellCountX, 2, F.cellCountZ, POS_Y);
  if (C.jSubgridEnd < C.cellCountY - 1) add(0, F.cellCountY - 2, 0, F.cellCountX, F.cellCountY, F.cellCountZ, NEG_Y);
  if (C.kSubgridStart > 0) add(0, 0, 0, F.cellCountX, F.cellCountY, 2, POS_Z);
  if (C.kSubgridEnd < C.cellCountZ - 1) add(0, 0, F.cellCountZ - 2, F.cellCountX, F.cellCountY, F.cellCountZ, NEG_Z);
  return to_fine<St>(c, f, Fb);
}

template <class St> struct CommFn {
  static int run(psl_grid* c, psl_grid* f, bool full) { return communicate<St>(c, f, full); }
};

static int comm_dispatch(psl_grid* c, psl_grid* f, bool full) {
  if (!c || !f) {
    set_error("null grid");
    return PSL_E_ARG;
  }
  if (c->storage != f->storage) {
    set_error("coarse and fine grids must use the same storage");
    return PSL_E_ARG;
  }
  const psl_info& C = c->info;
  const psl_info& F = f->info;
  if (F.cellCountX != 2 * (C.iSubgridEnd - C.iSubgridStart) || F.cellCountY != 2 * (C.jSubgridEnd - C.jSubgridStart) ||
      F.cellCountZ != 2 * (C.kSubgridEnd - C.kSubgridStart)) {
    set_error("fine grid counts must be twice the coarse sub-grid extent");
    return PSL_E_ARG;
  }
  if (C.iSubgridStart < 0 || C.jSubgridStart < 0 || C.kSubgridStart < 0 || C.iSubgridEnd > C.cellCountX ||
      C.jSubgridEnd > C.cellCountY || C.kSubgridEnd > C.cellCountZ) {
    set_error("sub-grid outside the coarse grid");
    return PSL_E_ARG;
  }
  if (c->stream != f->stream) {
    // order the fine grid's pending work before the transfer on the coarse stream
    hipEvent_t ev;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return PSL_E_HIP;
    (void)hipEventRecord(ev, f->stream);
    (void)hipStreamWaitEvent(c->stream, ev, 0);
    (void)hipEventDestroy(ev);
  }
  switch (c->storage) {
    case PSL_F32: return CommFn<StF32>::run(c, f, full);
    case PSL_F16DEV: return CommFn<StF16>::run(c, f, full);
    case PSL_BF16DEV: return CommFn<StBF16>::run(c, f, full);
  }
  return PSL_E_ARG;
}

static int update_level(psl_grid* const* levels, int nlevels, int level) {
  psl_grid* g = levels[level];
  int rc = psl_apply_streaming(g);
  if (!rc && (g->periodic_copies & 1)) rc = psl_apply_bc_periodic_x(g);
  if (!rc && (g->periodic_copies & 2)) rc = psl_apply_bc_periodic_y(g);
  if (!rc) rc = psl_apply_local_cell_update(g);
  if (rc) return rc;
  if (level < nlevels - 1) {
    for (int i = 0; i < 2; ++i) {
      rc = update_level(levels, nlevels, level + 1);
      if (rc) return rc;
    }
    rc = comm_dispatch(levels[level], levels[level + 1], false);
  }
  return rc;
}

}  // namespace psl

using namespace psl;

extern "C" {

int psl_coarse_fine_communication(psl_grid* coarse, psl_grid* fine) { return comm_dispatch(coarse, fine, false); }

int psl_fill_coarse_from_fine(psl_grid* coarse, psl_grid* fine) { return comm_dispatch(coarse, fine, true); }

int psl_update_grid(psl_grid* const* levels, int nlevels, int nsteps) {
  PSL_CHECK_ARG(levels && nlevels > 0 && nsteps >= 0, "bad level list");
  for (int l = 0; l < nlevels; ++l) PSL_CHECK_ARG(levels[l], "null level");
  for (int s = 0; s < nsteps; ++s) {
    const int rc = update_level(levels, nlevels, 0);
    if (rc) return rc;
  }
  return PSL_OK;
}

}  // extern "C"
