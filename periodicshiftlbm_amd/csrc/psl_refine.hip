// Block refinement on periodic-shift grids: fine->coarse averaging,
// coarse->fine gradient interpolation, both with central-moment rescaling,
// and the updateGrid level recursion.
//
// include/gridRefinementFunctions.h:283-536 and paper section 3 (eqs 3.21-3.26).
// Both transfers work on faces of coarse cells.  A workgroup owns a run of
// coarse cells along the face's first axis with more than one cell and all
// 2x2x2 fine children of each: lane = 2 * cell + (fine x parity), wave =
// (fine y parity) + 2 * (fine z parity), so that a wave touches 64
// consecutive fine cells of one row wherever the face extends in x.  The six
// faces of a direction set are one launch (blockIdx.y = face): faces overlap
// only on edges/corners, where every writer stores identical values, so the
// concurrent faces equal the reference's sequential ones.  Fine->coarse runs
// before coarse->fine because the interpolation stencil reads the coarse
// ring that the averaging writes.
#include "psl_internal.h"

#include <algorithm>

namespace psl {

struct GridView {
  Shift27 sh;
  uint32_t n;
  int nx, ny, nz;
};

// A face in coarse cells: cell (a, b, c) of the face is base + a*step_a +
// b*step_b + c*step_c, a < na, b < nb; workgroup w covers cells
// [(w % wa) * CELLS, +CELLS) along a at b = (w / wa) % nb, c = w / (wa * nb).
struct Face {
  int base[3];
  int step_a[3], step_b[3], step_c[3];
  int na, nb;
  uint32_t wa;
  uint32_t blocks;
  uint32_t ghost_pops;  // bit q: population q is written to ghost fine cells
};
struct Faces {
  Face f[6];
  int count;
};

struct FaceCell {
  int x, y, z;
  bool valid;
};

__device__ __forceinline__ FaceCell face_cell(const Face& F, int cells_per_block, int cell) {
  const uint32_t wa = blockIdx.x % F.wa, rest = blockIdx.x / F.wa;
  const int b = (int)(rest % (uint32_t)F.nb), c = (int)(rest / (uint32_t)F.nb);
  const int a = (int)wa * cells_per_block + cell;
  FaceCell r;
  r.x = F.base[0] + a * F.step_a[0] + b * F.step_b[0] + c * F.step_c[0];
  r.y = F.base[1] + a * F.step_a[1] + b * F.step_b[1] + c * F.step_c[1];
  r.z = F.base[2] + a * F.step_a[2] + b * F.step_b[2] + c * F.step_c[2];
  r.valid = a < F.na;
  return r;
}

__device__ __forceinline__ uint32_t cell_index(const GridView& g, int x, int y, int z) {
  return ((uint32_t)z * (uint32_t)g.ny + (uint32_t)y) * (uint32_t)g.nx + (uint32_t)x;
}

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

// Value of the neighbouring lane of the pair (lane ^ 1), a DPP quad permutation.
__device__ __forceinline__ float pair_swap(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

constexpr int TC_CELLS = 32;  // coarse cells per workgroup, fine->coarse
constexpr int TF_CELLS = 64;  // coarse cells per workgroup, coarse->fine

// writeToCoarseGridInterface (lines 283-336): each thread loads one fine
// child, pairs add in registers, the four waves add through LDS, and one lane
// per coarse cell rescales and stores.
template <class St>
__global__ __launch_bounds__(256) void k_to_coarse(typename St::T* __restrict__ fc, GridView gc, int iss, int jss,
                                                   int kss, const typename St::T* __restrict__ ff, GridView gf,
                                                   Faces faces) {
  __shared__ float part[4][27][TC_CELLS];
  const Face& F = faces.f[blockIdx.y];
  if (blockIdx.x >= F.blocks) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const FaceCell cc = face_cell(F, TC_CELLS, lane >> 1);
  float v[27];
  if (cc.valid) {
    const int fx = 2 * (cc.x - iss) + (lane & 1), fy = 2 * (cc.y - jss) + (w & 1), fz = 2 * (cc.z - kss) + (w >> 1);
    load_cell<St>(ff, gf, cell_index(gf, fx, fy, fz), v);
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = 0.f;
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] += pair_swap(v[q]);
  if (!(lane & 1)) {
#pragma unroll
    for (int q = 0; q < 27; ++q) part[w][q][lane >> 1] = v[q];
  }
  __syncthreads();
  if (threadIdx.x >= TC_CELLS) return;
  const int cell = threadIdx.x;
  const FaceCell oc = face_cell(F, TC_CELLS, cell);
  if (!oc.valid) return;
  float acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q)
    acc[q] = ((part[0][q][cell] + part[1][q][cell]) + (part[2][q][cell] + part[3][q][cell])) * 0.125f;
  rescale(acc, false);
  const uint32_t c = cell_index(gc, oc.x, oc.y, oc.z);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const float o = St::W_OFF ? acc[q] - weight(q) : acc[q];
    fc[slot_of(gc, c, q)] = St::store(o);
  }
}

// writeToFineGridInterface (lines 338-431).  The interpolant of child
// (sx, sy, sz) in {-1, 1}^3 of coarse cell c is
//   f_c + sx hx (f_{i+1} - f_{i-1}) + sy hy (...) + sz hz (...),
// h = 1/8 for the central difference and 1/4 for the one-sided difference at
// the grid edge.  Phase 1 (one coarse cell per lane, populations split over
// the four waves) puts f_c and the three scaled differences in LDS; phase 2
// (one fine cell per thread, two halves of the coarse run) interpolates,
// rescales and stores, only the inward populations in ghost cells.
template <class St>
__global__ __launch_bounds__(256) void k_to_fine(const typename St::T* __restrict__ fc, GridView gc, int iss, int jss,
                                                 int kss, typename St::T* __restrict__ ff, GridView gf,
                                                 const uint8_t* __restrict__ fine_flags, Faces faces) {
  __shared__ float st[4][27][TF_CELLS];
  const Face& F = faces.f[blockIdx.y];
  if (blockIdx.x >= F.blocks) return;
  {
    const int cell = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const FaceCell cc = face_cell(F, TF_CELLS, cell);
    if (cc.valid) {
      int ip = cc.x + 1, im = cc.x - 1, jp = cc.y + 1, jm = cc.y - 1, kp = cc.z + 1, km = cc.z - 1;
      float hx = 0.125f, hy = 0.125f, hz = 0.125f;
      if (ip >= gc.nx) { ip = cc.x; hx = 0.25f; }
      if (im < 0) { im = 0; hx = 0.25f; }
      if (jp >= gc.ny) { jp = cc.y; hy = 0.25f; }
      if (jm < 0) { jm = 0; hy = 0.25f; }
      if (kp >= gc.nz) { kp = cc.z; hz = 0.25f; }
      if (km < 0) { km = 0; hz = 0.25f; }
      const uint32_t c0 = cell_index(gc, cc.x, cc.y, cc.z);
      const uint32_t cxp = cell_index(gc, ip, cc.y, cc.z), cxm = cell_index(gc, im, cc.y, cc.z);
      const uint32_t cyp = cell_index(gc, cc.x, jp, cc.z), cym = cell_index(gc, cc.x, jm, cc.z);
      const uint32_t czp = cell_index(gc, cc.x, cc.y, kp), czm = cell_index(gc, cc.x, cc.y, km);
      for (int q = w; q < 27; q += 4) {
        const float wq = St::W_OFF ? weight(q) : 0.f;
        const float c = St::load(fc[slot_of(gc, c0, q)]) + wq;
        const float xp = St::load(fc[slot_of(gc, cxp, q)]), xm = St::load(fc[slot_of(gc, cxm, q)]);
        const float yp = St::load(fc[slot_of(gc, cyp, q)]), ym = St::load(fc[slot_of(gc, cym, q)]);
        const float zp = St::load(fc[slot_of(gc, czp, q)]), zm = St::load(fc[slot_of(gc, czm, q)]);
        st[0][q][cell] = c;
        st[1][q][cell] = hx * (xp - xm);
        st[2][q][cell] = hy * (yp - ym);
        st[3][q][cell] = hz * (zp - zm);
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ia = lane & 1, ja = w & 1, ka = w >> 1;
  const float sx = ia ? 1.f : -1.f, sy = ja ? 1.f : -1.f, sz = ka ? 1.f : -1.f;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int cell = h * (TF_CELLS / 2) + (lane >> 1);
    const FaceCell cc = face_cell(F, TF_CELLS, cell);
    if (!cc.valid) continue;
    const uint32_t cellf = cell_index(gf, 2 * (cc.x - iss) + ia, 2 * (cc.y - jss) + ja, 2 * (cc.z - kss) + ka);
    float f[27];
#pragma unroll
    for (int q = 0; q < 27; ++q)
      f[q] = st[0][q][cell] + sx * st[1][q][cell] + sy * st[2][q][cell] + sz * st[3][q][cell];
    rescale(f, true);
    const bool ghost = fine_flags && ((fine_flags[cellf] & 7) == KIND_GHOST);
    const uint32_t keep = ghost ? F.ghost_pops : 0x7ffffffu;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      if (!((keep >> q) & 1u)) continue;
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

// Appends the coarse box [s, e) as a face; the first axis with more than one
// cell runs along the workgroup.
static void add_face(Faces& F, const int s[3], const int e[3], int cells_per_block, uint32_t ghost_pops) {
  int n[3];
  for (int d = 0; d < 3; ++d) n[d] = std::max(0, e[d] - s[d]);
  if (n[0] == 0 || n[1] == 0 || n[2] == 0) return;
  const int a = n[0] > 1 ? 0 : n[1] > 1 ? 1 : n[2] > 1 ? 2 : 0;
  const int b = a == 0 ? 1 : 0;
  const int c = 3 - a - b;
  Face& f = F.f[F.count++];
  for (int d = 0; d < 3; ++d) {
    f.base[d] = s[d];
    f.step_a[d] = d == a;
    f.step_b[d] = d == b;
    f.step_c[d] = d == c;
  }
  f.na = n[a];
  f.nb = n[b];
  f.wa = (uint32_t)((n[a] + cells_per_block - 1) / cells_per_block);
  f.blocks = f.wa * (uint32_t)n[b] * (uint32_t)n[c];
  f.ghost_pops = ghost_pops;
}

static uint32_t face_blocks(const Faces& F) {
  uint32_t m = 0;
  for (int i = 0; i < F.count; ++i) m = std::max(m, F.f[i].blocks);
  return m;
}

static uint32_t pop_mask(const int (&dirs)[9]) {
  uint32_t m = 0;
  for (int t = 0; t < 9; ++t) m |= 1u << dirs[t];
  return m;
}

static const int POS_X[9] = {1, 7, 9, 12, 16, 20, 22, 23, 26};
static const int NEG_X[9] = {2, 8, 10, 11, 15, 19, 21, 24, 25};
static const int POS_Y[9] = {6, 12, 13, 15, 17, 19, 22, 24, 26};
static const int NEG_Y[9] = {5, 11, 14, 16, 18, 20, 21, 23, 25};
static const int POS_Z[9] = {4, 8, 9, 14, 17, 20, 21, 24, 26};
static const int NEG_Z[9] = {3, 7, 10, 13, 18, 19, 22, 23, 25};

template <class St>
static int to_coarse(psl_grid* c, psl_grid* f, const Faces& F) {
  const uint32_t nb = face_blocks(F);
  if (nb == 0) return PSL_OK;
  dim3 grid(nb, F.count);
  k_to_coarse<St><<<grid, 256, 0, c->stream>>>((typename St::T*)c->f, view_of(c), c->info.iSubgridStart,
                                               c->info.jSubgridStart, c->info.kSubgridStart,
                                               (const typename St::T*)f->f, view_of(f), F);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int to_fine(psl_grid* c, psl_grid* f, const Faces& F) {
  const uint32_t nb = face_blocks(F);
  if (nb == 0) return PSL_OK;
  dim3 grid(nb, F.count);
  k_to_fine<St><<<grid, 256, 0, c->stream>>>((const typename St::T*)c->f, view_of(c), c->info.iSubgridStart,
                                             c->info.jSubgridStart, c->info.kSubgridStart, (typename St::T*)f->f,
                                             view_of(f), f->flags, F);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int communicate(psl_grid* c, psl_grid* f, bool full_fill) {
  const psl_info& C = c->info;
  const int is = C.iSubgridStart, js = C.jSubgridStart, ks = C.kSubgridStart;
  const int ie = C.iSubgridEnd, je = C.jSubgridEnd, ke = C.kSubgridEnd;
  Faces B{};
  if (full_fill) {
    const int s[3] = {is + 1, js + 1, ks + 1}, e[3] = {ie - 1, je - 1, ke - 1};
    add_face(B, s, e, TC_CELLS, 0);
    return to_coarse<St>(c, f, B);
  }
  // fine -> coarse ring (lines 447-486)
  if (C.cellCountX > 1) {
    const int s0[3] = {is + 1, js + 1, ks + 1}, e0[3] = {is + 2, je - 1, ke - 1};
    const int s1[3] = {ie - 2, js + 1, ks + 1}, e1[3] = {ie - 1, je - 1, ke - 1};
    add_face(B, s0, e0, TC_CELLS, 0);
    add_face(B, s1, e1, TC_CELLS, 0);
  }
  if (C.cellCountY > 1) {
    const int s0[3] = {is + 1, js + 1, ks + 1}, e0[3] = {ie - 1, js + 2, ke - 1};
    const int s1[3] = {is + 1, je - 2, ks + 1}, e1[3] = {ie - 1, je - 1, ke - 1};
    add_face(B, s0, e0, TC_CELLS, 0);
    add_face(B, s1, e1, TC_CELLS, 0);
  }
  if (C.cellCountZ > 1) {
    const int s0[3] = {is + 1, js + 1, ks + 1}, e0[3] = {ie - 1, je - 1, ks + 2};
    const int s1[3] = {is + 1, js + 1, ke - 2}, e1[3] = {ie - 1, je - 1, ke - 1};
    add_face(B, s0, e0, TC_CELLS, 0);
    add_face(B, s1, e1, TC_CELLS, 0);
  }
  int rc = to_coarse<St>(c, f, B);
  if (rc) return rc;
  // coarse -> fine layers (lines 488-525): two fine cells deep, i.e. one
  // coarse layer of the sub-grid, on each side that is not the grid edge
  Faces Fb{};
  auto add = [&](int s0, int s1, int s2, int e0, int e1, int e2, const int(&dirs)[9]) {
    const int s[3] = {s0, s1, s2}, e[3] = {e0, e1, e2};
    add_face(Fb, s, e, TF_CELLS, pop_mask(dirs));
  };
  if (is > 0) add(is, js, ks, is + 1, je, ke, POS_X);
  if (ie < C.cellCountX - 1) add(ie - 1, js, ks, ie, je, ke, NEG_X);
  if (js > 0) add(is, js, ks, ie, js + 1, ke, POS_Y);
  if (je < C.cellCountY - 1) add(is, je - 1, ks, ie, je, ke, NEG_Y);
  if (ks > 0) add(is, js, ks, ie, je, ks + 1, POS_Z);
  if (ke < C.cellCountZ - 1) add(is, js, ke - 1, ie, je, ke, NEG_Z);
  return to_fine<St>(c, f, Fb);
}

template <class St> struct CommFn {
  static int run(psl_grid* c, psl_grid* f, bool full) { return communicate<St>(c, f, full); }
};

// Refuses a coarse/fine pair the transfers cannot serve; launches nothing.
static int comm_check(const psl_grid* c, const psl_grid* f) {
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
  // the fine-to-coarse ring is one coarse cell inside each face of the
  // sub-grid, so along every axis the coarse grid extends in the sub-grid must
  // span at least two coarse cells
  if ((C.cellCountX > 1 && C.iSubgridEnd - C.iSubgridStart < 2) ||
      (C.cellCountY > 1 && C.jSubgridEnd - C.jSubgridStart < 2) ||
      (C.cellCountZ > 1 && C.kSubgridEnd - C.kSubgridStart < 2)) {
    set_error("sub-grid must span at least two coarse cells along every axis of the coarse grid");
    return PSL_E_ARG;
  }
  return PSL_OK;
}

static int comm_dispatch(psl_grid* c, psl_grid* f, bool full) {
  const int rc = comm_check(c, f);
  if (rc) return rc;
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
  // every coarse/fine pair is checked before the first launch
  for (int l = 0; l + 1 < nlevels; ++l) {
    const int rc = comm_check(levels[l], levels[l + 1]);
    if (rc) return rc;
  }
  for (int s = 0; s < nsteps; ++s) {
    const int rc = update_level(levels, nlevels, 0);
    if (rc) return rc;
  }
  return PSL_OK;
}

}  // extern "C"
