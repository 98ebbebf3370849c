// Section-cut and history output in the reference's plotter formats.
//
// include/plotter/exportSectionCutPlot.h: a plane of a grid (XY at k = cut,
// ZY at i = cut, ZX at j = cut) as rows of (p, u_horizontal, u_vertical,
// u_normal, bounceback marker[, gridID]) with the pressure
// p = (rho - 1) rho_nominal c_s^2 and velocities in m/s (convertToPhysical*,
// include/cellFunctions.h:200-219), and the binary files the reference's
// Python plotters read: sim_data.bin = int header {plotNumber, rows, cols,
// components} + float rows, historyData.bin = int count + count floats.
//
// The fields are evaluated on the device: the macroscopic fields of the grid,
// then one thread per pixel of the plane.  For DIAD levels every level paints
// its cells on the plane in level order, coarse to fine, at the resolution of
// the finest level whose plane has fewer than max_pixels pixels (coarse cells
// fill their whole footprint; where several finer cells fall on one pixel,
// the one with the largest cell index wins).
#include "psl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

namespace psl {
namespace {

struct CutParams {
  int plane;  // 0 XY, 1 ZY, 2 ZX
  int cut;
  int h, v;   // pixels horizontally / vertically
  float vel_scale;  // res / 1000 / dtPhys
  float p_scale;    // rho_nominal * c_s^2
  int h0 = 0, v0 = 0;  // first pixel of a cropped window
  int raw = 0;         // 1: rho, ux, uy, uz in lattice units (flow reports)
  // unrolled cylinder about z (plane ignored): radius r (mm), the finest
  // level's spacing and the finest-level indices of the origin
  int cyl = 0;
  float r = 0.f, fres = 0.f;
  int io = 0, jo = 0;
  int tscale = 1;  // finest-level cells per pixel
};

__device__ __forceinline__ void plane_components(int plane, float ux, float uy, float uz, float& uh, float& uv,
                                                 float& un) {
  if (plane == 0) { uh = ux; uv = uy; un = uz; }
  else if (plane == 1) { uh = uz; uv = uy; un = ux; }
  else { uh = uz; uv = ux; un = uy; }
}

// exportSectionCutPlotGeneral(GridStruct&) (lines 2-89), one thread per pixel
// export3DPlot (lines 591-672): every cell in (i, j, k) order with k fastest,
// five floats (p, ux, uy, uz in m/s, bounceback marker)
__global__ void k_volume(const float* __restrict__ macro, const uint8_t* __restrict__ flags, int nx, int ny, int nz,
                         float vel_scale, float p_scale, float* __restrict__ out) {
  const size_t n = (size_t)nx * ny * nz;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(o % nz);
    const int j = (int)((o / nz) % ny);
    const int i = (int)(o / ((size_t)nz * ny));
    const size_t cell = ((size_t)k * ny + j) * nx + i;
    float* d = out + 5 * o;
    d[0] = (macro[cell] - 1.f) * p_scale;
    d[1] = macro[n + cell] * vel_scale;
    d[2] = macro[2 * n + cell] * vel_scale;
    d[3] = macro[3 * n + cell] * vel_scale;
    d[4] = (flags && (flags[cell] & 7) == KIND_BOUNCEBACK) ? 1.f : 0.f;
  }
}

// exportSectionCutPlotToiletPaperZ(GridStruct&) (lines 674-787): the cylinder of
// radius r about the z axis unrolled, rows along the circumference (row v at
// angle v res / r, cell column col[v] = j nx + i and cos / sin of that angle
// from the host), columns along k; five floats (p, uz, u_tangential,
// u_radial, bounceback marker).
__global__ void k_cut_cyl(const float* __restrict__ macro, const uint8_t* __restrict__ flags, size_t n, size_t plane,
                          int nz, const int* __restrict__ col, const float2* __restrict__ rot, int rows,
                          float vel_scale, float p_scale, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (k >= nz || v >= rows) return;
  const size_t cell = (size_t)k * plane + (size_t)col[v];
  const float ux = macro[n + cell] * vel_scale, uy = macro[2 * n + cell] * vel_scale,
              uz = macro[3 * n + cell] * vel_scale;
  const float c = rot[v].x, s = rot[v].y;
  float* o = out + 5 * ((size_t)v * nz + k);
  o[0] = (macro[cell] - 1.f) * p_scale;
  o[1] = uz;
  o[2] = -ux * s + uy * c;
  o[3] = ux * c + uy * s;
  o[4] = (flags && (flags[cell] & 7) == KIND_BOUNCEBACK) ? 1.f : 0.f;
}

// With `scalar` (n floats) a sixth component, the scalar of the cell (lines 108-204).
__global__ void k_cut_grid(const float* __restrict__ macro, const uint8_t* __restrict__ flags, int nx, int ny,
                           size_t n, CutParams c, float* __restrict__ out, const float* __restrict__ scalar) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= c.h || y >= c.v) return;
  int i, j, k;
  if (c.plane == 0) { i = x; j = y; k = c.cut; }
  else if (c.plane == 1) { i = c.cut; j = y; k = x; }
  else { i = y; j = c.cut; k = x; }
  const size_t cell = ((size_t)k * ny + j) * nx + i;
  const float rho = macro[cell], ux = macro[n + cell] * c.vel_scale, uy = macro[2 * n + cell] * c.vel_scale,
              uz = macro[3 * n + cell] * c.vel_scale;
  float uh, uv, un;
  plane_components(c.plane, ux, uy, uz, uh, uv, un);
  const int comps = scalar ? 6 : 5;
  float* o = out + comps * ((size_t)y * c.h + x);
  o[0] = (rho - 1.f) * c.p_scale;
  o[1] = uh;
  o[2] = uv;
  o[3] = un;
  o[4] = (flags && (flags[cell] & 7) == KIND_BOUNCEBACK) ? 1.f : 0.f;
  if (scalar) o[5] = scalar[cell];
}

// exportSectionCutPlotGeneral(std::vector<DIADGridStruct>&) (lines 377-573):
// the cells of one level on the plane claim their pixels (largest cell index wins)
__global__ void k_cut_claim(const int* __restrict__ ijk, int64_t n, int cell_scale, int target_scale, CutParams c,
                            int* __restrict__ winner) {
  for (int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < n;
       cell += (int64_t)gridDim.x * blockDim.x) {
    const int i = ijk[cell] * cell_scale, j = ijk[n + cell] * cell_scale, k = ijk[2 * n + cell] * cell_scale;
    int ih, iv, ic;
    if (c.plane == 0) { ic = k; ih = i; iv = j; }
    else if (c.plane == 1) { ic = i; iv = j; ih = k; }
    else { ic = j; iv = i; ih = k; }
    if (c.cut < ic || c.cut >= ic + cell_scale) continue;
    const int y0 = iv / target_scale - c.v0, x0 = ih / target_scale - c.h0;
    const int span = max(1, cell_scale / target_scale);
    for (int dy = 0; dy < span; ++dy) {
      const int y = y0 + dy;
      if (y < 0) continue;
      if (y >= c.v) break;
      for (int dx = 0; dx < span; ++dx) {
        const int x = x0 + dx;
        if (x < 0) continue;
        if (x >= c.h) break;
        atomicMax(&winner[(size_t)y * c.h + x], (int)cell);
      }
    }
  }
}

// exportSectionCutPlotToiletPaperZ(std::vector<DIADGridStruct>&) (lines 789-993):
// the cells of one level near the cylinder of radius r claim the pixels of
// their unrolled footprint (rows wrap around the circumference; largest cell
// index wins), positions on the finest level's lattice.
__global__ void k_cyl_claim(const int* __restrict__ ijk, int64_t n, int cell_scale, int target_scale, CutParams c,
                            int* __restrict__ winner) {
  const float PI = 3.14159265359f;
  for (int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < n;
       cell += (int64_t)gridDim.x * blockDim.x) {
    const int i = ijk[cell] * cell_scale, j = ijk[n + cell] * cell_scale, k = ijk[2 * n + cell] * cell_scale;
    const float xc = ((float)i + 0.5f * cell_scale - (float)c.io) * c.fres;
    const float yc = ((float)j + 0.5f * cell_scale - (float)c.jo) * c.fres;
    const float dist = sqrtf(xc * xc + yc * yc);
    const float half = cell_scale * c.fres * 0.5f;
    if (dist < c.r - half * 1.5f || dist > c.r + half * 1.5f) continue;
    float theta = atan2f(yc, xc);
    if (theta < 0.f) theta += 2.0f * PI;
    const int fv = (int)(theta * c.r / c.fres);
    const int span = max(1, cell_scale / target_scale);
    const int y0 = (fv - (span * target_scale) / 2) / target_scale;  // C division, towards zero
    const int x0 = k / target_scale;
    for (int dy = 0; dy < span; ++dy) {
      int y = (y0 + dy) % c.v;
      if (y < 0) y += c.v;
      for (int dx = 0; dx < span; ++dx) {
        const int x = x0 + dx;
        if (x >= c.h) continue;
        atomicMax(&winner[(size_t)y * c.h + x], (int)cell);
      }
    }
  }
}

// the claimed pixels take the winning cell's fields (converted with its level)
// `moving` (n bytes, or null) marks the cells flagged bounceback here that the
// reference's marker leaves open: moving walls, and forced cells in the plots
// (its flow reports count forced cells as blocked, include/flowReportFunctions.h:163).
__global__ void k_cut_paint(const int* __restrict__ winner, const float* __restrict__ macro,
                            const uint8_t* __restrict__ flags, const uint8_t* __restrict__ moving, int64_t n,
                            CutParams c, int level, float* __restrict__ out) {
  const size_t pix = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (pix >= (size_t)c.h * c.v) return;
  const int cell = winner[pix];
  if (cell < 0) return;
  const float rho = macro[cell], ux = macro[n + cell] * c.vel_scale, uy = macro[2 * n + cell] * c.vel_scale,
              uz = macro[3 * n + cell] * c.vel_scale;
  float* o = out + 6 * pix;
  if (c.cyl) {
    // p, uz, tangential and radial velocity at the row's angle (lines 971-978)
    const float theta = ((float)(pix / c.h) * (float)c.tscale * c.fres) / c.r;
    const float sn = sinf(theta), cs = cosf(theta);
    o[0] = (rho - 1.f) * c.p_scale;
    o[1] = uz;
    o[2] = -ux * sn + uy * cs;
    o[3] = ux * cs + uy * sn;
  } else if (c.raw) {
    o[0] = rho;
    o[1] = ux;
    o[2] = uy;
    o[3] = uz;
  } else {
    float uh, uv, un;
    plane_components(c.plane, ux, uy, uz, uh, uv, un);
    o[0] = (rho - 1.f) * c.p_scale;
    o[1] = uh;
    o[2] = uv;
    o[3] = un;
  }
  o[4] = (flags && (flags[cell] & 7) == KIND_BOUNCEBACK && !(moving && moving[cell])) ? 1.f : 0.f;
  o[5] = (float)level;
}

__global__ void k_mark_list(const int* __restrict__ cells, int64_t m, uint8_t* __restrict__ mask) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < m) mask[cells[t]] = 1;
}

__global__ void k_cut_fill(float* __restrict__ out, size_t pixels, int raw) {
  const size_t pix = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (pix >= pixels) return;
  float* o = out + 6 * pix;  // rho = 1, u = 0, marker 1, gridID 0 (lines 414-419)
  o[0] = raw ? 1.f : 0.f; o[1] = 0.f; o[2] = 0.f; o[3] = 0.f; o[4] = 1.f; o[5] = 0.f;
}

void plane_dims(int plane, int nx, int ny, int nz, int& h, int& v) {
  if (plane == 0) { h = nx; v = ny; }
  else if (plane == 1) { h = nz; v = ny; }
  else { h = nz; v = nx; }
}

}  // namespace
}  // namespace psl

using namespace psl;

extern "C" {

static int grid_cut(psl_grid* g, int plane, int cut, float rho_nominal, float soundspeed, float* host_out, int* dims,
                    bool with_scalar) {
  PSL_CHECK_ARG(g && dims && plane >= 0 && plane <= 2, "bad section-cut arguments");
  PSL_CHECK_ARG(!with_scalar || g->scalar_T, "scalar transport not enabled");
  const psl_info& I = g->info;
  int h, v;
  plane_dims(plane, I.cellCountX, I.cellCountY, I.cellCountZ, h, v);
  const int depth = plane == 0 ? I.cellCountZ : (plane == 1 ? I.cellCountX : I.cellCountY);
  PSL_CHECK_ARG(cut >= 0 && cut < depth, "cut index outside the grid");
  PSL_CHECK_ARG(I.dtPhys > 0.f, "dtPhys must be positive for physical velocities");
  dims[0] = v;
  dims[1] = h;
  if (!host_out) return PSL_OK;
  const size_t n = (size_t)I.cellCount;
  const int comps = with_scalar ? 6 : 5;
  float* macro = nullptr;
  float* out = nullptr;
  int rc = PSL_OK;
  if (hipMalloc(&macro, (with_scalar ? 5 : 4) * n * sizeof(float)) != hipSuccess ||
      hipMalloc(&out, comps * (size_t)h * v * sizeof(float)) != hipSuccess) {
    set_error("section cut: allocation failed");
    rc = PSL_E_NOMEM;
  }
  if (rc == PSL_OK) rc = psl_get_macroscopic_device(g, macro);
  if (rc == PSL_OK && with_scalar) rc = scalar_field_device(g, macro + 4 * n);
  if (rc == PSL_OK) {
    CutParams c{plane, cut, h, v, (I.res / 1000.f) / I.dtPhys, rho_nominal * soundspeed * soundspeed};
    dim3 grid((h + 255) / 256, v);
    k_cut_grid<<<grid, 256, 0, g->stream>>>(macro, g->flags, I.cellCountX, I.cellCountY, n, c, out,
                                            with_scalar ? macro + 4 * n : nullptr);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(host_out, out, comps * (size_t)h * v * sizeof(float), hipMemcpyDeviceToHost, g->stream) !=
            hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess) {
      set_error("section cut: kernel or copy failed");
      rc = PSL_E_HIP;
    }
  }
  if (macro) (void)hipFree(macro);
  if (out) (void)hipFree(out);
  return rc;
}

int psl_volume_export(psl_grid* g, float rho_nominal, float soundspeed, float* host_out) {
  PSL_CHECK_ARG(g && host_out, "null argument");
  const psl_info& I = g->info;
  PSL_CHECK_ARG(I.dtPhys > 0.f, "dtPhys must be positive for physical velocities");
  const size_t n = (size_t)I.cellCount;
  float* macro = nullptr;
  float* out = nullptr;
  int rc = PSL_OK;
  if (hipMalloc(&macro, 4 * n * sizeof(float)) != hipSuccess || hipMalloc(&out, 5 * n * sizeof(float)) != hipSuccess) {
    set_error("volume export: allocation failed");
    rc = PSL_E_NOMEM;
  }
  if (rc == PSL_OK) rc = psl_get_macroscopic_device(g, macro);
  if (rc == PSL_OK) {
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 1u << 20);
    k_volume<<<nb, 256, 0, g->stream>>>(macro, g->flags, I.cellCountX, I.cellCountY, I.cellCountZ,
                                        (I.res / 1000.f) / I.dtPhys, rho_nominal * soundspeed * soundspeed, out);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(host_out, out, 5 * n * sizeof(float), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess) {
      set_error("volume export: kernel or copy failed");
      rc = PSL_E_HIP;
    }
  }
  if (macro) (void)hipFree(macro);
  if (out) (void)hipFree(out);
  return rc;
}

int psl_cylinder_cut(psl_grid* g, float r, float rho_nominal, float soundspeed, float* host_out, int* dims) {
  PSL_CHECK_ARG(g && dims && r > 0.f, "bad cylinder-cut arguments");
  const psl_info& I = g->info;
  PSL_CHECK_ARG(I.dtPhys > 0.f, "dtPhys must be positive for physical velocities");
  const float PI = 3.14159265359f;
  const int rows = (int)((2.0f * PI * r) / I.res);
  PSL_CHECK_ARG(rows > 0, "radius below one cell");
  dims[0] = rows;
  dims[1] = I.cellCountZ;
  // pixel rows -> cell columns, getIJKCellIndexFromXYZ (include/cellFunctions.h:32-37)
  std::vector<int> col((size_t)rows);
  std::vector<float2> rot((size_t)rows);
  for (int v = 0; v < rows; ++v) {
    const float theta = ((float)v * I.res) / r;
    const float x = r * cosf(theta), y = r * sinf(theta);
    const float fi = (x - I.ox) / I.res + 0.5f, fj = (y - I.oy) / I.res + 0.5f;
    const int i = (int)fi, j = (int)fj;
    PSL_CHECK_ARG(fi > -1.f && fj > -1.f && i < I.cellCountX && j < I.cellCountY, "cylinder leaves the grid");
    col[v] = j * I.cellCountX + i;
    rot[v] = make_float2(cosf(theta), sinf(theta));
  }
  if (!host_out) return PSL_OK;
  const size_t n = (size_t)I.cellCount;
  float* macro = nullptr;
  float* out = nullptr;
  int* dcol = nullptr;
  float2* drot = nullptr;
  int rc = PSL_OK;
  if (hipMalloc(&macro, 4 * n * sizeof(float)) != hipSuccess ||
      hipMalloc(&out, 5 * (size_t)rows * I.cellCountZ * sizeof(float)) != hipSuccess ||
      hipMalloc(&dcol, (size_t)rows * sizeof(int)) != hipSuccess ||
      hipMalloc(&drot, (size_t)rows * sizeof(float2)) != hipSuccess) {
    set_error("cylinder cut: allocation failed");
    rc = PSL_E_NOMEM;
  }
  if (rc == PSL_OK && (hipMemcpy(dcol, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                       hipMemcpy(drot, rot.data(), rot.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("cylinder cut: upload failed");
    rc = PSL_E_HIP;
  }
  if (rc == PSL_OK) rc = psl_get_macroscopic_device(g, macro);
  if (rc == PSL_OK) {
    dim3 grid((I.cellCountZ + 255) / 256, rows);
    k_cut_cyl<<<grid, 256, 0, g->stream>>>(macro, g->flags, n, (size_t)I.cellCountX * I.cellCountY, I.cellCountZ,
                                           dcol, drot, rows, (I.res / 1000.f) / I.dtPhys,
                                           rho_nominal * soundspeed * soundspeed, out);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(host_out, out, 5 * (size_t)rows * I.cellCountZ * sizeof(float), hipMemcpyDeviceToHost,
                       g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess) {
      set_error("cylinder cut: kernel or copy failed");
      rc = PSL_E_HIP;
    }
  }
  for (void* a : {(void*)macro, (void*)out, (void*)dcol, (void*)drot})
    if (a) (void)hipFree(a);
  return rc;
}

int psl_section_cut(psl_grid* g, int plane, int cut, float rho_nominal, float soundspeed, float* host_out,
                    int* dims) {
  return grid_cut(g, plane, cut, rho_nominal, soundspeed, host_out, dims, false);
}

int psl_scalar_section_cut(psl_grid* g, int plane, int cut, float rho_nominal, float soundspeed, float* host_out,
                           int* dims) {
  return grid_cut(g, plane, cut, rho_nominal, soundspeed, host_out, dims, true);
}

}  // extern "C"

namespace psl {
namespace {
// Paints the levels onto a window of h x v pixels at 1 / target_scale of the
// finest resolution, starting at pixel (h0, v0); out: 6 floats per pixel.
int paint_levels(psl_diad* const* levels, int nlevels, const CutParams& base, int target_scale, float* host_out) {
  const size_t pixels = (size_t)base.h * base.v;
  hipStream_t st = levels[0]->stream;
  float* out = nullptr;
  int* winner = nullptr;
  int rc = PSL_OK;
  if (hipMalloc(&out, 6 * std::max<size_t>(pixels, 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&winner, std::max<size_t>(pixels, 1) * sizeof(int)) != hipSuccess) {
    set_error("section cut: allocation failed");
    rc = PSL_E_NOMEM;
  }
  if (rc == PSL_OK && pixels) {
    k_cut_fill<<<(unsigned)((pixels + 255) / 256), 256, 0, st>>>(out, pixels, base.raw);
    if (hipGetLastError() != hipSuccess) rc = PSL_E_HIP;
  }
  for (int l = 0; l < nlevels && rc == PSL_OK && pixels; ++l) {
    psl_diad* d = levels[l];
    const psl_info& I = d->info;
    float* macro = nullptr;
    uint8_t* moving = nullptr;
    const bool open_forced = !base.raw && d->n_forced;
    if (hipMalloc(&macro, 4 * (size_t)d->n * sizeof(float)) != hipSuccess ||
        ((d->n_moving || open_forced) && hipMalloc(&moving, (size_t)d->n) != hipSuccess)) {
      if (macro) (void)hipFree(macro);
      set_error("section cut: allocation failed");
      rc = PSL_E_NOMEM;
      break;
    }
    if (hipStreamSynchronize(d->stream) != hipSuccess) rc = PSL_E_HIP;
    if (rc == PSL_OK) rc = diad_macro_device(d, macro);
    if (rc == PSL_OK && hipStreamSynchronize(d->stream) != hipSuccess) rc = PSL_E_HIP;
    if (rc == PSL_OK) {
      CutParams c = base;
      c.vel_scale = base.raw ? 1.f : (I.res / 1000.f) / I.dtPhys;
      const int cell_scale = 1 << (nlevels - l - 1);
      if (hipMemsetAsync(winner, 0xff, pixels * sizeof(int), st) != hipSuccess) rc = PSL_E_HIP;
      if (moving) {
        if (hipMemsetAsync(moving, 0, (size_t)d->n, st) != hipSuccess) rc = PSL_E_HIP;
        if (d->n_moving)
          k_mark_list<<<(unsigned)((d->n_moving + 255) / 256), 256, 0, st>>>(d->moving, d->n_moving, moving);
        if (open_forced)
          k_mark_list<<<(unsigned)((d->n_forced + 255) / 256), 256, 0, st>>>(d->forced, d->n_forced, moving);
      }
      const unsigned nb = (unsigned)std::min<int64_t>((d->n + 255) / 256, 1 << 20);
      if (c.cyl) k_cyl_claim<<<nb, 256, 0, st>>>(d->ijk, d->n, cell_scale, target_scale, c, winner);
      else k_cut_claim<<<nb, 256, 0, st>>>(d->ijk, d->n, cell_scale, target_scale, c, winner);
      k_cut_paint<<<(unsigned)((pixels + 255) / 256), 256, 0, st>>>(winner, macro, d->flags, moving, d->n, c, l,
                                                                    out);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = PSL_E_HIP;
    }
    (void)hipFree(macro);
    if (moving) (void)hipFree(moving);
    if (rc == PSL_E_HIP) set_error("section cut: kernel failed");
  }
  if (rc == PSL_OK && pixels &&
      (hipMemcpy(host_out, out, 6 * pixels * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)) {
    set_error("section cut: copy failed");
    rc = PSL_E_HIP;
  }
  if (out) (void)hipFree(out);
  if (winner) (void)hipFree(winner);
  return rc;
}
}  // namespace
}  // namespace psl

extern "C" {

int psl_diad_section_cut(psl_diad* const* levels, int nlevels, int plane, int cut, float rho_nominal,
                         float soundspeed, int64_t max_pixels, float* host_out, int* dims3) {
  PSL_CHECK_ARG(levels && nlevels >= 1 && dims3 && plane >= 0 && plane <= 2, "bad section-cut arguments");
  for (int l = 0; l < nlevels; ++l) PSL_CHECK_ARG(levels[l] && levels[l]->info.dtPhys > 0.f, "bad level");
  // the finest level whose plane fits (lines 380-399)
  int target = nlevels, h = 0, v = 0;
  while (target > 0) {
    const psl_info& I = levels[target - 1]->info;
    plane_dims(plane, I.cellCountX, I.cellCountY, I.cellCountZ, h, v);
    if ((int64_t)h * v < max_pixels || target == 1) break;
    --target;
  }
  dims3[0] = v;
  dims3[1] = h;
  dims3[2] = target;
  if (!host_out) return PSL_OK;
  CutParams c{plane, cut, h, v, 1.f, rho_nominal * soundspeed * soundspeed};
  return paint_levels(levels, nlevels, c, 1 << (nlevels - target), host_out);
}

// exportSectionCutPlotToiletPaperZ over DIAD levels (lines 789-993).
int psl_diad_cylinder_cut(psl_diad* const* levels, int nlevels, float r, float rho_nominal, float soundspeed,
                          int64_t max_pixels, float* host_out, int* dims3) {
  PSL_CHECK_ARG(levels && nlevels >= 1 && dims3 && r > 0.f, "bad cylinder-cut arguments");
  for (int l = 0; l < nlevels; ++l) PSL_CHECK_ARG(levels[l] && levels[l]->info.dtPhys > 0.f, "bad level");
  const float PI = 3.14159265359f;
  int target = nlevels, h = 0, v = 0;
  while (target > 1) {
    const psl_info& I = levels[target - 1]->info;
    h = I.cellCountZ;
    v = (int)((2.0f * PI * r) / I.res);
    if ((int64_t)h * v < max_pixels) break;
    --target;
  }
  const psl_info& T = levels[target - 1]->info;
  h = T.cellCountZ;
  v = (int)((2.0f * PI * r) / T.res);
  PSL_CHECK_ARG(v > 0, "radius below one cell");
  dims3[0] = v;
  dims3[1] = h;
  dims3[2] = target;
  if (!host_out) return PSL_OK;
  const psl_info& F = levels[nlevels - 1]->info;
  CutParams c{0, 0, h, v, 1.f, rho_nominal * soundspeed * soundspeed};
  c.cyl = 1;
  c.tscale = 1 << (nlevels - target);
  c.r = r;
  c.fres = F.res;
  c.io = (int)((0.f - F.ox) / F.res + 0.5f);  // getIJKCellIndexFromXYZ(0, 0, 0) on the finest level
  c.jo = (int)((0.f - F.oy) / F.res + 0.5f);
  return paint_levels(levels, nlevels, c, 1 << (nlevels - target), host_out);
}

// getFlowReportXY/ZY/ZX (include/flowReportFunctions.h:2-263): the plane cut
// clipped to bounds6 = (xmin, ymin, zmin, xmax, ymax, zmax) in physical units,
// then over its non-solid pixels the mean ux, uy, uz, rho (lattice units) and
// the area in mm^2 (pixels x res^2 of the level the window was painted at):
// out5 = ux, uy, uz, rho, area.  Sums in FP64.
int psl_diad_flow_report(psl_diad* const* levels, int nlevels, int plane, int cut, const float* bounds6,
                         double* out5) {
  PSL_CHECK_ARG(levels && nlevels >= 1 && bounds6 && out5 && plane >= 0 && plane <= 2, "bad flow-report arguments");
  for (int l = 0; l < nlevels; ++l) PSL_CHECK_ARG(levels[l], "bad level");
  const psl_info& F = levels[nlevels - 1]->info;
  const int iMin = std::max(0, (int)std::floor((bounds6[0] - F.ox) / F.res));
  const int iMax = std::min(F.cellCountX, (int)std::ceil((bounds6[3] - F.ox) / F.res));
  const int jMin = std::max(0, (int)std::floor((bounds6[1] - F.oy) / F.res));
  const int jMax = std::min(F.cellCountY, (int)std::ceil((bounds6[4] - F.oy) / F.res));
  const int kMin = std::max(0, (int)std::floor((bounds6[2] - F.oz) / F.res));
  const int kMax = std::min(F.cellCountZ, (int)std::ceil((bounds6[5] - F.oz) / F.res));
  int h0f, h1f, v0f, v1f;
  if (plane == 0) { h0f = iMin; h1f = iMax; v0f = jMin; v1f = jMax; }
  else if (plane == 1) { h0f = kMin; h1f = kMax; v0f = jMin; v1f = jMax; }
  else { h0f = kMin; h1f = kMax; v0f = iMin; v1f = iMax; }
  int target = nlevels, scale = 1, h0 = 0, v0 = 0, w = 0, ht = 0;
  while (target > 0) {  // lines 36-58
    scale = 1 << (nlevels - target);
    h0 = h0f / scale;
    v0 = v0f / scale;
    w = std::max(0, (h1f + scale - 1) / scale - h0);
    ht = std::max(0, (v1f + scale - 1) / scale - v0);
    if ((int64_t)w * ht < 20000000 || target == 1) break;
    --target;
  }
  std::vector<float> px(6 * (size_t)w * ht);
  CutParams c{plane, cut, w, ht, 1.f, 1.f};
  c.h0 = h0;
  c.v0 = v0;
  c.raw = 1;
  const int rc = paint_levels(levels, nlevels, c, scale, px.data());
  if (rc != PSL_OK) return rc;
  int64_t cells = 0;
  double sum[4] = {0, 0, 0, 0};
  for (size_t p = 0; p < (size_t)w * ht; ++p) {
    const float* o = &px[6 * p];
    if (o[4] != 0.f) continue;
    ++cells;
    sum[0] += o[1];
    sum[1] += o[2];
    sum[2] += o[3];
    sum[3] += (double)o[0] - 1.0;
  }
  if (cells > 0) {
    const double res = levels[target - 1]->info.res;
    out5[0] = sum[0] / (double)cells;
    out5[1] = sum[1] / (double)cells;
    out5[2] = sum[2] / (double)cells;
    out5[3] = sum[3] / (double)cells + 1.0;
    out5[4] = (double)cells * res * res;
  } else {
    out5[0] = out5[1] = out5[2] = 0.0;
    out5[3] = 1.0;
    out5[4] = 0.0;
  }
  return PSL_OK;
}

// sim_data.bin (exportSectionCutPlot.h:64-86, 543-569): int header {plotNumber,
// rows, cols, components} and rows x cols x components floats.
int psl_write_sim_data(const char* path, int plot_number, int rows, int cols, int components, const float* data) {
  PSL_CHECK_ARG(path && data && rows > 0 && cols > 0 && components > 0, "bad sim_data arguments");
  FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("cannot open ") + path);
    return PSL_E_IO;
  }
  const int header[4] = {plot_number, rows, cols, components};
  const size_t count = (size_t)rows * cols * components;
  const bool ok = std::fwrite(header, sizeof(int), 4, fp) == 4 && std::fwrite(data, sizeof(float), count, fp) == count;
  std::fclose(fp);
  if (!ok) {
    set_error(std::string("short write to ") + path);
    return PSL_E_IO;
  }
  return PSL_OK;
}

// The 3-D sim_data.bin of export3DPlot (exportSectionCutPlot.h:640-670): int
// plot number, nx, ny, nz, components, then the floats.
int psl_write_volume_data(const char* path, int plot_number, int nx, int ny, int nz, int components,
                          const float* data) {
  PSL_CHECK_ARG(path && data && nx > 0 && ny > 0 && nz > 0 && components > 0, "bad volume arguments");
  FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("cannot open ") + path);
    return PSL_E_IO;
  }
  const int header[5] = {plot_number, nx, ny, nz, components};
  const size_t count = (size_t)nx * ny * nz * components;
  const bool ok = std::fwrite(header, sizeof(int), 5, fp) == 5 && std::fwrite(data, sizeof(float), count, fp) == count;
  std::fclose(fp);
  if (!ok) {
    set_error(std::string("short write to ") + path);
    return PSL_E_IO;
  }
  return PSL_OK;
}

// historyData.bin (exportSectionCutPlot.h:994-1001): int count and count floats.
int psl_write_history(const char* path, const float* values, int count) {
  PSL_CHECK_ARG(path && count >= 0 && (count == 0 || values), "bad history arguments");
  FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("cannot open ") + path);
    return PSL_E_IO;
  }
  const bool ok = std::fwrite(&count, sizeof(int), 1, fp) == 1 &&
                  std::fwrite(values, sizeof(float), (size_t)count, fp) == (size_t)count;
  std::fclose(fp);
  if (!ok) {
    set_error(std::string("short write to ") + path);
    return PSL_E_IO;
  }
  return PSL_OK;
}

}  // extern "C"
