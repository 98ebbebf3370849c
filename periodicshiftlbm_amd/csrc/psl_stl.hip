// STL input, watertightness check, ray-parity voxelizer and mask algebra.
//
// include/STLFunctions.h:3-479, 811-846.  The voxelizer casts +z rays through
// the cell centres of every (i, j) column on a lattice 100x finer than the
// grid, with triangle vertices on odd and rays on even integer coordinates so
// that no ray meets a vertex; the hit predicate is exact 64-bit integer
// arithmetic and the hit height is FP32.  Hits are collected per column with
// atomics, sorted per column (insertion sort, so the result does not depend
// on the atomic order) and the column is filled by alternating parity.
//
// This translation unit is compiled with -ffp-contract=off: the FP32 hit
// height and the lattice rounding must not be fused, so that masks are
// bit-identical to the CPU restatement.
#include "psl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>

// The shared headers end in contraction mode fast; from here on nothing is
// fused, whatever the command line says.
#pragma clang fp contract(off)

namespace psl {

__host__ __device__ inline bool ray_hit(long long ax, long long ay, long long bx, long long by, long long cx,
                                        long long cy) {
  const long long abx = bx - ax, aby = by - ay, bcx = cx - bx, bcy = cy - by, cax = ax - cx, cay = ay - cy;
  const long long area = abx * bcy - aby * bcx;
  if (area == 0) return false;
  const long long qz = area < 0 ? -1 : 1;
  const long long wab = qz * (abx * (-by) - aby * (-bx));
  const long long wbc = qz * (bcx * (-cy) - bcy * (-cx));
  const long long wca = qz * (cax * (-ay) - cay * (-ax));
  if (wab > 0 && wbc > 0 && wca > 0) return true;
  if (wab < 0 || wbc < 0 || wca < 0) return false;
  // ray on an edge: the edge belongs to the triangle on one fixed side
  if (wab == 0) {
    if (abx == 0) return cx < 0;
    if (bx > 0 && cy * bx > by * cx) return true;
    if (bx < 0 && cy * bx < by * cx) return true;
  }
  if (wbc == 0) {
    if (bcx == 0) return ax < 0;
    if (cx > 0 && ay * cx > cy * ax) return true;
    if (cx < 0 && ay * cx < cy * ax) return true;
  }
  if (wca == 0) {
    if (cax == 0) return bx < 0;
    if (ax > 0 && by * ax > ay * bx) return true;
    if (ax < 0 && by * ax < ay * bx) return true;
  }
  return false;
}

__host__ __device__ inline float ray_hit_z(float ax, float ay, float az, float bx, float by, float bz, float cx,
                                           float cy, float cz, float rx, float ry) {
  const float v1x = bx - ax, v1y = by - ay, v1z = bz - az;
  const float v2x = cx - ax, v2y = cy - ay, v2z = cz - az;
  const float nx = v1y * v2z - v1z * v2y;
  const float ny = v1z * v2x - v1x * v2z;
  const float nz = v1x * v2y - v1y * v2x;
  const float zmax = fmaxf(az, fmaxf(bz, cz)), zmin = fminf(az, fminf(bz, cz));
  float z;
  if (nz != 0.0f) {
    z = az - (nx * (rx - ax) + ny * (ry - ay)) / nz;
    if (z > zmax) z = zmax;
    else if (z < zmin) z = zmin;
  } else {
    z = 0.5f * (zmax + zmin);
  }
  return z;
}

struct VoxParams {
  float sox, soy, soz;  // STL origin
  float gox, goy, goz;  // grid origin
  float res;
  int nx, ny, nz;
};

struct TriLattice {
  float ax, ay, az, bx, by, bz, cx, cy, cz;
  long long ak, al, bk, bl, ck, cl;
  long long imin, imax, jmin, jmax;
};

__device__ inline TriLattice tri_lattice(const float* t, const VoxParams& p) {
  TriLattice L;
  // (vertex + STL origin) - grid origin, in that order (lines 300-305)
  L.ax = t[0] + p.sox - p.gox; L.ay = t[1] + p.soy - p.goy; L.az = t[2] + p.soz - p.goz;
  L.bx = t[3] + p.sox - p.gox; L.by = t[4] + p.soy - p.goy; L.bz = t[5] + p.soz - p.goz;
  L.cx = t[6] + p.sox - p.gox; L.cy = t[7] + p.soy - p.goy; L.cz = t[8] + p.soz - p.goz;
  const float scale = 50.0f / p.res;
  L.ak = (long long)roundf(L.ax * scale) * 2 + 1;
  L.al = (long long)roundf(L.ay * scale) * 2 + 1;
  L.bk = (long long)roundf(L.bx * scale) * 2 + 1;
  L.bl = (long long)roundf(L.by * scale) * 2 + 1;
  L.ck = (long long)roundf(L.cx * scale) * 2 + 1;
  L.cl = (long long)roundf(L.cy * scale) * 2 + 1;
  const long long XM = (long long)(p.nx - 1) * 100, YM = (long long)(p.ny - 1) * 100;
  const long long kmin = max(0LL, min(min(L.ak, L.bk), min(L.ck, XM)));
  const long long kmax = min(XM, max(max(L.ak, L.bk), max(L.ck, 0LL)));
  const long long lmin = max(0LL, min(min(L.al, L.bl), min(L.cl, YM)));
  const long long lmax = min(YM, max(max(L.al, L.bl), max(L.cl, 0LL)));
  L.imin = (kmin + 99) / 100;
  L.imax = kmax / 100;
  L.jmin = (lmin + 99) / 100;
  L.jmax = lmax / 100;
  return L;
}

// One wave per triangle: the lanes stride over the columns of its footprint
// (a large triangle covers hundreds of columns, which one thread would walk
// alone).  The per-column counts and the sorted layers do not depend on the
// order of the atomics.
constexpr int VOX_WAVES = 4;  // waves (triangles) per 256-thread block

// K14: hits per column
__global__ void k_count_hits(const float* __restrict__ tri, int ntri, VoxParams p, int* __restrict__ counter) {
  const int t = blockIdx.x * VOX_WAVES + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= ntri) return;
  const TriLattice L = tri_lattice(tri + 9 * (size_t)t, p);
  const long long w = L.imax - L.imin + 1, h = L.jmax - L.jmin + 1;
  if (w <= 0 || h <= 0) return;
  for (long long f = lane; f < w * h; f += 64) {
    const long long j = L.jmin + f / w, i = L.imin + f % w;
    const long long rk = i * 100, rl = j * 100;
    if (ray_hit(L.ak - rk, L.al - rl, L.bk - rk, L.bl - rl, L.ck - rk, L.cl - rl))
      atomicAdd(&counter[(size_t)j * p.nx + i], 1);
  }
}

__global__ void k_max_count(const int* __restrict__ counter, size_t n, int* out) {
  int m = 0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
    m = max(m, counter[c]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// K16: hit cell index per triangle and column
__global__ void k_hit_index(const float* __restrict__ tri, int ntri, VoxParams p, int* __restrict__ counter,
                            int* __restrict__ layers, int maxc) {
  const int t = blockIdx.x * VOX_WAVES + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= ntri) return;
  const TriLattice L = tri_lattice(tri + 9 * (size_t)t, p);
  const size_t ncol = (size_t)p.nx * p.ny;
  const long long fw = L.imax - L.imin + 1, fh = L.jmax - L.jmin + 1;
  if (fw <= 0 || fh <= 0) return;
  for (long long f = lane; f < fw * fh; f += 64) {
    const long long j = L.jmin + f / fw, i = L.imin + f % fw;
    const long long rk = i * 100, rl = j * 100;
    if (!ray_hit(L.ak - rk, L.al - rl, L.bk - rk, L.bl - rl, L.ck - rk, L.cl - rl)) continue;
    const size_t col = (size_t)j * p.nx + i;
    const int w = atomicAdd(&counter[col], 1);
    const float rx = (int)i * p.res;
    const float ry = (int)j * p.res;
    const float rz = ray_hit_z(L.ax, L.ay, L.az, L.bx, L.by, L.bz, L.cx, L.cy, L.cz, rx, ry);
    int k = (int)fmaxf(0.0f, ceilf(rz / p.res));
    k = min(k, p.nz);
    if (w < maxc) layers[(size_t)w * ncol + col] = k;
  }
}

// K17: per-column insertion sort (the result does not depend on the atomic order)
__global__ void k_sort_columns(int* __restrict__ layers, int maxc, VoxParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= p.nx || j >= p.ny) return;
  const size_t ncol = (size_t)p.nx * p.ny;
  const size_t col = (size_t)j * p.nx + i;
  for (int layer = 1; layer < maxc; ++layer) {
    const int key = layers[(size_t)layer * ncol + col];
    int s = layer - 1;
    while (s >= 0 && layers[(size_t)s * ncol + col] > key) {
      layers[(size_t)(s + 1) * ncol + col] = layers[(size_t)s * ncol + col];
      --s;
    }
    layers[(size_t)(s + 1) * ncol + col] = key;
  }
}

// parity fill of whole columns (lines 371-389)
__global__ void k_fill_columns(const int* __restrict__ layers, int maxc, VoxParams p, int inside_value,
                               uint8_t* __restrict__ mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= p.nx || j >= p.ny) return;
  const size_t ncol = (size_t)p.nx * p.ny;
  const size_t col = (size_t)j * p.nx + i;
  bool val = !inside_value;
  for (int iv = 0; iv <= maxc; ++iv) {
    int s0, e0;
    if (iv == 0) { s0 = 0; e0 = layers[col]; }
    else if (iv == maxc) { s0 = layers[(size_t)(maxc - 1) * ncol + col]; e0 = p.nz; }
    else { s0 = layers[(size_t)(iv - 1) * ncol + col]; e0 = layers[(size_t)iv * ncol + col]; }
    for (int k = s0; k < e0; ++k) mask[(size_t)k * ncol + col] = val ? 1 : 0;
    val = !val;
  }
}

// cell-list form (lines 449-477): cells given as packed (k, j, i) keys inside
// the nx x ny x nz box; sets out[c] = 1 where the cell lies inside
__global__ void k_mark_keys(const int* __restrict__ layers, int maxc, VoxParams p, const uint64_t* __restrict__ keys,
                            int64_t n, uint8_t* __restrict__ out) {
  const size_t ncol = (size_t)p.nx * p.ny;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[c];
    const int i = (int)(key & KEY_MASK), j = (int)((key >> KEY_BITS) & KEY_MASK), k = (int)(key >> (2 * KEY_BITS));
    const size_t col = (size_t)j * p.nx + i;
    bool val = false;  // starting outside
    for (int iv = 0; iv <= maxc; ++iv) {
      int s0, e0;
      if (iv == 0) { s0 = 0; e0 = layers[col]; }
      else if (iv == maxc) { s0 = layers[(size_t)(maxc - 1) * ncol + col]; e0 = p.nz; }
      else { s0 = layers[(size_t)(iv - 1) * ncol + col]; e0 = layers[(size_t)iv * ncol + col]; }
      if (k >= s0 && k < e0) {
        if (val) out[c] = 1;
        break;
      }
      val = !val;
    }
  }
}

// packed (k, j, i) keys of a DIAD grid's cells from its ijk[3][n] table
__global__ void k_pack_ijk(const int* __restrict__ ijk, int64_t n, uint64_t* __restrict__ keys) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    keys[c] = pack_key(ijk[c], ijk[n + c], ijk[2 * n + c]);
}

__global__ void k_fill_int(int* a, size_t n, int v) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) a[c] = v;
}

// K13: O(T^2) shared-edge count; triangles of the outer loop tile through LDS
__global__ __launch_bounds__(256) void k_edge_counts(const float* __restrict__ tri, int ntri, int* __restrict__ counts) {
  __shared__ float tile[256 * 9];
  const int t1 = blockIdx.x * 256 + threadIdx.x;
  float a[9];
  if (t1 < ntri)
    for (int c = 0; c < 9; ++c) a[c] = tri[9 * (size_t)t1 + c];
  int cnt[3] = {0, 0, 0};
  for (int base = 0; base < ntri; base += 256) {
    __syncthreads();
    for (int c = threadIdx.x; c < 256 * 9; c += 256) {
      const int t = base + c / 9;
      tile[c] = t < ntri ? tri[9 * (size_t)base + c] : NAN;
    }
    __syncthreads();
    if (t1 >= ntri) continue;
    const int lim = min(256, ntri - base);
    for (int u = 0; u < lim; ++u) {
      const float* b = tile + 9 * u;
      for (int e1 = 0; e1 < 3; ++e1) {
        const float* a0 = a + 3 * e1;
        const float* a1 = a + 3 * ((e1 + 1) % 3);
        for (int e2 = 0; e2 < 3; ++e2) {
          const float* b0 = b + 3 * e2;
          const float* b1 = b + 3 * ((e2 + 1) % 3);
          if (a0[0] == b0[0] && a0[1] == b0[1] && a0[2] == b0[2] && a1[0] == b1[0] && a1[1] == b1[1] && a1[2] == b1[2])
            cnt[e1]++;
          if (a0[0] == b1[0] && a0[1] == b1[1] && a0[2] == b1[2] && a1[0] == b0[0] && a1[1] == b0[1] && a1[2] == b0[2])
            cnt[e1]++;
        }
      }
    }
  }
  if (t1 < ntri)
    for (int e = 0; e < 3; ++e) counts[3 * (size_t)t1 + e] = cnt[e];
}

__global__ void k_mask_binop(const uint8_t* a, const uint8_t* b, uint8_t* out, size_t n, int op) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
    out[c] = op == 0 ? (uint8_t)(a[c] && b[c]) : (uint8_t)(a[c] || b[c]);
}
__global__ void k_mask_not(uint8_t* a, size_t n) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
    a[c] = (uint8_t)!a[c];
}

static int upload_stl(psl_stl* s) {
  if (s->d_tri) PSL_HIP_TRY(hipFree(s->d_tri));
  s->d_tri = nullptr;
  if (s->count == 0) return PSL_OK;
  PSL_HIP_TRY(hipMalloc(&s->d_tri, s->tri.size() * sizeof(float)));
  PSL_HIP_TRY(hipMemcpy(s->d_tri, s->tri.data(), s->tri.size() * sizeof(float), hipMemcpyHostToDevice));
  return PSL_OK;
}

static void stl_bounds(psl_stl* s) {
  s->xmin = s->ymin = s->zmin = std::numeric_limits<float>::max();
  s->xmax = s->ymax = s->zmax = std::numeric_limits<float>::lowest();
  for (int t = 0; t < s->count; ++t)
    for (int v = 0; v < 3; ++v) {
      const float* p = &s->tri[9 * (size_t)t + 3 * v];
      s->xmin = std::min(s->xmin, p[0]); s->xmax = std::max(s->xmax, p[0]);
      s->ymin = std::min(s->ymin, p[1]); s->ymax = std::max(s->ymax, p[1]);
      s->zmin = std::min(s->zmin, p[2]); s->zmax = std::max(s->zmax, p[2]);
    }
}

static unsigned grid1d(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 65535); }

VoxParams vox_params(const psl_stl* s, float gox, float goy, float goz, float res, int nx, int ny, int nz) {
  VoxParams p;
  p.sox = s->ox; p.soy = s->oy; p.soz = s->oz;
  p.gox = gox; p.goy = goy; p.goz = goz;
  p.res = res;
  p.nx = nx; p.ny = ny; p.nz = nz;
  return p;
}

// Sorted hit layers per column (maxc x nx*ny ints, unused layers = nz), the
// counter, hit-index and sort passes of include/STLFunctions.h:296-369.
int stl_layers(const psl_stl* s, const VoxParams& p, int** layers_out, int* maxc_out) {
  const size_t ncol = (size_t)p.nx * p.ny;
  int* counter = nullptr;
  int* dmax = nullptr;
  int* layers = nullptr;
  int rc = PSL_OK;
  int maxc = 1;
  auto fail = [&](const char* what) {
    set_error(std::string("voxelizer: ") + what);
    rc = PSL_E_HIP;
  };
  if (hipMalloc(&counter, ncol * sizeof(int)) != hipSuccess || hipMalloc(&dmax, sizeof(int)) != hipSuccess) fail("alloc");
  if (rc == PSL_OK) {
    (void)hipMemset(counter, 0, ncol * sizeof(int));
    (void)hipMemset(dmax, 0, sizeof(int));
    if (s->count > 0) k_count_hits<<<(s->count + VOX_WAVES - 1) / VOX_WAVES, 256>>>(s->d_tri, s->count, p, counter);
    k_max_count<<<grid1d(ncol), 256>>>(counter, ncol, dmax);
    if (hipGetLastError() != hipSuccess || hipMemcpy(&maxc, dmax, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      fail("count");
    if (maxc < 1) maxc = 1;
  }
  if (rc == PSL_OK && hipMalloc(&layers, (size_t)maxc * ncol * sizeof(int)) != hipSuccess) fail("alloc layers");
  if (rc == PSL_OK) {
    k_fill_int<<<grid1d((size_t)maxc * ncol), 256>>>(layers, (size_t)maxc * ncol, p.nz);
    (void)hipMemset(counter, 0, ncol * sizeof(int));
    if (s->count > 0)
      k_hit_index<<<(s->count + VOX_WAVES - 1) / VOX_WAVES, 256>>>(s->d_tri, s->count, p, counter, layers, maxc);
    dim3 g2((p.nx + 255) / 256, p.ny);
    k_sort_columns<<<g2, 256>>>(layers, maxc, p);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) fail("hits");
  }
  if (counter) (void)hipFree(counter);
  if (dmax) (void)hipFree(dmax);
  if (rc != PSL_OK && layers) {
    (void)hipFree(layers);
    layers = nullptr;
  }
  *layers_out = layers;
  *maxc_out = maxc;
  return rc;
}

int stl_mark_keys(const psl_stl* s, float gox, float goy, float goz, float res, int nx, int ny, int nz,
                  const uint64_t* keys, int64_t n, uint8_t* out) {
  const VoxParams p = vox_params(s, gox, goy, goz, res, nx, ny, nz);
  int* layers = nullptr;
  int maxc = 1;
  int rc = stl_layers(s, p, &layers, &maxc);
  if (rc == PSL_OK && n > 0) {
    k_mark_keys<<<grid1d((size_t)n), 256>>>(layers, maxc, p, keys, n, out);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      set_error("voxelizer: cell list");
      rc = PSL_E_HIP;
    }
  }
  if (layers) (void)hipFree(layers);
  return rc;
}

}  // namespace psl

using namespace psl;

extern "C" {

int psl_stl_from_triangles(const float* tri9, int n, psl_stl** out) {
  PSL_CHECK_ARG(out && n >= 0 && (tri9 || n == 0), "bad triangle list");
  psl_stl* s = new psl_stl();
  s->count = n;
  s->tri.assign(tri9, tri9 + 9 * (size_t)n);
  stl_bounds(s);
  int rc = upload_stl(s);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return PSL_OK;
}

// Binary STL: 80-byte header, uint32 count, then per triangle a normal
// (3 floats, skipped), three vertices (9 floats) and a uint16 attribute.
int psl_stl_read(const char* path, psl_stl** out) {
  PSL_CHECK_ARG(path && out, "null argument");
  std::ifstream file(path, std::ios::binary);
  if (!file.is_open()) {
    set_error(std::string("failed to open STL file ") + path);
    return PSL_E_IO;
  }
  char header[80];
  uint32_t count = 0;
  file.read(header, 80);
  file.read(reinterpret_cast<char*>(&count), 4);
  if (!file) {
    set_error("truncated STL header");
    return PSL_E_IO;
  }
  // the header's count must match the file (84 + 50 bytes per triangle)
  // before anything is allocated for it
  file.seekg(0, std::ios::end);
  const std::streamoff size = file.tellg();
  if (count > (uint32_t)std::numeric_limits<int>::max() || size != (std::streamoff)(84 + 50 * (uint64_t)count)) {
    set_error("STL file " + std::string(path) + ": " + std::to_string(count) + " triangles in the header, " +
              std::to_string((long long)size) + " bytes on disk (binary STL needs 84 + 50 per triangle)");
    return PSL_E_IO;
  }
  file.seekg(84, std::ios::beg);
  std::vector<float> tri;
  try {
    tri.resize(9 * (size_t)count);
  } catch (const std::bad_alloc&) {
    set_error("out of host memory for the STL triangles");
    return PSL_E_NOMEM;
  }
  for (uint32_t t = 0; t < count; ++t) {
    char rec[50];
    file.read(rec, 50);
    if (!file) {
      set_error("truncated STL triangle records");
      return PSL_E_IO;
    }
    memcpy(&tri[9 * (size_t)t], rec + 12, 36);
  }
  return psl_stl_from_triangles(tri.data(), (int)count, out);
}

int psl_stl_destroy(psl_stl* s) {
  if (!s) return PSL_OK;
  if (s->d_tri) (void)hipFree(s->d_tri);
  delete s;
  return PSL_OK;
}

int psl_stl_info(const psl_stl* s, int* count, float* b6) {
  PSL_CHECK_ARG(s, "null stl");
  if (count) *count = s->count;
  if (b6) {
    b6[0] = s->xmin; b6[1] = s->ymin; b6[2] = s->zmin;
    b6[3] = s->xmax; b6[4] = s->ymax; b6[5] = s->zmax;
  }
  return PSL_OK;
}

int psl_stl_set_origin(psl_stl* s, float ox, float oy, float oz) {
  PSL_CHECK_ARG(s, "null stl");
  s->ox = ox; s->oy = oy; s->oz = oz;
  return PSL_OK;
}

// rotateSTLAlongX / Y / Z (include/STLFunctions.h:698-809): every vertex
// rotated by `radians` about the axis through the STL's own origin,
// right-handed, in FP32 with sinf / cosf of the angle; bounds follow.
int psl_stl_rotate(psl_stl* s, int axis, float radians) {
  PSL_CHECK_ARG(s && axis >= 0 && axis <= 2, "bad rotation");
  const float sn = sinf(radians), cs = cosf(radians);
  const int a = axis == 0 ? 1 : axis == 1 ? 2 : 0;  // the pair (a, b) turns a -> b
  const int b = axis == 0 ? 2 : axis == 1 ? 0 : 1;
  for (size_t v = 0; v < s->tri.size(); v += 3) {
    const float pa = s->tri[v + a], pb = s->tri[v + b];
    s->tri[v + a] = pa * cs - pb * sn;
    s->tri[v + b] = pa * sn + pb * cs;
  }
  stl_bounds(s);
  return upload_stl(s);
}

// ApplyMarkersInsideSTL (include/STLFunctions.h:482-696) on a DIAD grid's
// cells: out[c] = 1 inside, 0 outside (the reference writes the value and its
// complement), host array of the grid's n cells.
int psl_diad_markers_inside_stl(const psl_diad* d, const psl_stl* s, uint8_t* out) {
  PSL_CHECK_ARG(d && s && out, "null argument");
  if (d->n == 0) return PSL_OK;
  uint64_t* keys = nullptr;
  uint8_t* m = nullptr;
  int rc = PSL_OK;
  if (hipMalloc(&keys, (size_t)d->n * sizeof(uint64_t)) != hipSuccess || hipMalloc(&m, (size_t)d->n) != hipSuccess ||
      hipMemset(m, 0, (size_t)d->n) != hipSuccess) {
    set_error("inside markers: allocation");
    rc = PSL_E_HIP;
  }
  if (rc == PSL_OK) {
    k_pack_ijk<<<grid1d((size_t)d->n), 256>>>(d->ijk, d->n, keys);
    if (hipGetLastError() != hipSuccess) {
      set_error("inside markers: key launch");
      rc = PSL_E_HIP;
    }
  }
  const psl_info& I = d->info;
  if (rc == PSL_OK) rc = stl_mark_keys(s, I.ox, I.oy, I.oz, I.res, I.cellCountX, I.cellCountY, I.cellCountZ, keys, d->n, m);
  if (rc == PSL_OK && hipMemcpy(out, m, (size_t)d->n, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("inside markers: copy");
    rc = PSL_E_HIP;
  }
  if (keys) (void)hipFree(keys);
  if (m) (void)hipFree(m);
  return rc;
}

int psl_stl_get_triangles(const psl_stl* s, float* tri9) {
  PSL_CHECK_ARG(s && tri9, "null argument");
  memcpy(tri9, s->tri.data(), s->tri.size() * sizeof(float));
  return PSL_OK;
}

int psl_stl_check_edges(const psl_stl* s, int* bad, int* counts) {
  PSL_CHECK_ARG(s && bad, "null argument");
  *bad = 0;
  if (s->count == 0) return PSL_OK;
  int* d = nullptr;
  PSL_HIP_TRY(hipMalloc(&d, 3 * (size_t)s->count * sizeof(int)));
  k_edge_counts<<<(s->count + 255) / 256, 256>>>(s->d_tri, s->count, d);
  hipError_t e = hipGetLastError();
  std::vector<int> h(3 * (size_t)s->count);
  if (e == hipSuccess) e = hipMemcpy(h.data(), d, h.size() * sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) {
    set_error(std::string("edge check: ") + hipGetErrorString(e));
    return PSL_E_HIP;
  }
  int b = 0;
  for (int t = 0; t < s->count; ++t)
    if (h[3 * t] != 2 || h[3 * t + 1] != 2 || h[3 * t + 2] != 2) ++b;
  *bad = b;
  if (counts) memcpy(counts, h.data(), h.size() * sizeof(int));
  return PSL_OK;
}

int psl_mask_create(int64_t n, int value, psl_mask** out) {
  PSL_CHECK_ARG(out && n > 0, "bad mask size");
  psl_mask* m = new psl_mask();
  m->n = n;
  if (hipMalloc(&m->d, (size_t)n) != hipSuccess || hipMemset(m->d, value ? 1 : 0, (size_t)n) != hipSuccess) {
    set_error("mask allocation failed");
    if (m->d) (void)hipFree(m->d);
    delete m;
    return PSL_E_NOMEM;
  }
  *out = m;
  return PSL_OK;
}

int psl_mask_destroy(psl_mask* m) {
  if (!m) return PSL_OK;
  if (m->d) (void)hipFree(m->d);
  delete m;
  return PSL_OK;
}

int psl_mask_get(const psl_mask* m, uint8_t* host) {
  PSL_CHECK_ARG(m && host, "null argument");
  PSL_HIP_TRY(hipMemcpy(host, m->d, (size_t)m->n, hipMemcpyDeviceToHost));
  return PSL_OK;
}

int psl_mask_set(psl_mask* m, const uint8_t* host) {
  PSL_CHECK_ARG(m && host, "null argument");
  PSL_HIP_TRY(hipMemcpy(m->d, host, (size_t)m->n, hipMemcpyHostToDevice));
  return PSL_OK;
}

int psl_mask_and(const psl_mask* a, const psl_mask* b, psl_mask* out) {
  PSL_CHECK_ARG(a && b && out && a->n == b->n && a->n == out->n, "mask sizes differ");
  k_mask_binop<<<grid1d(a->n), 256>>>(a->d, b->d, out->d, (size_t)a->n, 0);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

int psl_mask_or(const psl_mask* a, const psl_mask* b, psl_mask* out) {
  PSL_CHECK_ARG(a && b && out && a->n == b->n && a->n == out->n, "mask sizes differ");
  k_mask_binop<<<grid1d(a->n), 256>>>(a->d, b->d, out->d, (size_t)a->n, 1);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

int psl_mask_not(psl_mask* m) {
  PSL_CHECK_ARG(m, "null mask");
  k_mask_not<<<grid1d(m->n), 256>>>(m->d, (size_t)m->n);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

int psl_markers_inside_stl(psl_mask* mask, const psl_stl* s, int inside_value, const psl_info* info) {
  PSL_CHECK_ARG(mask && s && info, "null argument");
  const int nx = info->cellCountX, ny = info->cellCountY, nz = info->cellCountZ;
  PSL_CHECK_ARG(nx > 0 && ny > 0 && nz > 0, "bad grid counts");
  PSL_CHECK_ARG(mask->n == (int64_t)nx * ny * nz, "mask size does not match the grid");
  const VoxParams p = vox_params(s, info->ox, info->oy, info->oz, info->res, nx, ny, nz);
  int* layers = nullptr;
  int maxc = 1;
  int rc = stl_layers(s, p, &layers, &maxc);
  if (rc == PSL_OK) {
    dim3 g2((nx + 255) / 256, ny);
    k_fill_columns<<<g2, 256>>>(layers, maxc, p, inside_value, mask->d);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      set_error("voxelizer: fill");
      rc = PSL_E_HIP;
    }
  }
  if (layers) (void)hipFree(layers);
  return rc;
}

}  // extern "C"
