// DIAD grid builder on the device.
//
// buildDIADGrids: include/DIADFunctions.h:640-1046, with its helpers from the
// same file (sortIJK 56-123, findMatchingIJKIndex 296-357,
// findPeriodicNeighbourIJKIndex 360-429, markFinestFluidBounceback 500-555,
// getDIADNeighbours 557-574, getDIADEsotwistNbrArray 576-612,
// markDIADNeighbours 615-638).
//
// Every level is a list of cells held as packed (k, j, i) keys (pack_key) in
// ascending order, so the reference's comparator sorts become one integer
// order: neighbour lookup is a binary search over 64-bit keys, the
// (keep desc, k, j, i) sort-and-truncate steps are order-preserving stream
// compactions, and only the 2x2x2 children of each level need a radix sort.
// The marker callbacks of the drivers (getMarkers) arrive as an ordered rule
// list (psl_marker_rule) evaluated per cell on the device.
//
// This translation unit is compiled with -ffp-contract=off: rule positions
// (iCell * res + ox) and the shifted origins of markFinestFluidBounceback
// must round exactly as written, host and device.
#include "psl_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

// The shared headers end in contraction mode fast; from here on nothing is
// fused, whatever the command line says.
#pragma clang fp contract(off)

namespace psl {
namespace {

constexpr int TPB = 256;
inline unsigned blocks_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, 1 << 20));
}

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    n = count;
    return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
  }
};

#define DB_TRY(expr)                    \
  do {                                  \
    const int _rc = (expr);             \
    if (_rc != PSL_OK) return _rc;      \
  } while (0)

__constant__ int DN26[26][3] = {
    {1, 0, 0},   {-1, 0, 0}, {0, 0, -1}, {0, 0, 1},   {0, -1, 0},  {0, 1, 0},   {1, 0, -1},  {-1, 0, 1},  {1, 0, 1},
    {-1, 0, -1}, {-1, -1, 0}, {1, 1, 0},  {0, 1, -1},  {0, -1, 1},  {-1, 1, 0},  {1, -1, 0},  {0, 1, 1},   {0, -1, -1},
    {-1, 1, -1}, {1, -1, 1}, {-1, -1, 1}, {1, 1, -1}, {1, -1, -1}, {-1, 1, 1}, {-1, -1, -1}, {1, 1, 1}};

__device__ __forceinline__ void unpack(uint64_t key, int& i, int& j, int& k) {
  i = (int)(key & KEY_MASK);
  j = (int)((key >> KEY_BITS) & KEY_MASK);
  k = (int)(key >> (2 * KEY_BITS));
}

// first index with key[idx] >= q
__device__ __forceinline__ int64_t lower_bound_key(const uint64_t* __restrict__ key, int64_t n, uint64_t q) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (key[mid] < q) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int find_key(const uint64_t* __restrict__ key, int64_t n, uint64_t q) {
  const int64_t lo = lower_bound_key(key, n, q);
  return (lo < n && key[lo] == q) ? (int)lo : -1;
}

// buildIJKFromInfo (lines 5-23): the whole box in cell order
__global__ void k_box_keys(uint64_t* __restrict__ key, int nx, int ny, int64_t n) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = c / nx;
    key[c] = pack_key((int)(c - r * nx), (int)(r % ny), (int)(r / ny));
  }
}

// getDIADNeighbours: the 26 neighbours, -1 where no cell exists (no wrap)
__global__ void k_nbr26(const uint64_t* __restrict__ key, int64_t n, int* __restrict__ nb) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    unpack(key[c], i, j, k);
    for (int d = 0; d < 26; ++d) {
      const int a = i + DN26[d][0], b = j + DN26[d][1], e = k + DN26[d][2];
      nb[(size_t)d * n + c] = (a < 0 || b < 0 || e < 0) ? -1 : find_key(key, n, pack_key(a, b, e));
    }
  }
}

// markDIADNeighbours: tgt |= some existing neighbour is set in src
__global__ void k_mark_nbrs(const int* __restrict__ nb, int64_t n, const uint8_t* __restrict__ src,
                            uint8_t* __restrict__ tgt) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    if (tgt[c]) continue;
    for (int d = 0; d < 26; ++d) {
      const int m = nb[(size_t)d * n + c];
      if (m >= 0 && src[m]) {
        tgt[c] = 1;
        break;
      }
    }
  }
}

enum { OP_OR, OP_AND, OP_ANDNOT, OP_NOT, OP_ZERO, OP_COPY };
// sumBoolArrays / multiplyBoolArrays / invertBoolArray (include/STLFunctions.h:811-846)
__global__ void k_ew(uint8_t* __restrict__ out, const uint8_t* a, const uint8_t* b, int64_t n, int op) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    uint8_t v;
    switch (op) {
      case OP_OR: v = a[c] | b[c]; break;
      case OP_AND: v = a[c] & b[c]; break;
      case OP_ANDNOT: v = a[c] & (uint8_t)!b[c]; break;
      case OP_NOT: v = (uint8_t)!a[c]; break;
      case OP_COPY: v = a[c]; break;
      default: v = 0; break;
    }
    out[c] = v;
  }
}

struct RuleInfo {
  int gridID;
  float res, ox, oy, oz;
  int nx, ny, nz;
};

__device__ bool rule_contains(const psl_marker_rule& r, const RuleInfo& I, int i, int j, int k) {
  if (r.shape == PSL_SHAPE_INDEX_BOX) {
    auto in = [](int v, int lo, int hi, int n) {
      if (lo < 0) lo += n;
      if (hi < 0) hi += n;
      return v >= lo && v <= hi;
    };
    return in(i, r.idx[0], r.idx[1], I.nx) && in(j, r.idx[2], r.idx[3], I.ny) && in(k, r.idx[4], r.idx[5], I.nz);
  }
  const float x = (float)i * I.res + I.ox, y = (float)j * I.res + I.oy, z = (float)k * I.res + I.oz;
  if (r.shape == PSL_SHAPE_BOX) return x > r.p[0] && x < r.p[3] && y > r.p[1] && y < r.p[4] && z > r.p[2] && z < r.p[5];
  const float dx = x - r.p[0], dy = y - r.p[1], dz = z - r.p[2];
  const float w2 = r.p[3] * r.p[3];
  if (r.shape == PSL_SHAPE_XCYLINDER) {
    const float a2 = dy * dy + dz * dz;
    return dx > 0.f && dx < r.p[4] && a2 < w2;
  }
  const float r2 = dx * dx + dy * dy + dz * dz;
  return r.shape == PSL_SHAPE_SPHERE_LE ? r2 <= w2 : r2 < w2;
}

__device__ __forceinline__ uint8_t apply_rules(const psl_marker_rule* __restrict__ rules, int nrules, int what,
                                               const RuleInfo& I, int i, int j, int k, uint8_t v) {
  for (int r = 0; r < nrules; ++r) {
    const psl_marker_rule& R = rules[r];
    if (R.what == what && ((R.level_mask >> I.gridID) & 1u) && rule_contains(R, I, i, j, k)) v = R.value ? 1 : 0;
  }
  return v;
}

// apply{Bounceback,Refinement,Outlet}MarkerFromFunction (lines 432-498)
__global__ void k_rules(const uint64_t* __restrict__ key, int64_t n, RuleInfo I,
                        const psl_marker_rule* __restrict__ rules, int nrules, int what, uint8_t* __restrict__ m) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    unpack(key[c], i, j, k);
    m[c] = apply_rules(rules, nrules, what, I, i, j, k, m[c]);
  }
}

// markFinestFluidBounceback without meshes: all shifted origins in one pass.
// boxes[6 b]: bounds of the value-1 bounceback rules of the level; a cell
// whose extent (all shifted positions lie within half a cell of its centre)
// misses every box cannot be marked by any shift, so its result (fluid, no
// wall) is written without the loop.  nbox < 0 disables the test.
__global__ void k_finest_rules(const uint64_t* __restrict__ key, int64_t n, RuleInfo I0,
                               const float* __restrict__ origins, int nshift, const psl_marker_rule* __restrict__ rules,
                               int nrules, const float* __restrict__ boxes, int nbox, uint8_t* __restrict__ fluid,
                               uint8_t* __restrict__ fbb) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    unpack(key[c], i, j, k);
    if (nbox >= 0) {
      const float xc[3] = {(float)i * I0.res + I0.ox, (float)j * I0.res + I0.oy, (float)k * I0.res + I0.oz};
      bool near = false;
      for (int b = 0; b < nbox && !near; ++b) {
        bool hit = true;
        for (int a = 0; a < 3; ++a) {
          const float h = 0.5f * I0.res + 1e-4f * (fabsf(xc[a]) + I0.res);
          hit = hit && xc[a] + h >= boxes[6 * b + a] && xc[a] - h <= boxes[6 * b + 3 + a];
        }
        near = hit;
      }
      if (!near) {
        fluid[c] = nshift > 0 ? 1 : 0;
        fbb[c] = 0;
        continue;
      }
    }
    uint8_t f = 0, b = 0;
    RuleInfo I = I0;
    for (int s = 0; s < nshift; ++s) {
      I.ox = origins[3 * s];
      I.oy = origins[3 * s + 1];
      I.oz = origins[3 * s + 2];
      const uint8_t m = apply_rules(rules, nrules, PSL_RULE_BOUNCEBACK, I, i, j, k, 0);
      b |= m;
      f |= (uint8_t)!m;
    }
    fluid[c] = f;
    fbb[c] = b;
  }
}

__global__ void k_accum_shift(const uint8_t* __restrict__ m, int64_t n, uint8_t* __restrict__ fluid,
                              uint8_t* __restrict__ fbb) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    fbb[c] |= m[c];
    fluid[c] |= (uint8_t)!m[c];
  }
}

// children of the listed coarse cells, offsets nested i-j-k (lines 872-899)
__global__ void k_children(const uint64_t* __restrict__ ckey, const int* __restrict__ list, int64_t m,
                           const uint8_t* __restrict__ bb_under, const uint8_t* __restrict__ fl_under,
                           uint64_t* __restrict__ fkey, uint8_t* __restrict__ enf_b, uint8_t* __restrict__ enf_f) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < 8 * m; t += (int64_t)gridDim.x * blockDim.x) {
    const int c = list[t >> 3];
    const int o = (int)(t & 7);
    int i, j, k;
    unpack(ckey[c], i, j, k);
    fkey[t] = pack_key(2 * i + (o >> 2), 2 * j + ((o >> 1) & 1), 2 * k + (o & 1));
    enf_b[t] = bb_under[c];
    enf_f[t] = fl_under[c];
  }
}

__global__ void k_gather_u8(const uint8_t* __restrict__ src, const int* __restrict__ perm, int64_t n,
                            uint8_t* __restrict__ dst) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    dst[c] = src[perm[c]];
}

struct EsoBounds {
  int64_t steps[10];
};

// occupancy of 16^3 bricks of the level's box, one bit each
constexpr int BRICK_SHIFT = 4;
__global__ void k_brick_mark(const uint64_t* __restrict__ key, int64_t n, int bx, int by,
                             uint32_t* __restrict__ words) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    // one atomic per run of a row inside a brick: the first cell of the run
    const uint64_t kc = key[c];
    if (c > 0 && (key[c - 1] >> BRICK_SHIFT) == (kc >> BRICK_SHIFT)) continue;
    int i, j, k;
    unpack(kc, i, j, k);
    const int64_t b = ((int64_t)(k >> BRICK_SHIFT) * by + (j >> BRICK_SHIFT)) * bx + (i >> BRICK_SHIFT);
    const uint32_t bit = 1u << (b & 31);
    if (!(words[b >> 5] & bit)) atomicOr(&words[b >> 5], bit);
  }
}

// getDIADEsotwistNbrArray / findPeriodicNeighbourIJKIndex: the first existing
// cell along the periodic ray +i +j +k +ij +ik +jk +ijk -i -j -k (the cell
// itself at the latest).  Along +-i the cells of a (j, k) row are contiguous
// in key order, so the answer is the adjacent entry or the row's wrap-around
// end.  The other directions walk the ray, bounded by its period: a point in
// an empty brick is skipped together with every following point of the ray
// inside that brick (up to the brick face or the box edge, whichever comes
// first); a point in an occupied brick costs one binary search.
__global__ void k_eso_nbrs(const uint64_t* __restrict__ key, int64_t n, int nx, int ny, int nz, EsoBounds bounds,
                           const uint32_t* __restrict__ bricks, int bx, int by, int* __restrict__ out,
                           int* __restrict__ fail) {
  const int D[10][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1},
                        {0, 1, 1}, {1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  constexpr int BM = (1 << BRICK_SHIFT) - 1;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t kc = key[c];
    const uint64_t row = kc >> KEY_BITS;
    int i, j, k;
    unpack(kc, i, j, k);
    for (int d = 0; d < 10; ++d) {
      int r = -1;
      if (d == 0) {
        r = (c + 1 < n && (key[c + 1] >> KEY_BITS) == row) ? (int)(c + 1)
                                                             : (int)lower_bound_key(key, n, row << KEY_BITS);
      } else if (d == 7) {
        r = (c > 0 && (key[c - 1] >> KEY_BITS) == row) ? (int)(c - 1)
                                                       : (int)lower_bound_key(key, n, (row + 1) << KEY_BITS) - 1;
      } else {
        const int dx = D[d][0], dy = D[d][1], dz = D[d][2];
        auto wrap = [](int v, int m) { return v >= m ? v - m : (v < 0 ? v + m : v); };
        int a = wrap(i + dx, nx), b = wrap(j + dy, ny), e = wrap(k + dz, nz);
        int64_t s = 1;  // points of the ray visited so far, this one included
        while (s <= bounds.steps[d]) {
          const int64_t br = ((int64_t)(e >> BRICK_SHIFT) * by + (b >> BRICK_SHIFT)) * bx + (a >> BRICK_SHIFT);
          if (!((bricks[br >> 5] >> (br & 31)) & 1u)) {
            int jump = 1 << 30;
            if (dx > 0) jump = min(jump, min(BM + 1 - (a & BM), nx - a));
            if (dx < 0) jump = min(jump, (a & BM) + 1);
            if (dy > 0) jump = min(jump, min(BM + 1 - (b & BM), ny - b));
            if (dy < 0) jump = min(jump, (b & BM) + 1);
            if (dz > 0) jump = min(jump, min(BM + 1 - (e & BM), nz - e));
            if (dz < 0) jump = min(jump, (e & BM) + 1);
            a = wrap(a + dx * jump, nx);
            b = wrap(b + dy * jump, ny);
            e = wrap(e + dz * jump, nz);
            s += jump;
            continue;
          }
          r = find_key(key, n, pack_key(a, b, e));
          if (r >= 0) break;
          a = wrap(a + dx, nx);
          b = wrap(b + dy, ny);
          e = wrap(e + dz, nz);
          ++s;
        }
        if (r < 0) atomicAdd(fail, 1);
      }
      out[(size_t)d * n + c] = r;
    }
  }
}

// mapInterfaceLambda (lines 964-1043): the fine cell at (2i, 2j, 2k)
__global__ void k_map_children(const uint64_t* __restrict__ ckey, const int* __restrict__ list, int64_t m,
                               const uint64_t* __restrict__ fkey, int64_t nf, int* __restrict__ out,
                               int* __restrict__ fail) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    unpack(ckey[list[t]], i, j, k);
    const int r = find_key(fkey, nf, pack_key(2 * i, 2 * j, 2 * k));
    out[t] = r;
    if (r < 0) atomicAdd(fail, 1);
  }
}

__global__ void k_flags_from_bb(const uint8_t* __restrict__ bb, int64_t n, uint8_t* __restrict__ flags) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    flags[c] = bb[c] ? (uint8_t)KIND_BOUNCEBACK : (uint8_t)KIND_FLUID;
}

__global__ void k_unpack_ijk(const uint64_t* __restrict__ key, int64_t n, int* __restrict__ ijk) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    unpack(key[c], i, j, k);
    ijk[c] = i;
    ijk[(size_t)n + c] = j;
    ijk[2 * (size_t)n + c] = k;
  }
}

int64_t gcd64(int64_t a, int64_t b) { return b ? gcd64(b, a % b) : a; }
int64_t lcm64(int64_t a, int64_t b) { return a / gcd64(a, b) * b; }

}  // namespace

// arrays kept per level (indices of psl_diad_levels_get beyond the cell tables)
enum { A_C2F_READ, A_C2F_WRITE, A_F2C_WRITE, A_F2C_READ, A_OUTLET, A_COUNT };

struct DiadLevel {
  psl_info info{};
  int64_t n = 0;
  DBuf<uint64_t> key;
  DBuf<uint8_t> bb, iface, c2f, f2c, enf_b, enf_f;
  DBuf<int> nbr;  // 10 x n
  DBuf<int> arr[A_COUNT];
  int64_t cnt[A_COUNT] = {0, 0, 0, 0, 0};
};

}  // namespace psl

struct psl_diad_levels {
  int levels = 0;
  std::vector<psl::DiadLevel> lv;
  double build_ms = 0.0;
};

namespace psl {
namespace {

struct Builder {
  int levels = 1, span = 0;
  float nu_phys = 0.f;
  std::vector<const psl_stl*> stls;
  DBuf<psl_marker_rule> rules;
  std::vector<psl_marker_rule> host_rules;
  int nrules = 0;
  DBuf<uint8_t> tmp;
  size_t tmp_bytes = 0;
  DBuf<int> dnum;  // selected count / failure counter
  psl_diad_levels* out = nullptr;
};

int launch_check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("diad build: ") + what + ": " + hipGetErrorString(e));
    return PSL_E_HIP;
  }
  return PSL_OK;
}

int ew(uint8_t* out, const uint8_t* a, const uint8_t* b, int64_t n, int op) {
  if (n == 0) return PSL_OK;
  k_ew<<<blocks_for(n), TPB>>>(out, a, b, n, op);
  return launch_check("elementwise");
}

int alloc_u8(DBuf<uint8_t>& b, int64_t n, bool zero) {
  PSL_HIP_TRY(b.alloc((size_t)n));
  if (zero && n) PSL_HIP_TRY(hipMemset(b.p, 0, (size_t)n));
  return PSL_OK;
}

int tmp_storage(Builder& B, size_t bytes) {
  if (bytes > B.tmp_bytes) {
    PSL_HIP_TRY(B.tmp.alloc(bytes));
    B.tmp_bytes = bytes;
  }
  return PSL_OK;
}

// order-preserving compaction of in[] where flags[] is set
template <class It, class T>
int select_flagged(Builder& B, It in, const uint8_t* flags, DBuf<T>& out, int64_t n, int64_t* count) {
  PSL_HIP_TRY(out.alloc((size_t)std::max<int64_t>(n, 1)));
  size_t bytes = 0;
  PSL_HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, in, flags, out.p, B.dnum.p, (int)n));
  DB_TRY(tmp_storage(B, bytes));
  bytes = B.tmp_bytes;
  PSL_HIP_TRY(hipcub::DeviceSelect::Flagged(B.tmp.p, bytes, in, flags, out.p, B.dnum.p, (int)n));
  int h = 0;
  PSL_HIP_TRY(hipMemcpy(&h, B.dnum.p, sizeof(int), hipMemcpyDeviceToHost));
  *count = h;
  return PSL_OK;
}

int select_index(Builder& B, const uint8_t* flags, DBuf<int>& out, int64_t n, int64_t* count) {
  return select_flagged(B, hipcub::CountingInputIterator<int>(0), flags, out, n, count);
}

RuleInfo rule_info(const psl_info& I, int level) {
  RuleInfo R;
  R.gridID = level;
  R.res = I.res;
  R.ox = I.ox; R.oy = I.oy; R.oz = I.oz;
  R.nx = I.cellCountX; R.ny = I.cellCountY; R.nz = I.cellCountZ;
  return R;
}

int run_rules(Builder& B, const uint64_t* key, int64_t n, const RuleInfo& I, int what, uint8_t* m) {
  if (n == 0 || B.nrules == 0) return PSL_OK;
  k_rules<<<blocks_for(n), TPB>>>(key, n, I, B.rules.p, B.nrules, what, m);
  return launch_check("rules");
}

// OR of the inside markers of all meshes (ApplyMarkersInsideSTL + sumBoolArrays)
int stl_union(Builder& B, const uint64_t* key, int64_t n, const RuleInfo& I, uint8_t* m) {
  DB_TRY(ew(m, nullptr, nullptr, n, OP_ZERO));
  for (const psl_stl* s : B.stls) DB_TRY(stl_mark_keys(s, I.ox, I.oy, I.oz, I.res, I.nx, I.ny, I.nz, key, n, m));
  return PSL_OK;
}

int mark_nbrs(const DBuf<int>& nb, int64_t n, const uint8_t* src, uint8_t* tgt) {
  if (n == 0) return PSL_OK;
  k_mark_nbrs<<<blocks_for(n), TPB>>>(nb.p, n, src, tgt);
  return launch_check("mark neighbours");
}

int esotwist_nbrs(Builder& B, DiadLevel& L) {
  const int nx = L.info.cellCountX, ny = L.info.cellCountY, nz = L.info.cellCountZ;
  const int D[10][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1},
                        {0, 1, 1}, {1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  EsoBounds eb;
  for (int d = 0; d < 10; ++d) {
    int64_t p = 1;
    if (D[d][0]) p = lcm64(p, nx);
    if (D[d][1]) p = lcm64(p, ny);
    if (D[d][2]) p = lcm64(p, nz);
    eb.steps[d] = p;  // the ray is back at its cell after one period
  }
  PSL_HIP_TRY(L.nbr.alloc(10 * (size_t)std::max<int64_t>(L.n, 1)));
  PSL_HIP_TRY(hipMemset(B.dnum.p, 0, sizeof(int)));
  const int bx = (nx + 15) >> 4, by = (ny + 15) >> 4, bz = (nz + 15) >> 4;
  const size_t nwords = ((size_t)bx * by * bz + 31) / 32;
  DBuf<uint32_t> bricks;
  PSL_HIP_TRY(bricks.alloc(nwords));
  PSL_HIP_TRY(hipMemset(bricks.p, 0, nwords * sizeof(uint32_t)));
  if (L.n) {
    k_brick_mark<<<blocks_for(L.n), TPB>>>(L.key.p, L.n, bx, by, bricks.p);
    k_eso_nbrs<<<blocks_for(L.n), TPB>>>(L.key.p, L.n, nx, ny, nz, eb, bricks.p, bx, by, L.nbr.p, B.dnum.p);
    DB_TRY(launch_check("esotwist neighbours"));
  }
  int fail = 0;
  PSL_HIP_TRY(hipMemcpy(&fail, B.dnum.p, sizeof(int), hipMemcpyDeviceToHost));
  if (fail) {
    set_error("diad build: esotwist neighbour not found");
    return PSL_E_HIP;
  }
  return PSL_OK;
}

// markFinestFluidBounceback (lines 500-555)
int finest_fluid_bb(Builder& B, const DiadLevel& L, int level, DBuf<uint8_t>& fluid, DBuf<uint8_t>& fbb) {
  const int64_t n = L.n;
  const RuleInfo I0 = rule_info(L.info, level);
  std::vector<float> origins;
  for (int tl = level; tl < B.levels; ++tl) {
    const int shifts = 1 << (tl - level);
    const float target_res = I0.res / shifts;
    const float start = -0.5f * I0.res + 0.5f * target_res;
    for (int sx = 0; sx < shifts; ++sx)
      for (int sy = 0; sy < shifts; ++sy)
        for (int sz = 0; sz < shifts; ++sz) {
          origins.push_back(I0.ox + start + sx * target_res);
          origins.push_back(I0.oy + start + sy * target_res);
          origins.push_back(I0.oz + start + sz * target_res);
        }
  }
  const int nshift = (int)(origins.size() / 3);
  DB_TRY(alloc_u8(fluid, n, true));
  DB_TRY(alloc_u8(fbb, n, true));
  if (n == 0) return PSL_OK;
  if (B.stls.empty()) {
    DBuf<float> dor;
    PSL_HIP_TRY(dor.alloc(origins.size()));
    PSL_HIP_TRY(hipMemcpy(dor.p, origins.data(), origins.size() * sizeof(float), hipMemcpyHostToDevice));
    // bounds of the rules that can mark a wall on this level
    std::vector<float> boxes;
    int nbox = 0;
    for (const psl_marker_rule& r : B.host_rules) {
      if (r.what != PSL_RULE_BOUNCEBACK || !r.value || !((r.level_mask >> level) & 1u)) continue;
      if (r.shape == PSL_SHAPE_INDEX_BOX) {
        nbox = -1;
        break;
      }
      float lo[3], hi[3];
      if (r.shape == PSL_SHAPE_BOX) {
        for (int a = 0; a < 3; ++a) { lo[a] = r.p[a]; hi[a] = r.p[3 + a]; }
      } else if (r.shape == PSL_SHAPE_XCYLINDER) {
        const float w = std::fabs(r.p[3]);
        lo[0] = r.p[0]; hi[0] = r.p[0] + std::fabs(r.p[4]);
        for (int a = 1; a < 3; ++a) { lo[a] = r.p[a] - w; hi[a] = r.p[a] + w; }
      } else {
        const float w = std::fabs(r.p[3]);
        for (int a = 0; a < 3; ++a) { lo[a] = r.p[a] - w; hi[a] = r.p[a] + w; }
      }
      for (int a = 0; a < 3; ++a) boxes.push_back(lo[a]);
      for (int a = 0; a < 3; ++a) boxes.push_back(hi[a]);
      ++nbox;
    }
    DBuf<float> dbox;
    PSL_HIP_TRY(dbox.alloc(std::max<size_t>(boxes.size(), 1)));
    if (!boxes.empty())
      PSL_HIP_TRY(hipMemcpy(dbox.p, boxes.data(), boxes.size() * sizeof(float), hipMemcpyHostToDevice));
    k_finest_rules<<<blocks_for(n), TPB>>>(L.key.p, n, I0, dor.p, nshift, B.rules.p, B.nrules, dbox.p,
                                           nbox < 0 ? -1 : nbox, fluid.p, fbb.p);
    DB_TRY(launch_check("finest fluid"));
    PSL_HIP_TRY(hipDeviceSynchronize());
    return PSL_OK;
  }
  DBuf<uint8_t> m;
  DB_TRY(alloc_u8(m, n, false));
  for (int s = 0; s < nshift; ++s) {
    RuleInfo I = I0;
    I.ox = origins[3 * s];
    I.oy = origins[3 * s + 1];
    I.oz = origins[3 * s + 2];
    DB_TRY(stl_union(B, L.key.p, n, I, m.p));
    DB_TRY(run_rules(B, L.key.p, n, I, PSL_RULE_BOUNCEBACK, m.p));
    k_accum_shift<<<blocks_for(n), TPB>>>(m.p, n, fluid.p, fbb.p);
    DB_TRY(launch_check("finest fluid"));
  }
  return PSL_OK;
}

int build_finest(Builder& B, int level) {
  DiadLevel& L = B.out->lv[level];
  const int64_t n = L.n;
  const RuleInfo I = rule_info(L.info, level);
  const bool multi = B.levels > 1;
  DBuf<int> nb;
  PSL_HIP_TRY(nb.alloc(26 * (size_t)std::max<int64_t>(n, 1)));
  if (n) {
    k_nbr26<<<blocks_for(n), TPB>>>(L.key.p, n, nb.p);
    DB_TRY(launch_check("neighbours"));
  }
  DBuf<uint8_t> bb, fluid, keep;
  DB_TRY(alloc_u8(bb, n, false));
  DB_TRY(alloc_u8(fluid, n, false));
  DB_TRY(alloc_u8(keep, n, false));
  DB_TRY(stl_union(B, L.key.p, n, I, bb.p));
  DB_TRY(run_rules(B, L.key.p, n, I, PSL_RULE_BOUNCEBACK, bb.p));
  if (multi) DB_TRY(ew(bb.p, bb.p, L.enf_b.p, n, OP_OR));
  DB_TRY(ew(fluid.p, bb.p, nullptr, n, OP_NOT));
  if (multi) DB_TRY(ew(fluid.p, fluid.p, L.enf_f.p, n, OP_OR));
  DB_TRY(ew(keep.p, fluid.p, nullptr, n, OP_COPY));
  DB_TRY(mark_nbrs(nb, n, fluid.p, keep.p));
  nb.reset();
  DB_TRY(ew(bb.p, fluid.p, nullptr, n, OP_NOT));
  // sortFinestGrid + resize
  DBuf<uint64_t> key2;
  DBuf<uint8_t> bb2;
  int64_t n2 = 0, chk = 0;
  DB_TRY(select_flagged(B, L.key.p, keep.p, key2, n, &n2));
  DB_TRY(select_flagged(B, bb.p, keep.p, bb2, n, &chk));
  L.key = std::move(key2);
  L.bb = std::move(bb2);
  L.n = n2;
  L.info.cellCount = (int)n2;
  DB_TRY(alloc_u8(L.iface, n2, true));
  L.enf_b.reset();
  L.enf_f.reset();
  return esotwist_nbrs(B, L);
}

int build_coarse(Builder& B, int level) {
  DiadLevel& L = B.out->lv[level];
  DiadLevel& F = B.out->lv[level + 1];
  const int64_t n = L.n;
  const RuleInfo I = rule_info(L.info, level);
  DBuf<int> nb;
  PSL_HIP_TRY(nb.alloc(26 * (size_t)std::max<int64_t>(n, 1)));
  if (n) {
    k_nbr26<<<blocks_for(n), TPB>>>(L.key.p, n, nb.p);
    DB_TRY(launch_check("neighbours"));
  }
  DBuf<uint8_t> fluid, fbb, border, keep, ref, nref, f2c, c2f, cbb, bbu, flu, t;
  DB_TRY(finest_fluid_bb(B, L, level, fluid, fbb));
  if (level) DB_TRY(ew(fluid.p, fluid.p, L.enf_f.p, n, OP_OR));
  DB_TRY(alloc_u8(border, n, true));
  DB_TRY(mark_nbrs(nb, n, fluid.p, border.p));
  DB_TRY(ew(border.p, border.p, fluid.p, n, OP_ANDNOT));
  DB_TRY(alloc_u8(keep, n, false));
  DB_TRY(ew(keep.p, fluid.p, border.p, n, OP_OR));
  DB_TRY(alloc_u8(ref, n, false));
  DB_TRY(ew(ref.p, fbb.p, border.p, n, OP_OR));
  DB_TRY(alloc_u8(nref, n, true));
  const int thickness = B.span + (B.levels - level - 1);
  for (int layer = 0; layer < thickness; ++layer) {
    DB_TRY(mark_nbrs(nb, n, ref.p, nref.p));
    DB_TRY(ew(nref.p, ref.p, nref.p, n, OP_OR));
    DB_TRY(ew(ref.p, nref.p, keep.p, n, OP_AND));
  }
  DB_TRY(run_rules(B, L.key.p, n, I, PSL_RULE_REFINEMENT, ref.p));
  DB_TRY(alloc_u8(f2c, n, true));
  DB_TRY(mark_nbrs(nb, n, ref.p, f2c.p));
  DB_TRY(ew(f2c.p, keep.p, f2c.p, n, OP_AND));
  DB_TRY(ew(f2c.p, f2c.p, ref.p, n, OP_ANDNOT));
  DB_TRY(alloc_u8(c2f, n, true));
  DB_TRY(mark_nbrs(nb, n, f2c.p, c2f.p));
  DB_TRY(ew(c2f.p, keep.p, c2f.p, n, OP_AND));
  DB_TRY(ew(c2f.p, c2f.p, f2c.p, n, OP_ANDNOT));
  DB_TRY(ew(c2f.p, c2f.p, ref.p, n, OP_ANDNOT));
  nb.reset();
  DB_TRY(ew(ref.p, ref.p, f2c.p, n, OP_OR));
  DB_TRY(ew(ref.p, ref.p, c2f.p, n, OP_OR));
  DB_TRY(alloc_u8(cbb, n, true));
  DB_TRY(alloc_u8(t, n, false));
  DB_TRY(run_rules(B, L.key.p, n, I, PSL_RULE_BOUNCEBACK, cbb.p));
  DB_TRY(stl_union(B, L.key.p, n, I, t.p));
  DB_TRY(ew(cbb.p, cbb.p, t.p, n, OP_OR));
  DB_TRY(alloc_u8(bbu, n, false));
  DB_TRY(alloc_u8(flu, n, false));
  DB_TRY(ew(t.p, c2f.p, f2c.p, n, OP_OR));  // the interface
  DB_TRY(ew(bbu.p, t.p, cbb.p, n, OP_AND));
  DB_TRY(ew(flu.p, t.p, cbb.p, n, OP_ANDNOT));
  DB_TRY(ew(keep.p, fluid.p, border.p, n, OP_OR));
  DB_TRY(ew(keep.p, keep.p, ref.p, n, OP_ANDNOT));
  DB_TRY(ew(keep.p, keep.p, c2f.p, n, OP_OR));
  DB_TRY(ew(keep.p, keep.p, f2c.p, n, OP_OR));
  if (level) DB_TRY(ew(keep.p, keep.p, L.enf_f.p, n, OP_OR));
  fluid.reset(); fbb.reset(); border.reset(); nref.reset(); cbb.reset();

  // next level: children of the refined cells, sorted by (k, j, i)
  {
    DBuf<int> list;
    int64_t m = 0;
    DB_TRY(select_index(B, ref.p, list, n, &m));
    if (8 * m > (int64_t)INT32_MAX) {
      set_error("diad build: too many cells on the next level");
      return PSL_E_ARG;
    }
    const int64_t nf = 8 * m;
    DBuf<uint64_t> ukey, skey;
    DBuf<uint8_t> ub, uf;
    DBuf<int> perm_in, perm;
    PSL_HIP_TRY(ukey.alloc((size_t)std::max<int64_t>(nf, 1)));
    PSL_HIP_TRY(skey.alloc((size_t)std::max<int64_t>(nf, 1)));
    DB_TRY(alloc_u8(ub, nf, false));
    DB_TRY(alloc_u8(uf, nf, false));
    PSL_HIP_TRY(perm_in.alloc((size_t)std::max<int64_t>(nf, 1)));
    PSL_HIP_TRY(perm.alloc((size_t)std::max<int64_t>(nf, 1)));
    if (nf) {
      k_children<<<blocks_for(nf), TPB>>>(L.key.p, list.p, m, bbu.p, flu.p, ukey.p, ub.p, uf.p);
      DB_TRY(launch_check("children"));
      std::vector<int> iota((size_t)nf);
      std::iota(iota.begin(), iota.end(), 0);
      PSL_HIP_TRY(hipMemcpy(perm_in.p, iota.data(), (size_t)nf * sizeof(int), hipMemcpyHostToDevice));
      size_t bytes = 0;
      PSL_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ukey.p, skey.p, perm_in.p, perm.p, (int)nf, 0,
                                                     3 * KEY_BITS));
      DB_TRY(tmp_storage(B, bytes));
      bytes = B.tmp_bytes;
      PSL_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(B.tmp.p, bytes, ukey.p, skey.p, perm_in.p, perm.p, (int)nf, 0,
                                                     3 * KEY_BITS));
    }
    DB_TRY(alloc_u8(F.enf_b, nf, false));
    DB_TRY(alloc_u8(F.enf_f, nf, false));
    if (nf) {
      k_gather_u8<<<blocks_for(nf), TPB>>>(ub.p, perm.p, nf, F.enf_b.p);
      k_gather_u8<<<blocks_for(nf), TPB>>>(uf.p, perm.p, nf, F.enf_f.p);
      DB_TRY(launch_check("gather"));
    }
    F.key = std::move(skey);
    F.n = nf;
  }
  bbu.reset(); flu.reset(); ref.reset();

  // sortCoarseGrid + resize
  int64_t n2 = 0, chk = 0;
  {
    DBuf<uint64_t> key2;
    DBuf<uint8_t> f2c2, c2f2, eb2, ef2;
    DB_TRY(select_flagged(B, L.key.p, keep.p, key2, n, &n2));
    DB_TRY(select_flagged(B, f2c.p, keep.p, f2c2, n, &chk));
    DB_TRY(select_flagged(B, c2f.p, keep.p, c2f2, n, &chk));
    if (level) {
      DB_TRY(select_flagged(B, L.enf_b.p, keep.p, eb2, n, &chk));
      DB_TRY(select_flagged(B, L.enf_f.p, keep.p, ef2, n, &chk));
    }
    L.key = std::move(key2);
    f2c = std::move(f2c2);
    c2f = std::move(c2f2);
    L.enf_b = std::move(eb2);
    L.enf_f = std::move(ef2);
    L.n = n2;
    L.info.cellCount = (int)n2;
  }
  keep.reset();
  DB_TRY(alloc_u8(L.iface, n2, false));
  DB_TRY(ew(L.iface.p, f2c.p, c2f.p, n2, OP_OR));
  // final bounceback pass
  DB_TRY(alloc_u8(L.bb, n2, true));
  DB_TRY(alloc_u8(t, n2, false));
  DB_TRY(run_rules(B, L.key.p, n2, I, PSL_RULE_BOUNCEBACK, L.bb.p));
  DB_TRY(stl_union(B, L.key.p, n2, I, t.p));
  DB_TRY(ew(L.bb.p, L.bb.p, t.p, n2, OP_OR));
  if (level) {
    DB_TRY(ew(L.bb.p, L.bb.p, L.enf_b.p, n2, OP_OR));
    DB_TRY(ew(L.bb.p, L.bb.p, L.enf_f.p, n2, OP_ANDNOT));
  }
  L.enf_b.reset();
  L.enf_f.reset();
  DB_TRY(ew(c2f.p, c2f.p, L.bb.p, n2, OP_ANDNOT));
  DB_TRY(ew(f2c.p, f2c.p, L.bb.p, n2, OP_ANDNOT));
  L.c2f = std::move(c2f);
  L.f2c = std::move(f2c);
  DB_TRY(esotwist_nbrs(B, L));
  // outlet cells
  DB_TRY(ew(t.p, nullptr, nullptr, n2, OP_ZERO));
  DB_TRY(run_rules(B, L.key.p, n2, I, PSL_RULE_OUTLET, t.p));
  DB_TRY(select_index(B, t.p, L.arr[A_OUTLET], n2, &L.cnt[A_OUTLET]));

  // Info of the next level (lines 950-958)
  psl_info& FI = F.info;
  FI = L.info;
  FI.gridID = L.info.gridID + 1;
  FI.res = L.info.res * 0.5f;
  FI.dtPhys = L.info.dtPhys * 0.5f;
  FI.nu = (FI.dtPhys * B.nu_phys) / ((FI.res / 1000.f) * (FI.res / 1000.f));
  FI.ox = L.info.ox - FI.res * 0.5f;
  FI.oy = L.info.oy - FI.res * 0.5f;
  FI.oz = L.info.oz - FI.res * 0.5f;
  FI.cellCountX = L.info.cellCountX * 2;
  FI.cellCountY = L.info.cellCountY * 2;
  FI.cellCountZ = L.info.cellCountZ * 2;
  FI.cellCount = (int)F.n;
  return PSL_OK;
}

int map_interface(Builder& B, int level, const DBuf<uint8_t>& marker, int coarse_slot, int fine_slot) {
  DiadLevel& L = B.out->lv[level];
  const DiadLevel& F = B.out->lv[level + 1];
  int64_t m = 0;
  DB_TRY(select_index(B, marker.p, L.arr[coarse_slot], L.n, &m));
  L.cnt[coarse_slot] = m;
  L.cnt[fine_slot] = m;
  PSL_HIP_TRY(L.arr[fine_slot].alloc((size_t)std::max<int64_t>(m, 1)));
  PSL_HIP_TRY(hipMemset(B.dnum.p, 0, sizeof(int)));
  if (m) {
    k_map_children<<<blocks_for(m), TPB>>>(L.key.p, L.arr[coarse_slot].p, m, F.key.p, F.n, L.arr[fine_slot].p,
                                           B.dnum.p);
    DB_TRY(launch_check("interface map"));
  }
  int fail = 0;
  PSL_HIP_TRY(hipMemcpy(&fail, B.dnum.p, sizeof(int), hipMemcpyDeviceToHost));
  if (fail) {
    set_error("DIAD Grid Generation Error: " + std::to_string(fail) + " invalid interface connections found on level " +
              std::to_string(level));
    return PSL_E_ARG;
  }
  return PSL_OK;
}

int run_build(Builder& B, const psl_info& base) {
  psl_diad_levels* out = B.out;
  out->lv.resize(B.levels);
  DiadLevel& L0 = out->lv[0];
  L0.info = base;
  L0.info.gridID = 0;
  L0.n = (int64_t)base.cellCountX * base.cellCountY * base.cellCountZ;
  L0.info.cellCount = (int)L0.n;
  PSL_HIP_TRY(L0.key.alloc((size_t)L0.n));
  k_box_keys<<<blocks_for(L0.n), TPB>>>(L0.key.p, base.cellCountX, base.cellCountY, L0.n);
  DB_TRY(launch_check("box"));
  for (int l = 0; l < B.levels; ++l) {
    if (l == B.levels - 1) DB_TRY(build_finest(B, l));
    else DB_TRY(build_coarse(B, l));
  }
  for (int l = B.levels - 2; l >= 0; --l) {
    DiadLevel& L = out->lv[l];
    DB_TRY(map_interface(B, l, L.c2f, A_C2F_READ, A_C2F_WRITE));
    DB_TRY(map_interface(B, l, L.f2c, A_F2C_WRITE, A_F2C_READ));
    L.c2f.reset();
    L.f2c.reset();
  }
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

}  // namespace
}  // namespace psl

using namespace psl;

extern "C" {

int psl_diad_build(const psl_info* base, int levels, int wall_refinement_span, float nu_phys,
                   const psl_stl* const* stls, int nstl, const psl_marker_rule* rules, int nrules,
                   psl_diad_levels** out) {
  PSL_CHECK_ARG(base && out, "null argument");
  PSL_CHECK_ARG(levels >= 1 && levels <= 16, "levels must be in [1, 16]");
  PSL_CHECK_ARG(wall_refinement_span >= 0, "negative wall refinement span");
  PSL_CHECK_ARG(nstl >= 0 && (nstl == 0 || stls), "bad mesh list");
  PSL_CHECK_ARG(nrules >= 0 && (nrules == 0 || rules), "bad rule list");
  PSL_CHECK_ARG(base->cellCountX > 0 && base->cellCountY > 0 && base->cellCountZ > 0, "bad base cell counts");
  PSL_CHECK_ARG(base->res > 0.f, "bad resolution");
  const int64_t top = (int64_t)std::max(base->cellCountX, std::max(base->cellCountY, base->cellCountZ))
                      << (levels - 1);
  PSL_CHECK_ARG(top <= (int64_t)KEY_MASK, "finest box exceeds 2^21 - 1 cells along an axis");
  PSL_CHECK_ARG((int64_t)base->cellCountX * base->cellCountY * base->cellCountZ < (int64_t)INT32_MAX,
                "base box too large");
  for (int r = 0; r < nrules; ++r) {
    PSL_CHECK_ARG(rules[r].what >= PSL_RULE_BOUNCEBACK && rules[r].what <= PSL_RULE_OUTLET, "bad rule kind");
    PSL_CHECK_ARG(rules[r].shape >= PSL_SHAPE_SPHERE_LE && rules[r].shape <= PSL_SHAPE_INDEX_BOX, "bad rule shape");
  }
  for (int s = 0; s < nstl; ++s) PSL_CHECK_ARG(stls[s], "null mesh");
  const auto t0 = std::chrono::steady_clock::now();
  psl_diad_levels* res = new psl_diad_levels;
  res->levels = levels;
  Builder B;
  B.levels = levels;
  B.span = wall_refinement_span;
  B.nu_phys = nu_phys;
  B.stls.assign(stls, stls + nstl);
  B.nrules = nrules;
  B.host_rules.assign(rules, rules + nrules);
  B.out = res;
  int rc = PSL_OK;
  if (B.rules.alloc((size_t)std::max(nrules, 1)) != hipSuccess || B.dnum.alloc(1) != hipSuccess) {
    set_error("diad build: alloc");
    rc = PSL_E_HIP;
  }
  if (rc == PSL_OK && nrules &&
      hipMemcpy(B.rules.p, rules, (size_t)nrules * sizeof(psl_marker_rule), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("diad build: rule upload");
    rc = PSL_E_HIP;
  }
  if (rc == PSL_OK) rc = run_build(B, *base);
  if (rc != PSL_OK) {
    delete res;
    return rc;
  }
  res->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = res;
  return PSL_OK;
}

int psl_diad_levels_destroy(psl_diad_levels* levels) {
  delete levels;
  return PSL_OK;
}

int psl_diad_levels_build_ms(const psl_diad_levels* levels, double* ms) {
  PSL_CHECK_ARG(levels && ms, "null argument");
  *ms = levels->build_ms;
  return PSL_OK;
}

int psl_diad_levels_info(const psl_diad_levels* levels, int level, psl_info* info, int64_t counts[PSL_DIAD_ARRAYS]) {
  PSL_CHECK_ARG(levels && level >= 0 && level < levels->levels, "bad level");
  const DiadLevel& L = levels->lv[level];
  if (info) {
    *info = L.info;
    info->gridID = (float)level;
  }
  if (counts) {
    counts[PSL_DIAD_IJK] = 3 * L.n;
    counts[PSL_DIAD_BOUNCEBACK] = L.n;
    counts[PSL_DIAD_INTERFACE] = L.n;
    counts[PSL_DIAD_NBR] = 10 * L.n;
    counts[PSL_DIAD_C2F_READ] = L.cnt[A_C2F_READ];
    counts[PSL_DIAD_C2F_WRITE] = L.cnt[A_C2F_WRITE];
    counts[PSL_DIAD_F2C_WRITE] = L.cnt[A_F2C_WRITE];
    counts[PSL_DIAD_F2C_READ] = L.cnt[A_F2C_READ];
    counts[PSL_DIAD_OUTLET] = L.cnt[A_OUTLET];
  }
  return PSL_OK;
}

int psl_diad_levels_get(const psl_diad_levels* levels, int level, int which, void* host) {
  PSL_CHECK_ARG(levels && level >= 0 && level < levels->levels && host, "bad argument");
  PSL_CHECK_ARG(which >= 0 && which < PSL_DIAD_ARRAYS, "bad array");
  const DiadLevel& L = levels->lv[level];
  const size_t n = (size_t)L.n;
  switch (which) {
    case PSL_DIAD_IJK: {
      if (!n) return PSL_OK;
      DBuf<int> ijk;
      PSL_HIP_TRY(ijk.alloc(3 * n));
      k_unpack_ijk<<<blocks_for((int64_t)n), TPB>>>(L.key.p, (int64_t)n, ijk.p);
      PSL_LAUNCH_CHECK();
      PSL_HIP_TRY(hipMemcpy(host, ijk.p, 3 * n * sizeof(int), hipMemcpyDeviceToHost));
      return PSL_OK;
    }
    case PSL_DIAD_BOUNCEBACK:
      if (n) PSL_HIP_TRY(hipMemcpy(host, L.bb.p, n, hipMemcpyDeviceToHost));
      return PSL_OK;
    case PSL_DIAD_INTERFACE:
      if (n) PSL_HIP_TRY(hipMemcpy(host, L.iface.p, n, hipMemcpyDeviceToHost));
      return PSL_OK;
    case PSL_DIAD_NBR:
      if (n) PSL_HIP_TRY(hipMemcpy(host, L.nbr.p, 10 * n * sizeof(int), hipMemcpyDeviceToHost));
      return PSL_OK;
    default: {
      const int slot = which - PSL_DIAD_C2F_READ;
      if (L.cnt[slot])
        PSL_HIP_TRY(hipMemcpy(host, L.arr[slot].p, (size_t)L.cnt[slot] * sizeof(int), hipMemcpyDeviceToHost));
      return PSL_OK;
    }
  }
}

int psl_diad_create_level(const psl_diad_levels* levels, int level, int storage, psl_diad** out) {
  PSL_CHECK_ARG(levels && out && level >= 0 && level < levels->levels, "bad level");
  PSL_CHECK_ARG(storage == PSL_F32 || storage == PSL_F16DEV || storage == PSL_BF16DEV, "unknown storage");
  const DiadLevel& L = levels->lv[level];
  const int64_t n = L.n;
  PSL_CHECK_ARG(n > 0, "empty level");
  psl_diad* d = new psl_diad();
  d->rec_version = diad_next_version();
  d->info = L.info;
  d->info.gridID = (float)level;
  d->info.cellCount = (int)n;
  d->storage = storage;
  d->n = n;
  auto fail = [&](const char* what) {
    set_error(std::string("DIAD level grid: ") + what);
    psl_diad_destroy(d);
    return PSL_E_HIP;
  };
  const size_t fb = 27ull * (size_t)n * storage_bytes(storage);
  if (hipMalloc(&d->f, fb) != hipSuccess || hipMalloc(&d->bc_dev, 4 * sizeof(float4)) != hipSuccess ||
      hipMemset(d->bc_dev, 0, 4 * sizeof(float4)) != hipSuccess ||
      hipMalloc(&d->ijk, 3ull * (size_t)n * sizeof(int)) != hipSuccess ||
      hipMalloc(&d->nbr, 7ull * (size_t)n * sizeof(int)) != hipSuccess || hipMalloc(&d->flags, (size_t)n) != hipSuccess)
    return fail("allocation");
  k_unpack_ijk<<<blocks_for(n), TPB>>>(L.key.p, n, d->ijk);
  k_flags_from_bb<<<blocks_for(n), TPB>>>(L.bb.p, n, d->flags);
  if (hipGetLastError() != hipSuccess) return fail("tables");
  // the first seven neighbour columns (+i +j +k +ij +ik +jk +ijk)
  if (hipMemcpy(d->nbr, L.nbr.p, 7ull * (size_t)n * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess ||
      diad_build_adjacency(d) != PSL_OK)
    return fail("neighbours");
  const int64_t nc = L.cnt[A_C2F_READ], nf = L.cnt[A_F2C_WRITE];
  if (nc) {
    if (hipMalloc(&d->c2f_read, (size_t)nc * sizeof(int)) != hipSuccess ||
        hipMalloc(&d->c2f_write, (size_t)nc * sizeof(int)) != hipSuccess ||
        hipMemcpy(d->c2f_read, L.arr[A_C2F_READ].p, (size_t)nc * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(d->c2f_write, L.arr[A_C2F_WRITE].p, (size_t)nc * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess)
      return fail("interface");
  }
  if (nf) {
    if (hipMalloc(&d->f2c_write, (size_t)nf * sizeof(int)) != hipSuccess ||
        hipMalloc(&d->f2c_read, (size_t)nf * sizeof(int)) != hipSuccess ||
        hipMemcpy(d->f2c_write, L.arr[A_F2C_WRITE].p, (size_t)nf * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(d->f2c_read, L.arr[A_F2C_READ].p, (size_t)nf * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess)
      return fail("interface");
  }
  d->n_c2f = nc;
  d->n_f2c = nf;
  const int64_t no = L.cnt[A_OUTLET];
  if (no) {
    if (hipMalloc(&d->outlet, (size_t)no * sizeof(int)) != hipSuccess ||
        hipMalloc(&d->kmin, (size_t)n * sizeof(int)) != hipSuccess ||
        hipMemcpy(d->outlet, L.arr[A_OUTLET].p, (size_t)no * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(d->kmin, L.nbr.p + 9 * (size_t)n, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess)
      return fail("outlet");
  }
  d->n_outlet = no;
  if (hipDeviceSynchronize() != hipSuccess) return fail("sync");
  *out = d;
  return PSL_OK;
}

}  // extern "C"
