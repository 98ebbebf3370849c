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

#include <algorithm>
#include <atomic>
#include <cmath>
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
  const uint8_t* adj;  // per cell: which of +i +ij +ik +ijk are the next cell (see eso_slots_adj)
  const float4* bc;  // 4 given-value slots (device)
  // interface records the local update writes (null: none): child records,
  // read by the transfer to the coarser level, and source records, read by
  // the transfer to the finer level
  void* rec;                  // child records
  const uint64_t* rec_mask;   // per 64 cells: cells with a child record
  const uint32_t* rec_base;   // per 64 cells: child records below
  void* srec;                 // source records
  const uint64_t* srec_mask;
  const uint32_t* srec_base;
};

// Interface records: the post-collision storage words of one cell, in q
// order, padded to ESO_REC_STRIDE elements (64 B in 16-bit storage, 128 B in
// FP32) and written with 16-byte stores.
constexpr int ESO_REC_STRIDE = 32;

template <class St>
__device__ __forceinline__ void rec_store(void* rec, uint32_t idx, const float (&v)[27]) {
  typedef typename St::T T;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  u4* r = (u4*)((T*)rec + (size_t)idx * ESO_REC_STRIDE);
  if constexpr (sizeof(T) == 2) {
    uint32_t w[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t lo = 2 * k < 27 ? (uint32_t)__builtin_bit_cast(uint16_t, St::store(v[2 * k])) : 0u;
      const uint32_t hi = 2 * k + 1 < 27 ? (uint32_t)__builtin_bit_cast(uint16_t, St::store(v[2 * k + 1])) : 0u;
      w[k] = lo | (hi << 16);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = u4{w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]};
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      u4 x;
#pragma unroll
      for (int c = 0; c < 4; ++c) x[c] = 4 * k + c < 27 ? __builtin_bit_cast(uint32_t, St::store(v[4 * k + c])) : 0u;
      r[k] = x;
    }
  }
}

// Full values of a record (the lattice weights added back for 16-bit storage).
template <class St>
__device__ __forceinline__ void rec_load(const void* rec, uint32_t idx, float (&v)[27]) {
  typedef typename St::T T;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4* r = (const u4*)((const T*)rec + (size_t)idx * ESO_REC_STRIDE);
  if constexpr (sizeof(T) == 2) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const u4 x = r[k];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int q = 8 * k + 2 * c;
        if (q < 27) v[q] = St::load(__builtin_bit_cast(T, (uint16_t)(x[c] & 0xffffu)));
        if (q + 1 < 27) v[q + 1] = St::load(__builtin_bit_cast(T, (uint16_t)(x[c] >> 16)));
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const u4 x = r[k];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        // the element goes through a scalar first: a bit cast applied to a
        // vector element directly compiles to a read of element 0 here
        const uint32_t w = x[c];
        if (4 * k + c < 27) v[4 * k + c] = St::load(__builtin_bit_cast(T, w));
      }
    }
  }
  if (St::W_OFF) {
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] += weight(q);
  }
}

// Record index of `cell`, or -1 when it has none.
__device__ __forceinline__ int rec_index(const uint64_t* __restrict__ mask, const uint32_t* __restrict__ base,
                                         uint32_t cell) {
  const uint64_t m = mask[cell >> 6];
  const uint32_t lane = cell & 63u;
  if (!((m >> lane) & 1u)) return -1;
  return (int)(base[cell >> 6] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)));
}

template <class St>
__device__ __forceinline__ void eso_slots(const EsoParams& p, uint32_t cell, uint32_t (&slot)[8]) {
  slot[0] = cell;
#pragma unroll
  for (int t = 0; t < 7; ++t) slot[t + 1] = (uint32_t)p.nbr[(size_t)t * p.n + cell];
}

// The eight slots of a cell for the local update from three neighbour columns
// instead of seven.  Cells are sorted by (k, j, i), so the +i, +ij, +ik and
// +ijk neighbours are nearly always the cell after the cell itself, +j, +k and
// +jk; adjacency bit t says so.  Where a bit is clear the neighbour is read
// from its column; the lanes whose bits are set read the column's first entry
// instead (one shared line), so no lane waits on a branch.
__device__ __forceinline__ void eso_slots_adj(const EsoParams& p, uint32_t cell, uint32_t (&slot)[8]) {
  const size_t n = p.n;
  const uint32_t m = p.adj[cell];
  const uint32_t j = (uint32_t)p.nbr[n + cell], k = (uint32_t)p.nbr[2 * n + cell], jk = (uint32_t)p.nbr[5 * n + cell];
  const uint32_t xi = (uint32_t)p.nbr[(m & 1u) ? 0 : cell];
  const uint32_t xij = (uint32_t)p.nbr[(m & 2u) ? 3 * n : 3 * n + cell];
  const uint32_t xik = (uint32_t)p.nbr[(m & 4u) ? 4 * n : 4 * n + cell];
  const uint32_t xijk = (uint32_t)p.nbr[(m & 8u) ? 6 * n : 6 * n + cell];
  slot[0] = cell;
  slot[1] = (m & 1u) ? cell + 1 : xi;
  slot[2] = j;
  slot[3] = k;
  slot[4] = (m & 2u) ? j + 1 : xij;
  slot[5] = (m & 4u) ? k + 1 : xik;
  slot[6] = jk;
  slot[7] = (m & 8u) ? jk + 1 : xijk;
}

__global__ void k_eso_adjacency(const int* __restrict__ nbr, uint32_t n, uint8_t* __restrict__ adj) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= n) return;
  const size_t N = n;
  const int j = nbr[N + c], k = nbr[2 * N + c], jk = nbr[5 * N + c];
  uint8_t m = 0;
  if (nbr[c] == (int)c + 1) m |= 1;
  if (nbr[3 * N + c] == j + 1) m |= 2;
  if (nbr[4 * N + c] == k + 1) m |= 4;
  if (nbr[6 * N + c] == jk + 1) m |= 8;
  adj[c] = m;
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
  eso_slots_adj(p, cell, slot);
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
  if (p.rec) {
    const int r = rec_index(p.rec_mask, p.rec_base, cell);
    if (r >= 0) rec_store<St>(p.rec, (uint32_t)r, v);
  }
  if (p.srec) {
    const int r = rec_index(p.srec_mask, p.srec_base, cell);
    if (r >= 0) rec_store<St>(p.srec, (uint32_t)r, v);
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

// Slot of population q in the write maps (ESO_WRITE) for a q that varies
// across lanes: the table packed three bits per entry.
constexpr uint64_t pack3(const int (&t)[27], int lo, int hi) {
  uint64_t r = 0;
  for (int q = hi - 1; q >= lo; --q) r = (r << 3) | (uint64_t)t[q];
  return r;
}
constexpr uint64_t ESO_WRITE_LO = pack3(ESO_WRITE, 0, 21), ESO_WRITE_HI = pack3(ESO_WRITE, 21, 27);
__device__ __forceinline__ int write_slot_of(int q) {
  return q < 21 ? (int)((ESO_WRITE_LO >> (3 * q)) & 7u) : (int)((ESO_WRITE_HI >> (3 * (q - 21))) & 7u);
}

// Storage index of population q (any lane-varying q) of `cell` as its last
// update left it (the element eso_get reads).
template <bool FLIP>
__device__ __forceinline__ size_t eso_at(const EsoParams& p, uint32_t cell, int q) {
  const int ws = write_slot_of(q);
  const uint32_t slot = ws == 0 ? cell : (uint32_t)p.nbr[(size_t)(ws - 1) * p.n + cell];
  return (size_t)(FLIP ? opp(q) : q) * p.n + slot;
}

// writeToCoarseGridInterface(DIAD): include/gridRefinementFunctions.h:540-622.
// Eight lanes per interface cell.  Lane w owns populations 4w .. 4w+3 (lane
// 7 none): it sums them over the eight fine children (the read cell and its
// +i +j +k +ij +ik +jk +ijk neighbours) in child order, the reference's
// sequential order, reading each child's record (one 8- or 16-byte load) or,
// for a child without one, its slots.  27 lane shuffles then give every lane
// the 27 averages, each lane rescales fine -> coarse (identical in every
// lane) and writes its own populations of the coarse cell.
template <class St, bool FC, bool FF>
__global__ __launch_bounds__(256) void k_eso_to_coarse(typename St::T* __restrict__ fc, EsoParams pc,
                                                       const typename St::T* __restrict__ ff, EsoParams pf,
                                                       const int* __restrict__ write, const int* __restrict__ read,
                                                       uint32_t m, const int* __restrict__ child_rec) {
  typedef typename St::T T;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t g = t >> 3, w = t & 7;
  const bool valid = g < m;  // the eight lanes of a group agree
  const int q0 = 4 * (int)w;
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  // the coarse cell and its neighbour slots, loaded ahead of the children so
  // that the stores at the end wait for nothing
  const uint32_t wc = valid ? (uint32_t)write[g] : 0u;
  uint32_t slot[8];
  if (valid) eso_slots<St>(pc, wc, slot);
  if (valid && w < 7) {
    // the eight children's records (two 16-byte loads), -1 for a child without one
    int r[8];
    if (child_rec) {
      typedef int i4 __attribute__((ext_vector_type(4)));
      const i4 a = *(const i4*)(child_rec + 8 * (size_t)g);
      const i4 b = *(const i4*)(child_rec + 8 * (size_t)g + 4);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int ac = a[c], bc = b[c];
        r[c] = ac;
        r[c + 4] = bc;
      }
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) r[c] = -1;
    }
    bool all = true;
#pragma unroll
    for (int c = 0; c < 8; ++c) all &= r[c] >= 0;
    float x[8][4];
    if (all) {
      // every child has a record: eight independent loads, all in flight
      // (elements q0 .. q0+3; the padding beyond 26 is never used)
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const T* e = (const T*)pf.rec + (size_t)r[c] * ESO_REC_STRIDE + q0;
        if constexpr (sizeof(T) == 2) {
          typedef uint32_t u2 __attribute__((ext_vector_type(2)));
          const u2 y = *(const u2*)e;
          const uint32_t y0 = y[0], y1 = y[1];
          x[c][0] = St::load(__builtin_bit_cast(T, (uint16_t)(y0 & 0xffffu)));
          x[c][1] = St::load(__builtin_bit_cast(T, (uint16_t)(y0 >> 16)));
          x[c][2] = St::load(__builtin_bit_cast(T, (uint16_t)(y1 & 0xffffu)));
          x[c][3] = St::load(__builtin_bit_cast(T, (uint16_t)(y1 >> 16)));
        } else {
          typedef uint32_t u4 __attribute__((ext_vector_type(4)));
          const u4 y = *(const u4*)e;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t yk = y[k];
            x[c][k] = St::load(__builtin_bit_cast(T, yk));
          }
        }
      }
    } else {
      // some child without a record: every child from its slots
      const uint32_t rc = (uint32_t)read[g];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const uint32_t child = c == 0 ? rc : (uint32_t)pf.nbr[(size_t)(c - 1) * pf.n + rc];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[c][k] = q0 + k < 27 ? St::load(ff[eso_at<FF>(pf, child, q0 + k)]) : 0.f;
      }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k) s[k] += St::W_OFF ? x[c][k] + weight(q0 + k) : x[c][k];
  }
  float acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q] = __shfl(s[q & 3], q >> 2, 8) * 0.125f;
  if (!valid || w == 7) return;
  rescale(acc, false);
  // population q is stored by lane q / 4; q is a constant in each store, so
  // acc stays in registers
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if (w == (uint32_t)(q >> 2)) {
      float x = acc[q];
      if (St::W_OFF) x -= weight(q);
      fc[(size_t)(FC ? opp(q) : q) * pc.n + slot[ESO_WRITE[q]]] = St::store(x);
    }
  }
}

// writeToFineGridInterface(DIAD): include/gridRefinementFunctions.h:624-702.
// Eight lanes per interface cell: each reads the coarse cell (its record, or
// its slots; the same lines for all eight), rescales coarse -> fine
// (identical in every lane) and writes one child: the write cell or its +i +j
// +k +ij +ik +jk +ijk neighbour.
template <class St, bool FC, bool FF>
__global__ __launch_bounds__(256) void k_eso_to_fine(const typename St::T* __restrict__ fc, EsoParams pc,
                                                     typename St::T* __restrict__ ff, EsoParams pf,
                                                     const int* __restrict__ write, const int* __restrict__ read,
                                                     uint32_t m, const int* __restrict__ src_rec,
                                                     const uint32_t* __restrict__ child_slots) {
  typedef typename St::T T;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t g = t >> 3, w = t & 7;
  if (g >= m) return;
  // the child's slot cells: one contiguous 32-byte read from the pair's table
  // (issued first, in flight with the source), else through the columns
  uint32_t slot[8];
  if (child_slots) {
    const uint4* cs = (const uint4*)(child_slots + (size_t)t * 8);
    const uint4 a = cs[0], b = cs[1];
    slot[0] = a.x; slot[1] = a.y; slot[2] = a.z; slot[3] = a.w;
    slot[4] = b.x; slot[5] = b.y; slot[6] = b.z; slot[7] = b.w;
  }
  float v[27];
  const int r = src_rec ? src_rec[g] : -1;
  if (r >= 0) rec_load<St>(pc.srec, (uint32_t)r, v);
  else eso_get<St, FC>(fc, pc, (uint32_t)read[g], v);
  rescale(v, true);
  if (!child_slots) {
    const uint32_t wc = (uint32_t)write[g];
    const uint32_t child = w == 0 ? wc : (uint32_t)pf.nbr[(size_t)(w - 1) * pf.n + wc];
    eso_slots<St>(pf, child, slot);
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int fi = FF ? opp(q) : q;
    float x = v[q];
    if (St::W_OFF) x -= weight(q);
    ff[(size_t)fi * pf.n + slot[ESO_WRITE[q]]] = St::store(x);
  }
}

// The table k_eso_to_fine reads: for entry g and child w (child 0 the entry's
// fine cell, children 1-7 its +i, +j, +k, +ij, +ik, +jk, +ijk neighbours), the
// child's eight slot cells, in (entry, child, slot) order.
__global__ __launch_bounds__(256) void k_eso_c2f_slots(EsoParams pf, const int* __restrict__ write, uint32_t m,
                                                       uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t g = t >> 3, w = t & 7;
  if (g >= m) return;
  const uint32_t wc = (uint32_t)write[g];
  const uint32_t child = w == 0 ? wc : (uint32_t)pf.nbr[(size_t)(w - 1) * pf.n + wc];
  uint32_t slot[8];
  eso_slots<StF32>(pf, child, slot);
  uint4* o = (uint4*)(out + (size_t)t * 8);
  o[0] = make_uint4(slot[0], slot[1], slot[2], slot[3]);
  o[1] = make_uint4(slot[4], slot[5], slot[6], slot[7]);
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

// ------------------------------------------------------- interface records
// Which cells get a record: the children of every fine-to-coarse entry on
// the finer level (mark bit 0: child records) and the source of every
// coarse-to-fine entry on the coarser level (bit 1: source records).  Cells a
// transfer writes after the level's update (the level's own fine-to-coarse
// targets) are left to the slot path, as are cells the local update skips
// (bounceback).
__global__ void k_rec_mark_children(const int* __restrict__ read, int64_t m, const int* __restrict__ nbr, int64_t n,
                                    uint8_t* __restrict__ mark) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 8 * m) return;
  const int64_t g = t >> 3;
  const int w = (int)(t & 7);
  const int rc = read[g];
  const int child = w == 0 ? rc : nbr[(int64_t)(w - 1) * n + rc];
  mark[child] |= 1;
}

__global__ void k_rec_mark_list(const int* __restrict__ cells, int64_t m, uint8_t* __restrict__ mark, uint8_t bit,
                                bool set) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= m) return;
  if (set) mark[cells[t]] |= bit;
  else mark[cells[t]] = 0;
}

// one wave per 64 cells: the record mask and the record count of the group
// for the cells whose mark has `bit`
__global__ __launch_bounds__(256) void k_rec_mask(const uint8_t* __restrict__ mark, uint8_t bit,
                                                  const uint8_t* __restrict__ flags, int64_t n,
                                                  uint64_t* __restrict__ mask, uint32_t* __restrict__ count) {
  const int64_t cell = blockIdx.x * 256ll + threadIdx.x;
  bool keep = false;
  if (cell < n) keep = (mark[cell] & bit) && (!flags || (flags[cell] & 7) != KIND_BOUNCEBACK);
  const uint64_t b = __ballot(keep);
  if ((threadIdx.x & 63) == 0 && cell < n) {
    mask[cell >> 6] = b;
    count[cell >> 6] = (uint32_t)__popcll(b);
  }
}

// exclusive prefix sum of nw counts in one workgroup of 1024 threads, each
// owning a contiguous chunk; total[0] = the sum
__global__ __launch_bounds__(1024) void k_rec_scan(const uint32_t* __restrict__ count, int64_t nw,
                                                   uint32_t* __restrict__ base, uint32_t* __restrict__ total) {
  __shared__ uint32_t part[1024];
  const int64_t chunk = (nw + 1023) / 1024;
  const int64_t lo = threadIdx.x * chunk, hi = min(nw, lo + chunk);
  uint32_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += count[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < 1024; ++i) {
      const uint32_t x = part[i];
      part[i] = acc;
      acc += x;
    }
    total[0] = acc;
  }
  __syncthreads();
  uint32_t acc = part[threadIdx.x];
  for (int64_t i = lo; i < hi; ++i) {
    base[i] = acc;
    acc += count[i];
  }
}

__global__ void k_rec_index_children(const int* __restrict__ read, int64_t m, const int* __restrict__ nbr, int64_t n,
                                     const uint64_t* __restrict__ mask, const uint32_t* __restrict__ base,
                                     int* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 8 * m) return;
  const int64_t g = t >> 3;
  const int w = (int)(t & 7);
  const int rc = read[g];
  const int child = w == 0 ? rc : nbr[(int64_t)(w - 1) * n + rc];
  out[t] = rec_index(mask, base, (uint32_t)child);
}

__global__ void k_rec_index_list(const int* __restrict__ cells, int64_t m, const uint64_t* __restrict__ mask,
                                 const uint32_t* __restrict__ base, int* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= m) return;
  out[t] = rec_index(mask, base, (uint32_t)cells[t]);
}

// Records that differ from the populations the cell's slots hold now (both
// in the storage words, so a fresh record matches bit for bit).
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_rec_verify(const typename St::T* __restrict__ f, EsoParams p,
                                                    const void* __restrict__ rec, const uint64_t* __restrict__ mask,
                                                    const uint32_t* __restrict__ base,
                                                    unsigned long long* __restrict__ bad) {
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  if (cell >= p.n) return;
  const int r = rec_index(mask, base, cell);
  if (r < 0) return;
  float a[27], b[27];
  eso_get<St, FLIP>(f, p, cell, a);
  rec_load<St>(rec, (uint32_t)r, b);
  int nbad = 0;
#pragma unroll
  for (int q = 0; q < 27; ++q) nbad += __builtin_bit_cast(uint32_t, a[q]) != __builtin_bit_cast(uint32_t, b[q]);
  if (nbad) atomicAdd(bad, 1ull);
}

// Per block: cells whose populations (as their last update left them) hold a
// non-finite value or sum to rho <= 0, and the least and greatest density of
// the other cells.
template <class St, bool FLIP>
__global__ __launch_bounds__(256) void k_eso_health(const typename St::T* __restrict__ f, EsoParams p,
                                                    uint32_t* __restrict__ bad, float2* __restrict__ range) {
  __shared__ uint32_t sb[256];
  __shared__ float slo[256], shi[256];
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  uint32_t b = 0;
  float lo = INFINITY, hi = -INFINITY;
  if (cell < p.n) {
    float v[27];
    eso_get<St, FLIP>(f, p, cell, v);
    float rho = 0.f;
    bool finite = true;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      finite = finite && __builtin_isfinite(v[q]);
      rho += v[q];
    }
    if (!finite || !(rho > 0.f)) {
      b = 1;
    } else {
      lo = rho;
      hi = rho;
    }
  }
  sb[threadIdx.x] = b;
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sb[threadIdx.x] += sb[threadIdx.x + w];
      slo[threadIdx.x] = fminf(slo[threadIdx.x], slo[threadIdx.x + w]);
      shi[threadIdx.x] = fmaxf(shi[threadIdx.x], shi[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bad[blockIdx.x] = sb[0];
    range[blockIdx.x] = make_float2(slo[0], shi[0]);
  }
}

}  // namespace psl

using namespace psl;

// child_records: whether a local update with these parameters writes the
// child records (the source records are written whenever they exist)
static EsoParams eso_params(const psl_diad* d, bool child_records = true) {
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
  p.adj = d->adj;
  p.bc = d->bc_dev;
  p.rec = child_records && d->child.n ? d->child.rec : nullptr;
  p.rec_mask = d->child.mask;
  p.rec_base = d->child.base;
  p.srec = d->source.n ? d->source.rec : nullptr;
  p.srec_mask = d->source.mask;
  p.srec_base = d->source.base;
  return p;
}

int psl::diad_build_adjacency(psl_diad* d) {
  if (d->adj) (void)hipFree(d->adj);
  d->adj = nullptr;
  PSL_HIP_TRY(hipMalloc(&d->adj, (size_t)d->n));
  k_eso_adjacency<<<(unsigned)((d->n + 255) / 256), 256>>>(d->nbr, (uint32_t)d->n, d->adj);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

// Record versions are unique across all grids of the process, so a key can
// never match a level that was destroyed and replaced at the same address.
uint64_t psl::diad_next_version() {
  static std::atomic<uint64_t> next{1};
  return next.fetch_add(1) + 1;
}

// The populations changed by something other than a local update: no record
// of the level holds them until the next update.
static void records_stale(psl_diad* d) {
  d->child.fresh = false;
  d->source.fresh = false;
}

static void records_changed(psl_diad* d) {
  d->rec_version = diad_next_version();
  records_stale(d);
  d->rec_key = 0;  // no transfer uses this level's records until they are rebuilt
}

static void free_record_set(psl_diad::Records& r) {
  for (void* a : {(void*)r.mask, (void*)r.base, r.rec})
    if (a) (void)hipFree(a);
  r = psl_diad::Records();
}

static void free_records(psl_diad* d) {
  free_record_set(d->child);
  free_record_set(d->source);
  for (void* a : {(void*)d->f2c_child_rec, (void*)d->c2f_src_rec})
    if (a) (void)hipFree(a);
  d->f2c_child_rec = nullptr;
  d->c2f_src_rec = nullptr;
  d->rec_key = 0;
  d->rec_fine = nullptr;
}

// Mask, record bases and record storage of the cells of one level whose mark
// has `bit` (stream-ordered on st; synchronizes st for the record count).
static int build_record_set(psl_diad* d, const uint8_t* mark, uint8_t bit, hipStream_t st, psl_diad::Records& r) {
  const int64_t nw = (d->n + 63) / 64;
  uint32_t* count = nullptr;
  uint32_t* total = nullptr;
  uint32_t ntot = 0;
  int rc = PSL_OK;
  if (hipMalloc(&r.mask, (size_t)nw * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&r.base, (size_t)nw * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&count, (size_t)nw * sizeof(uint32_t)) != hipSuccess || hipMalloc(&total, sizeof(uint32_t)) != hipSuccess) {
    rc = PSL_E_HIP;
  } else {
    k_rec_mask<<<(unsigned)((d->n + 255) / 256), 256, 0, st>>>(mark, bit, d->flags, d->n, r.mask, count);
    k_rec_scan<<<1, 1024, 0, st>>>(count, nw, r.base, total);
    if (hipMemcpyAsync(&ntot, total, sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = PSL_E_HIP;
  }
  if (count) (void)hipFree(count);
  if (total) (void)hipFree(total);
  r.n = ntot;
  if (rc == PSL_OK && ntot && hipMalloc(&r.rec, (size_t)ntot * ESO_REC_STRIDE * storage_bytes(d->storage)) != hipSuccess)
    rc = PSL_E_HIP;
  return rc;
}

// What the records of a level list depend on: every level's version and the
// level order.
static uint64_t records_key(psl_diad* const* lv, int nl, int l) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) { h = (h ^ x) * 1099511628211ull; };
  mix((uint64_t)nl);
  mix((uint64_t)l);
  for (int k = 0; k < nl; ++k) {
    mix((uint64_t)(uintptr_t)lv[k]);
    mix(lv[k]->rec_version);
  }
  return h;
}

// (Re)build the interface records of every level of the list when anything
// they depend on changed since they were built.
static int diad_prepare_records(psl_diad* const* lv, int nl) {
  if (nl < 2) return PSL_OK;
  bool stale = false;
  for (int l = 0; l < nl; ++l) stale |= lv[l]->rec_key != records_key(lv, nl, l);
  if (!stale) return PSL_OK;
  const hipStream_t st = lv[0]->stream;
  std::vector<uint8_t*> mark(nl, nullptr);
  int rc = PSL_OK;
  auto grid1d = [](int64_t k) { return (unsigned)((k + 255) / 256); };
  for (int l = 0; l < nl && rc == PSL_OK; ++l) {
    free_records(lv[l]);
    if (hipMalloc(&mark[l], (size_t)lv[l]->n) != hipSuccess || hipMemsetAsync(mark[l], 0, (size_t)lv[l]->n, st) != hipSuccess)
      rc = PSL_E_HIP;
  }
  for (int l = 0; l + 1 < nl && rc == PSL_OK; ++l) {
    psl_diad *c = lv[l], *f = lv[l + 1];
    if (c->n_f2c) k_rec_mark_children<<<grid1d(8 * c->n_f2c), 256, 0, st>>>(c->f2c_read, c->n_f2c, f->nbr, f->n, mark[l + 1]);
    if (c->n_c2f) k_rec_mark_list<<<grid1d(c->n_c2f), 256, 0, st>>>(c->c2f_read, c->n_c2f, mark[l], 2, true);
  }
  for (int l = 0; l < nl && rc == PSL_OK; ++l) {
    psl_diad* d = lv[l];
    if (l + 1 < nl && d->n_f2c) k_rec_mark_list<<<grid1d(d->n_f2c), 256, 0, st>>>(d->f2c_write, d->n_f2c, mark[l], 0, false);
    rc = build_record_set(d, mark[l], 1, st, d->child);
    if (rc == PSL_OK) rc = build_record_set(d, mark[l], 2, st, d->source);
  }
  for (int l = 0; l + 1 < nl && rc == PSL_OK; ++l) {
    psl_diad *c = lv[l], *f = lv[l + 1];
    c->rec_fine = f;
    if (c->n_f2c) {
      if (hipMalloc(&c->f2c_child_rec, 8 * (size_t)c->n_f2c * sizeof(int)) != hipSuccess) rc = PSL_E_HIP;
      else
        k_rec_index_children<<<grid1d(8 * c->n_f2c), 256, 0, st>>>(c->f2c_read, c->n_f2c, f->nbr, f->n, f->child.mask,
                                                                  f->child.base, c->f2c_child_rec);
    }
    if (c->n_c2f && rc == PSL_OK) {
      if (hipMalloc(&c->c2f_src_rec, (size_t)c->n_c2f * sizeof(int)) != hipSuccess) rc = PSL_E_HIP;
      else
        k_rec_index_list<<<grid1d(c->n_c2f), 256, 0, st>>>(c->c2f_read, c->n_c2f, c->source.mask, c->source.base,
                                                           c->c2f_src_rec);
    }
  }
  if (rc == PSL_OK && hipGetLastError() != hipSuccess) rc = PSL_E_HIP;
  if (rc == PSL_OK && hipStreamSynchronize(st) != hipSuccess) rc = PSL_E_HIP;
  for (uint8_t* m : mark)
    if (m) (void)hipFree(m);
  if (rc) {
    for (int l = 0; l < nl; ++l) free_records(lv[l]);
    set_error("DIAD interface records: setup failed");
    return rc;
  }
  for (int l = 0; l < nl; ++l) lv[l]->rec_key = records_key(lv, nl, l);
  return PSL_OK;
}

template <class St>
static int eso_update(psl_diad* d, bool child_records) {
  const EsoParams p = eso_params(d, child_records);
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  if (d->flipper) k_eso_update<St, true><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p);
  else k_eso_update<St, false><<<nb, 256, 0, d->stream>>>((typename St::T*)d->f, p);
  PSL_LAUNCH_CHECK();
  // the records this update wrote hold its values; the others are behind
  d->child.fresh = p.rec != nullptr;
  d->source.fresh = p.srec != nullptr;
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

// The coarse level's table of child slot cells for the fine level f (256 B per
// coarse-to-fine entry), built on first use and again whenever either level's
// tables change.  Returns null where it cannot be allocated: the transfer then
// reads the neighbour columns.
static const uint32_t* c2f_child_slots(psl_diad* c, psl_diad* f) {
  const uint64_t key = (c->rec_version * 0x9E3779B97F4A7C15ull) ^ (f->rec_version << 1) ^ (uint64_t)(uintptr_t)f;
  if (c->c2f_slots && c->c2f_slots_key == key) return c->c2f_slots;
  if (c->c2f_slots) (void)hipFree(c->c2f_slots);
  c->c2f_slots = nullptr;
  c->c2f_slots_key = 0;
  if (hipMalloc(&c->c2f_slots, (size_t)c->n_c2f * 64 * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    c->c2f_slots = nullptr;
    return nullptr;
  }
  k_eso_c2f_slots<<<(unsigned)((8 * c->n_c2f + 255) / 256), 256, 0, c->stream>>>(
      eso_params(f, false), c->c2f_write, (uint32_t)c->n_c2f, c->c2f_slots);
  if (hipGetLastError() != hipSuccess) {
    (void)hipFree(c->c2f_slots);
    c->c2f_slots = nullptr;
    return nullptr;
  }
  c->c2f_slots_key = key;
  return c->c2f_slots;
}

template <class St, bool FC, bool FF>
static int eso_comm_t(psl_diad* c, psl_diad* f) {
  typedef typename St::T T;
  const EsoParams pc = eso_params(c), pf = eso_params(f);
  // records are read only when both levels' last updates wrote them for
  // tables built with this pair
  const bool built = c->rec_key && f->rec_key && c->rec_fine == f;
  const int* child_rec = built && f->child.fresh ? c->f2c_child_rec : nullptr;
  const int* src_rec = built && c->source.fresh ? c->c2f_src_rec : nullptr;
  if (c->n_f2c) {
    k_eso_to_coarse<St, FC, FF><<<(unsigned)((8 * c->n_f2c + 255) / 256), 256, 0, c->stream>>>(
        (T*)c->f, pc, (const T*)f->f, pf, c->f2c_write, c->f2c_read, (uint32_t)c->n_f2c, child_rec);
    PSL_LAUNCH_CHECK();
  }
  if (c->n_c2f) {
    k_eso_to_fine<St, FC, FF><<<(unsigned)((8 * c->n_c2f + 255) / 256), 256, 0, c->stream>>>(
        (const T*)c->f, pc, (T*)f->f, pf, c->c2f_write, c->c2f_read, (uint32_t)c->n_c2f, src_rec,
        c2f_child_slots(c, f));
    PSL_LAUNCH_CHECK();
    records_stale(f);  // fine cells were rewritten after the fine update
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

template <class St>
static int rec_verify(psl_diad* d, unsigned long long* bad) {
  const EsoParams p = eso_params(d);
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  for (const psl_diad::Records* r : {&d->child, &d->source}) {
    if (!r->n) continue;
    if (d->flipper)
      k_rec_verify<St, true><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, r->rec, r->mask, r->base, bad);
    else
      k_rec_verify<St, false><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, r->rec, r->mask, r->base, bad);
  }
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

template <class St>
static int eso_health(psl_diad* d, uint32_t nb, uint32_t* bad, float2* range) {
  const EsoParams p = eso_params(d);
  if (d->flipper) k_eso_health<St, true><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, bad, range);
  else k_eso_health<St, false><<<nb, 256, 0, d->stream>>>((const typename St::T*)d->f, p, bad, range);
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
  d->rec_version = diad_next_version();
  d->info = *info;
  d->storage = storage;
  d->n = n;
  const size_t fb = 27ull * n * storage_bytes(storage);
  if (hipMalloc(&d->f, fb) != hipSuccess || hipMalloc(&d->bc_dev, 4 * sizeof(float4)) != hipSuccess ||
      hipMemset(d->bc_dev, 0, 4 * sizeof(float4)) != hipSuccess || hipMalloc(&d->ijk, 3ull * n * sizeof(int)) != hipSuccess ||
      hipMalloc(&d->nbr, 7ull * n * sizeof(int)) != hipSuccess ||
      hipMemcpy(d->ijk, ijk3, 3ull * n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->nbr, nbr7, 7ull * n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      diad_build_adjacency(d) != PSL_OK) {
    set_error("DIAD allocation / upload failed");
    if (d->f) (void)hipFree(d->f);
    if (d->bc_dev) (void)hipFree(d->bc_dev);
    if (d->ijk) (void)hipFree(d->ijk);
    if (d->nbr) (void)hipFree(d->nbr);
    if (d->adj) (void)hipFree(d->adj);
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
  if (d->adj) (void)hipFree(d->adj);
  if (d->c2f_slots) (void)hipFree(d->c2f_slots);
  if (d->flags) (void)hipFree(d->flags);
  if (d->bc_dev) (void)hipFree(d->bc_dev);
  for (int* a : {d->c2f_read, d->c2f_write, d->f2c_write, d->f2c_read, d->outlet, d->kmin, d->moving, d->forced})
    if (a) (void)hipFree(a);
  if (d->moving_u) (void)hipFree(d->moving_u);
  if (d->forced_u) (void)hipFree(d->forced_u);
  if (d->partials) (void)hipFree(d->partials);
  free_records(d);
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
  records_changed(d);
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
  records_stale(d);
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
  int rc = PSL_ESO_DISPATCH(eso_update, d, true);
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
  records_stale(d);
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
  records_changed(coarse);
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
  records_changed(d);
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
  // the outlet stores where the next step reads, which includes populations
  // of the outlet cells' neighbours: the level's records no longer hold them
  records_stale(d);
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
// (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:200-209).  `last`:
// this is the second of the level's two steps, after which the coarser
// level's transfer reads the child records, so only this update writes them.
static int diad_recurse(psl_diad* const* lv, int nl, int l, bool last) {
  psl_diad* d = lv[l];
  d->flipper = !d->flipper;
  int rc = PSL_ESO_DISPATCH(eso_update, d, last);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_moving, d);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_forced, d);
  if (rc != PSL_OK || l == nl - 1) return rc;
  for (int s = 0; s < 2 && rc == PSL_OK; ++s) rc = diad_recurse(lv, nl, l + 1, s == 1);
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_comm, d, lv[l + 1]);
  return rc;
}

int psl_diad_verify_interface_records(psl_diad* d, int64_t* mismatches) {
  PSL_CHECK_ARG(d && mismatches, "null argument");
  *mismatches = 0;
  if (!d->rec_key || !(d->child.n + d->source.n)) return PSL_OK;
  unsigned long long* bad = nullptr;
  PSL_HIP_TRY(hipMalloc(&bad, sizeof(unsigned long long)));
  int rc = hipMemsetAsync(bad, 0, sizeof(unsigned long long), d->stream) == hipSuccess ? PSL_OK : PSL_E_HIP;
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(rec_verify, d, bad);
  unsigned long long h = 0;
  if (rc == PSL_OK && (hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
                       hipStreamSynchronize(d->stream) != hipSuccess))
    rc = PSL_E_HIP;
  (void)hipFree(bad);
  *mismatches = (int64_t)h;
  return rc;
}

int psl_diad_state_bytes(const psl_diad* d, int64_t* bytes) {
  PSL_CHECK_ARG(d && bytes, "null argument");
  *bytes = (int64_t)sizeof(StateHeader) + 27 * d->n * (int64_t)storage_bytes(d->storage);
  return PSL_OK;
}

int psl_diad_save_state(psl_diad* d, void* host, int64_t bytes) {
  PSL_CHECK_ARG(d && host, "null argument");
  int64_t want = 0;
  (void)psl_diad_state_bytes(d, &want);
  PSL_CHECK_ARG(bytes == want, "state buffer size differs from psl_diad_state_bytes");
  StateHeader h;
  state_header_init(h, STATE_DIAD, d->storage, d->n, d->info);
  h.flipper = d->flipper ? 1 : 0;
  memcpy(host, &h, sizeof(h));
  PSL_HIP_TRY(hipStreamSynchronize(d->stream));
  PSL_HIP_TRY(hipMemcpy((char*)host + sizeof(h), d->f, (size_t)(want - (int64_t)sizeof(h)), hipMemcpyDeviceToHost));
  return PSL_OK;
}

int psl_diad_load_state(psl_diad* d, const void* host, int64_t bytes) {
  PSL_CHECK_ARG(d && host && bytes >= (int64_t)sizeof(StateHeader), "bad state arguments");
  StateHeader h;
  memcpy(&h, host, sizeof(h));
  int64_t want = 0;
  (void)psl_diad_state_bytes(d, &want);
  const int rc = state_header_check(h, STATE_DIAD, d->storage, d->n, d->info, bytes, want);
  if (rc) return rc;
  PSL_CHECK_ARG(h.sections == 0 && (h.flipper == 0 || h.flipper == 1), "malformed esotwist state");
  PSL_HIP_TRY(hipStreamSynchronize(d->stream));
  PSL_HIP_TRY(hipMemcpy(d->f, (const char*)host + sizeof(h), (size_t)(want - (int64_t)sizeof(h)),
                        hipMemcpyHostToDevice));
  d->flipper = h.flipper != 0;
  records_stale(d);
  return PSL_OK;
}

int psl_diad_check_health(psl_diad* d, int64_t* bad_cells, double* rho_min, double* rho_max) {
  PSL_CHECK_ARG(d && bad_cells, "null argument");
  const uint32_t nb = (uint32_t)((d->n + 255) / 256);
  uint32_t* bad = nullptr;
  float2* range = nullptr;
  PSL_HIP_TRY(hipMalloc(&bad, (size_t)nb * sizeof(uint32_t)));
  int rc = hipMalloc(&range, (size_t)nb * sizeof(float2)) == hipSuccess ? PSL_OK : PSL_E_HIP;
  if (rc == PSL_OK) rc = PSL_ESO_DISPATCH(eso_health, d, nb, bad, range);
  std::vector<uint32_t> hb(rc == PSL_OK ? nb : 0);
  std::vector<float2> hr(rc == PSL_OK ? nb : 0);
  if (rc == PSL_OK &&
      (hipMemcpyAsync(hb.data(), bad, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
       hipMemcpyAsync(hr.data(), range, nb * sizeof(float2), hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
       hipStreamSynchronize(d->stream) != hipSuccess))
    rc = PSL_E_HIP;
  (void)hipFree(bad);
  if (range) (void)hipFree(range);
  if (rc) {
    set_error("health check failed");
    return rc;
  }
  int64_t nbad = 0;
  double lo = INFINITY, hi = -INFINITY;
  for (uint32_t b = 0; b < nb; ++b) {
    nbad += hb[b];
    lo = std::min(lo, (double)hr[b].x);
    hi = std::max(hi, (double)hr[b].y);
  }
  *bad_cells = nbad;
  if (rho_min) *rho_min = lo;
  if (rho_max) *rho_max = hi;
  return PSL_OK;
}

int psl_diad_interface_records(const psl_diad* d, int64_t* count, int* fresh) {
  PSL_CHECK_ARG(d && count, "null argument");
  *count = d->rec_key ? d->child.n + d->source.n : 0;
  // every record of the level holds the last local update's values
  if (fresh)
    *fresh = d->rec_key && (!d->child.n || d->child.fresh) && (!d->source.n || d->source.fresh) ? 1 : 0;
  return PSL_OK;
}

int psl_diad_update_grid(psl_diad* const* levels, int nlevels, int nsteps) {
  PSL_CHECK_ARG(levels && nlevels >= 1 && nsteps >= 0, "bad level list");
  for (int l = 0; l < nlevels; ++l) {
    PSL_CHECK_ARG(levels[l], "null level");
    PSL_CHECK_ARG(levels[l]->storage == levels[0]->storage, "levels must share the storage type");
    PSL_CHECK_ARG(levels[l]->stream == levels[0]->stream, "levels must share the stream");
  }
  const int rc0 = diad_prepare_records(levels, nlevels);
  if (rc0 != PSL_OK) return rc0;
  for (int s = 0; s < nsteps; ++s) {
    const int rc = diad_recurse(levels, nlevels, 0, true);
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
