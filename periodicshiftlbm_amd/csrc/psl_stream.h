// Shifted-stream addressing shared by the single-grid and batched step kernels.
#pragma once

#include "psl_internal.h"

namespace psl {

// --------------------------------------------------------------- addressing
// Population q of cell c lives at element (c + shift[q]) mod N of row q.  For
// each q the host passes two row bases, lo[q] = row + shift[q] (cells below
// the wrap threshold thr[q] = N - shift[q]) and hi[q] = lo[q] - N (cells at
// or above it), so that the element is base + c either way.  A wave whose 64
// cells lie on one side of every threshold picks each base with one scalar
// compare/select and addresses all 27 populations with the same per-lane byte
// offset c * sizeof(T) (SGPR base + VGPR offset loads, no per-q vector
// arithmetic).  At most one wave per q straddles a threshold; those waves
// select the base per lane.
struct StepParams {
  const void* lo[27];
  const void* hi[27];
  uint32_t thr[27];
  uint32_t n;
  uint32_t xcd_remap;  // non-zero: consecutive cell ranges stay on one XCD
  float omega;  // 1/tau, used when cs == 0 and there is no per-cell cs
  float tau;
  float cs;
  const float* cs_cell;
  const uint8_t* flags;
};

__device__ __forceinline__ uint32_t wave_first_cell(uint32_t block_first) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  return block_first + wave * 64u;
}

// True when some population's wrap threshold falls inside this wave: one
// lane per population tests it and a ballot makes the answer wave-uniform.
template <class Tab>
__device__ __forceinline__ bool wave_straddles(const Tab& p, uint32_t w_first) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w_last = min(w_first + 63u, p.n - 1u);
  bool s = false;
  if (lane < 27u) {
    const uint32_t t = p.thr[lane];
    s = (w_first < t) && (w_last >= t);
  }
  return __ballot(s) != 0;
}

// Tuning bits of the bulk kernel (chosen per grid by psl_grid_autotune).
enum : int { TUNE_NT_LOAD = 1, TUNE_NT_STORE = 2 };

template <class T, bool NT>
__device__ __forceinline__ T ld(const T* p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <class T, bool NT>
__device__ __forceinline__ void st(T* p, T v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <class St, int TUNE, class Tab>
__device__ __forceinline__ void load27(const Tab& p, uint32_t w_first, bool straddle, uint32_t cell,
                                       float (&v)[27]) {
  typedef typename St::T T;
  constexpr bool NT = (TUNE & TUNE_NT_LOAD) != 0;
  if (!straddle) {
    const uint32_t off = cell * (uint32_t)sizeof(T);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const char* base = (const char*)(w_first >= p.thr[q] ? p.hi[q] : p.lo[q]);
      v[q] = St::load(ld<T, NT>((const T*)(base + off)));
    }
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const T* base = (const T*)(cell >= p.thr[q] ? p.hi[q] : p.lo[q]);
      v[q] = St::load(ld<T, NT>(base + cell));
    }
  }
}

template <class St, int TUNE, class Tab>
__device__ __forceinline__ void store27(const Tab& p, uint32_t w_first, bool straddle, uint32_t cell,
                                        const float (&v)[27]) {
  typedef typename St::T T;
  constexpr bool NT = (TUNE & TUNE_NT_STORE) != 0;
  if (!straddle) {
    const uint32_t off = cell * (uint32_t)sizeof(T);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      char* base = (char*)(w_first >= p.thr[q] ? p.hi[q] : p.lo[q]);
      st<T, NT>((T*)(base + off), St::store(v[q]));
    }
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      T* base = (T*)(cell >= p.thr[q] ? p.hi[q] : p.lo[q]);
      st<T, NT>(base + cell, St::store(v[q]));
    }
  }
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8),
// each with its own L2.  The PS streams are misaligned, so the cache line at
// every block boundary of every stream is shared by neighbouring blocks;
// remapping gives each XCD one contiguous range of blocks, so those lines are
// fetched into one L2 only.  The map is a bijection of [0, nb).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t xcd = b & 7u, idx = b >> 3, per = nb >> 3, rem = nb & 7u;
  return xcd * per + min(xcd, rem) + idx;
}

}  // namespace psl
