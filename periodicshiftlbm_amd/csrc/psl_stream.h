// Shifted-stream addressing and the per-cell update paths shared by the
// single-grid and batched step kernels.
#pragma once

#include "psl_internal.h"

#include <type_traits>

// FMA formation fixed by the source expressions (see psl_math.h).
#pragma clang fp contract(on)

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
  // learned closure held between network evaluations: per-cell network
  // output y (FP16 bits); fluid cells relax with tau = 3 (nu + nut_scale
  // max(y, 0)) + 1/2 and no Smagorinsky rate (read by the HELD update only).
  const uint16_t* nut_y;
  float nu;
  float nut_scale;
};

// tau of a cell under the learned closure's eddy viscosity (network output y)
__device__ __forceinline__ Relax closure_relax(float nu, float nut_scale, float y) {
  const float tau = 3.f * (nu + nut_scale * fmaxf(y, 0.f)) + 0.5f;
  return Relax{1.f / tau, tau, 0.f};
}

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

// The stream bases may come from device memory (batched grids), where the
// compiler cannot see that they point to global memory; the explicit global
// address space keeps global_load / global_store instead of flat accesses.
#define PSL_GLOBAL __attribute__((address_space(1)))

template <class T, bool NT>
__device__ __forceinline__ T ld(const T* p) {
  const PSL_GLOBAL T* g = (const PSL_GLOBAL T*)p;
  if constexpr (NT) return __builtin_nontemporal_load(g);
  else return *g;
}
template <class T, bool NT>
__device__ __forceinline__ void st(T* p, T v) {
  PSL_GLOBAL T* g = (PSL_GLOBAL T*)p;
  if constexpr (NT) __builtin_nontemporal_store(v, g);
  else *g = v;
}

// A storage word as loaded: 16-bit words zero-extended to 32 bits by the load
// itself (so no masking is needed where only the low half is read).
template <class St>
using Word = typename std::conditional<sizeof(typename St::T) == 2, uint32_t, typename St::T>::type;

// The 27 storage words of a cell, unconverted.
template <class St, int TUNE, class Tab>
__device__ __forceinline__ void load27_raw(const Tab& p, uint32_t w_first, bool straddle, uint32_t cell,
                                           Word<St> (&r)[27]) {
  typedef typename St::T T;
  constexpr bool NT = (TUNE & TUNE_NT_LOAD) != 0;
  if (!straddle) {
    const uint32_t off = cell * (uint32_t)sizeof(T);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const char* base = (const char*)(w_first >= p.thr[q] ? p.hi[q] : p.lo[q]);
      r[q] = (Word<St>)ld<T, NT>((const T*)(base + off));
    }
    // keeps the loads in this branch: merged with the per-lane branch they
    // would all take 64-bit per-lane addresses
    __asm__ volatile("" ::: "memory");
    if constexpr (sizeof(T) == 2) {
      // the words leave the branch as 32-bit values (the load's zero
      // extension); otherwise the merge re-masks every 16-bit word
#pragma unroll
      for (int q = 0; q < 27; ++q) __asm__("" : "+v"(r[q]));
    }
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const T* base = (const T*)(cell >= p.thr[q] ? p.hi[q] : p.lo[q]);
      r[q] = (Word<St>)ld<T, NT>(base + cell);
    }
    __asm__ volatile("" ::: "memory");
    if constexpr (sizeof(T) == 2) {
#pragma unroll
      for (int q = 0; q < 27; ++q) __asm__("" : "+v"(r[q]));
    }
  }
}

template <class St>
__device__ __forceinline__ void words_to_float(const Word<St> (&r)[27], float (&v)[27]) {
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] = St::load((typename St::T)r[q]);
}

// v_fma_mix_f32 with f16 operands (the low 16 bits of their registers; the
// high bits are not read) and a +-1 multiplier: one rounding of the exact sum,
// the same value as adding the converted operands, without the conversions.
__device__ __forceinline__ float mix_add_hh(uint32_t a, uint32_t c) {  // a + c
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(c));
  return r;
}
__device__ __forceinline__ float mix_sub_hh(uint32_t c, uint32_t a) {  // c - a
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(c));
  return r;
}
__device__ __forceinline__ float mix_add_hf(uint32_t b, float z) {  // b + z
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(b), "v"(z));
  return r;
}

// raw_moments of the converted storage words; for FP16 storage the z pass
// reads the f16 words directly (bit-identical to converting first).
template <class St>
__device__ __forceinline__ Moments raw_moments_words(const Word<St> (&r)[27]) {
  if constexpr (std::is_same<St, StF16>::value) {
    float z0[3][3], z1[3][3], z2[3][3];
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) {
        const uint32_t a = r[LINE[x][y][0]], b = r[LINE[x][y][1]], c = r[LINE[x][y][2]];
        z2[x][y] = mix_add_hh(a, c);
        z0[x][y] = mix_add_hf(b, z2[x][y]);
        z1[x][y] = mix_sub_hh(c, a);
      }
    return moments_from_z(z0, z1, z2);
  } else {
    float v[27];
    words_to_float<St>(r, v);
    return raw_moments(v);
  }
}

template <class St, int TUNE, class Tab>
__device__ __forceinline__ void store27(const Tab& p, uint32_t w_first, bool straddle, uint32_t cell,
                                        const float (&v)[27]) {
  typedef typename St::T T;
  constexpr bool NT = (TUNE & TUNE_NT_STORE) != 0;
  // Storage words in 32-bit registers: for FP16, pairs (q, q + 1) converted by
  // one v_cvt_pk_f16_f32 (round to nearest even, as the single conversion);
  // word(q) takes the pair's half inside each store loop, so that the shift
  // of an odd q has one use and folds into global_store_short_d16_hi.
  uint32_t pk[27];
  if constexpr (std::is_same<St, StF16>::value) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int q = 0; q + 1 < 27; q += 2) {
      const f2 x = {v[q], v[q + 1]};
      pk[q] = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, h2));
    }
    pk[26] = St::store(v[26]);
  }
  auto word = [&](int q) -> T {
    if constexpr (std::is_same<St, StF16>::value) return (q & 1) ? (T)(pk[q - 1] >> 16) : (T)pk[q];
    else return St::store(v[q]);
  };
  if (!straddle) {
    const uint32_t off = cell * (uint32_t)sizeof(T);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      char* base = (char*)(w_first >= p.thr[q] ? p.hi[q] : p.lo[q]);
      st<T, NT>((T*)(base + off), word(q));
    }
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      T* base = (T*)(cell >= p.thr[q] ? p.hi[q] : p.lo[q]);
      st<T, NT>(base + cell, word(q));
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

// The same in chunks of K consecutive blocks dealt round-robin to the XCDs
// (blocks past the last whole round of 8 K keep their order): neighbouring
// blocks still share one L2, and every XCD gets a slice of every part of the
// launch.  Used by the batch, whose grids differ in boundary work, so that
// one contiguous eighth per XCD would load the XCDs unevenly.  A bijection
// of [0, nb).
template <uint32_t K>
__device__ __forceinline__ uint32_t xcd_block_chunked(uint32_t b, uint32_t nb) {
  const uint32_t full = nb / (8u * K) * (8u * K);
  if (b >= full) return b;
  const uint32_t x = b & 7u, i = b >> 3;
  return (x + 8u * (i / K)) * K + (i % K);
}

// ------------------------------------------------------- per-cell updates
// Box-face cells of one grid (include/applyLocalCellUpdate.h:46-67).
struct BoundaryParams {
  const uint32_t* list;  // box-face cells in cell order
  uint32_t n;            // list length
  int nx, ny, nz;
  const float4* bc;      // 4 given-value slots (rho, ux, uy, uz)
};

// Staging of box-face cells.  The bulk kernel already reads the lines that
// hold a box-face cell's 27 slots (its row neighbours share them), so with
// staging it also loads the listed cells and writes their full f as one
// contiguous record of STAGE_STRIDE floats at the cell's list index (the
// loaded values as the bulk path holds them: f, or f - w for the reduced
// storages, to which the boundary kernel adds w exactly as when it loads the
// slots itself, so both paths compute identical results).  The
// boundary kernel then reads the records instead of 27 strided slots, whose
// lines it would otherwise fetch again for 2-4 useful bytes each.  Words 27
// and 28 of a record hold the cell index and its flag byte (integer bits) and
// the record is padded to 128 bytes, one cache line: the boundary kernel's
// only per-cell load is its record.  base[w] is the number of listed cells
// below cell 64 w.
constexpr int STAGE_STRIDE = 32;
struct Stage {
  float* buf;
  const uint32_t* base;
};

// Arguments of the staged boundary kernel, kept compact (the slot of
// population q of cell c is (c + s[q]) mod n in row q of f) so that its few
// scalar loads are issued together at the start; with the 27 row pointers and
// thresholds of StepParams they were fetched one cache line after another.
struct StagedParams {
  void* f;
  const float* stage;
  const float* cs_cell;
  const float4* bc;
  uint32_t n;
  uint32_t nlist;
  int s[27];
  float omega, tau, cs;
  int nx, ny, nz;
};

// Restore + MBBC + collision of listed cell `cell` (flags fl) whose full f is
// in v, and the in-place store through store(q, word) (include/
// applyLocalCellUpdate.h:46-71).
template <class St, class Store>
__device__ __forceinline__ void boundary_core(uint32_t cell, int fl, float (&v)[27], float4 given, int nx, int ny,
                                              int nz, const Relax& rx, float cs, Store&& store) {
  const int kind = fl & 7;
  const uint32_t nxy = (uint32_t)nx * (uint32_t)ny;
  const int k = (int)(cell / nxy);
  const uint32_t rem = cell - (uint32_t)k * nxy;
  const int j = (int)(rem / (uint32_t)nx);
  const int i = (int)(rem - (uint32_t)j * (uint32_t)nx);
  const int code = outer_normal_code(i, j, k, nx, ny, nz, fl);
  boundary_cell(v, kind, code, given.x, given.y, given.z, given.w);
  collide<false>(v, rx, cs);
#pragma unroll
  for (int q = 0; q < 27; ++q) store(q, St::store(St::W_OFF ? v[q] - weight(q) : v[q]));
}

template <class St>
__device__ __forceinline__ void boundary_finish_given(const StepParams& p, const BoundaryParams& b, uint32_t cell,
                                                      int fl, float (&v)[27], float4 given) {
  typedef typename St::T T;
  const float cs = p.cs_cell ? ld<float, false>(p.cs_cell + cell) : p.cs;
  boundary_core<St>(cell, fl, v, given, b.nx, b.ny, b.nz, Relax{p.omega, p.tau, p.cs}, cs, [&](int q, T w) {
    T* base = (T*)(cell >= p.thr[q] ? p.hi[q] : p.lo[q]);
    st<T, false>(base + cell, w);
  });
}

// Given (rho, u) of the cell's slot: one indexed load from the device table
// (a select chain over the four slots was miscompiled for divergent slots, so
// no branch or select is used here).
template <class St>
__device__ __forceinline__ void boundary_finish(const StepParams& p, const BoundaryParams& b, uint32_t cell, int fl,
                                                float (&v)[27]) {
  const float* g = (const float*)(b.bc + ((fl >> 6) & 3));
  boundary_finish_given<St>(p, b, cell, fl, v,
                            make_float4(ld<float, false>(g), ld<float, false>(g + 1), ld<float, false>(g + 2),
                                        ld<float, false>(g + 3)));
}

// Listed cell t, loaded from its 27 slots.
template <class St>
__device__ __forceinline__ void boundary_update(const StepParams& p, const BoundaryParams& b, uint32_t t) {
  typedef typename St::T T;
  if (t >= b.n) return;
  const uint32_t cell = ld<uint32_t, false>(b.list + t);
  const int fl = ld<uint8_t, false>(p.flags + cell);
  float v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const T* base = (const T*)(cell >= p.thr[q] ? p.hi[q] : p.lo[q]);
    v[q] = St::load(ld<T, false>(base + cell));
    if (St::W_OFF) v[q] += weight(q);
  }
  boundary_finish<St>(p, b, cell, fl, v);
}

// Listed cell t, loaded from its staged record (populations, cell index and
// flag byte in one line).  The caller's workgroup has the given-value table in
// `given` (LDS), so the record is the cell's only global load.
template <class St>
__device__ __forceinline__ void load_staged(const float* stage, uint32_t t, float (&v)[27], uint32_t& cell, int& fl) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const PSL_GLOBAL f4* r = (const PSL_GLOBAL f4*)(stage + (size_t)t * STAGE_STRIDE);
  f4 x[STAGE_STRIDE / 4];
#pragma unroll
  for (int i = 0; i < STAGE_STRIDE / 4; ++i) x[i] = r[i];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    v[q] = x[q >> 2][q & 3];
    if (St::W_OFF) v[q] += weight(q);
  }
  // copied to scalars before the casts (a cast of a vector element in place
  // reads the wrong element)
  const float wc = x[6][3], wf = x[7][0];
  cell = __builtin_bit_cast(uint32_t, wc);
  fl = (int)__builtin_bit_cast(uint32_t, wf);
}

// Fluid and bounceback cells of one block in cell order; with STAGE the
// listed box-face cells of the block are loaded too and written to their
// staging records (their update is the boundary kernel's).  HELD: a step
// between evaluations of a learned closure, relaxing with the network output
// stored per cell (p.nut_y) and no Smagorinsky term.
template <class St, bool FLAGS, bool LES, int TUNE, bool STAGE = false, bool HELD = false>
__device__ __forceinline__ void bulk_update(const StepParams& p, uint32_t block, const Stage& sg = Stage{}) {
  const uint32_t block_first = block * 256u;
  const uint32_t cell = block_first + threadIdx.x;
  const uint32_t w_first = wave_first_cell(block_first);
  if (w_first >= p.n) return;  // whole wave past the grid (wave-uniform)
  // all 64 lanes take part in the ballot, so every population is tested
  const bool straddle = wave_straddles(p, w_first);
  if (cell >= p.n) return;
  int fl = 0;
  if (FLAGS) fl = ld<uint8_t, false>(p.flags + cell);
  const int kind = FLAGS ? fl & 7 : KIND_FLUID;
  if (kind > (STAGE ? KIND_GIVEN_RHOU : KIND_BOUNCEBACK)) return;
  // the stored network output is loaded ahead of the populations, so that it
  // is in flight with them rather than after them
  uint16_t y_held = 0;
  if constexpr (HELD) y_held = ld<uint16_t, false>(p.nut_y + cell);
  Word<St> r[27];
  load27_raw<St, TUNE>(p, w_first, straddle, cell, r);
  float v[27];
  if (STAGE) {
    const bool listed = kind >= KIND_GIVEN_RHO;
    const uint64_t m = __ballot(listed);
    if (m) {
      if (listed) {
        words_to_float<St>(r, v);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t t = ld<uint32_t, false>(sg.base + (w_first >> 6)) + rank;
        // eight 16-byte stores per record (the listed lanes are active)
        typedef float f4 __attribute__((ext_vector_type(4)));
        PSL_GLOBAL f4* rec = (PSL_GLOBAL f4*)(sg.buf + (size_t)t * STAGE_STRIDE);
        const float wc = __builtin_bit_cast(float, cell), wf = __builtin_bit_cast(float, (uint32_t)fl);
#pragma unroll
        for (int i = 0; i < STAGE_STRIDE / 4; ++i) {
          f4 x;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int q = 4 * i + c;
            x[c] = q < 27 ? v[q] : q == 27 ? wc : q == 28 ? wf : 0.f;
          }
          rec[i] = x;
        }
        return;
      }
    }
  }
  if (kind == KIND_BOUNCEBACK) {
    // the swap commutes with "f - w" (opposite directions share a weight)
    words_to_float<St>(r, v);
    bounceback(v);
  } else if constexpr (HELD) {
    // the rate from the stored network output; the constant zero removes the
    // Smagorinsky term from this instantiation
    const float y = (float)__builtin_bit_cast(_Float16, y_held);
    collide_with_moments<St::W_OFF>(v, raw_moments_words<St>(r), closure_relax(p.nu, p.nut_scale, y), 0.f);
  } else {
    Relax rx{p.omega, p.tau, p.cs};
    float cs = p.cs;
    if (LES && p.cs_cell) cs = ld<float, false>(p.cs_cell + cell);
    // the collision reads only the moments; v receives the post-collision values
    collide_with_moments<St::W_OFF>(v, raw_moments_words<St>(r), rx, LES ? cs : 0.f);
  }
  store27<St, TUNE>(p, w_first, straddle, cell, v);
}

}  // namespace psl

#pragma clang fp contract(fast)
