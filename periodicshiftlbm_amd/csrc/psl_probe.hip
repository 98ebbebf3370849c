// Data-movement probes of the periodic-shift access pattern (no collision).
//
// They measure what bounds the local update's HBM rate on this device: the
// number of concurrent population streams, reads against writes, the
// alignment of the streams to cache lines and the width of each lane's
// access.  Each launch touches rows of the layout fArray(q, slot) at
// f[q * n + slot] (the rows a grid of n cells holds), with the shifts a grid
// has after 1000 streaming steps, or those shifts rounded down to whole
// 256-byte runs.  Every slot index is taken modulo n and every access of
// cells_per_lane consecutive slots stays inside its row, so all accesses stay
// inside the 27 rows.  No load result is used before every load of the lane
// is issued (no branch between them), as in the step kernels.
#include "psl_internal.h"

#include <algorithm>

namespace psl {
namespace {

struct ProbeShift {
  uint32_t s[27];
};

// S streams, CPL consecutive cells per lane as one access of CPL * sizeof(T)
// bytes declared ALIGN-byte aligned (the element's own size for FP32 at any
// shift: multi-dword accesses need 4-byte alignment only; the whole width for
// line-aligned shifts); MODE 0: read then write in place (+1), 1: read only,
// 2: write only.
template <class T, int CPL, int MODE, int ALIGN, int S>
__global__ __launch_bounds__(256) void k_probe(T* __restrict__ f, ProbeShift sh, uint32_t n, T* __restrict__ sink) {
  typedef T W __attribute__((ext_vector_type(CPL), aligned(ALIGN)));
  const uint32_t c = (blockIdx.x * 256u + threadIdx.x) * CPL;
  if (c >= n) return;
  // run start of each stream, kept inside its row without branches: a run
  // that would cross n (possible only for several FP32 cells per lane at
  // unaligned shifts, a few lanes per launch) is moved back to end at n
  uint32_t at[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    uint32_t s = c + sh.s[q];
    if (s >= n) s -= n;
    at[q] = min(s, n - CPL);
  }
  W v[S];
  if (MODE != 2) {
#pragma unroll
    for (int q = 0; q < S; ++q) v[q] = *(const W*)(f + (size_t)q * n + at[q]);
  }
  if (MODE == 1) {
    T acc = 0;
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int e = 0; e < CPL; ++e) acc += v[q][e];
    // never true for the zero-filled probe rows: keeps the loads alive
    if (acc == (T)12345) sink[c] = acc;
    return;
  }
#pragma unroll
  for (int q = 0; q < S; ++q) {
    W w;
#pragma unroll
    for (int e = 0; e < CPL; ++e) w[e] = MODE == 0 ? (T)(v[q][e] + (T)1) : (T)q;
    *(W*)(f + (size_t)q * n + at[q]) = w;
  }
}

template <class T, int CPL, int ALIGN, int S>
void launch_probe_s(void* f, const ProbeShift& sh, uint32_t n, int mode, void* sink) {
  const uint32_t blocks = (uint32_t)(((uint64_t)n + 256ull * CPL - 1) / (256ull * CPL));
  if (mode == 0) k_probe<T, CPL, 0, ALIGN, S><<<blocks, 256>>>((T*)f, sh, n, (T*)sink);
  else if (mode == 1) k_probe<T, CPL, 1, ALIGN, S><<<blocks, 256>>>((T*)f, sh, n, (T*)sink);
  else k_probe<T, CPL, 2, ALIGN, S><<<blocks, 256>>>((T*)f, sh, n, (T*)sink);
}

// the stream counts are template arguments, so every loop over them unrolls
// and the values stay in registers
template <class T, int CPL, int ALIGN>
void launch_probe(void* f, const ProbeShift& sh, uint32_t n, int streams, int mode, void* sink) {
  switch (streams) {
    case 1: launch_probe_s<T, CPL, ALIGN, 1>(f, sh, n, mode, sink); break;
    case 2: launch_probe_s<T, CPL, ALIGN, 2>(f, sh, n, mode, sink); break;
    case 4: launch_probe_s<T, CPL, ALIGN, 4>(f, sh, n, mode, sink); break;
    case 8: launch_probe_s<T, CPL, ALIGN, 8>(f, sh, n, mode, sink); break;
    case 16: launch_probe_s<T, CPL, ALIGN, 16>(f, sh, n, mode, sink); break;
    default: launch_probe_s<T, CPL, ALIGN, 27>(f, sh, n, mode, sink); break;
  }
}

template <class T>
void launch_probe_cpl(int cpl, bool aligned, void* f, const ProbeShift& sh, uint32_t n, int streams, int mode,
                      void* sink) {
  constexpr int E = (int)sizeof(T);
  if (cpl == 1) launch_probe<T, 1, E>(f, sh, n, streams, mode, sink);
  else if (cpl == 2 && aligned) launch_probe<T, 2, 2 * E>(f, sh, n, streams, mode, sink);
  else if (cpl == 2) launch_probe<T, 2, 4>(f, sh, n, streams, mode, sink);
  else if (aligned) launch_probe<T, 4, 4 * E>(f, sh, n, streams, mode, sink);
  else launch_probe<T, 4, 4>(f, sh, n, streams, mode, sink);
}

}  // namespace
}  // namespace psl

using namespace psl;

extern "C" int psl_probe_ps_access(int nx, int ny, int nz, int storage, int streams, int mode, int cells_per_lane,
                                   int line_aligned, int launches, double* gbs) {
  PSL_CHECK_ARG(nx > 0 && ny > 0 && nz > 0 && launches > 0 && gbs, "bad arguments");
  PSL_CHECK_ARG(storage == PSL_F32 || storage == PSL_F16DEV || storage == PSL_BF16DEV, "unknown storage");
  PSL_CHECK_ARG(streams == 1 || streams == 2 || streams == 4 || streams == 8 || streams == 16 || streams == 27,
                "streams must be 1, 2, 4, 8, 16 or 27");
  PSL_CHECK_ARG(mode >= 0 && mode <= 2, "mode must be 0 (in place), 1 (read only) or 2 (write only)");
  PSL_CHECK_ARG(cells_per_lane == 1 || cells_per_lane == 2 || cells_per_lane == 4, "cells_per_lane must be 1, 2 or 4");
  const uint64_t n64 = (uint64_t)nx * ny * nz;
  PSL_CHECK_ARG(n64 < (1ull << 31) - 1024, "cellCount must be below 2^31");
  // wide accesses: rows start on a multiple of 4 elements, and 16-bit words
  // are paired only where the shifts keep them aligned
  PSL_CHECK_ARG(cells_per_lane == 1 || n64 % 4 == 0, "several cells per lane need a cell count divisible by 4");
  PSL_CHECK_ARG(cells_per_lane == 1 || storage == PSL_F32 || line_aligned,
                "16-bit storage with several cells per lane needs line-aligned shifts");
  const uint32_t n = (uint32_t)n64;
  const size_t esz = storage == PSL_F32 ? 4 : 2;
  // shifts after 1000 streaming steps (include/applyStreaming.h:11-31)
  ProbeShift sh;
  const int64_t run = 256 / (int64_t)esz;
  for (int q = 0; q < 27; ++q) {
    const int64_t d = (int64_t)CX[q] + (int64_t)nx * CY[q] + (int64_t)nx * ny * CZ[q];
    int64_t v = (-(d * 1000)) % (int64_t)n;
    if (v < 0) v += n;
    if (line_aligned) v -= v % run;
    sh.s[q] = (uint32_t)v;
  }
  void* f = nullptr;
  void* sink = nullptr;
  PSL_HIP_TRY(hipMalloc(&f, 27ull * n * esz));
  int rc = PSL_OK;
  if (hipMalloc(&sink, (size_t)n * esz) != hipSuccess) rc = PSL_E_HIP;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == PSL_OK && (hipMemset(f, 0, 27ull * n * esz) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                       hipEventCreate(&e1) != hipSuccess))
    rc = PSL_E_HIP;
  auto launch = [&]() {
    if (esz == 4) launch_probe_cpl<float>(cells_per_lane, line_aligned != 0, f, sh, n, streams, mode, sink);
    else launch_probe_cpl<uint16_t>(cells_per_lane, line_aligned != 0, f, sh, n, streams, mode, sink);
  };
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
  if (sink) (void)hipFree(sink);
  if (rc) {
    set_error("access probe failed");
    return rc;
  }
  const double per_launch = (double)streams * n * esz * (mode == 0 ? 2.0 : 1.0);
  *gbs = per_launch * launches / (ms * 1e-3) / 1e9;
  return PSL_OK;
}
