// Restart state and health checks of periodic-shift grids.
//
// No reference counterpart: the reference never writes fArray and has no
// guard against a diverging run.  The restart state is what a local update
// reads: the storage words of f exactly as held on the device (FP32, or the
// 16-bit f - w words, so that a reloaded grid continues bit for bit, which a
// round trip through FP32 f = w + d would not guarantee), the shifter, and the
// scalar populations and the held learned-closure output where the grid has
// them.  The health check counts cells with a non-finite population or a
// density <= 0 and reports the density range of the others.
#include "psl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace psl {

struct ShiftTab {
  uint32_t s[27];
};

// Per block: cells with a non-finite population or rho <= 0, and the least
// and greatest density of the other cells (+inf / -inf when there are none).
template <class St>
__global__ __launch_bounds__(256) void k_health(const typename St::T* __restrict__ f, ShiftTab sh, uint32_t n,
                                                uint32_t* __restrict__ bad, float2* __restrict__ range) {
  __shared__ uint32_t sb[256];
  __shared__ float slo[256], shi[256];
  const uint32_t cell = blockIdx.x * 256u + threadIdx.x;
  uint32_t b = 0;
  float lo = INFINITY, hi = -INFINITY;
  if (cell < n) {
    float rho = 0.f;
    bool finite = true;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      uint32_t s = cell + sh.s[q];
      if (s >= n) s -= n;
      float x = St::load(f[(size_t)q * n + s]);
      if (St::W_OFF) x += weight(q);
      finite = finite && __builtin_isfinite(x);
      rho += x;
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

void state_header_init(StateHeader& h, uint32_t kind, int storage, int64_t n, const psl_info& info) {
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "PSLSTATE", 8);
  h.version = STATE_VERSION;
  h.kind = kind;
  h.storage = storage;
  h.n = n;
  h.nx = info.cellCountX;
  h.ny = info.cellCountY;
  h.nz = info.cellCountZ;
  h.closure_every = 1;
}

int state_header_check(const StateHeader& h, uint32_t kind, int storage, int64_t n, const psl_info& info,
                       int64_t bytes, int64_t expect_bytes) {
  PSL_CHECK_ARG(memcmp(h.magic, "PSLSTATE", 8) == 0, "not a grid state (bad magic)");
  PSL_CHECK_ARG(h.version == STATE_VERSION, "grid state of another format version");
  PSL_CHECK_ARG(h.kind == kind, kind == STATE_PS ? "state of an esotwist grid, not a periodic-shift grid"
                                                 : "state of a periodic-shift grid, not an esotwist grid");
  PSL_CHECK_ARG(h.storage == storage, "state of another storage type");
  PSL_CHECK_ARG(h.n == n && h.nx == info.cellCountX && h.ny == info.cellCountY && h.nz == info.cellCountZ,
                "state of a grid of another size");
  PSL_CHECK_ARG(bytes == expect_bytes, "state size does not match its header");
  return PSL_OK;
}

}  // namespace psl

using namespace psl;

namespace {

int32_t grid_sections(const psl_grid* g) {
  return (g->scalar_T ? STATE_SCALAR : 0) | (g->closure_y ? STATE_CLOSURE_Y : 0);
}

int64_t grid_state_bytes(const psl_grid* g, int32_t sections) {
  const int64_t n = (int64_t)g->info.cellCount;
  int64_t b = (int64_t)sizeof(StateHeader) + 27 * n * (int64_t)storage_bytes(g->storage);
  if (sections & STATE_SCALAR) b += 27 * n * 4;
  if (sections & STATE_CLOSURE_Y) b += n * 2;
  return b;
}

template <class St>
int health_launch(psl_grid* g, uint32_t nb, uint32_t* bad, float2* range) {
  ShiftTab sh;
  for (int q = 0; q < 27; ++q) sh.s[q] = (uint32_t)g->shift[q];
  k_health<St><<<nb, 256, 0, g->stream>>>((const typename St::T*)g->f, sh, (uint32_t)g->info.cellCount, bad, range);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

}  // namespace

extern "C" {

int psl_grid_state_bytes(const psl_grid* g, int64_t* bytes) {
  PSL_CHECK_ARG(g && bytes, "null argument");
  *bytes = grid_state_bytes(g, grid_sections(g));
  return PSL_OK;
}

int psl_grid_save_state(psl_grid* g, void* host, int64_t bytes) {
  PSL_CHECK_ARG(g && host, "null argument");
  const int32_t sec = grid_sections(g);
  PSL_CHECK_ARG(bytes == grid_state_bytes(g, sec), "state buffer size differs from psl_grid_state_bytes");
  StateHeader h;
  state_header_init(h, STATE_PS, g->storage, (int64_t)g->info.cellCount, g->info);
  h.sections = sec;
  for (int q = 0; q < 27; ++q) h.shift[q] = g->shift[q];
  h.closure_every = g->closure_every;
  h.closure_count = g->closure_count;
  char* out = (char*)host;
  memcpy(out, &h, sizeof(h));
  size_t at = sizeof(h);
  const size_t n = (size_t)g->info.cellCount;
  const size_t fb = 27 * n * storage_bytes(g->storage);
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  PSL_HIP_TRY(hipMemcpy(out + at, g->f, fb, hipMemcpyDeviceToHost));
  at += fb;
  if (sec & STATE_SCALAR) {
    PSL_HIP_TRY(hipMemcpy(out + at, g->scalar_T, 27 * n * 4, hipMemcpyDeviceToHost));
    at += 27 * n * 4;
  }
  if (sec & STATE_CLOSURE_Y) PSL_HIP_TRY(hipMemcpy(out + at, g->closure_y, n * 2, hipMemcpyDeviceToHost));
  return PSL_OK;
}

int psl_grid_load_state(psl_grid* g, const void* host, int64_t bytes) {
  PSL_CHECK_ARG(g && host && bytes >= (int64_t)sizeof(StateHeader), "bad state arguments");
  StateHeader h;
  memcpy(&h, host, sizeof(h));
  const int32_t sec = grid_sections(g);
  int rc = state_header_check(h, STATE_PS, g->storage, (int64_t)g->info.cellCount, g->info, bytes,
                              grid_state_bytes(g, h.sections));
  if (rc) return rc;
  PSL_CHECK_ARG(h.sections == sec, (h.sections & STATE_SCALAR) != (sec & STATE_SCALAR)
                                       ? "the state's scalar transport differs from the grid's (psl_scalar_enable)"
                                       : "the state's held closure output differs from the grid's closure setting");
  PSL_CHECK_ARG(!(sec & STATE_CLOSURE_Y) || h.closure_every == g->closure_every,
                "the state's closure evaluation interval differs from the grid's");
  for (int q = 0; q < 27; ++q)
    PSL_CHECK_ARG(h.shift[q] >= 0 && (int64_t)h.shift[q] < h.n, "state shifter out of range");
  const char* in = (const char*)host;
  size_t at = sizeof(h);
  const size_t n = (size_t)g->info.cellCount;
  const size_t fb = 27 * n * storage_bytes(g->storage);
  PSL_HIP_TRY(hipStreamSynchronize(g->stream));
  PSL_HIP_TRY(hipMemcpy(g->f, in + at, fb, hipMemcpyHostToDevice));
  at += fb;
  if (sec & STATE_SCALAR) {
    PSL_HIP_TRY(hipMemcpy(g->scalar_T, in + at, 27 * n * 4, hipMemcpyHostToDevice));
    at += 27 * n * 4;
  }
  if (sec & STATE_CLOSURE_Y) PSL_HIP_TRY(hipMemcpy(g->closure_y, in + at, n * 2, hipMemcpyHostToDevice));
  for (int q = 0; q < 27; ++q) g->shift[q] = h.shift[q];
  if (sec & STATE_CLOSURE_Y) g->closure_count = h.closure_count;
  return PSL_OK;
}

int psl_grid_check_health(psl_grid* g, int64_t* bad_cells, double* rho_min, double* rho_max) {
  PSL_CHECK_ARG(g && bad_cells, "null argument");
  const uint32_t n = (uint32_t)g->info.cellCount;
  const uint32_t nb = (n + 255u) / 256u;
  uint32_t* bad = nullptr;
  float2* range = nullptr;
  PSL_HIP_TRY(hipMalloc(&bad, (size_t)nb * sizeof(uint32_t)));
  int rc = hipMalloc(&range, (size_t)nb * sizeof(float2)) == hipSuccess ? PSL_OK : PSL_E_HIP;
  if (rc == PSL_OK) {
    switch (g->storage) {
      case PSL_F32: rc = health_launch<StF32>(g, nb, bad, range); break;
      case PSL_F16DEV: rc = health_launch<StF16>(g, nb, bad, range); break;
      default: rc = health_launch<StBF16>(g, nb, bad, range); break;
    }
  }
  std::vector<uint32_t> hb(rc == PSL_OK ? nb : 0);
  std::vector<float2> hr(rc == PSL_OK ? nb : 0);
  if (rc == PSL_OK && (hipMemcpyAsync(hb.data(), bad, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                       hipMemcpyAsync(hr.data(), range, nb * sizeof(float2), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                       hipStreamSynchronize(g->stream) != hipSuccess))
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

}  // extern "C"
