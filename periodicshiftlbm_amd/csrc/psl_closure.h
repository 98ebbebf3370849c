// Learned eddy-viscosity closure: closure inputs of a cell and the network
// evaluated with bf16 matrix-core instructions, one wavefront of 64 cells at
// a time.
//
// No reference counterpart: it replaces the Smagorinsky relaxation rate of
// include/cellFunctions.h:157-198 (getOmegaLES) by
//   tau_eff = 3 (nu + nu_t) + 1/2,  nu_t = NU_SCALE * max(0, y),
// with y the output of a 9 -> 32 (tanh) -> 32 (tanh) -> 1 network whose inputs
// are the standardised features of the non-equilibrium second moments
// Pi = kappa - rho/3 I (kappa: central second moments of the populations).
//
// Inference layout (per wavefront, 64 cells, v_mfma_f32_32x32x16_bf16):
// weights are the A operand, activations the B operand, one 32-cell column
// block at a time (two blocks per wavefront), so hidden units are rows and
// cells are columns:
//   layer 1: D1[h][c] = sum_k W1[k][h] x[c][k]     (K = 16: 9 inputs + zeros)
//   layer 2: D2[g][c] = sum_h W2[h][g] tanh(D1[h][c] + b1[h])   (K = 32)
//   output : y[c] = sum_g W3[g] tanh(D2[g][c] + b2[g]) + b3.
// For a 32x32 result, lane l holds column l % 32 and, in accumulator r, row
// 8 (r / 4) + 4 (l / 32) + r % 4.  The B operand of one 16-deep K step takes
// k = 8 (l / 32) + j (j < 8) in lane l, so accumulators 8 s .. 8 s + 7 of a
// layer are exactly the B operand of K step s of the next layer once the
// hidden units are numbered in that order; the host packs W2 accordingly and
// no data moves between lanes inside the network.
#pragma once

#include "psl_internal.h"
#include "psl_stream.h"

#pragma clang fp contract(on)

namespace psl {

constexpr int MLP_IN = 9;
constexpr int MLP_HID = 32;
constexpr int MLP_PARAMS = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID + 1;  // 1409
constexpr float MLP_NU_SCALE = 1e-3f;

// Device image of a network, packed per lane by the host (closure_pack).
struct MlpNet {
  uint16_t a1[64][8];     // layer-1 A operand (bf16): W1[8 (l/32) + j][l % 32], 0 for inputs >= 9
  uint16_t a2[2][64][8];  // layer-2 A operand (bf16) for K steps 0, 1 (hidden units in accumulator order)
  float b1[64][16];       // b1 of the 16 accumulator rows of lane l
  float b2[64][16];
  float w3[64][16];
  float b3;
  float mean[MLP_IN];     // input standardisation
  float inv_std[MLP_IN];
};

// Row of accumulator r in lane half `half` (lane / 32) of a 32x32 result.
PSL_HD PSL_INLINE int acc_row(int r, int half) { return 8 * (r / 4) + 4 * half + (r % 4); }

// Raw closure inputs (before standardisation) from density and the central
// second moments: |dev Pi|/rho, tr Pi/rho, the unit deviatoric direction
// (xx, yy, zz, xy, xz, yz) and the normalised third invariant
// 3 sqrt(6) det(dev Pi) / |dev Pi|^3.
PSL_HD PSL_INLINE void closure_features(float rho, float k200, float k020, float k002, float k110, float k101,
                                        float k011, float (&x)[MLP_IN]) {
  const float ir = 1.f / rho;
  const float third = 1.f / 3.f;
  const float pxx = k200 * ir - third, pyy = k020 * ir - third, pzz = k002 * ir - third;
  const float pxy = k110 * ir, pxz = k101 * ir, pyz = k011 * ir;
  const float tr = pxx + pyy + pzz;
  const float dxx = pxx - tr * third, dyy = pyy - tr * third, dzz = pzz - tr * third;
  const float mag = sqrtf(dxx * dxx + dyy * dyy + dzz * dzz + 2.f * (pxy * pxy + pxz * pxz + pyz * pyz));
  const float inv = mag > 0.f ? 1.f / mag : 0.f;
  const float det = dxx * (dyy * dzz - pyz * pyz) - pxy * (pxy * dzz - pyz * pxz) + pxz * (pxy * pyz - dyy * pxz);
  x[0] = mag;
  x[1] = tr;
  x[2] = dxx * inv;
  x[3] = dyy * inv;
  x[4] = dzz * inv;
  x[5] = pxy * inv;
  x[6] = pxz * inv;
  x[7] = pyz * inv;
  x[8] = 7.348469228349534f * det * inv * inv * inv;  // 3 sqrt(6)
}

// Raw closure inputs of a cell from the raw moments of its populations
// (m = raw_moments of f, or of f - w when W_OFF).
template <bool W_OFF>
PSL_HD PSL_INLINE void cell_closure_features(Moments m, float (&x)[MLP_IN]) {
  if (W_OFF) {
    m.m000 += 1.f;
    m.m200 += 1.f / 3.f;
    m.m020 += 1.f / 3.f;
    m.m002 += 1.f / 3.f;
  }
  const float rho = m.m000;
  const float inv_rho = 1.f / rho;
  const float ux = m.m100 * inv_rho, uy = m.m010 * inv_rho, uz = m.m001 * inv_rho;
  closure_features(rho, m.m200 - ux * m.m100, m.m020 - uy * m.m010, m.m002 - uz * m.m001, m.m110 - ux * m.m010,
                   m.m101 - ux * m.m001, m.m011 - uy * m.m001, x);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// tanh by linear interpolation in a table held in LDS: TANH_N intervals on
// [-TANH_X, TANH_X], entry i = (tanh(x_i), tanh(x_{i+1}) - tanh(x_i)); beyond
// the range the value is tanh(+-TANH_X) = +-0.99975.  Interpolation error
// <= h^2/8 max|tanh''| = 3e-5, clamp error 2.5e-4: both below the bf16
// resolution (2^-8) the hidden activations are rounded to.  A table read and
// two full-rate operations replace the two quarter-rate transcendentals of
// 1 - 2 / (e^{2x} + 1).
constexpr int TANH_N = 512;
constexpr float TANH_X = 4.5f;

struct TanhTable {
  float2 e[TANH_N + 1];
};

// Fill by the whole workgroup (blockDim.x threads), then barrier: call before
// any thread of the workgroup can leave.
__device__ __forceinline__ void tanh_table_fill(TanhTable& t) {
  const float h = 2.f * TANH_X / TANH_N;
  for (int i = threadIdx.x; i <= TANH_N; i += blockDim.x) {
    const float a = tanhf(-TANH_X + h * i);
    const float b = tanhf(-TANH_X + h * (i + 1));
    t.e[i] = make_float2(a, i < TANH_N ? b - a : 0.f);
  }
  __syncthreads();
}

__device__ __forceinline__ float table_tanh(const TanhTable& t, float x) {
  const float u = (fminf(fmaxf(x, -TANH_X), TANH_X) + TANH_X) * (TANH_N / (2.f * TANH_X));
  // u >= 0: truncation is the floor, and fract(u) = u - floor(u)
  const float2 e = t.e[(int)u];
  return e.x + __builtin_amdgcn_fractf(u) * e.y;
}

__device__ __forceinline__ bf16x8 load_a(const uint16_t (&a)[8]) {
  bf16x8 r;
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = __builtin_bit_cast(__bf16, a[j]);
  return r;
}

// Network output y of this lane's cell.  x: raw features of the cell (any
// value on lanes whose result is not used).  Must be called by all 64 lanes.
__device__ __forceinline__ float closure_wave(const MlpNet* __restrict__ net, const TanhTable& tt,
                                             const float (&raw)[MLP_IN]) {
#if defined(__HIP_DEVICE_COMPILE__)
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5;
  float x[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) x[i] = (raw[i] - net->mean[i]) * net->inv_std[i];
  // partner lane (lane ^ 32) features: column block 0 needs input 8 of cell
  // lane-32 in the upper half, block 1 inputs 0..7 of cell lane+32 in the lower half
  float xp[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) xp[i] = __shfl_xor(x[i], 32);
  const bf16x8 A1 = load_a(net->a1[lane]);
  const bf16x8 A20 = load_a(net->a2[0][lane]);
  const bf16x8 A21 = load_a(net->a2[1][lane]);
  float y_own = 0.f;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    // the per-lane bias / output-weight rows are re-read for each column block
    // (L1 hits) rather than held in 48 registers across both
    const MlpNet* ntp = net;
    asm volatile("" : "+s"(ntp));
    // the laundered pointer is global memory (otherwise flat loads)
    const PSL_GLOBAL MlpNet* nt = (const PSL_GLOBAL MlpNet*)ntp;
    bf16x8 B;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float v;
      if (half == 0) v = nb == 0 ? x[j] : xp[j];
      else v = j == 0 ? (nb == 0 ? xp[8] : x[8]) : 0.f;
      B[j] = (__bf16)v;
    }
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = nt->b1[lane][r];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B, acc, 0, 0, 0);
    bf16x8 H0, H1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      H0[j] = (__bf16)table_tanh(tt, acc[j]);
      H1[j] = (__bf16)table_tanh(tt, acc[8 + j]);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = nt->b2[lane][r];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A20, H0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A21, H1, acc, 0, 0, 0);
    float part = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) part += table_tanh(tt, acc[r]) * nt->w3[lane][r];
    const float y = part + __shfl_xor(part, 32) + net->b3;
    if (nb == half) y_own = y;  // block nb holds cells 32 nb .. 32 nb + 31
  }
  return y_own;
#else
  return 0.f;
#endif
}

// Bulk cells with the learned closure (include/applyLocalCellUpdate.h:36-39,
// 68-69 with getOmegaLES replaced).  Every lane of a wavefront reaches the
// network, including lanes past the grid or on skipped / boundary-list cells
// (they run it on rest-state inputs and discard the result).
template <class St, bool FLAGS, int TUNE>
__device__ __forceinline__ void bulk_update_closure(const StepParams& p, uint32_t block,
                                                    const MlpNet* __restrict__ net, float nu, TanhTable& tt,
                                                    uint16_t* __restrict__ y_out) {
  tanh_table_fill(tt);  // before any lane may leave (barrier)
  const uint32_t block_first = block * 256u;
  const uint32_t cell = block_first + threadIdx.x;
  const uint32_t w_first = wave_first_cell(block_first);
  if (w_first >= p.n) return;  // whole wave past the grid (wave-uniform)
  const bool straddle = wave_straddles(p, w_first);
  const bool valid = cell < p.n;
  int kind = valid ? KIND_FLUID : KIND_GHOST;
  if (FLAGS && valid) kind = ld<uint8_t, false>(p.flags + cell) & 7;
  const bool active = kind <= KIND_BOUNCEBACK;
  Word<St> r[27];
  if (active) {
    load27_raw<St, TUNE>(p, w_first, straddle, cell, r);
  } else {
    // rest state: f = w, stored as 0 in the reduced storages
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      if constexpr (St::W_OFF) r[q] = 0;
      else r[q] = weight(q);
    }
  }
  const Moments m = raw_moments_words<St>(r);
  float raw[MLP_IN];
  cell_closure_features<St::W_OFF>(m, raw);
  float y = closure_wave(net, tt, raw);
  if (!active) return;
  if (y_out) {
    // held for the next steps as FP16; this step uses the same rounded value
    const _Float16 h = (_Float16)y;
    y_out[cell] = __builtin_bit_cast(uint16_t, h);
    y = (float)h;
  }
  float v[27];
  if (kind == KIND_BOUNCEBACK) {
    words_to_float<St>(r, v);
    bounceback(v);
  } else {
    collide_with_moments<St::W_OFF>(v, m, closure_relax(nu, MLP_NU_SCALE, y), 0.f);
  }
  store27<St, TUNE>(p, w_first, straddle, cell, v);
}

}  // namespace psl

#pragma clang fp contract(fast)
