// Learned eddy-viscosity closure: training (FP32), data preparation from a
// fine-grid run, inference (bf16 matrix cores) and the C ABI.
//
// No reference counterpart (the reference's LES model is Smagorinsky,
// include/cellFunctions.h:157-198).  Pipeline:
//   1. a fine periodic grid (this solver) gives rho and u per cell
//      (psl_get_macroscopic);
//   2. k_les_filter: box filter of width r -> coarse rho, Favre velocity and
//      subgrid stress tau = <rho u u>/<rho> - u~ u~;
//   3. k_les_targets: coarse strain rate S by periodic central differences;
//      target nu_t* = max(0, -dev(tau):S / (2 S:S)); the non-equilibrium
//      second moments a perfect closure would produce on the coarse grid,
//      Pi = -(2/3) rho tau_eff S with tau_eff = 3 (nu + nu_t*) + 1/2, and from
//      them the raw closure inputs (closure_features);
//   4. training: mean-squared error of y against nu_t* / NU_SCALE, Adam,
//      mini-batches drawn by a counter-based generator; per-sample forward and
//      backward in k_mlp_fwd_bwd, deterministic gradient reduction in
//      k_mlp_grad_partial + k_mlp_adam (fixed summation order);
//   5. inference inside the step (closure_wave, psl_closure.h).
#include "psl_closure.h"
#include "psl_stream.h"

#include <cmath>
#include <cstring>

namespace psl {

// ------------------------------------------------------------ training
struct TrainBufs {
  float* h1;   // B x 32
  float* h2;   // B x 32
  float* dz1;  // B x 32
  float* dz2;  // B x 32
  float* dy;   // B
  float* sq;   // B squared errors
};

// Per-sample forward pass and deltas.  x: standardised inputs (n x 9) of the
// whole set, idx: the batch (B indices into x), t: targets.
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ x,
                                                     const float* __restrict__ t, const uint32_t* __restrict__ idx,
                                                     int B, TrainBufs tb) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= B) return;
  const uint32_t s = idx ? idx[b] : (uint32_t)b;
  const float* W1 = P;
  const float* b1 = W1 + MLP_IN * MLP_HID;
  const float* W2 = b1 + MLP_HID;
  const float* b2 = W2 + MLP_HID * MLP_HID;
  const float* W3 = b2 + MLP_HID;
  const float b3 = W3[MLP_HID];
  float xi[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) xi[i] = x[(size_t)s * MLP_IN + i];
  float h1[MLP_HID], h2[MLP_HID];
  for (int h = 0; h < MLP_HID; ++h) {
    float z = b1[h];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) z += xi[i] * W1[i * MLP_HID + h];
    h1[h] = tanhf(z);
  }
  float y = b3;
  for (int g = 0; g < MLP_HID; ++g) {
    float z = b2[g];
    for (int h = 0; h < MLP_HID; ++h) z += h1[h] * W2[h * MLP_HID + g];
    h2[g] = tanhf(z);
    y += h2[g] * W3[g];
  }
  const float e = y - t[s];
  const float dy = 2.f * e / (float)B;
  tb.dy[b] = dy;
  tb.sq[b] = e * e;
  float dz2[MLP_HID];
  for (int g = 0; g < MLP_HID; ++g) {
    dz2[g] = dy * W3[g] * (1.f - h2[g] * h2[g]);
    tb.h1[(size_t)b * MLP_HID + g] = h1[g];
    tb.h2[(size_t)b * MLP_HID + g] = h2[g];
    tb.dz2[(size_t)b * MLP_HID + g] = dz2[g];
  }
  for (int h = 0; h < MLP_HID; ++h) {
    float d = 0.f;
    for (int g = 0; g < MLP_HID; ++g) d += dz2[g] * W2[h * MLP_HID + g];
    tb.dz1[(size_t)b * MLP_HID + h] = d * (1.f - h1[h] * h1[h]);
  }
}

// Partial gradient sums over a chunk of CHUNK samples per block (fixed order),
// plus the partial squared error in slot MLP_PARAMS.
constexpr int CHUNK = 64;
__global__ __launch_bounds__(256) void k_mlp_grad_partial(const float* __restrict__ x,
                                                          const uint32_t* __restrict__ idx, int B, TrainBufs tb,
                                                          float* __restrict__ partial) {
  __shared__ float sx[CHUNK][MLP_IN];
  __shared__ float sh1[CHUNK][MLP_HID];
  __shared__ float sh2[CHUNK][MLP_HID];
  __shared__ float sz1[CHUNK][MLP_HID];
  __shared__ float sz2[CHUNK][MLP_HID];
  __shared__ float sdy[CHUNK];
  __shared__ float ssq[CHUNK];
  const int b0 = blockIdx.x * CHUNK;
  const int nb = min(CHUNK, B - b0);
  for (int e = threadIdx.x; e < CHUNK * MLP_HID; e += 256) {
    const int r = e / MLP_HID, c = e % MLP_HID;
    const bool ok = r < nb;
    const size_t o = (size_t)(b0 + r) * MLP_HID + c;
    sh1[r][c] = ok ? tb.h1[o] : 0.f;
    sh2[r][c] = ok ? tb.h2[o] : 0.f;
    sz1[r][c] = ok ? tb.dz1[o] : 0.f;
    sz2[r][c] = ok ? tb.dz2[o] : 0.f;
  }
  for (int e = threadIdx.x; e < CHUNK * MLP_IN; e += 256) {
    const int r = e / MLP_IN, c = e % MLP_IN;
    sx[r][c] = r < nb ? x[(size_t)(idx ? idx[b0 + r] : (uint32_t)(b0 + r)) * MLP_IN + c] : 0.f;
  }
  for (int r = threadIdx.x; r < CHUNK; r += 256) {
    sdy[r] = r < nb ? tb.dy[b0 + r] : 0.f;
    ssq[r] = r < nb ? tb.sq[b0 + r] : 0.f;
  }
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * (MLP_PARAMS + 1);
  for (int p = threadIdx.x; p < MLP_PARAMS + 1; p += 256) {
    float s = 0.f;
    int q = p;
    if (q < MLP_IN * MLP_HID) {  // W1[i][h]
      const int i = q / MLP_HID, h = q % MLP_HID;
      for (int r = 0; r < CHUNK; ++r) s += sx[r][i] * sz1[r][h];
    } else if ((q -= MLP_IN * MLP_HID) < MLP_HID) {  // b1
      for (int r = 0; r < CHUNK; ++r) s += sz1[r][q];
    } else if ((q -= MLP_HID) < MLP_HID * MLP_HID) {  // W2[h][g]
      const int h = q / MLP_HID, g = q % MLP_HID;
      for (int r = 0; r < CHUNK; ++r) s += sh1[r][h] * sz2[r][g];
    } else if ((q -= MLP_HID * MLP_HID) < MLP_HID) {  // b2
      for (int r = 0; r < CHUNK; ++r) s += sz2[r][q];
    } else if ((q -= MLP_HID) < MLP_HID) {  // W3
      for (int r = 0; r < CHUNK; ++r) s += sh2[r][q] * sdy[r];
    } else if (q == MLP_HID) {  // b3
      for (int r = 0; r < CHUNK; ++r) s += sdy[r];
    } else {  // squared error
      for (int r = 0; r < CHUNK; ++r) s += ssq[r];
    }
    out[p] = s;
  }
}

// Sum the block partials in order; Adam update when lr > 0 (step from 1).
__global__ __launch_bounds__(256) void k_mlp_adam(const float* __restrict__ partial, int nblk,
                                                  float* __restrict__ P, float* __restrict__ m,
                                                  float* __restrict__ v, float* __restrict__ grad_out, float lr,
                                                  int step, float* __restrict__ loss_out, int B) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p > MLP_PARAMS) return;
  float g = 0.f;
  for (int k = 0; k < nblk; ++k) g += partial[(size_t)k * (MLP_PARAMS + 1) + p];
  if (p == MLP_PARAMS) {
    if (loss_out) *loss_out = g / (float)B;
    return;
  }
  if (grad_out) grad_out[p] = g;
  if (lr > 0.f) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step));
    const float vh = vv / (1.f - powf(b2, (float)step));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
  }
}

// Mini-batch indices: splitmix64 of (seed, step, b) mod n.
__global__ void k_batch_indices(uint32_t* __restrict__ idx, int B, uint32_t n, uint64_t seed, int step) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= B) return;
  uint64_t z = seed + ((uint64_t)step * (uint64_t)B + (uint64_t)b + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[b] = (uint32_t)(z % n);
}

__global__ void k_standardise(const float* __restrict__ raw, float* __restrict__ x, uint32_t n,
                              const float* __restrict__ mean, const float* __restrict__ inv_std) {
  const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (uint64_t)n * MLP_IN) return;
  const int i = (int)(e % MLP_IN);
  x[e] = (raw[e] - mean[i]) * inv_std[i];
}

// Per-input sums of x and x^2 (FP64) of chunks of raw features.
__global__ __launch_bounds__(256) void k_feature_moments(const float* __restrict__ raw, uint32_t n,
                                                         double* __restrict__ part) {
  __shared__ double s1[256][MLP_IN];
  __shared__ double s2[256][MLP_IN];
  double a[MLP_IN] = {0}, q[MLP_IN] = {0};
  for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u)
    for (int i = 0; i < MLP_IN; ++i) {
      const double v = raw[(size_t)c * MLP_IN + i];
      a[i] += v;
      q[i] += v * v;
    }
  for (int i = 0; i < MLP_IN; ++i) {
    s1[threadIdx.x][i] = a[i];
    s2[threadIdx.x][i] = q[i];
  }
  __syncthreads();
  if (threadIdx.x < 2 * MLP_IN) {
    const int i = threadIdx.x % MLP_IN;
    const bool sq = threadIdx.x >= MLP_IN;
    double s = 0;
    for (int k = 0; k < 256; ++k) s += sq ? s2[k][i] : s1[k][i];
    part[(size_t)blockIdx.x * 2 * MLP_IN + threadIdx.x] = s;
  }
}

// ------------------------------------------------------------ inference
// Network output of raw feature rows (n x 9) with the matrix-core path of the
// step kernel; whole wavefronts run it, rows past n read row n - 1.
__global__ __launch_bounds__(256) void k_mlp_eval(const MlpNet* __restrict__ net, const float* __restrict__ raw,
                                                  uint32_t n, float* __restrict__ y) {
  __shared__ TanhTable tt;
  tanh_table_fill(tt);
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  const uint32_t r = c < n ? c : n - 1u;
  float f[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) f[i] = raw[(size_t)r * MLP_IN + i];
  const float out = closure_wave(net, tt, f);
  if (c < n) y[c] = out;
}

// ------------------------------------------------------------ data
// Coarse cell (I, J, K) of an m^3 periodic coarse grid averages the fine
// cells r I .. r I + r - 1 (each axis) of the fine grid's unique n^3 region
// (the fine layout may carry duplicated x / y boundary layers: strides fx, fxy).
__global__ __launch_bounds__(256) void k_les_filter(const float* __restrict__ rho, const float* __restrict__ ux,
                                                    const float* __restrict__ uy, const float* __restrict__ uz,
                                                    int fx, int fxy, int r, int m, float* __restrict__ out) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  const uint32_t mc = (uint32_t)m * m * m;
  if (c >= mc) return;
  const int I = c % m, J = (c / m) % m, K = c / (m * m);
  double sr = 0, sx = 0, sy = 0, sz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
  for (int k = 0; k < r; ++k)
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < r; ++i) {
        const size_t f = (size_t)(K * r + k) * fxy + (size_t)(J * r + j) * fx + (I * r + i);
        const double d = rho[f], a = ux[f], b = uy[f], e = uz[f];
        sr += d;
        sx += d * a;
        sy += d * b;
        sz += d * e;
        sxx += d * a * a;
        syy += d * b * b;
        szz += d * e * e;
        sxy += d * a * b;
        sxz += d * a * e;
        syz += d * b * e;
      }
  const double inv = 1.0 / (double)(r * r * r);
  const double R = sr * inv;
  const double U = sx * inv / R, V = sy * inv / R, W = sz * inv / R;
  float* o = out;
  o[0 * (size_t)mc + c] = (float)R;
  o[1 * (size_t)mc + c] = (float)U;
  o[2 * (size_t)mc + c] = (float)V;
  o[3 * (size_t)mc + c] = (float)W;
  o[4 * (size_t)mc + c] = (float)(sxx * inv / R - U * U);
  o[5 * (size_t)mc + c] = (float)(syy * inv / R - V * V);
  o[6 * (size_t)mc + c] = (float)(szz * inv / R - W * W);
  o[7 * (size_t)mc + c] = (float)(sxy * inv / R - U * V);
  o[8 * (size_t)mc + c] = (float)(sxz * inv / R - U * W);
  o[9 * (size_t)mc + c] = (float)(syz * inv / R - V * W);
}

__global__ __launch_bounds__(256) void k_les_targets(const float* __restrict__ co, int m, float nu,
                                                     float* __restrict__ feat, float* __restrict__ target) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  const uint32_t mc = (uint32_t)m * m * m;
  if (c >= mc) return;
  const int I = c % m, J = (c / m) % m, K = c / (m * m);
  auto at = [&](int comp, int i, int j, int k) -> float {
    i = (i + m) % m;
    j = (j + m) % m;
    k = (k + m) % m;
    return co[(size_t)comp * mc + ((size_t)k * m + j) * m + i];
  };
  // g[a][b] = d_a u_b, central differences
  float g[3][3];
  for (int b = 0; b < 3; ++b) {
    g[0][b] = 0.5f * (at(1 + b, I + 1, J, K) - at(1 + b, I - 1, J, K));
    g[1][b] = 0.5f * (at(1 + b, I, J + 1, K) - at(1 + b, I, J - 1, K));
    g[2][b] = 0.5f * (at(1 + b, I, J, K + 1) - at(1 + b, I, J, K - 1));
  }
  const float S[6] = {g[0][0], g[1][1], g[2][2], 0.5f * (g[0][1] + g[1][0]), 0.5f * (g[0][2] + g[2][0]),
                      0.5f * (g[1][2] + g[2][1])};
  const float rho = co[c];
  const float t0 = co[4 * (size_t)mc + c], t1 = co[5 * (size_t)mc + c], t2 = co[6 * (size_t)mc + c];
  const float tr = (t0 + t1 + t2) / 3.f;
  const float dev[6] = {t0 - tr, t1 - tr, t2 - tr, co[7 * (size_t)mc + c], co[8 * (size_t)mc + c],
                        co[9 * (size_t)mc + c]};
  float num = 0.f, den = 0.f;
  for (int a = 0; a < 6; ++a) {
    const float w = a < 3 ? 1.f : 2.f;
    num += w * dev[a] * S[a];
    den += w * S[a] * S[a];
  }
  den *= 2.f;
  const float nut = den > 0.f ? fmaxf(0.f, -num / den) : 0.f;
  const float te = 3.f * (nu + nut) + 0.5f;
  const float k = -(2.f / 3.f) * rho * te;
  // central second moments kappa = Pi + rho/3 I
  float x[MLP_IN];
  closure_features(rho, k * S[0] + rho / 3.f, k * S[1] + rho / 3.f, k * S[2] + rho / 3.f, k * S[3], k * S[4],
                   k * S[5], x);
  for (int i = 0; i < MLP_IN; ++i) feat[(size_t)c * MLP_IN + i] = x[i];
  target[c] = nut / MLP_NU_SCALE;
}

// Host packing of a flat parameter vector into the device image.
static uint16_t bf16_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}

void closure_pack(const float* p, const float* mean, const float* std_, MlpNet& net) {
  const float* W1 = p;
  const float* b1 = W1 + MLP_IN * MLP_HID;
  const float* W2 = b1 + MLP_HID;
  const float* b2 = W2 + MLP_HID * MLP_HID;
  const float* W3 = b2 + MLP_HID;
  memset(&net, 0, sizeof(net));
  for (int l = 0; l < 64; ++l) {
    const int half = l / 32, row = l % 32;
    for (int j = 0; j < 8; ++j) {
      const int k = 8 * half + j;
      net.a1[l][j] = bf16_bits(k < MLP_IN ? W1[k * MLP_HID + row] : 0.f);
      for (int s = 0; s < 2; ++s) net.a2[s][l][j] = bf16_bits(W2[acc_row(8 * s + j, half) * MLP_HID + row]);
    }
    for (int r = 0; r < 16; ++r) {
      const int h = acc_row(r, half);
      net.b1[l][r] = b1[h];
      net.b2[l][r] = b2[h];
      net.w3[l][r] = W3[h];
    }
  }
  net.b3 = W3[MLP_HID];
  for (int i = 0; i < MLP_IN; ++i) {
    net.mean[i] = mean[i];
    net.inv_std[i] = 1.f / std_[i];
  }
}

}  // namespace psl

using namespace psl;

struct psl_mlp {
  float params[MLP_PARAMS];
  float mean[MLP_IN];
  float std_[MLP_IN];
  MlpNet* d_net = nullptr;  // packed device image (inference)
};

const MlpNet* psl_mlp_device_net(const psl_mlp* m) { return m ? m->d_net : nullptr; }

namespace {

int upload_net(psl_mlp* m) {
  MlpNet net;
  closure_pack(m->params, m->mean, m->std_, net);
  PSL_HIP_TRY(hipMemcpy(m->d_net, &net, sizeof(net), hipMemcpyHostToDevice));
  return PSL_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, bytes) != hipSuccess) {
      set_error("hipMalloc failed in the closure trainer");
      return PSL_E_NOMEM;
    }
    return PSL_OK;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

// One forward/backward over `B` samples (idx or the first B rows) and either
// an Adam step (lr > 0) or only the gradient; loss to *loss.
int train_pass(const float* x, const float* t, const uint32_t* idx, int B, float* P, float* mo, float* ve,
               float* grad, float lr, int step, float* loss_dev, const TrainBufs& tb, float* partial) {
  k_mlp_fwd_bwd<<<(B + 255) / 256, 256>>>(P, x, t, idx, B, tb);
  PSL_LAUNCH_CHECK();
  const int nblk = (B + CHUNK - 1) / CHUNK;
  k_mlp_grad_partial<<<nblk, 256>>>(x, idx, B, tb, partial);
  PSL_LAUNCH_CHECK();
  k_mlp_adam<<<(MLP_PARAMS + 1 + 255) / 256, 256>>>(partial, nblk, P, mo, ve, grad, lr, step, loss_dev, B);
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

int alloc_train(int B, DevBuf (&bufs)[7], TrainBufs& tb, float*& partial) {
  const size_t h = (size_t)B * MLP_HID * sizeof(float);
  int rc = bufs[0].alloc(h);
  if (!rc) rc = bufs[1].alloc(h);
  if (!rc) rc = bufs[2].alloc(h);
  if (!rc) rc = bufs[3].alloc(h);
  if (!rc) rc = bufs[4].alloc((size_t)B * sizeof(float));
  if (!rc) rc = bufs[5].alloc((size_t)B * sizeof(float));
  if (!rc) rc = bufs[6].alloc((size_t)((B + CHUNK - 1) / CHUNK) * (MLP_PARAMS + 1) * sizeof(float));
  if (rc) return rc;
  tb = {bufs[0].as<float>(), bufs[1].as<float>(), bufs[2].as<float>(), bufs[3].as<float>(), bufs[4].as<float>(),
        bufs[5].as<float>()};
  partial = bufs[6].as<float>();
  return PSL_OK;
}

}  // namespace

extern "C" {

int psl_mlp_create(psl_mlp** out) {
  PSL_CHECK_ARG(out, "null argument");
  psl_mlp* m = new psl_mlp();
  memset(m->params, 0, sizeof(m->params));
  for (int i = 0; i < MLP_IN; ++i) {
    m->mean[i] = 0.f;
    m->std_[i] = 1.f;
  }
  if (hipMalloc(&m->d_net, sizeof(MlpNet)) != hipSuccess) {
    delete m;
    set_error("hipMalloc of the network image failed");
    return PSL_E_NOMEM;
  }
  int rc = upload_net(m);
  if (rc) {
    (void)hipFree(m->d_net);
    delete m;
    return rc;
  }
  *out = m;
  return PSL_OK;
}

int psl_mlp_destroy(psl_mlp* m) {
  if (!m) return PSL_OK;
  (void)hipDeviceSynchronize();
  if (m->d_net) (void)hipFree(m->d_net);
  delete m;
  return PSL_OK;
}

int psl_mlp_set_params(psl_mlp* m, const float* params, const float* mean, const float* std_) {
  PSL_CHECK_ARG(m && params, "null argument");
  memcpy(m->params, params, sizeof(m->params));
  if (mean) memcpy(m->mean, mean, sizeof(m->mean));
  if (std_) {
    for (int i = 0; i < MLP_IN; ++i) PSL_CHECK_ARG(std_[i] > 0.f, "standard deviations must be positive");
    memcpy(m->std_, std_, sizeof(m->std_));
  }
  return upload_net(m);
}

int psl_mlp_get_params(const psl_mlp* m, float* params, float* mean, float* std_) {
  PSL_CHECK_ARG(m, "null network");
  if (params) memcpy(params, m->params, sizeof(m->params));
  if (mean) memcpy(mean, m->mean, sizeof(m->mean));
  if (std_) memcpy(std_, m->std_, sizeof(m->std_));
  return PSL_OK;
}

int psl_mlp_loss_grad(const psl_mlp* m, const float* x_dev, const float* t_dev, int n, double* loss, float* grad) {
  PSL_CHECK_ARG(m && x_dev && t_dev && n > 0, "bad arguments");
  DevBuf bufs[7], dP, dG, dL;
  TrainBufs tb;
  float* partial = nullptr;
  int rc = alloc_train(n, bufs, tb, partial);
  if (!rc) rc = dP.alloc(sizeof(m->params));
  if (!rc) rc = dG.alloc(sizeof(m->params));
  if (!rc) rc = dL.alloc(sizeof(float));
  if (rc) return rc;
  PSL_HIP_TRY(hipMemcpy(dP.p, m->params, sizeof(m->params), hipMemcpyHostToDevice));
  rc = train_pass(x_dev, t_dev, nullptr, n, dP.as<float>(), nullptr, nullptr, dG.as<float>(), 0.f, 1, dL.as<float>(),
                  tb, partial);
  if (rc) return rc;
  float l = 0.f;
  PSL_HIP_TRY(hipMemcpy(&l, dL.p, sizeof(float), hipMemcpyDeviceToHost));
  if (loss) *loss = l;
  if (grad) PSL_HIP_TRY(hipMemcpy(grad, dG.p, sizeof(m->params), hipMemcpyDeviceToHost));
  return PSL_OK;
}

int psl_mlp_train(psl_mlp* m, const float* raw_dev, const float* t_dev, int n, int batch, int iters, float lr,
                  uint64_t seed, float* loss_hist) {
  PSL_CHECK_ARG(m && raw_dev && t_dev && n > 0 && batch > 0 && iters > 0 && lr > 0.f, "bad arguments");
  // standardisation from the data (FP64 moments)
  const int nblk_m = 256;
  DevBuf dpart, dx, didx, dP, dm, dv, dloss, dmean, dinv;
  int rc = dpart.alloc((size_t)nblk_m * 2 * MLP_IN * sizeof(double));
  if (!rc) rc = dx.alloc((size_t)n * MLP_IN * sizeof(float));
  if (!rc) rc = didx.alloc((size_t)batch * sizeof(uint32_t));
  if (!rc) rc = dP.alloc(sizeof(m->params));
  if (!rc) rc = dm.alloc(sizeof(m->params));
  if (!rc) rc = dv.alloc(sizeof(m->params));
  if (!rc) rc = dloss.alloc((size_t)iters * sizeof(float));
  if (!rc) rc = dmean.alloc(MLP_IN * sizeof(float));
  if (!rc) rc = dinv.alloc(MLP_IN * sizeof(float));
  if (rc) return rc;
  k_feature_moments<<<nblk_m, 256>>>(raw_dev, (uint32_t)n, dpart.as<double>());
  PSL_LAUNCH_CHECK();
  std::vector<double> part((size_t)nblk_m * 2 * MLP_IN);
  PSL_HIP_TRY(hipMemcpy(part.data(), dpart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  float inv[MLP_IN];
  for (int i = 0; i < MLP_IN; ++i) {
    double s = 0, q = 0;
    for (int k = 0; k < nblk_m; ++k) {
      s += part[(size_t)k * 2 * MLP_IN + i];
      q += part[(size_t)k * 2 * MLP_IN + MLP_IN + i];
    }
    const double mu = s / n, var = std::max(q / n - mu * mu, 0.0);
    m->mean[i] = (float)mu;
    m->std_[i] = (float)std::max(std::sqrt(var), 1e-6);
    inv[i] = 1.f / m->std_[i];
  }
  PSL_HIP_TRY(hipMemcpy(dmean.p, m->mean, sizeof(m->mean), hipMemcpyHostToDevice));
  PSL_HIP_TRY(hipMemcpy(dinv.p, inv, sizeof(inv), hipMemcpyHostToDevice));
  k_standardise<<<(uint32_t)(((uint64_t)n * MLP_IN + 255) / 256), 256>>>(raw_dev, dx.as<float>(), (uint32_t)n,
                                                                        dmean.as<float>(), dinv.as<float>());
  PSL_LAUNCH_CHECK();
  DevBuf bufs[7];
  TrainBufs tb;
  float* partial = nullptr;
  rc = alloc_train(batch, bufs, tb, partial);
  if (rc) return rc;
  PSL_HIP_TRY(hipMemcpy(dP.p, m->params, sizeof(m->params), hipMemcpyHostToDevice));
  PSL_HIP_TRY(hipMemset(dm.p, 0, sizeof(m->params)));
  PSL_HIP_TRY(hipMemset(dv.p, 0, sizeof(m->params)));
  for (int it = 0; it < iters; ++it) {
    k_batch_indices<<<(batch + 255) / 256, 256>>>(didx.as<uint32_t>(), batch, (uint32_t)n, seed, it);
    PSL_LAUNCH_CHECK();
    rc = train_pass(dx.as<float>(), t_dev, didx.as<uint32_t>(), batch, dP.as<float>(), dm.as<float>(),
                    dv.as<float>(), nullptr, lr, it + 1, dloss.as<float>() + it, tb, partial);
    if (rc) return rc;
  }
  PSL_HIP_TRY(hipMemcpy(m->params, dP.p, sizeof(m->params), hipMemcpyDeviceToHost));
  if (loss_hist) PSL_HIP_TRY(hipMemcpy(loss_hist, dloss.p, (size_t)iters * sizeof(float), hipMemcpyDeviceToHost));
  return upload_net(m);
}

int psl_mlp_eval(const psl_mlp* m, const float* raw_dev, int n, float* y_dev) {
  PSL_CHECK_ARG(m && raw_dev && y_dev && n > 0, "bad arguments");
  k_mlp_eval<<<(n + 255) / 256, 256>>>(m->d_net, raw_dev, (uint32_t)n, y_dev);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

int psl_les_filter(const float* rho, const float* ux, const float* uy, const float* uz, int fine_nx, int fine_ny,
                   int n, int r, float* coarse) {
  PSL_CHECK_ARG(rho && ux && uy && uz && coarse, "null argument");
  PSL_CHECK_ARG(r > 0 && n > 0 && n % r == 0 && fine_nx >= n && fine_ny >= n, "bad filter geometry");
  const int m = n / r;
  const uint32_t mc = (uint32_t)m * m * m;
  k_les_filter<<<(mc + 255) / 256, 256>>>(rho, ux, uy, uz, fine_nx, fine_nx * fine_ny, r, m, coarse);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

int psl_les_targets(const float* coarse, int m, float nu, float* feat, float* target) {
  PSL_CHECK_ARG(coarse && feat && target && m > 2, "bad arguments");
  const uint32_t mc = (uint32_t)m * m * m;
  k_les_targets<<<(mc + 255) / 256, 256>>>(coarse, m, nu, feat, target);
  PSL_LAUNCH_CHECK();
  PSL_HIP_TRY(hipDeviceSynchronize());
  return PSL_OK;
}

}  // extern "C"
