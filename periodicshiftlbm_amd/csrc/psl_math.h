// Per-cell D3Q27 numerics shared by every kernel of the library.
//
// Written for gfx950 VALU throughput: the whole cumulant collision works on
// the ten moments up to second order plus the post-collision central moments
// that the "AllOne" closure leaves non-zero, instead of the 27-entry feq /
// fneq arrays and full chimera chains of the reference
// (include/applyCollision.h:6-295).  The arithmetic is the same model:
//
//   * raw moments m_abc (a+b+c <= 2) by three axis passes over the 27 values;
//   * kappa_ab = m_ab - rho u_a u_b (second-order central moments);
//   * Smagorinsky omega (include/cellFunctions.h:157-198) with
//     Pi^neq_ab = kappa_ab - rho/3 delta_ab, which equals
//     sum_q c_a c_b (f_q - feq_q) exactly for the second-order D3Q27 feq;
//   * relaxation of the second-order cumulants with the Galilean-corrected
//     Dxu, Dyv, Dzw terms (Geier 2015 eqs 55-65, applyCollision.h:117-148);
//   * odd post-collision central moments zero, fourth/sixth order from the
//     cumulant products (Geier 2015 eqs 81-84, applyCollision.h:173-191);
//   * backward chimera transform along x, y, z (eqs 88-96, lines 197-294).
//
// The first-order post-collision central moments are the negated roundoff of
// rho*u - u*rho in the reference (eqs 85-87, lines 168-171); they are exactly
// zero here.  Results therefore agree with the reference to FP32 roundoff.
//
// When the storage holds deviations d_q = f_q - w_q (FP16 / BF16 storage),
// the moments are taken of d and the lattice-weight moments are added
// analytically (sum w = 1, sum w c_a^2 = 1/3), which keeps the small
// non-equilibrium parts at full FP32 precision.
//
// Contraction: a * b + c is fused into one FMA only inside a single source
// expression (FP_CONTRACT ON), never across statements.  Which FMAs form is
// then fixed by the source, not by how a kernel's control flow splits basic
// blocks, so every kernel that runs these functions (single grid, fused,
// batched) produces the same bits for the same cell.
#pragma once

#include <stdint.h>

#pragma clang fp contract(on)

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PSL_HD __host__ __device__
#define PSL_INLINE __attribute__((always_inline)) inline
#else
#include <math.h>
#define PSL_HD
#define PSL_INLINE inline
#endif

namespace psl {

// D3Q27 ordering of the reference (include/applyStreaming.h:1-8).
// id:  0  1  2  3  4  5  6  7  8  9 10 11 12 13 14 15 16 17 18 19 20 21 22 23 24 25 26
// cx:  0  1 -1  0  0  0  0  1 -1  1 -1 -1  1  0  0 -1  1  0  0 -1  1 -1  1  1 -1 -1  1
// cy:  0  0  0  0  0 -1  1  0  0  0  0 -1  1  1 -1  1 -1  1 -1  1 -1 -1  1 -1  1 -1  1
// cz:  0  0  0 -1  1  0  0 -1  1  1 -1  0  0 -1  1  0  0  1 -1 -1  1  1 -1 -1  1 -1  1
constexpr int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1};
constexpr int CY[27] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1};
constexpr int CZ[27] = {0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1};

// lattice weight of direction q (tensor product of 2/3, 1/6, 1/6)
PSL_HD constexpr float weight(int q) {
  return q == 0 ? 8.f / 27.f : q <= 6 ? 2.f / 27.f : q <= 18 ? 1.f / 54.f : 1.f / 216.f;
}

// Index of population with velocity (x,y,z), each in {-1,0,1} encoded as 0,1,2.
// LINE[x][y][z] follows the reference's chimera naming a=-1, b=0, c=+1.
constexpr int LINE[3][3][3] = {
    // x = a (-1)
    {{25, 11, 21}, {10, 2, 8}, {19, 15, 24}},
    // x = b (0)
    {{18, 5, 14}, {3, 0, 4}, {13, 6, 17}},
    // x = c (+1)
    {{23, 16, 20}, {7, 1, 9}, {22, 12, 26}},
};

// Cell flag byte: bits 0-2 kind, bits 3-5 periodic x/y/z (normal component
// zeroed), bits 6-7 slot of the given (rho,ux,uy,uz) table.
enum CellKind : int {
  KIND_FLUID = 0,
  KIND_BOUNCEBACK = 1,
  KIND_GIVEN_RHO = 2,   // restoreUxUyUz + MBBC
  KIND_GIVEN_U = 3,     // restoreRho + MBBC
  KIND_GIVEN_NONE = 4,  // restoreRhoUxUyUz + MBBC
  KIND_GIVEN_RHOU = 5,  // MBBC with the given values
  KIND_GHOST = 6,       // not updated
  KIND_COVERED = 7,     // inside a finer grid's sub-box: not updated
};
constexpr int FLAG_PERIODIC_X = 8, FLAG_PERIODIC_Y = 16, FLAG_PERIODIC_Z = 32;

// Second-order raw moments of 27 values v (ordering above).
struct Moments {
  float m000, m100, m010, m001, m200, m020, m002, m110, m101, m011;
};

// y and x passes of raw_moments, from the z pass's zeroth, first and second
// moments of each (x, y) line.
PSL_HD PSL_INLINE Moments moments_from_z(const float (&z0)[3][3], const float (&z1)[3][3], const float (&z2)[3][3]) {
  // y pass: (y-order, z-order) with total <= 2 per x plane
  float y00[3], y10[3], y20[3], y01[3], y11[3], y02[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    y20[x] = z0[x][0] + z0[x][2];
    y00[x] = y20[x] + z0[x][1];
    y10[x] = z0[x][2] - z0[x][0];
    y01[x] = (z1[x][0] + z1[x][2]) + z1[x][1];
    y11[x] = z1[x][2] - z1[x][0];
    y02[x] = (z2[x][0] + z2[x][2]) + z2[x][1];
  }
  Moments m;
  m.m200 = y00[0] + y00[2];
  m.m000 = m.m200 + y00[1];
  m.m100 = y00[2] - y00[0];
  m.m010 = (y10[0] + y10[2]) + y10[1];
  m.m110 = y10[2] - y10[0];
  m.m001 = (y01[0] + y01[2]) + y01[1];
  m.m101 = y01[2] - y01[0];
  m.m020 = (y20[0] + y20[2]) + y20[1];
  m.m002 = (y02[0] + y02[2]) + y02[1];
  m.m011 = (y11[0] + y11[2]) + y11[1];
  return m;
}

PSL_HD PSL_INLINE Moments raw_moments(const float (&v)[27]) {
  // z pass: per (x,y) line: zeroth, first, second moment in z
  float z0[3][3], z1[3][3], z2[3][3];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 3; ++y) {
      const float a = v[LINE[x][y][0]], b = v[LINE[x][y][1]], c = v[LINE[x][y][2]];
      z2[x][y] = a + c;
      z0[x][y] = z2[x][y] + b;
      z1[x][y] = c - a;
    }
  return moments_from_z(z0, z1, z2);
}

// Backward chimera step along one axis for a line whose central moments
// along that axis are (k0, k1, k2), velocity u: outputs the three
// populations with c = -1, 0, +1 (Geier 2015 eqs 88-96).
struct Axis {
  float a0, a1, b0, b1, c0, c1;  // a = a0 k0 + a1 k1 + k2/2, b = b0 k0 + b1 k1 - k2, c = c0 k0 + c1 k1 + k2/2
  PSL_HD PSL_INLINE explicit Axis(float u) {
    const float u2 = u * u;
    a0 = 0.5f * (u2 - u);
    c0 = 0.5f * (u2 + u);
    a1 = u - 0.5f;
    c1 = u + 0.5f;
    b0 = 1.f - u2;
    b1 = -2.f * u;
  }
  PSL_HD PSL_INLINE void general(float k0, float k1, float k2, float& a, float& b, float& c) const {
    const float h = 0.5f * k2;
    a = a0 * k0 + (a1 * k1 + h);
    c = c0 * k0 + (c1 * k1 + h);
    b = b0 * k0 + (b1 * k1 - k2);
  }
  // k1 == 0
  PSL_HD PSL_INLINE void even(float k0, float k2, float& a, float& b, float& c) const {
    const float h = 0.5f * k2;
    a = a0 * k0 + h;
    c = c0 * k0 + h;
    b = b0 * k0 - k2;
  }
  // k0 == k2 == 0
  PSL_HD PSL_INLINE void odd(float k1, float& a, float& b, float& c) const {
    a = a1 * k1;
    c = c1 * k1;
    b = b1 * k1;
  }
};

// Post-collision state -> populations.  ks_* are the non-zero central moments
// of the post-collision distribution; writes v (full f, before any weight
// offset is removed).  With FIRST the first-order central moments ks100,
// ks010, ks001 enter as well (non-zero only under a force).
template <bool FIRST = false>
PSL_HD PSL_INLINE void central_to_populations(float rho, float inv_rho, float ux, float uy, float uz, float ks200,
                                              float ks020, float ks002, float ks110, float ks101, float ks011,
                                              float (&v)[27], float ks100 = 0.f, float ks010 = 0.f,
                                              float ks001 = 0.f) {
  // fourth and sixth order from the factorised cumulants (eqs 81-84)
  const float ks211 = (ks200 * ks011 + 2.f * ks101 * ks110) * inv_rho;
  const float ks121 = (ks020 * ks101 + 2.f * ks110 * ks011) * inv_rho;
  const float ks112 = (ks002 * ks110 + 2.f * ks011 * ks101) * inv_rho;
  const float ks220 = (ks020 * ks200 + 2.f * ks110 * ks110) * inv_rho;
  const float ks022 = (ks002 * ks020 + 2.f * ks011 * ks011) * inv_rho;
  const float ks202 = (ks200 * ks002 + 2.f * ks101 * ks101) * inv_rho;
  const float ks222 = (ks200 * ks022 + ks020 * ks202 + ks002 * ks220 +
                       4.f * (ks011 * ks211 + ks101 * ks121 + ks110 * ks112)) * inv_rho -
                      (16.f * ks110 * ks101 * ks011 +
                       4.f * (ks101 * ks101 * ks020 + ks011 * ks011 * ks200 + ks110 * ks110 * ks002) +
                       2.f * ks200 * ks020 * ks002) * (inv_rho * inv_rho);

  const Axis X(ux), Y(uy), Z(uz);
  // x pass: p[x][b][c] for x in {a,b,c}, (b,c) = (y order, z order)
  float p[3][3][3];
  if (FIRST) {
    X.general(rho, ks100, ks200, p[0][0][0], p[1][0][0], p[2][0][0]);
    X.general(ks001, ks101, 0.f, p[0][0][1], p[1][0][1], p[2][0][1]);
  } else {
    X.even(rho, ks200, p[0][0][0], p[1][0][0], p[2][0][0]);
    X.odd(ks101, p[0][0][1], p[1][0][1], p[2][0][1]);
  }
  X.even(ks002, ks202, p[0][0][2], p[1][0][2], p[2][0][2]);
  if (FIRST) X.general(ks010, ks110, 0.f, p[0][1][0], p[1][1][0], p[2][1][0]);
  else X.odd(ks110, p[0][1][0], p[1][1][0], p[2][1][0]);
  X.even(ks011, ks211, p[0][1][1], p[1][1][1], p[2][1][1]);
  X.odd(ks112, p[0][1][2], p[1][1][2], p[2][1][2]);
  X.even(ks020, ks220, p[0][2][0], p[1][2][0], p[2][2][0]);
  X.odd(ks121, p[0][2][1], p[1][2][1], p[2][2][1]);
  X.even(ks022, ks222, p[0][2][2], p[1][2][2], p[2][2][2]);
  // y pass: r[x][y][c]
  float r[3][3][3];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int c = 0; c < 3; ++c) Y.general(p[x][0][c], p[x][1][c], p[x][2][c], r[x][0][c], r[x][1][c], r[x][2][c]);
  // z pass
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 3; ++y)
      Z.general(r[x][y][0], r[x][y][1], r[x][y][2], v[LINE[x][y][0]], v[LINE[x][y][1]], v[LINE[x][y][2]]);
}

// Relaxation parameters of one update.
struct Relax {
  float omega;      // used when cs == 0
  float tau;        // 3 nu + 1/2
  float cs;         // Smagorinsky constant (0: plain BGK-rate cumulant)
};

// Smagorinsky relaxation rate from the second-order central moments
// (include/cellFunctions.h:157-198).
PSL_HD PSL_INLINE float omega_les(float rho, float inv_rho, float k200, float k020, float k002, float k110,
                                  float k101, float k011, float tau, float cs) {
  const float third = rho * (1.f / 3.f);
  const float xx = k200 - third, yy = k020 - third, zz = k002 - third;
  float P = xx * xx + yy * yy + zz * zz + 2.f * (k110 * k110 + k101 * k101 + k011 * k011);
  P = sqrtf(P);
  const float tau_les = 0.5f * tau + 0.5f * sqrtf(tau * tau + 18.f * cs * inv_rho * P);
  return 1.f / tau_les;
}

// The cumulant collision on values v, where v holds f (w_off = false) or
// f - w (w_off = true), given m = raw_moments(v).  On return v holds the
// post-collision values in the same representation.
template <bool W_OFF>
PSL_HD PSL_INLINE void collide_with_moments(float (&v)[27], Moments m, const Relax& rx, float cs) {
  if (W_OFF) {
    m.m000 += 1.f;
    m.m200 += 1.f / 3.f;
    m.m020 += 1.f / 3.f;
    m.m002 += 1.f / 3.f;
  }
  const float rho = m.m000;
  const float inv_rho = 1.f / rho;
  const float ux = m.m100 * inv_rho, uy = m.m010 * inv_rho, uz = m.m001 * inv_rho;
  const float k200 = m.m200 - ux * m.m100;
  const float k020 = m.m020 - uy * m.m010;
  const float k002 = m.m002 - uz * m.m001;
  const float k110 = m.m110 - ux * m.m010;
  const float k101 = m.m101 - ux * m.m001;
  const float k011 = m.m011 - uy * m.m001;

  const float omega = cs == 0.f ? rx.omega : omega_les(rho, inv_rho, k200, k020, k002, k110, k101, k011, rx.tau, cs);

  // eqs 58-65 (applyCollision.h:119-148)
  const float h_rho = 0.5f * inv_rho;
  const float Dxu = -omega * h_rho * (2.f * k200 - k020 - k002) - h_rho * (k200 + k020 + k002 - rho);
  const float Dyv = Dxu + 3.f * omega * h_rho * (k200 - k020);
  const float Dzw = Dxu + 3.f * omega * h_rho * (k200 - k002);
  const float om1 = 1.f - omega;
  const float ks110 = om1 * k110, ks101 = om1 * k101, ks011 = om1 * k011;
  const float ux2Dxu = ux * ux * Dxu, uy2Dyv = uy * uy * Dyv, uz2Dzw = uz * uz * Dzw;
  const float c3 = 3.f * rho * (1.f - 0.5f * omega);
  const float e61 = om1 * (k200 - k020) - c3 * (ux2Dxu - uy2Dyv);
  const float e64 = om1 * (k200 - k002) - c3 * (ux2Dxu - uz2Dzw);
  const float e65 = rho - 1.5f * rho * (ux2Dxu + uy2Dyv + uz2Dzw);
  const float ks200 = (1.f / 3.f) * (e61 + e64 + e65);
  const float ks020 = (1.f / 3.f) * (e64 - 2.f * e61 + e65);
  const float ks002 = (1.f / 3.f) * (e61 - 2.f * e64 + e65);

  central_to_populations(rho, inv_rho, ux, uy, uz, ks200, ks020, ks002, ks110, ks101, ks011, v);
  if (W_OFF) {
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] -= weight(q);
  }
}

// The collision of one cell (include/applyCollision.h:6-295).
template <bool W_OFF>
PSL_HD PSL_INLINE void collide(float (&v)[27], const Relax& rx, float cs) {
  collide_with_moments<W_OFF>(v, raw_moments(v), rx, cs);
}

// The forcing collision of one cell (M35Full/applyCollisionCustomized.h:5-17)
// on full f with force (gx, gy, gz): the velocity moves by half the force,
// u' = (rho u + g/2) / rho, the central moments are taken about u' (their
// first order is then -g/2), the Smagorinsky rate sees f - feq(rho, u'), the
// relaxation is the unforced one with u', and the backward step restores the
// first-order central moments as +g/2, so the momentum gains g.
PSL_HD PSL_INLINE void collide_forced(float (&v)[27], const Relax& rx, float cs, float gx, float gy, float gz) {
  const Moments m = raw_moments(v);
  const float rho = m.m000;
  const float inv_rho = 1.f / rho;
  const float ux = (m.m100 + 0.5f * gx) * inv_rho, uy = (m.m010 + 0.5f * gy) * inv_rho,
              uz = (m.m001 + 0.5f * gz) * inv_rho;
  const float k100 = m.m100 - rho * ux, k010 = m.m010 - rho * uy, k001 = m.m001 - rho * uz;
  // second-order central moments about u' from the raw ones
  const float k200 = m.m200 - ux * (2.f * m.m100 - rho * ux);
  const float k020 = m.m020 - uy * (2.f * m.m010 - rho * uy);
  const float k002 = m.m002 - uz * (2.f * m.m001 - rho * uz);
  const float k110 = m.m110 - ux * m.m010 - uy * m.m100 + rho * ux * uy;
  const float k101 = m.m101 - ux * m.m001 - uz * m.m100 + rho * ux * uz;
  const float k011 = m.m011 - uy * m.m001 - uz * m.m010 + rho * uy * uz;
  // f - feq(rho, u') has second moments m2 - rho (I/3 + u' u'^T)
  const float omega = cs == 0.f ? rx.omega
                                : omega_les(rho, inv_rho, m.m200 - rho * ux * ux, m.m020 - rho * uy * uy,
                                            m.m002 - rho * uz * uz, m.m110 - rho * ux * uy, m.m101 - rho * ux * uz,
                                            m.m011 - rho * uy * uz, rx.tau, cs);
  const float h_rho = 0.5f * inv_rho;
  const float Dxu = -omega * h_rho * (2.f * k200 - k020 - k002) - h_rho * (k200 + k020 + k002 - rho);
  const float Dyv = Dxu + 3.f * omega * h_rho * (k200 - k020);
  const float Dzw = Dxu + 3.f * omega * h_rho * (k200 - k002);
  const float om1 = 1.f - omega;
  const float ks110 = om1 * k110, ks101 = om1 * k101, ks011 = om1 * k011;
  const float ux2Dxu = ux * ux * Dxu, uy2Dyv = uy * uy * Dyv, uz2Dzw = uz * uz * Dzw;
  const float c3 = 3.f * rho * (1.f - 0.5f * omega);
  const float e61 = om1 * (k200 - k020) - c3 * (ux2Dxu - uy2Dyv);
  const float e64 = om1 * (k200 - k002) - c3 * (ux2Dxu - uz2Dzw);
  const float e65 = rho - 1.5f * rho * (ux2Dxu + uy2Dyv + uz2Dzw);
  const float ks200 = (1.f / 3.f) * (e61 + e64 + e65);
  const float ks020 = (1.f / 3.f) * (e64 - 2.f * e61 + e65);
  const float ks002 = (1.f / 3.f) * (e61 - 2.f * e64 + e65);
  central_to_populations<true>(rho, inv_rho, ux, uy, uz, ks200, ks020, ks002, ks110, ks101, ks011, v, -k100, -k010,
                               -k001);
}

// Central-moment rescaling between grid levels (include/gridRefinementFunctions.h:6-281,
// paper eqs 3.17-3.20): operates on full f.
PSL_HD PSL_INLINE void rescale(float (&f)[27], bool coarse_to_fine) {
  const Moments m = raw_moments(f);
  const float rho = m.m000;
  const float inv_rho = 1.f / rho;
  const float ux = m.m100 * inv_rho, uy = m.m010 * inv_rho, uz = m.m001 * inv_rho;
  const float k200 = m.m200 - ux * m.m100;
  const float k020 = m.m020 - uy * m.m010;
  const float k002 = m.m002 - uz * m.m001;
  const float k110 = m.m110 - ux * m.m010;
  const float k101 = m.m101 - ux * m.m001;
  const float k011 = m.m011 - uy * m.m001;
  float s200, s020, s002, s110, s101, s011;
  if (coarse_to_fine) {
    s200 = (2.f / 3.f) * k200 + (1.f / 6.f) * k020 + (1.f / 6.f) * k002;
    s020 = (1.f / 6.f) * k200 + (2.f / 3.f) * k020 + (1.f / 6.f) * k002;
    s002 = (1.f / 6.f) * k200 + (1.f / 6.f) * k020 + (2.f / 3.f) * k002;
    s110 = 0.5f * k110;
    s101 = 0.5f * k101;
    s011 = 0.5f * k011;
  } else {
    s200 = (5.f / 3.f) * k200 - (1.f / 3.f) * k020 - (1.f / 3.f) * k002;
    s020 = -(1.f / 3.f) * k200 + (5.f / 3.f) * k020 - (1.f / 3.f) * k002;
    s002 = -(1.f / 3.f) * k200 - (1.f / 3.f) * k020 + (5.f / 3.f) * k002;
    s110 = 2.f * k110;
    s101 = 2.f * k101;
    s011 = 2.f * k011;
  }
  central_to_populations(rho, inv_rho, ux, uy, uz, s200, s020, s002, s110, s101, s011, f);
}

// Second-order equilibrium (include/cellFunctions.h:60-128).
PSL_HD PSL_INLINE float feq(int q, float rho, float ux, float uy, float uz) {
  const float cu = CX[q] * ux + CY[q] * uy + CZ[q] * uz;
  const float u2 = ux * ux + uy * uy + uz * uz;
  return rho * weight(q) * (1.f + 3.f * cu + 4.5f * cu * cu - 1.5f * u2);
}

// Swap of opposite populations (include/boundaryConditions/applyBounceback.h:8-27).
PSL_HD PSL_INLINE void bounceback(float (&f)[27]) {
#pragma unroll
  for (int q = 1; q < 27; q += 2) {
    const float t = f[q];
    f[q] = f[q + 1];
    f[q + 1] = t;
  }
}

PSL_HD PSL_INLINE void rho_u(const float (&f)[27], float& rho, float& ux, float& uy, float& uz) {
  const Moments m = raw_moments(f);
  rho = m.m000;
  const float inv = 1.f / rho;
  ux = m.m100 * inv;
  uy = m.m010 * inv;
  uz = m.m001 * inv;
}

#include "bc_generated.inc"

// Box-face boundary treatment of one cell (include/applyLocalCellUpdate.h:46-67):
// outer normal with periodic components removed, restore the missing
// macroscopic quantities, then impose them through MBBC.  f is full f.
PSL_HD PSL_INLINE void boundary_cell(float (&f)[27], int kind, int ncode, float rho, float ux, float uy, float uz) {
  if (kind == KIND_GIVEN_RHO) bc_restore_u<float>(ncode, f, rho, ux, uy, uz);
  else if (kind == KIND_GIVEN_U) bc_restore_rho<float>(ncode, f, rho, ux, uy, uz);
  else if (kind == KIND_GIVEN_NONE) bc_restore_rhou<float>(ncode, f, rho, ux, uy, uz);
  bc_mbbc<float>(ncode, f, rho, ux, uy, uz);
}

PSL_HD PSL_INLINE int outer_normal_code(int i, int j, int k, int nx, int ny, int nz, int flag) {
  int a = 0, b = 0, c = 0;
  if (i == 0) a = -1;
  else if (i == nx - 1) a = 1;
  if (j == 0) b = -1;
  else if (j == ny - 1) b = 1;
  if (k == 0) c = -1;
  else if (k == nz - 1) c = 1;
  if (flag & FLAG_PERIODIC_X) a = 0;
  if (flag & FLAG_PERIODIC_Y) b = 0;
  if (flag & FLAG_PERIODIC_Z) c = 0;
  return (a + 5) * 100 + (b + 5) * 10 + (c + 5);
}

}  // namespace psl

#pragma clang fp contract(fast)
