// PeriodicShiftLBM CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++17 restatement of the reference's D3Q27 periodic-shift LBM hot
// path, used by tests/, by __graft_entry__.smoke() and by bench.py's CPU
// baseline leg as the checker.  Nothing in the shipped library links it.
//
// Every function names the reference file:line it follows.  The FP32
// instantiation (Real = float) reproduces the reference's mixed precision:
// where the reference writes a double literal (3.0, 16.0, 0.5 * tau ...), the
// same sub-expression is evaluated in double here (type W).  The FP64
// instantiation (Real = double) is the FP64 oracle used to state errors.
//
// Parity pins (see tests/):
//   velocity set / weights ... literal tables + moment identities (test_oracle_core)
//   PS shifter ............... paper Table 1 D2Q9 map + brute-force streaming
//   feq / moments ............ exact moment identities
//   collision ................ mass/momentum conservation, rest-state fixed point,
//                              cube-symmetry equivariance, shear-wave decay rate;
//                              non-linear part: every post-collision cumulant of
//                              order 3-6, computed from its definition (log of
//                              the moment generating function), is zero;
//                              second-order central moments relax by (1 - omega)
//                              (off-diagonal at any u, deviatoric at u = 0, trace
//                              to rho); diagonal shear and sound waves decay at
//                              the same rate with and without a uniform
//                              background flow (the Galilean corrections)
//   LES omega ................ closed form (Cs=0, Pi_neq=0)
//   boundary conditions ...... bit-exact vs the reference's compiled headers
//                              (tests/golden/bc_vectors.npz) + moment identities
//   bounceback ............... bit-exact vs the compiled reference header
//   rescaleF ................. round trip identity, invariants
//   refinement transfers ..... constant / linear fields
//   voxelizer ................ brute force, symmetry, volume, closed meshes
//   esotwist ................. equals brute-force A-B streaming on periodic boxes
//
// Build: make -C oracle   (produces oracle/libpsl_oracle.so)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifndef PSL_ORACLE_FLAGS
#define PSL_ORACLE_FLAGS "unknown"
#endif

namespace {

// ---------------------------------------------------------------- velocity set
// include/applyStreaming.h:6-8, include/cellFunctions.h:3-16
const int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1};
const int CY[27] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1};
const int CZ[27] = {0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1};
// opposite of q: pairs (1,2),(3,4),...,(25,26) -- include/boundaryConditions/applyBounceback.h:12-26
inline int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

template <typename R> struct Wide { typedef double type; };

// ---------------------------------------------------------------- cell helpers
// getFeq: include/cellFunctions.h:60-128
template <typename R>
void get_feq(R rho, R ux, R uy, R uz, R (&feq)[27]) {
  const R u2 = ux * ux + uy * uy + uz * uz;
  const R w0 = R(8) / R(27), w1 = R(2) / R(27), w2 = R(1) / R(54), w3 = R(1) / R(216);
  for (int q = 0; q < 27; ++q) {
    // cu in the reference is written out term by term; the sum over the
    // non-zero components in x,y,z order reproduces it exactly.
    R cu = R(0);
    bool first = true;
    const int c[3] = {CX[q], CY[q], CZ[q]};
    const R u[3] = {ux, uy, uz};
    for (int a = 0; a < 3; ++a) {
      if (c[a] == 0) continue;
      const R term = c[a] > 0 ? u[a] : -u[a];
      cu = first ? term : cu + term;
      first = false;
    }
    const int nnz = (CX[q] != 0) + (CY[q] != 0) + (CZ[q] != 0);
    const R w = nnz == 0 ? w0 : nnz == 1 ? w1 : nnz == 2 ? w2 : w3;
    feq[q] = rho * w * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * u2);
  }
}

// getRhoUxUyUz: include/cellFunctions.h:135-155 (summation order kept)
template <typename R>
void get_rho_u(const R (&f)[27], R& rho, R& ux, R& uy, R& uz) {
  rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
        f[14] + f[15] + f[16] + f[17] + f[18] + f[19] + f[20] + f[21] + f[22] + f[23] + f[24] + f[25] + f[26];
  ux = (+f[1] - f[2] + f[7] - f[8] + f[9] - f[10] - f[11] + f[12] - f[15] + f[16] - f[19] + f[20] - f[21] + f[22] +
        f[23] - f[24] - f[25] + f[26]) / rho;
  uy = (-f[5] + f[6] - f[11] + f[12] + f[13] - f[14] + f[15] - f[16] + f[17] - f[18] + f[19] - f[20] - f[21] +
        f[22] - f[23] + f[24] - f[25] + f[26]) / rho;
  uz = (-f[3] + f[4] - f[7] + f[8] + f[9] - f[10] - f[13] + f[14] + f[17] - f[18] - f[19] + f[20] + f[21] - f[22] -
        f[23] + f[24] - f[25] + f[26]) / rho;
}

// getOmegaLES: include/cellFunctions.h:157-198
template <typename R>
R get_omega_les(const R (&fneq)[27], R rho, R nu, R cs) {
  typedef typename Wide<R>::type W;
  const R tau = R(3) * nu + R(0.5);
  if (cs == R(0)) return R(1) / tau;
  R P = R(0);
  const R xx = (fneq[1] + fneq[2] + fneq[7] + fneq[8] + fneq[9] + fneq[10] + fneq[11] + fneq[12] + fneq[15] +
                fneq[16] + fneq[19] + fneq[20] + fneq[21] + fneq[22] + fneq[23] + fneq[24] + fneq[25] + fneq[26]);
  P += xx * xx;
  const R yy = (fneq[5] + fneq[6] + fneq[11] + fneq[12] + fneq[13] + fneq[14] + fneq[15] + fneq[16] + fneq[17] +
                fneq[18] + fneq[19] + fneq[20] + fneq[21] + fneq[22] + fneq[23] + fneq[24] + fneq[25] + fneq[26]);
  P += yy * yy;
  const R zz = (fneq[3] + fneq[4] + fneq[7] + fneq[8] + fneq[9] + fneq[10] + fneq[13] + fneq[14] + fneq[17] +
                fneq[18] + fneq[19] + fneq[20] + fneq[21] + fneq[22] + fneq[23] + fneq[24] + fneq[25] + fneq[26]);
  P += zz * zz;
  const R yz = (-fneq[13] - fneq[14] + fneq[17] + fneq[18] - fneq[19] - fneq[20] - fneq[21] - fneq[22] + fneq[23] +
                fneq[24] + fneq[25] + fneq[26]);
  P += R(2) * (yz * yz);
  const R xz = (-fneq[7] - fneq[8] + fneq[9] + fneq[10] + fneq[19] + fneq[20] - fneq[21] - fneq[22] - fneq[23] -
                fneq[24] + fneq[25] + fneq[26]);
  P += R(2) * (xz * xz);
  const R xy = (fneq[11] + fneq[12] - fneq[15] - fneq[16] - fneq[19] - fneq[20] + fneq[21] + fneq[22] - fneq[23] -
                fneq[24] + fneq[25] + fneq[26]);
  P += R(2) * (xy * xy);
  P = std::sqrt(P);
  const R cles = R(18) * cs * (R(1) / rho);
  // reference line 195: double literals promote the sum
  const R tau_les = (R)(W(0.5) * W(tau) + W(0.5) * W(std::sqrt(R(tau * tau + cles * P))));
  return R(1) / tau_les;
}

// Forward transformation to the central moments used by the collision and
// by rescaleF: include/applyCollision.h:17-84 (Geier 2015 eqs 43-45).
template <typename R>
struct Central {
  R k000, k100, k010, k001, k200, k020, k002, k110, k101, k011;
};
template <typename R>
Central<R> central_moments(const R (&f)[27], R ux, R uy, R uz) {
  const R k_aa0 = (f[21] + f[25]) + f[11];
  const R k_ab0 = (f[8] + f[10]) + f[2];
  const R k_ac0 = (f[24] + f[19]) + f[15];
  const R k_ba0 = (f[14] + f[18]) + f[5];
  const R k_bb0 = (f[4] + f[3]) + f[0];
  const R k_bc0 = (f[17] + f[13]) + f[6];
  const R k_ca0 = (f[20] + f[23]) + f[16];
  const R k_cb0 = (f[9] + f[7]) + f[1];
  const R k_cc0 = (f[26] + f[22]) + f[12];

  const R k_aa1 = (f[21] - f[25]) - uz * k_aa0;
  const R k_ab1 = (f[8] - f[10]) - uz * k_ab0;
  const R k_ac1 = (f[24] - f[19]) - uz * k_ac0;
  const R k_ba1 = (f[14] - f[18]) - uz * k_ba0;
  const R k_bb1 = (f[4] - f[3]) - uz * k_bb0;
  const R k_bc1 = (f[17] - f[13]) - uz * k_bc0;
  const R k_ca1 = (f[20] - f[23]) - uz * k_ca0;
  const R k_cb1 = (f[9] - f[7]) - uz * k_cb0;
  const R k_cc1 = (f[26] - f[22]) - uz * k_cc0;

  const R k_aa2 = (f[21] + f[25]) - R(2) * uz * (f[21] - f[25]) + uz * uz * k_aa0;
  const R k_ab2 = (f[8] + f[10]) - R(2) * uz * (f[8] - f[10]) + uz * uz * k_ab0;
  const R k_ac2 = (f[24] + f[19]) - R(2) * uz * (f[24] - f[19]) + uz * uz * k_ac0;
  const R k_ba2 = (f[14] + f[18]) - R(2) * uz * (f[14] - f[18]) + uz * uz * k_ba0;
  const R k_bb2 = (f[4] + f[3]) - R(2) * uz * (f[4] - f[3]) + uz * uz * k_bb0;
  const R k_bc2 = (f[17] + f[13]) - R(2) * uz * (f[17] - f[13]) + uz * uz * k_bc0;
  const R k_ca2 = (f[20] + f[23]) - R(2) * uz * (f[20] - f[23]) + uz * uz * k_ca0;
  const R k_cb2 = (f[9] + f[7]) - R(2) * uz * (f[9] - f[7]) + uz * uz * k_cb0;
  const R k_cc2 = (f[26] + f[22]) - R(2) * uz * (f[26] - f[22]) + uz * uz * k_cc0;

  const R k_a00 = (k_ac0 + k_aa0) + k_ab0;
  const R k_b00 = (k_bc0 + k_ba0) + k_bb0;
  const R k_c00 = (k_cc0 + k_ca0) + k_cb0;
  const R k_a01 = (k_ac1 + k_aa1) + k_ab1;
  const R k_b01 = (k_bc1 + k_ba1) + k_bb1;
  const R k_c01 = (k_cc1 + k_ca1) + k_cb1;
  const R k_a02 = (k_ac2 + k_aa2) + k_ab2;
  const R k_b02 = (k_bc2 + k_ba2) + k_bb2;
  const R k_c02 = (k_cc2 + k_ca2) + k_cb2;

  const R k_a10 = (k_ac0 - k_aa0) - uy * k_a00;
  const R k_b10 = (k_bc0 - k_ba0) - uy * k_b00;
  const R k_c10 = (k_cc0 - k_ca0) - uy * k_c00;
  const R k_a11 = (k_ac1 - k_aa1) - uy * k_a01;
  const R k_b11 = (k_bc1 - k_ba1) - uy * k_b01;
  const R k_c11 = (k_cc1 - k_ca1) - uy * k_c01;
  const R k_a20 = (k_ac0 + k_aa0) - R(2) * uy * (k_ac0 - k_aa0) + uy * uy * k_a00;
  const R k_b20 = (k_bc0 + k_ba0) - R(2) * uy * (k_bc0 - k_ba0) + uy * uy * k_b00;
  const R k_c20 = (k_cc0 + k_ca0) - R(2) * uy * (k_cc0 - k_ca0) + uy * uy * k_c00;

  Central<R> k;
  k.k000 = (k_c00 + k_a00) + k_b00;
  k.k001 = (k_c01 + k_a01) + k_b01;
  k.k002 = (k_c02 + k_a02) + k_b02;
  k.k010 = (k_c10 + k_a10) + k_b10;
  k.k011 = (k_c11 + k_a11) + k_b11;
  k.k020 = (k_c20 + k_a20) + k_b20;
  k.k100 = (k_c00 - k_a00) - ux * k.k000;
  k.k101 = (k_c01 - k_a01) - ux * k.k001;
  k.k110 = (k_c10 - k_a10) - ux * k.k010;
  k.k200 = (k_c00 + k_a00) - R(2) * ux * (k_c00 - k_a00) + ux * ux * k.k000;
  return k;
}

// Backward transformation from post-collision central moments (2nd order
// given, higher cumulants zero) to distributions:
// include/applyCollision.h:154-294 (Geier 2015 eqs 81-96).
template <typename R>
void central_to_f(R rho, R ux, R uy, R uz, R ks_000, R ks_100, R ks_010, R ks_001, R ks_200, R ks_020, R ks_002,
                  R ks_110, R ks_101, R ks_011, R (&f)[27]) {
  typedef typename Wide<R>::type W;
  const R ks_211 = (ks_200 * ks_011 + R(2) * ks_101 * ks_110) / rho;
  const R ks_121 = (ks_020 * ks_101 + R(2) * ks_110 * ks_011) / rho;
  const R ks_112 = (ks_002 * ks_110 + R(2) * ks_011 * ks_101) / rho;
  const R ks_220 = (ks_020 * ks_200 + R(2) * ks_110 * ks_110) / rho;
  const R ks_022 = (ks_002 * ks_020 + R(2) * ks_011 * ks_011) / rho;
  const R ks_202 = (ks_200 * ks_002 + R(2) * ks_101 * ks_101) / rho;
  // include/applyCollision.h:184-191: 16.0 promotes the second group to double
  const R first = (ks_200 * ks_022 + ks_020 * ks_202 + ks_002 * ks_220 +
                   R(4) * (ks_011 * ks_211 + ks_101 * ks_121 + ks_110 * ks_112)) / rho;
  const W second = (W(16.0) * W(ks_110) * W(ks_101) * W(ks_011) +
                    W(R(4) * (ks_101 * ks_101 * ks_020 + ks_011 * ks_011 * ks_200 + ks_110 * ks_110 * ks_002)) +
                    W(R(2) * ks_200 * ks_020 * ks_002)) / W(rho) / W(rho);
  const R ks_222 = (R)(W(first) - second);

  const R ks_b00 = ks_000 * (R(1) - ux * ux) - R(2) * ux * ks_100 - ks_200;
  const R ks_b01 = ks_001 * (R(1) - ux * ux) - R(2) * ux * ks_101;
  const R ks_b02 = ks_002 * (R(1) - ux * ux) - ks_202;
  const R ks_b10 = ks_010 * (R(1) - ux * ux) - R(2) * ux * ks_110;
  const R ks_b11 = ks_011 * (R(1) - ux * ux) - ks_211;
  const R ks_b12 = -R(2) * ux * ks_112;
  const R ks_b20 = ks_020 * (R(1) - ux * ux) - ks_220;
  const R ks_b21 = -R(2) * ux * ks_121;
  const R ks_b22 = ks_022 * (R(1) - ux * ux) - ks_222;

  const R h = R(0.5);
  const R ks_a00 = (ks_000 * (ux * ux - ux) + ks_100 * (R(2) * ux - R(1)) + ks_200) * h;
  const R ks_a01 = (ks_001 * (ux * ux - ux) + ks_101 * (R(2) * ux - R(1))) * h;
  const R ks_a02 = (ks_002 * (ux * ux - ux) + ks_202) * h;
  const R ks_a10 = (ks_010 * (ux * ux - ux) + ks_110 * (R(2) * ux - R(1))) * h;
  const R ks_a11 = (ks_011 * (ux * ux - ux) + ks_211) * h;
  const R ks_a12 = (ks_112 * (R(2) * ux - R(1))) * h;
  const R ks_a20 = (ks_020 * (ux * ux - ux) + ks_220) * h;
  const R ks_a21 = (ks_121 * (R(2) * ux - R(1))) * h;
  const R ks_a22 = (ks_022 * (ux * ux - ux) + ks_222) * h;

  const R ks_c00 = (ks_000 * (ux * ux + ux) + ks_100 * (R(2) * ux + R(1)) + ks_200) * h;
  const R ks_c01 = (ks_001 * (ux * ux + ux) + ks_101 * (R(2) * ux + R(1))) * h;
  const R ks_c02 = (ks_002 * (ux * ux + ux) + ks_202) * h;
  const R ks_c10 = (ks_010 * (ux * ux + ux) + ks_110 * (R(2) * ux + R(1))) * h;
  const R ks_c11 = (ks_011 * (ux * ux + ux) + ks_211) * h;
  const R ks_c12 = (ks_112 * (R(2) * ux + R(1))) * h;
  const R ks_c20 = (ks_020 * (ux * ux + ux) + ks_220) * h;
  const R ks_c21 = (ks_121 * (R(2) * ux + R(1))) * h;
  const R ks_c22 = (ks_022 * (ux * ux + ux) + ks_222) * h;

  const R ks_ab0 = ks_a00 * (R(1) - uy * uy) - R(2) * uy * ks_a10 - ks_a20;
  const R ks_ab1 = ks_a01 * (R(1) - uy * uy) - R(2) * uy * ks_a11 - ks_a21;
  const R ks_ab2 = ks_a02 * (R(1) - uy * uy) - R(2) * uy * ks_a12 - ks_a22;
  const R ks_bb0 = ks_b00 * (R(1) - uy * uy) - R(2) * uy * ks_b10 - ks_b20;
  const R ks_bb1 = ks_b01 * (R(1) - uy * uy) - R(2) * uy * ks_b11 - ks_b21;
  const R ks_bb2 = ks_b02 * (R(1) - uy * uy) - R(2) * uy * ks_b12 - ks_b22;
  const R ks_cb0 = ks_c00 * (R(1) - uy * uy) - R(2) * uy * ks_c10 - ks_c20;
  const R ks_cb1 = ks_c01 * (R(1) - uy * uy) - R(2) * uy * ks_c11 - ks_c21;
  const R ks_cb2 = ks_c02 * (R(1) - uy * uy) - R(2) * uy * ks_c12 - ks_c22;

  const R ks_aa0 = (ks_a00 * (uy * uy - uy) + ks_a10 * (R(2) * uy - R(1)) + ks_a20) * h;
  const R ks_aa1 = (ks_a01 * (uy * uy - uy) + ks_a11 * (R(2) * uy - R(1)) + ks_a21) * h;
  const R ks_aa2 = (ks_a02 * (uy * uy - uy) + ks_a12 * (R(2) * uy - R(1)) + ks_a22) * h;
  const R ks_ba0 = (ks_b00 * (uy * uy - uy) + ks_b10 * (R(2) * uy - R(1)) + ks_b20) * h;
  const R ks_ba1 = (ks_b01 * (uy * uy - uy) + ks_b11 * (R(2) * uy - R(1)) + ks_b21) * h;
  const R ks_ba2 = (ks_b02 * (uy * uy - uy) + ks_b12 * (R(2) * uy - R(1)) + ks_b22) * h;
  const R ks_ca0 = (ks_c00 * (uy * uy - uy) + ks_c10 * (R(2) * uy - R(1)) + ks_c20) * h;
  const R ks_ca1 = (ks_c01 * (uy * uy - uy) + ks_c11 * (R(2) * uy - R(1)) + ks_c21) * h;
  const R ks_ca2 = (ks_c02 * (uy * uy - uy) + ks_c12 * (R(2) * uy - R(1)) + ks_c22) * h;

  const R ks_ac0 = (ks_a00 * (uy * uy + uy) + ks_a10 * (R(2) * uy + R(1)) + ks_a20) * h;
  const R ks_ac1 = (ks_a01 * (uy * uy + uy) + ks_a11 * (R(2) * uy + R(1)) + ks_a21) * h;
  const R ks_ac2 = (ks_a02 * (uy * uy + uy) + ks_a12 * (R(2) * uy + R(1)) + ks_a22) * h;
  const R ks_bc0 = (ks_b00 * (uy * uy + uy) + ks_b10 * (R(2) * uy + R(1)) + ks_b20) * h;
  const R ks_bc1 = (ks_b01 * (uy * uy + uy) + ks_b11 * (R(2) * uy + R(1)) + ks_b21) * h;
  const R ks_bc2 = (ks_b02 * (uy * uy + uy) + ks_b12 * (R(2) * uy + R(1)) + ks_b22) * h;
  const R ks_cc0 = (ks_c00 * (uy * uy + uy) + ks_c10 * (R(2) * uy + R(1)) + ks_c20) * h;
  const R ks_cc1 = (ks_c01 * (uy * uy + uy) + ks_c11 * (R(2) * uy + R(1)) + ks_c21) * h;
  const R ks_cc2 = (ks_c02 * (uy * uy + uy) + ks_c12 * (R(2) * uy + R(1)) + ks_c22) * h;

  f[11] = ks_aa0 * (R(1) - uz * uz) - R(2) * uz * ks_aa1 - ks_aa2;
  f[2] = ks_ab0 * (R(1) - uz * uz) - R(2) * uz * ks_ab1 - ks_ab2;
  f[15] = ks_ac0 * (R(1) - uz * uz) - R(2) * uz * ks_ac1 - ks_ac2;
  f[5] = ks_ba0 * (R(1) - uz * uz) - R(2) * uz * ks_ba1 - ks_ba2;
  f[0] = ks_bb0 * (R(1) - uz * uz) - R(2) * uz * ks_bb1 - ks_bb2;
  f[6] = ks_bc0 * (R(1) - uz * uz) - R(2) * uz * ks_bc1 - ks_bc2;
  f[16] = ks_ca0 * (R(1) - uz * uz) - R(2) * uz * ks_ca1 - ks_ca2;
  f[1] = ks_cb0 * (R(1) - uz * uz) - R(2) * uz * ks_cb1 - ks_cb2;
  f[12] = ks_cc0 * (R(1) - uz * uz) - R(2) * uz * ks_cc1 - ks_cc2;

  f[25] = (ks_aa0 * (uz * uz - uz) + ks_aa1 * (R(2) * uz - R(1)) + ks_aa2) * h;
  f[10] = (ks_ab0 * (uz * uz - uz) + ks_ab1 * (R(2) * uz - R(1)) + ks_ab2) * h;
  f[19] = (ks_ac0 * (uz * uz - uz) + ks_ac1 * (R(2) * uz - R(1)) + ks_ac2) * h;
  f[18] = (ks_ba0 * (uz * uz - uz) + ks_ba1 * (R(2) * uz - R(1)) + ks_ba2) * h;
  f[3] = (ks_bb0 * (uz * uz - uz) + ks_bb1 * (R(2) * uz - R(1)) + ks_bb2) * h;
  f[13] = (ks_bc0 * (uz * uz - uz) + ks_bc1 * (R(2) * uz - R(1)) + ks_bc2) * h;
  f[23] = (ks_ca0 * (uz * uz - uz) + ks_ca1 * (R(2) * uz - R(1)) + ks_ca2) * h;
  f[7] = (ks_cb0 * (uz * uz - uz) + ks_cb1 * (R(2) * uz - R(1)) + ks_cb2) * h;
  f[22] = (ks_cc0 * (uz * uz - uz) + ks_cc1 * (R(2) * uz - R(1)) + ks_cc2) * h;

  f[21] = (ks_aa0 * (uz * uz + uz) + ks_aa1 * (R(2) * uz + R(1)) + ks_aa2) * h;
  f[8] = (ks_ab0 * (uz * uz + uz) + ks_ab1 * (R(2) * uz + R(1)) + ks_ab2) * h;
  f[24] = (ks_ac0 * (uz * uz + uz) + ks_ac1 * (R(2) * uz + R(1)) + ks_ac2) * h;
  f[14] = (ks_ba0 * (uz * uz + uz) + ks_ba1 * (R(2) * uz + R(1)) + ks_ba2) * h;
  f[4] = (ks_bb0 * (uz * uz + uz) + ks_bb1 * (R(2) * uz + R(1)) + ks_bb2) * h;
  f[17] = (ks_bc0 * (uz * uz + uz) + ks_bc1 * (R(2) * uz + R(1)) + ks_bc2) * h;
  f[20] = (ks_ca0 * (uz * uz + uz) + ks_ca1 * (R(2) * uz + R(1)) + ks_ca2) * h;
  f[9] = (ks_cb0 * (uz * uz + uz) + ks_cb1 * (R(2) * uz + R(1)) + ks_cb2) * h;
  f[26] = (ks_cc0 * (uz * uz + uz) + ks_cc1 * (R(2) * uz + R(1)) + ks_cc2) * h;
}

// applyCollision: include/applyCollision.h:6-295 (cumulant "AllOne"); with a
// force g (three values, or null for none) the forcing version of
// M35Full/applyCollisionCustomized.h:5-17, which moves the velocity by half the
// force before the central moments.  The first-order central moments are then
// -g/2 and the backward step restores them as +g/2, so the cell's momentum
// gains g.
template <typename R>
void collide_g(R (&f)[27], R nu, R cs, const R* g) {
  typedef typename Wide<R>::type W;
  R rho, ux, uy, uz;
  get_rho_u(f, rho, ux, uy, uz);
  if (g) {
    ux = ((ux * rho) + g[0] / R(2)) / rho;
    uy = ((uy * rho) + g[1] / R(2)) / rho;
    uz = ((uz * rho) + g[2] / R(2)) / rho;
  }
  const Central<R> k = central_moments(f, ux, uy, uz);
  const R C_110 = k.k110, C_101 = k.k101, C_011 = k.k011;
  const R C_200 = k.k200, C_020 = k.k020, C_002 = k.k002;

  R feq[27], fneq[27];
  get_feq(rho, ux, uy, uz, feq);
  for (int q = 0; q < 27; ++q) fneq[q] = f[q] - feq[q];
  const R omega1 = get_omega_les(fneq, rho, nu, cs);

  // eqs 58-60 (lines 122-128); 3.0 promotes Dyv, Dzw to double
  const R Dxu = -omega1 * R(0.5) / rho * (R(2) * C_200 - C_020 - C_002) - R(0.5) / rho * (C_200 + C_020 + C_002 - k.k000);
  const R Dyv = (R)(W(Dxu) + W(3.0) * W(omega1) * W(R(0.5)) / W(rho) * W(C_200 - C_020));
  const R Dzw = (R)(W(Dxu) + W(3.0) * W(omega1) * W(R(0.5)) / W(rho) * W(C_200 - C_002));

  const R Cs_110 = (R(1) - omega1) * C_110;
  const R Cs_101 = (R(1) - omega1) * C_101;
  const R Cs_011 = (R(1) - omega1) * C_011;

  // eqs 61-65 (lines 142-148)
  const R Eq61 = (R)(W((R(1) - omega1) * (C_200 - C_020)) -
                     W(3.0) * W(rho) * W(R(1) - omega1 * R(0.5)) * W(ux * ux * Dxu - uy * uy * Dyv));
  const R Eq64 = (R)(W((R(1) - omega1) * (C_200 - C_002)) -
                     W(3.0) * W(rho) * W(R(1) - omega1 * R(0.5)) * W(ux * ux * Dxu - uz * uz * Dzw));
  const R Eq65 = (R)(W(k.k000) - W(3.0) * W(rho) * W(R(0.5)) * W(ux * ux * Dxu + uy * uy * Dyv + uz * uz * Dzw));
  const R Cs_200 = (R)(W(R(1)) / W(3.0) * W(Eq61 + Eq64 + Eq65));
  const R Cs_020 = (R)(W(R(1)) / W(3.0) * W(Eq64 - R(2) * Eq61 + Eq65));
  const R Cs_002 = (R)(W(R(1)) / W(3.0) * W(Eq61 - R(2) * Eq64 + Eq65));

  central_to_f(rho, ux, uy, uz, k.k000, -k.k100, -k.k010, -k.k001, Cs_200, Cs_020, Cs_002, Cs_110, Cs_101, Cs_011, f);
}
template <typename R>
void collide(R (&f)[27], R nu, R cs) {
  collide_g(f, nu, cs, (const R*)nullptr);
}

// rescaleF: include/gridRefinementFunctions.h:6-281 (paper eqs 3.17-3.20)
template <typename R>
void rescale_f(R (&f)[27], bool coarse_to_fine) {
  R rho, ux, uy, uz;
  get_rho_u(f, rho, ux, uy, uz);
  const Central<R> k = central_moments(f, ux, uy, uz);
  R k200, k020, k002, k011, k101, k110;
  if (coarse_to_fine) {
    k200 = (R(2) / R(3)) * k.k200 + (R(1) / R(6)) * k.k020 + (R(1) / R(6)) * k.k002;
    k020 = (R(1) / R(6)) * k.k200 + (R(2) / R(3)) * k.k020 + (R(1) / R(6)) * k.k002;
    k002 = (R(1) / R(6)) * k.k200 + (R(1) / R(6)) * k.k020 + (R(2) / R(3)) * k.k002;
    k011 = R(0.5) * k.k011;
    k101 = R(0.5) * k.k101;
    k110 = R(0.5) * k.k110;
  } else {
    k200 = (R(5) / R(3)) * k.k200 - (R(1) / R(3)) * k.k020 - (R(1) / R(3)) * k.k002;
    k020 = -(R(1) / R(3)) * k.k200 + (R(5) / R(3)) * k.k020 - (R(1) / R(3)) * k.k002;
    k002 = -(R(1) / R(3)) * k.k200 - (R(1) / R(3)) * k.k020 + (R(5) / R(3)) * k.k002;
    k011 = R(2) * k.k011;
    k101 = R(2) * k.k101;
    k110 = R(2) * k.k110;
  }
  central_to_f(rho, ux, uy, uz, k.k000, -k.k100, -k.k010, -k.k001, k200, k020, k002, k110, k101, k011, f);
}

// applyBounceback: include/boundaryConditions/applyBounceback.h:8-27
template <typename R>
void bounceback(R (&f)[27]) {
  for (int q = 1; q < 27; q += 2) std::swap(f[q], f[q + 1]);
}

// ------------------------------------------------------------ boundary conditions
struct OracleBcRecord {
  int normal[3];
  int code;
  int nk, nu;
  int known[27];
  int unknown[27];
  int mbbc_rows[27];
  int mbbc_smfk[27][27];
  int mbbc_inv[27][27][2];
  int rhou_full[4][27][2];
  int u_inv[3][3][2];
  int u_scmfk[3][27][2];
  int u_scq[3][2];
  int rho_scqrho[1][2];
  int rho_scmfk[27][2];
  int rho_scqu[3][2];
};
#include "bc_tables.inc"

const OracleBcRecord* find_bc(int code) {
  for (int i = 0; i < ORACLE_BC_COUNT; ++i)
    if (ORACLE_BC[i].code == code) return &ORACLE_BC[i];
  return nullptr;
}

// Accumulate sum_j c_j x_j left to right the way the reference's generated
// text is evaluated: skip zero coefficients, +-1 as add/subtract, other
// coefficients as (num/den)*x.
template <typename R>
R lincomb(const int (*coef)[2], const R* x, int n) {
  R acc = R(0);
  bool first = true;
  for (int j = 0; j < n; ++j) {
    const int num = coef[j][0], den = coef[j][1];
    if (num == 0) continue;
    R term;
    bool neg = false;
    if (den == 1 && num == 1) term = x[j];
    else if (den == 1 && num == -1) { term = x[j]; neg = true; }
    else term = (R(num) / R(den)) * x[j];
    if (first) acc = neg ? -term : term;
    else acc = neg ? acc - term : acc + term;
    first = false;
  }
  return acc;
}

// m_eq of the 23 moments (helperMain.py:30-53 ordering), same evaluation order
// as the generated reference text.
template <typename R>
R meq_moment(int m, R rho, R ux, R uy, R uz) {
  const R t = R(1) / R(3), n9 = R(1) / R(9);
  switch (m) {
    case 0: return rho;
    case 1: return rho * ux;
    case 2: return rho * uy;
    case 3: return rho * uz;
    case 4: return t * rho + rho * ux * ux;
    case 5: return t * rho + rho * uy * uy;
    case 6: return t * rho + rho * uz * uz;
    case 7: return rho * uy * uz;
    case 8: return rho * ux * uz;
    case 9: return rho * ux * uy;
    case 10: return R(0);
    case 11: return t * rho * uy + rho * ux * ux * uy;
    case 12: return t * rho * uz + rho * ux * ux * uz;
    case 13: return t * rho * ux + rho * ux * uy * uy;
    case 14: return t * rho * uz + rho * uy * uy * uz;
    case 15: return t * rho * ux + rho * ux * uz * uz;
    case 16: return t * rho * uy + rho * uy * uz * uz;
    case 17: return n9 * rho + t * rho * uy * uy + t * rho * uz * uz + rho * uy * uy * uz * uz;
    case 18: return n9 * rho + t * rho * ux * ux + t * rho * uz * uz + rho * ux * ux * uz * uz;
    case 19: return n9 * rho + t * rho * ux * ux + t * rho * uy * uy + rho * ux * ux * uy * uy;
    case 20: return t * rho * uy * uz;
    case 21: return t * rho * ux * uz;
    case 22: return t * rho * ux * uy;
  }
  return R(0);
}

// applyMBBC: include/boundaryConditions/applyMBBC.h (paper section 2.1)
template <typename R>
void bc_mbbc(int code, R (&f)[27], R rho, R ux, R uy, R uz) {
  const OracleBcRecord* b = find_bc(code);
  if (!b) return;
  R fk[27];
  for (int j = 0; j < b->nk; ++j) fk[j] = f[b->known[j]];
  R s[27];
  for (int i = 0; i < b->nu; ++i) {
    int c[27][2];
    for (int j = 0; j < b->nk; ++j) { c[j][0] = b->mbbc_smfk[i][j]; c[j][1] = 1; }
    const R smf = lincomb<R>(c, fk, b->nk);
    s[i] = meq_moment<R>(b->mbbc_rows[i], rho, ux, uy, uz) - smf;
  }
  for (int i = 0; i < b->nu; ++i) f[b->unknown[i]] = lincomb<R>(b->mbbc_inv[i], s, b->nu);
}

// restoreRhoUxUyUz: include/boundaryConditions/restoreRhoUxUyUz.h
template <typename R>
void bc_restore_rhou(int code, const R (&f)[27], R& rho, R& ux, R& uy, R& uz) {
  const OracleBcRecord* b = find_bc(code);
  if (!b) return;
  R fk[27];
  for (int j = 0; j < b->nk; ++j) fk[j] = f[b->known[j]];
  R q[4];
  for (int i = 0; i < 4; ++i) q[i] = lincomb<R>(b->rhou_full[i], fk, b->nk);
  rho = q[0];
  ux = q[1] / rho;
  uy = q[2] / rho;
  uz = q[3] / rho;
}

// restoreUxUyUz: include/boundaryConditions/restoreUxUyUz.h
template <typename R>
void bc_restore_u(int code, const R (&f)[27], R rho, R& ux, R& uy, R& uz) {
  const OracleBcRecord* b = find_bc(code);
  if (!b) return;
  R fk[27];
  for (int j = 0; j < b->nk; ++j) fk[j] = f[b->known[j]];
  R s[3];
  for (int i = 0; i < 3; ++i) {
    const R scmf = lincomb<R>(b->u_scmfk[i], fk, b->nk) / rho;
    s[i] = scmf - R(b->u_scq[i][0]) / R(b->u_scq[i][1]);
  }
  ux = lincomb<R>(b->u_inv[0], s, 3);
  uy = lincomb<R>(b->u_inv[1], s, 3);
  uz = lincomb<R>(b->u_inv[2], s, 3);
}

// restoreRho: include/boundaryConditions/restoreRho.h
template <typename R>
void bc_restore_rho(int code, const R (&f)[27], R& rho, R ux, R uy, R uz) {
  const OracleBcRecord* b = find_bc(code);
  if (!b) return;
  R fk[27];
  for (int j = 0; j < b->nk; ++j) fk[j] = f[b->known[j]];
  // scqu is written with every coefficient explicit, "(0.f) * uy" included
  const R u[3] = {ux, uy, uz};
  R scqu = R(0);
  for (int d = 0; d < 3; ++d) {
    const R term = (R(b->rho_scqu[d][0]) / R(b->rho_scqu[d][1])) * u[d];
    scqu = d == 0 ? term : scqu + term;
  }
  const R s = R(b->rho_scqrho[0][0]) / R(b->rho_scqrho[0][1]) + scqu;
  const R scmf = lincomb<R>(b->rho_scmfk, fk, b->nk);
  rho = scmf / s;
}

// ------------------------------------------------------------------- grid level
// Cell flag byte shared with the product's C ABI:
//   bits 0-2 kind, bit 3/4/5 periodic x/y/z (zero that normal component),
//   bits 6-7 index into the 4-entry table of given (rho,ux,uy,uz).
enum { K_FLUID = 0, K_BOUNCEBACK = 1, K_GIVEN_RHO = 2, K_GIVEN_U = 3, K_GIVEN_NONE = 4, K_GIVEN_RHOU = 5,
       K_GHOST = 6, K_COVERED = 7 };

struct OracleInfo {
  int nx, ny, nz;
  int i_sub_start, j_sub_start, k_sub_start, i_sub_end, j_sub_end, k_sub_end;  // all zero: no fine grid
  double nu;
  double cs;  // used when cs_cell == nullptr
};

inline long long cell_count(const OracleInfo& I) { return (long long)I.nx * I.ny * I.nz; }

// getShiftedIndex: include/applyStreaming.h:33-41
inline long long shifted(long long cell, int shift, long long N) {
  long long s = cell + shift;
  if (s >= N) s -= N;
  return s;
}

// applyStreaming(GridStruct&): include/applyStreaming.h:11-31.  The reference
// wraps the new shift at most twice, which is the reduction modulo N whenever
// the step's offset |cx + nx cy + nx ny cz| <= 2N (every grid with ny nz >= 2);
// on thinner grids (ny = nz = 1, or a single cell) it leaves the shift outside
// [0, N) and would index outside fArray, so here the shift is reduced modulo N.
void stream_shift(int* shift, int nx, int ny, int nz) {
  const long long N = (long long)nx * ny * nz;
  for (int q = 0; q < 27; ++q) {
    long long s = shift[q];
    s -= CX[q];
    s -= (long long)nx * CY[q];
    s -= (long long)ny * nx * CZ[q];
    s %= N;
    if (s < 0) s += N;
    shift[q] = (int)s;
  }
}

// getOuterNormal: include/cellFunctions.h:46-58
inline void outer_normal(int i, int j, int k, const OracleInfo& I, int& nx, int& ny, int& nz) {
  nx = ny = nz = 0;
  if (i == 0) nx = -1;
  else if (i == I.nx - 1) nx = 1;
  if (j == 0) ny = -1;
  else if (j == I.ny - 1) ny = 1;
  if (k == 0) nz = -1;
  else if (k == I.nz - 1) nz = 1;
}

template <typename R>
void scalar_collide(const R (&f)[27], R (&T)[27], R tauT);

// applyLocalCellUpdate(GridStruct&) (include/applyLocalCellUpdate.h:1-74) and,
// with Ta, applyLocalCellUpdate(GridStruct&,
// ScalarTransportStruct&) (include/scalarTransportFunctions.h:10-103): the
// scalar populations stream with f, bounce back with f, take a given value
// where givenT is not NaN, and collide after f (with the post-collision f).
template <typename R>
void local_update(R* fa, const int* shift, const OracleInfo& I, const uint8_t* flags, const R* bcvals,
                  const R* cs_cell, R* Ta = nullptr, const R* givenT = nullptr, R tauT = 1) {
  const long long N = cell_count(I);
  const int nxy = I.nx * I.ny;
#pragma omp parallel for schedule(static)
  for (long long cell = 0; cell < N; ++cell) {
    const int k = (int)(cell / nxy);
    const int rem = (int)(cell % nxy);
    const int j = rem / I.nx, i = rem % I.nx;
    const uint8_t fl = flags ? flags[cell] : 0;
    const int kind = fl & 7;
    if (kind == K_GHOST || kind == K_COVERED) continue;
    // sub-box skip (lines 25-30), computed from the bounds as the reference does
    if ((i > I.i_sub_start + 1 && i < I.i_sub_end - 2) && (j > I.j_sub_start + 1 && j < I.j_sub_end - 2) &&
        (k > I.k_sub_start + 1 && k < I.k_sub_end - 2))
      continue;
    R f[27], T[27];
    long long sl[27];
    for (int q = 0; q < 27; ++q) {
      sl[q] = shifted(cell, shift[q], N);
      f[q] = fa[(long long)q * N + sl[q]];
      if (Ta) T[q] = Ta[(long long)q * N + sl[q]];
    }
    if (kind == K_BOUNCEBACK) {
      bounceback(f);
      if (Ta) bounceback(T);
    } else {
      if (kind != K_FLUID) {
        int nx, ny, nz;
        outer_normal(i, j, k, I, nx, ny, nz);
        if (fl & 8) nx = 0;
        if (fl & 16) ny = 0;
        if (fl & 32) nz = 0;
        const int code = (nx + 5) * 100 + (ny + 5) * 10 + (nz + 5);
        const int slot = (fl >> 6) & 3;
        R rho = bcvals[4 * slot + 0], ux = bcvals[4 * slot + 1], uy = bcvals[4 * slot + 2], uz = bcvals[4 * slot + 3];
        if (kind == K_GIVEN_RHO) bc_restore_u(code, f, rho, ux, uy, uz);
        else if (kind == K_GIVEN_U) bc_restore_rho(code, f, rho, ux, uy, uz);
        else if (kind == K_GIVEN_NONE) bc_restore_rhou(code, f, rho, ux, uy, uz);
        bc_mbbc(code, f, rho, ux, uy, uz);
      }
      if (Ta && givenT && givenT[cell] == givenT[cell])  // applyScalarTransportGivenT, lines 172-175
        for (int q = 0; q < 27; ++q) T[q] = givenT[cell];
      const R cs = cs_cell ? cs_cell[cell] : (R)I.cs;
      collide(f, (R)I.nu, cs);
      if (Ta) scalar_collide(f, T, tauT);
    }
    for (int q = 0; q < 27; ++q) fa[(long long)q * N + sl[q]] = f[q];
    if (Ta)
      for (int q = 0; q < 27; ++q) Ta[(long long)q * N + sl[q]] = T[q];
  }
}

// applyScalarTransportCollision: include/scalarTransportFunctions.h:106-170
// (Zhang, Fan, Chen 2011, eqs 2.5-2.7).  The momentum-like sums run over the
// populations with a non-zero velocity component in increasing q, as written.
template <typename R>
void scalar_collide(const R (&f)[27], R (&T)[27], R tauT) {
  R rho, ux, uy, uz;
  get_rho_u(f, rho, ux, uy, uz);
  R Tsum = 0;
  for (int q = 0; q < 27; ++q) Tsum += f[q] * T[q];
  Tsum = Tsum / rho;
  R d[27];
  for (int q = 0; q < 27; ++q) d[q] = f[q] * (T[q] - Tsum);
  R S[3];
  const int* C[3] = {CX, CY, CZ};
  for (int a = 0; a < 3; ++a) {
    bool first = true;
    R acc = 0;
    for (int q = 0; q < 27; ++q) {
      if (!C[a][q]) continue;
      const R term = C[a][q] > 0 ? d[q] : -d[q];
      acc = first ? term : acc + term;
      first = false;
    }
    S[a] = acc;
  }
  const R rhodiv = (R)3 / rho;
  const R bracket = (R)1 - (R)1 / tauT;
  for (int q = 0; q < 27; ++q) {
    const R Fi = rhodiv * (((R)CX[q] - ux) * S[0] + ((R)CY[q] - uy) * S[1] + ((R)CZ[q] - uz) * S[2]);
    T[q] = Tsum + bracket * Fi;
  }
}

// applyBCPeriodicX / Y: include/boundaryConditions/periodicXY.h:38-82, 84-128
const int POS_CX[9] = {1, 7, 9, 12, 16, 20, 22, 23, 26};
const int NEG_CX[9] = {2, 8, 10, 11, 15, 19, 21, 24, 25};
const int POS_CY[9] = {6, 12, 13, 15, 17, 19, 22, 24, 26};
const int NEG_CY[9] = {5, 11, 14, 16, 18, 20, 21, 23, 25};
const int POS_CZ[9] = {4, 8, 9, 14, 17, 20, 21, 24, 26};
const int NEG_CZ[9] = {3, 7, 10, 13, 18, 19, 22, 23, 25};

template <typename R>
void periodic_copy(R* fa, const int* shift, const OracleInfo& I, int axis) {
  const long long N = cell_count(I);
  const int* pos = axis == 0 ? POS_CX : POS_CY;
  const int* neg = axis == 0 ? NEG_CX : NEG_CY;
  const int na = axis == 0 ? I.ny : I.nx;
  for (int k = 0; k < I.nz; ++k)
    for (int a = 0; a < na; ++a) {
      long long left, right;
      if (axis == 0) {
        left = (long long)k * I.nx * I.ny + (long long)a * I.nx + 0;
        right = (long long)k * I.nx * I.ny + (long long)a * I.nx + (I.nx - 1);
      } else {
        left = (long long)k * I.nx * I.ny + 0 * (long long)I.nx + a;
        right = (long long)k * I.nx * I.ny + (long long)(I.ny - 1) * I.nx + a;
      }
      for (int t = 0; t < 9; ++t) {
        const int q = pos[t];
        fa[(long long)q * N + shifted(left, shift[q], N)] = fa[(long long)q * N + shifted(right, shift[q], N)];
      }
      for (int t = 0; t < 9; ++t) {
        const int q = neg[t];
        fa[(long long)q * N + shifted(right, shift[q], N)] = fa[(long long)q * N + shifted(left, shift[q], N)];
      }
    }
}

// fillEquilibriumFromFunction: include/fillEquilibrium.h:1-22 (written at the
// unshifted index, as the reference does; callers use it with shifter = 0)
template <typename R>
void fill_equilibrium(R* fa, const OracleInfo& I, const R* rho, const R* ux, const R* uy, const R* uz) {
  const long long N = cell_count(I);
  for (long long c = 0; c < N; ++c) {
    R feq[27];
    get_feq(rho ? rho[c] : R(1), ux ? ux[c] : R(0), uy ? uy[c] : R(0), uz ? uz[c] : R(0), feq);
    for (int q = 0; q < 27; ++q) fa[(long long)q * N + c] = feq[q];
  }
}

// Momentum-exchange force over bounceback cells of a box [lo, hi):
// flowAroundSphere/flowAroundSphereMainSingle.cu:99-156 (drag = x component),
// NACA0012/NACA0012Main.cu:96-133 (lift = y component).  Accumulated in
// double (the FP64 surface reduction); out[3] = sum rho*(u - u_bounced).
template <typename R>
void momentum_exchange(const R* fa, const int* shift, const OracleInfo& I, const uint8_t* flags, const int* lo,
                       const int* hi, double* out) {
  const long long N = cell_count(I);
  double gx = 0, gy = 0, gz = 0;
  for (int k = lo[2]; k < hi[2]; ++k)
    for (int j = lo[1]; j < hi[1]; ++j)
      for (int i = lo[0]; i < hi[0]; ++i) {
        const long long cell = (long long)k * I.nx * I.ny + (long long)j * I.nx + i;
        if (!flags || (flags[cell] & 7) != K_BOUNCEBACK) continue;
        R f[27];
        for (int q = 0; q < 27; ++q) f[q] = fa[(long long)q * N + shifted(cell, shift[q], N)];
        R rho, ux, uy, uz;
        get_rho_u(f, rho, ux, uy, uz);
        bounceback(f);
        R rp, uxp, uyp, uzp;
        get_rho_u(f, rp, uxp, uyp, uzp);
        gx += (double)(rho * (ux - uxp));
        gy += (double)(rho * (uy - uyp));
        gz += (double)(rho * (uz - uzp));
      }
  out[0] = gx;
  out[1] = gy;
  out[2] = gz;
}

// Total mass sum_cells sum_q f over cells whose kind is not ghost/covered/bounceback.
template <typename R>
double total_mass(const R* fa, const int* shift, const OracleInfo& I, const uint8_t* flags) {
  const long long N = cell_count(I);
  double m = 0;
  for (long long c = 0; c < N; ++c) {
    const int kind = flags ? (flags[c] & 7) : 0;
    if (kind == K_GHOST || kind == K_COVERED || kind == K_BOUNCEBACK) continue;
    for (int q = 0; q < 27; ++q) m += (double)fa[(long long)q * N + shifted(c, shift[q], N)];
  }
  return m;
}

// ----------------------------------------------------------- refinement transfers
// writeToCoarseGridInterface: include/gridRefinementFunctions.h:283-336
template <typename R>
void write_to_coarse(R* fc, const int* sc, const OracleInfo& Ic, const R* ff, const int* sf, const OracleInfo& If,
                     const int* start, const int* end) {
  const long long Nc = cell_count(Ic), Nf = cell_count(If);
  for (int kc = start[2]; kc < end[2]; ++kc)
    for (int jc = start[1]; jc < end[1]; ++jc)
      for (int ic = start[0]; ic < end[0]; ++ic) {
        const long long cellc = (long long)kc * Ic.nx * Ic.ny + (long long)jc * Ic.nx + ic;
        const int i0 = (ic - Ic.i_sub_start) * 2, j0 = (jc - Ic.j_sub_start) * 2, k0 = (kc - Ic.k_sub_start) * 2;
        R f[27] = {0};
        for (int ka = 0; ka <= 1; ++ka)
          for (int ja = 0; ja <= 1; ++ja)
            for (int ia = 0; ia <= 1; ++ia) {
              const long long cf = (long long)(k0 + ka) * If.nx * If.ny + (long long)(j0 + ja) * If.nx + (i0 + ia);
              for (int q = 0; q < 27; ++q) f[q] += ff[(long long)q * Nf + shifted(cf, sf[q], Nf)];
            }
        for (int q = 0; q < 27; ++q) f[q] = f[q] * R(0.125);
        rescale_f(f, false);
        for (int q = 0; q < 27; ++q) fc[(long long)q * Nc + shifted(cellc, sc[q], Nc)] = f[q];
      }
}

// writeToFineGridInterface: include/gridRefinementFunctions.h:338-431.
// fine_ghost: per fine cell, non-zero when the fine marker is ghost.
template <typename R>
void write_to_fine(const R* fc, const int* sc, const OracleInfo& Ic, R* ff, const int* sf, const OracleInfo& If,
                   const uint8_t* fine_flags, const int* start, const int* end, const int* dirs9) {
  const long long Nc = cell_count(Ic), Nf = cell_count(If);
  for (int kf = start[2]; kf < end[2]; ++kf)
    for (int jf = start[1]; jf < end[1]; ++jf)
      for (int i_f = start[0]; i_f < end[0]; ++i_f) {
        const long long cellf = (long long)kf * If.nx * If.ny + (long long)jf * If.nx + i_f;
        const int ic = i_f / 2 + Ic.i_sub_start, jc = jf / 2 + Ic.j_sub_start, kc = kf / 2 + Ic.k_sub_start;
        const R dx = (i_f % 2 > 0) ? R(0.25) : R(-0.25);
        const R dy = (jf % 2 > 0) ? R(0.25) : R(-0.25);
        const R dz = (kf % 2 > 0) ? R(0.25) : R(-0.25);
        const long long cellc = (long long)kc * Ic.nx * Ic.ny + (long long)jc * Ic.nx + ic;
        R f[27];
        for (int q = 0; q < 27; ++q) f[q] = fc[(long long)q * Nc + shifted(cellc, sc[q], Nc)];
        int ip = ic + 1, im = ic - 1, jp = jc + 1, jm = jc - 1, kp = kc + 1, km = kc - 1;
        R dix = R(0.5), diy = R(0.5), diz = R(0.5);
        if (ip >= Ic.nx) { ip = ic; dix = R(1); }
        if (im < 0) { im = 0; dix = R(1); }
        if (jp >= Ic.ny) { jp = jc; diy = R(1); }
        if (jm < 0) { jm = 0; diy = R(1); }
        if (kp >= Ic.nz) { kp = kc; diz = R(1); }
        if (km < 0) { km = 0; diz = R(1); }
        const int ni[6] = {ip, im, ic, ic, ic, ic};
        const int nj[6] = {jc, jc, jp, jm, jc, jc};
        const int nk[6] = {kc, kc, kc, kc, kp, km};
        const R dinv[6] = {dix, -dix, diy, -diy, diz, -diz};
        const R pos[6] = {dx, dx, dy, dy, dz, dz};
        for (int n = 0; n < 6; ++n) {
          const long long cn = (long long)nk[n] * Ic.nx * Ic.ny + (long long)nj[n] * Ic.nx + ni[n];
          for (int q = 0; q < 27; ++q) f[q] += fc[(long long)q * Nc + shifted(cn, sc[q], Nc)] * dinv[n] * pos[n];
        }
        rescale_f(f, true);
        const bool ghost = fine_flags && ((fine_flags[cellf] & 7) == K_GHOST);
        if (ghost) {
          for (int t = 0; t < 9; ++t) ff[(long long)dirs9[t] * Nf + shifted(cellf, sf[dirs9[t]], Nf)] = f[dirs9[t]];
        } else {
          for (int q = 0; q < 27; ++q) ff[(long long)q * Nf + shifted(cellf, sf[q], Nf)] = f[q];
        }
      }
}

// applyCoarseFineGridCommunication: include/gridRefinementFunctions.h:433-526
template <typename R>
void coarse_fine_comm(R* fc, const int* sc, const OracleInfo& Ic, R* ff, const int* sf, const OracleInfo& If,
                      const uint8_t* fine_flags) {
  int s[3], e[3];
  if (Ic.nx > 1) {
    s[0] = Ic.i_sub_start + 1; s[1] = Ic.j_sub_start + 1; s[2] = Ic.k_sub_start + 1;
    e[0] = Ic.i_sub_start + 2; e[1] = Ic.j_sub_end - 1; e[2] = Ic.k_sub_end - 1;
    write_to_coarse(fc, sc, Ic, ff, sf, If, s, e);
    s[0] = Ic.i_sub_end - 2; e[0] = Ic.i_sub_end - 1;
    write_to_coarse(fc, sc, Ic, ff, sf, If, s, e);
  }
  if (Ic.ny > 1) {
    s[0] = Ic.i_sub_start + 1; s[1] = Ic.j_sub_start + 1; s[2] = Ic.k_sub_start + 1;
    e[0] = Ic.i_sub_end - 1; e[1] = Ic.j_sub_start + 2; e[2] = Ic.k_sub_end - 1;
    write_to_coarse(fc, sc, Ic, ff, sf, If, s, e);
    s[1] = Ic.j_sub_end - 2; e[1] = Ic.j_sub_end - 1;
    write_to_coarse(fc, sc, Ic, ff, sf, If, s, e);
  }
  if (Ic.nz > 1) {
    s[0] = Ic.i_sub_start + 1; s[1] = Ic.j_sub_start + 1; s[2] = Ic.k_sub_start + 1;
    e[0] = Ic.i_sub_end - 1; e[1] = Ic.j_sub_end - 1; e[2] = Ic.k_sub_start + 2;
    write_to_coarse(fc, sc, Ic, ff, sf, If, s, e);
    s[2] = Ic.k_sub_end - 2; e[2] = Ic.k_sub_end - 1;
    write_to_coarse(fc, sc, Ic, ff, sf, If, s, e);
  }
  if (Ic.i_sub_start > 0) {
    s[0] = 0; s[1] = 0; s[2] = 0; e[0] = 2; e[1] = If.ny; e[2] = If.nz;
    write_to_fine(fc, sc, Ic, ff, sf, If, fine_flags, s, e, POS_CX);
  }
  if (Ic.i_sub_end < Ic.nx - 1) {
    s[0] = If.nx - 2; s[1] = 0; s[2] = 0; e[0] = If.nx; e[1] = If.ny; e[2] = If.nz;
    write_to_fine(fc, sc, Ic, ff, sf, If, fine_flags, s, e, NEG_CX);
  }
  if (Ic.j_sub_start > 0) {
    s[0] = 0; s[1] = 0; s[2] = 0; e[0] = If.nx; e[1] = 2; e[2] = If.nz;
    write_to_fine(fc, sc, Ic, ff, sf, If, fine_flags, s, e, POS_CY);
  }
  if (Ic.j_sub_end < Ic.ny - 1) {
    s[0] = 0; s[1] = If.ny - 2; s[2] = 0; e[0] = If.nx; e[1] = If.ny; e[2] = If.nz;
    write_to_fine(fc, sc, Ic, ff, sf, If, fine_flags, s, e, NEG_CY);
  }
  if (Ic.k_sub_start > 0) {
    s[0] = 0; s[1] = 0; s[2] = 0; e[0] = If.nx; e[1] = If.ny; e[2] = 2;
    write_to_fine(fc, sc, Ic, ff, sf, If, fine_flags, s, e, POS_CZ);
  }
  if (Ic.k_sub_end < Ic.nz - 1) {
    s[0] = 0; s[1] = 0; s[2] = If.nz - 2; e[0] = If.nx; e[1] = If.ny; e[2] = If.nz;
    write_to_fine(fc, sc, Ic, ff, sf, If, fine_flags, s, e, NEG_CZ);
  }
}

// ------------------------------------------------------------------ esotwist
// getEsotwistReadIndex / getEsotwistWriteIndex: include/applyStreaming.h:82-192.
// Neighbour slot per direction: 0 = cell, 1 = i, 2 = j, 3 = k, 4 = ij, 5 = ik, 6 = jk, 7 = ijk.
const int ESO_READ[27] = {0, 0, 1, 3, 0, 2, 0, 3, 1, 0, 5, 4, 0, 3, 2, 1, 2, 0, 6, 5, 2, 4, 3, 6, 1, 7, 0};
const int ESO_WRITE[27] = {0, 1, 0, 0, 3, 0, 2, 1, 3, 5, 0, 0, 4, 2, 3, 2, 1, 6, 0, 2, 5, 3, 4, 1, 6, 0, 7};

// applyLocalCellUpdate(DIADGridStruct&): include/applyLocalCellUpdate.h:77-171.
// nbr: 7 arrays (i, j, k, ij, ik, jk, ijk) of length n; ijk: 3 arrays.
template <typename R>
void esotwist_update(R* fa, long long n, const int* const* nbr, const int* ijk_i, const int* ijk_j, const int* ijk_k,
                     const OracleInfo& I, const uint8_t* flags, const R* bcvals, bool flipper) {
#pragma omp parallel for schedule(static)
  for (long long cell = 0; cell < n; ++cell) {
    const uint8_t fl = flags ? flags[cell] : 0;
    // The DIAD update tests only the bounceback marker (line 90); every other
    // non-fluid marker goes through the boundary branch.
    int kind = fl & 7;
    if (kind == K_BOUNCEBACK) continue;
    if (kind == K_GHOST || kind == K_COVERED) kind = K_GIVEN_NONE;
    long long slot[8];
    slot[0] = cell;
    for (int t = 0; t < 7; ++t) slot[t + 1] = nbr[t][cell];
    R f[27];
    for (int q = 0; q < 27; ++q) {
      const int fi = flipper ? q : opp(q);
      f[q] = fa[(long long)fi * n + slot[ESO_READ[q]]];
    }
    if (kind != K_FLUID) {
      int nx, ny, nz;
      outer_normal(ijk_i[cell], ijk_j[cell], ijk_k[cell], I, nx, ny, nz);
      if (fl & 8) nx = 0;
      if (fl & 16) ny = 0;
      if (fl & 32) nz = 0;
      const int code = (nx + 5) * 100 + (ny + 5) * 10 + (nz + 5);
      const int s = (fl >> 6) & 3;
      R rho = bcvals[4 * s + 0], ux = bcvals[4 * s + 1], uy = bcvals[4 * s + 2], uz = bcvals[4 * s + 3];
      if (kind == K_GIVEN_RHO) bc_restore_u(code, f, rho, ux, uy, uz);
      else if (kind == K_GIVEN_U) bc_restore_rho(code, f, rho, ux, uy, uz);
      else if (kind == K_GIVEN_NONE) bc_restore_rhou(code, f, rho, ux, uy, uz);
      bc_mbbc(code, f, rho, ux, uy, uz);
    }
    collide(f, (R)I.nu, (R)I.cs);
    for (int q = 0; q < 27; ++q) {
      const int fi = flipper ? opp(q) : q;
      fa[(long long)fi * n + slot[ESO_WRITE[q]]] = f[q];
    }
  }
}

// Populations of a DIAD cell as the last update left them, through the write
// maps of the current parity (getEsotwistWriteIndex, include/applyStreaming.h:138-192).
template <typename R>
void eso_get(const R* fa, long long n, const int* const* nbr, long long cell, bool flipper, R (&f)[27]) {
  long long slot[8];
  slot[0] = cell;
  for (int t = 0; t < 7; ++t) slot[t + 1] = nbr[t][cell];
  for (int q = 0; q < 27; ++q) f[q] = fa[(long long)(flipper ? opp(q) : q) * n + slot[ESO_WRITE[q]]];
}
template <typename R>
void eso_put(R* fa, long long n, const int* const* nbr, long long cell, bool flipper, const R (&f)[27]) {
  long long slot[8];
  slot[0] = cell;
  for (int t = 0; t < 7; ++t) slot[t + 1] = nbr[t][cell];
  for (int q = 0; q < 27; ++q) fa[(long long)(flipper ? opp(q) : q) * n + slot[ESO_WRITE[q]]] = f[q];
}

// writeToCoarseGridInterface(DIAD): include/gridRefinementFunctions.h:540-622.
// Mean of the read cell and its +i +j +k +ij +ik +jk +ijk fine neighbours
// (summed in that order), rescaled fine -> coarse, into the coarse write cell.
template <typename R>
void eso_to_coarse(R* fc, long long nc, const int* const* nbc, bool flip_c, const R* ff, long long nf,
                   const int* const* nbf, bool flip_f, const int* write, const int* read, long long m) {
  for (long long t = 0; t < m; ++t) {
    const long long rc = read[t];
    long long fine[8];
    fine[0] = rc;
    for (int a = 0; a < 7; ++a) fine[a + 1] = nbf[a][rc];
    R f[27] = {0};
    for (int w = 0; w < 8; ++w) {
      R g[27];
      eso_get(ff, nf, nbf, fine[w], flip_f, g);
      for (int q = 0; q < 27; ++q) f[q] += g[q];
    }
    for (int q = 0; q < 27; ++q) f[q] = f[q] * (R)0.125;
    rescale_f(f, false);
    eso_put(fc, nc, nbc, write[t], flip_c, f);
  }
}

// writeToFineGridInterface(DIAD): include/gridRefinementFunctions.h:624-702.
// The coarse read cell, rescaled coarse -> fine, into the fine write cell and
// its +i +j +k +ij +ik +jk +ijk neighbours.
template <typename R>
void eso_to_fine(const R* fc, long long nc, const int* const* nbc, bool flip_c, R* ff, long long nf,
                 const int* const* nbf, bool flip_f, const int* write, const int* read, long long m) {
  for (long long t = 0; t < m; ++t) {
    R f[27];
    eso_get(fc, nc, nbc, read[t], flip_c, f);
    rescale_f(f, true);
    const long long wc = write[t];
    long long fine[8];
    fine[0] = wc;
    for (int a = 0; a < 7; ++a) fine[a + 1] = nbf[a][wc];
    for (int w = 0; w < 8; ++w) eso_put(ff, nf, nbf, fine[w], flip_f, f);
  }
}

// applyNonReflectiveOutletZ: include/boundaryConditions/applyNonReflectiveOutlet.h:9-73.
// Outlet cells in list order; kmin: the -k neighbour of every cell.
template <typename R>
void eso_outlet_z(R* fa, long long n, const int* const* nbr, bool flipper, const int* outlet, long long m,
                  const int* kmin) {
  static const int QZ[9] = {3, 7, 10, 13, 18, 19, 22, 23, 25};
  const float inv_sqrt3 = 0.577350269f;
  for (long long t = 0; t < m; ++t) {
    const long long cell = outlet[t];
    R fc[27], fk[27];
    eso_get(fa, n, nbr, cell, flipper, fc);
    eso_get(fa, n, nbr, (long long)kmin[cell], flipper, fk);
    long long slot[8];
    slot[0] = cell;
    for (int a = 0; a < 7; ++a) slot[a + 1] = nbr[a][cell];
    for (int a = 0; a < 9; ++a) {
      const int q = QZ[a];
      const R v = (R)inv_sqrt3 * fk[q] + (R)(1.f - inv_sqrt3) * fc[q];
      // read index of the inverted parity
      const int fi = !flipper ? q : opp(q);
      fa[(long long)fi * n + slot[ESO_READ[q]]] = v;
    }
  }
}

// applyMovingBounceback: include/boundaryConditions/applyMovingBounceback.h:8-62, on
// listed cells read through the read maps and written through the write maps
// (M35Full/applyLocalCellUpdateCustomized.h:69-77).  u3: 3 per listed cell.
template <typename R>
void eso_moving_bounceback(R* fa, long long n, const int* const* nbr, bool flipper, const int* cells, long long m,
                           const float* u3) {
  for (long long t = 0; t < m; ++t) {
    const long long cell = cells[t];
    long long slot[8];
    slot[0] = cell;
    for (int a = 0; a < 7; ++a) slot[a + 1] = nbr[a][cell];
    R f[27];
    for (int q = 0; q < 27; ++q) f[q] = fa[(long long)(flipper ? q : opp(q)) * n + slot[ESO_READ[q]]];
    const float ux = u3[3 * t], uy = u3[3 * t + 1], uz = u3[3 * t + 2];
    const float w1 = 4.0f / 9.0f, w2 = 1.0f / 9.0f, w3 = 1.0f / 36.0f;
    // (odd member of the pair, c . u as written in the reference, weight)
    const struct { int q; float cu, w; } P[13] = {
        {1, ux, w1}, {3, -uz, w1}, {5, -uy, w1},
        {7, ux - uz, w2}, {9, ux + uz, w2}, {11, -ux - uy, w2}, {13, uy - uz, w2}, {15, -ux + uy, w2},
        {17, uy + uz, w2},
        {19, -ux + uy - uz, w3}, {21, -ux - uy + uz, w3}, {23, ux - uy - uz, w3}, {25, -ux - uy - uz, w3}};
    for (const auto& p : P) {
      const R tmp = f[p.q];
      f[p.q] = f[p.q + 1] + (R)(p.w * p.cu);
      f[p.q + 1] = tmp - (R)(p.w * p.cu);
    }
    for (int q = 0; q < 27; ++q) fa[(long long)(flipper ? opp(q) : q) * n + slot[ESO_WRITE[q]]] = f[q];
  }
}

// getSphereDrag(DIADGridStruct&): flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:132-198,
// all three components: sum over bounceback cells of rho (u - u after bounceback).
template <typename R>
void eso_momentum_exchange(const R* fa, long long n, const int* const* nbr, bool flipper, const uint8_t* flags,
                           double* out3) {
  double acc[3] = {0, 0, 0};
  for (long long c = 0; c < n; ++c) {
    if ((flags[c] & 7) != K_BOUNCEBACK) continue;
    R f[27];
    eso_get(fa, n, nbr, c, flipper, f);
    R r, a, b, d;
    get_rho_u(f, r, a, b, d);
    bounceback(f);
    R r2, a2, b2, d2;
    get_rho_u(f, r2, a2, b2, d2);
    acc[0] += (double)(r * (a - a2));
    acc[1] += (double)(r * (b - b2));
    acc[2] += (double)(r * (d - d2));
  }
  for (int a = 0; a < 3; ++a) out3[a] = acc[a];
}

// Forced-velocity cells: M35Full/applyLocalCellUpdateCustomized.h:80-86,114-121
// with getForcing (M35Full/m35Main.cu:100-118): on each listed cell, read
// through the read maps, force g = rho (u_target - u) from the cell's own
// rho and u, the forcing collision, written through the write maps.  No
// box-face treatment.  ut3: 3 target velocities per listed cell.
template <typename R>
void eso_forced_velocity(R* fa, long long n, const int* const* nbr, bool flipper, const int* cells, long long m,
                         const float* ut3, R nu, R cs) {
  for (long long t = 0; t < m; ++t) {
    const long long cell = cells[t];
    long long slot[8];
    slot[0] = cell;
    for (int a = 0; a < 7; ++a) slot[a + 1] = nbr[a][cell];
    R f[27];
    for (int q = 0; q < 27; ++q) f[q] = fa[(long long)(flipper ? q : opp(q)) * n + slot[ESO_READ[q]]];
    R rho, ux, uy, uz;
    get_rho_u(f, rho, ux, uy, uz);
    const R g[3] = {rho * ((R)ut3[3 * t] - ux), rho * ((R)ut3[3 * t + 1] - uy), rho * ((R)ut3[3 * t + 2] - uz)};
    collide_g(f, nu, cs, g);
    for (int q = 0; q < 27; ++q) fa[(long long)(flipper ? opp(q) : q) * n + slot[ESO_WRITE[q]]] = f[q];
  }
}

// getTorque (M35Full/m35Main.cu:176-260) over the listed cells, all six
// components: the force g = rho (u_target - u) of each cell from the
// populations the next update will read (read maps of the next parity, as
// getTorque reads them, m35Main.cu:240-245), in newtons through convertToPhysicalForce
// (include/cellFunctions.h:221-226), its sum, and the sum of r x g about the
// origin with r = (i res + ox, j res + oy, k res + oz) in millimetres, divided
// by 1000 (newton metres).  out6 = (Fx, Fy, Fz, Tx, Ty, Tz); Tz is the
// reference's torque.
template <typename R>
void eso_forcing_sum(const R* fa, long long n, const int* const* nbr, bool flipper, const int* cells, long long m,
                     const float* ut3, const int* ijk3, float res, float ox, float oy, float oz, float dt,
                     float rho_phys, double* out6) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const float h = res / 1000.f;
  const float scale = rho_phys * h * h * h * h / (dt * dt);
  for (long long t = 0; t < m; ++t) {
    const long long cell = cells[t];
    long long slot[8];
    slot[0] = cell;
    for (int a = 0; a < 7; ++a) slot[a + 1] = nbr[a][cell];
    R f[27];
    for (int q = 0; q < 27; ++q) f[q] = fa[(long long)(flipper ? opp(q) : q) * n + slot[ESO_READ[q]]];
    R rho, ux, uy, uz;
    get_rho_u(f, rho, ux, uy, uz);
    const R gx = rho * ((R)ut3[3 * t] - ux) * (R)scale;
    const R gy = rho * ((R)ut3[3 * t + 1] - uy) * (R)scale;
    const R gz = rho * ((R)ut3[3 * t + 2] - uz) * (R)scale;
    const R x = (R)(ijk3[cell] * res + ox), y = (R)(ijk3[n + cell] * res + oy), z = (R)(ijk3[2 * n + cell] * res + oz);
    acc[0] += (double)gx;
    acc[1] += (double)gy;
    acc[2] += (double)gz;
    acc[3] += (double)(y * gz - z * gy);
    acc[4] += (double)(z * gx - x * gz);
    acc[5] += (double)(-gx * y + gy * x);
  }
  for (int a = 0; a < 3; ++a) out6[a] = acc[a];
  for (int a = 3; a < 6; ++a) out6[a] = acc[a] / 1000.0;
}

// applyZeroGradientOutletScalarTransportBC (scalarTransport/scalarTransportMain.cu:95-135):
// every non-bounceback cell of the last x layer takes the 27 scalar
// populations of its -x neighbour, both read at their shifted slots.
template <typename R>
void scalar_outlet_x(R* Ta, const int* shift, const OracleInfo& I, const uint8_t* flags) {
  const long long n = (long long)I.nx * I.ny * I.nz;
  for (int k = 0; k < I.nz; ++k)
    for (int j = 0; j < I.ny; ++j) {
      const long long cw = ((long long)k * I.ny + j) * I.nx + (I.nx - 1), cr = cw - 1;
      if (flags && (flags[cw] & 7) == K_BOUNCEBACK) continue;
      for (int q = 0; q < 27; ++q) Ta[(long long)q * n + (cw + shift[q]) % n] = Ta[(long long)q * n + (cr + shift[q]) % n];
    }
}

// ------------------------------------------------------------------- STL voxelizer
// getRayHitYesNo: include/STLFunctions.h:184-245 (exact integer predicate)
bool ray_hit(long long ax, long long ay, long long bx, long long by, long long cx, long long cy) {
  const long long abx = bx - ax, aby = by - ay, bcx = cx - bx, bcy = cy - by, cax = ax - cx, cay = ay - cy;
  const long long area = abx * bcy - aby * bcx;
  if (area == 0) return false;
  const long long qz = area < 0 ? -1 : 1;
  const long long wab = qz * (abx * (-by) - aby * (-bx));
  const long long wbc = qz * (bcx * (-cy) - bcy * (-cx));
  const long long wca = qz * (cax * (-ay) - cay * (-ax));
  if (wab > 0 && wbc > 0 && wca > 0) return true;
  if (wab < 0 || wbc < 0 || wca < 0) return false;
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

// getRayHitZCoordinate: include/STLFunctions.h:247-275
float ray_hit_z(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz, float rx,
                float ry) {
  const float v1x = bx - ax, v1y = by - ay, v1z = bz - az;
  const float v2x = cx - ax, v2y = cy - ay, v2z = cz - az;
  const float nx = v1y * v2z - v1z * v2y;
  const float ny = v1z * v2x - v1x * v2z;
  const float nz = v1x * v2y - v1y * v2x;
  const float zmax = std::max(az, std::max(bz, cz)), zmin = std::min(az, std::min(bz, cz));
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

struct Tri { float ax, ay, az, bx, by, bz, cx, cy, cz; };

// Lattice-point footprint of one triangle on the 100x-finer integer grid
// (include/STLFunctions.h:297-320).
struct Footprint { long long ak, al, bk, bl, ck, cl; long long imin, imax, jmin, jmax; };
Footprint footprint(const float* ax, const float* ay, const float* bx, const float* by, const float* cx,
                    const float* cy, float res, int nx, int ny) {
  Footprint F;
  const float scale = 50.0f / res;
  F.ak = (long long)(std::round(*ax * scale)) * 2 + 1;
  F.al = (long long)(std::round(*ay * scale)) * 2 + 1;
  F.bk = (long long)(std::round(*bx * scale)) * 2 + 1;
  F.bl = (long long)(std::round(*by * scale)) * 2 + 1;
  F.ck = (long long)(std::round(*cx * scale)) * 2 + 1;
  F.cl = (long long)(std::round(*cy * scale)) * 2 + 1;
  const long long XM = (long long)(nx - 1) * 100, YM = (long long)(ny - 1) * 100;
  const long long kmin = std::max(0LL, std::min(std::min(F.ak, F.bk), std::min(F.ck, XM)));
  const long long kmax = std::min(XM, std::max(std::max(F.ak, F.bk), std::max(F.ck, 0LL)));
  const long long lmin = std::max(0LL, std::min(std::min(F.al, F.bl), std::min(F.cl, YM)));
  const long long lmax = std::min(YM, std::max(std::max(F.al, F.bl), std::max(F.cl, 0LL)));
  F.imin = (kmin + 99) / 100;
  F.imax = kmax / 100;
  F.jmin = (lmin + 99) / 100;
  F.jmax = lmax / 100;
  return F;
}

// Sorted hit layers per (i, j) column, padded to the largest count with
// cellCountZ (include/STLFunctions.h:296-389: counter pass, hit-index pass,
// per-column insertion sort).  Triangle vertices are moved into grid
// coordinates as (vertex + STL origin) - grid origin, in that order.
struct Columns {
  int nx = 0, ny = 0, nz = 0;
  size_t maxc = 1;
  std::vector<int> layers;  // maxc x (nx * ny), layer-major
  int at(size_t layer, int i, int j) const { return layers[layer * (size_t)nx * ny + (size_t)j * nx + i]; }
};

Columns stl_columns(const float* tri, int ntri, float sox, float soy, float soz, float gox, float goy, float goz,
                    float res, int nx, int ny, int nz) {
  std::vector<std::vector<int>> cols((size_t)nx * ny);
  for (int t = 0; t < ntri; ++t) {
    const float* p = tri + 9 * (size_t)t;
    const float ax = p[0] + sox - gox, ay = p[1] + soy - goy, az = p[2] + soz - goz;
    const float bx = p[3] + sox - gox, by = p[4] + soy - goy, bz = p[5] + soz - goz;
    const float cx = p[6] + sox - gox, cy = p[7] + soy - goy, cz = p[8] + soz - goz;
    const Footprint F = footprint(&ax, &ay, &bx, &by, &cx, &cy, res, nx, ny);
    for (long long j = F.jmin; j <= F.jmax; ++j)
      for (long long i = F.imin; i <= F.imax; ++i) {
        const long long rk = i * 100, rl = j * 100;
        if (!ray_hit(F.ak - rk, F.al - rl, F.bk - rk, F.bl - rl, F.ck - rk, F.cl - rl)) continue;
        const float rx = (int)i * res;
        const float ry = (int)j * res;
        const float rz = ray_hit_z(ax, ay, az, bx, by, bz, cx, cy, cz, rx, ry);
        int k = (int)std::max(0.0f, std::ceil(rz / res));
        k = std::min(k, nz);
        cols[(size_t)j * nx + i].push_back(k);
      }
  }
  Columns C;
  C.nx = nx; C.ny = ny; C.nz = nz;
  for (auto& c : cols) C.maxc = std::max(C.maxc, c.size());
  C.layers.assign(C.maxc * (size_t)nx * ny, nz);  // unused layers hold cellCountZ (line 369)
  for (size_t col = 0; col < cols.size(); ++col) {
    std::vector<int> v = cols[col];
    v.resize(C.maxc, nz);
    std::sort(v.begin(), v.end());
    for (size_t l = 0; l < C.maxc; ++l) C.layers[l * (size_t)nx * ny + col] = v[l];
  }
  return C;
}

// Parity of cell (i, j, k) in its column (include/STLFunctions.h:371-389 for
// whole grids, 449-477 for cell lists): the value of the interval holding k,
// starting outside.
bool column_value(const Columns& C, int i, int j, int k, int inside_value) {
  bool val = !inside_value;
  for (size_t iv = 0; iv <= C.maxc; ++iv) {
    int s0, e0;
    if (iv == 0) { s0 = 0; e0 = C.at(0, i, j); }
    else if (iv == C.maxc) { s0 = C.at(C.maxc - 1, i, j); e0 = C.nz; }
    else { s0 = C.at(iv - 1, i, j); e0 = C.at(iv, i, j); }
    if (k >= s0 && k < e0) return val;
    val = !val;
  }
  return false;  // not reached for 0 <= k < nz
}

// applyMarkersInsideSTL: include/STLFunctions.h:277-479.  tri: n x 9 floats
// (a,b,c vertices) already in STL coordinates; (sox,soy,soz) STL origin,
// (gox,goy,goz) grid origin.  Writes marker[cell] (0/1) for every cell.
void markers_inside_stl(const float* tri, int ntri, float sox, float soy, float soz, float gox, float goy, float goz,
                        float res, int nx, int ny, int nz, int inside_value, uint8_t* marker) {
  const Columns C = stl_columns(tri, ntri, sox, soy, soz, gox, goy, goz, res, nx, ny, nz);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      bool val = !inside_value;
      for (size_t iv = 0; iv <= C.maxc; ++iv) {
        int s0, e0;
        if (iv == 0) { s0 = 0; e0 = C.at(0, i, j); }
        else if (iv == C.maxc) { s0 = C.at(C.maxc - 1, i, j); e0 = nz; }
        else { s0 = C.at(iv - 1, i, j); e0 = C.at(iv, i, j); }
        for (int k = s0; k < e0; ++k) marker[(size_t)k * nx * ny + (size_t)j * nx + i] = val ? 1 : 0;
        val = !val;
      }
    }
}

// checkSTLEdges: include/STLFunctions.h:92-182 -- per-triangle edge share
// counts; returns the number of triangles with any count != 2.
int check_stl_edges(const float* tri, int ntri, int* counts) {
  for (int t = 0; t < 3 * ntri; ++t) counts[t] = 0;
  for (int t1 = 0; t1 < ntri; ++t1) {
    const float* a = tri + 9 * (size_t)t1;
    for (int t2 = 0; t2 < ntri; ++t2) {
      const float* b = tri + 9 * (size_t)t2;
      for (int e1 = 0; e1 < 3; ++e1)
        for (int e2 = 0; e2 < 3; ++e2) {
          const float* a0 = a + 3 * e1;
          const float* a1 = a + 3 * ((e1 + 1) % 3);
          const float* b0 = b + 3 * e2;
          const float* b1 = b + 3 * ((e2 + 1) % 3);
          if (a0[0] == b0[0] && a0[1] == b0[1] && a0[2] == b0[2] && a1[0] == b1[0] && a1[1] == b1[1] && a1[2] == b1[2])
            counts[3 * t1 + e1]++;
          if (a0[0] == b1[0] && a0[1] == b1[1] && a0[2] == b1[2] && a1[0] == b0[0] && a1[1] == b0[1] && a1[2] == b0[2])
            counts[3 * t1 + e1]++;
        }
    }
  }
  int bad = 0;
  for (int t = 0; t < ntri; ++t)
    if (counts[3 * t] != 2 || counts[3 * t + 1] != 2 || counts[3 * t + 2] != 2) ++bad;
  return bad;
}

// ------------------------------------------------------------ DIAD grid builder
// buildDIADGrids: include/DIADFunctions.h:640-1046, with its helpers from the
// same file (sortIJK 56-123, findMatchingIJKIndex 296-357,
// findPeriodicNeighbourIJKIndex 360-429, markFinestFluidBounceback 500-555,
// getDIADNeighbours 557-574, getDIADEsotwistNbrArray 576-612,
// markDIADNeighbours 615-638).  The drivers' getMarkers callbacks are given as
// an ordered rule list: every rule for the queried marker whose level mask
// holds the level's gridID and whose shape contains the cell sets the marker
// to the rule's value (a driver's sequence of `if (...) Marker.x = v;`).
// Positions are iCell * res + ox etc. in FP32, as in the drivers.
struct DRule { int what, value; unsigned level_mask; int shape; float p[8]; int idx[6]; };
enum { R_BOUNCEBACK = 0, R_REFINEMENT = 1, R_OUTLET = 2 };
enum { S_SPHERE_LE = 0, S_SPHERE_LT = 1, S_XCYLINDER = 2, S_BOX = 3, S_INDEX_BOX = 4 };

struct DInfo { int gridID; float res, ox, oy, oz, nu, dt; int nx, ny, nz; };

bool rule_contains(const DRule& r, const DInfo& I, int i, int j, int k) {
  if (r.shape == S_INDEX_BOX) {
    auto in = [](int v, int lo, int hi, int n) {
      if (lo < 0) lo += n;
      if (hi < 0) hi += n;
      return v >= lo && v <= hi;
    };
    return in(i, r.idx[0], r.idx[1], I.nx) && in(j, r.idx[2], r.idx[3], I.ny) && in(k, r.idx[4], r.idx[5], I.nz);
  }
  const float x = (float)i * I.res + I.ox, y = (float)j * I.res + I.oy, z = (float)k * I.res + I.oz;
  if (r.shape == S_BOX) return x > r.p[0] && x < r.p[3] && y > r.p[1] && y < r.p[4] && z > r.p[2] && z < r.p[5];
  const float dx = x - r.p[0], dy = y - r.p[1], dz = z - r.p[2];
  const float w2 = r.p[3] * r.p[3];
  if (r.shape == S_XCYLINDER) {
    const float a2 = dy * dy + dz * dz;
    return dx > 0.f && dx < r.p[4] && a2 < w2;
  }
  const float r2 = dx * dx + dy * dy + dz * dz;
  return r.shape == S_SPHERE_LE ? r2 <= w2 : r2 < w2;
}

struct DStl { const float* tri; int n; float ox, oy, oz; };

struct DLevel {
  DInfo info{};
  std::vector<int> I, J, K;
  std::vector<uint8_t> bb, iface, enfB, enfF, c2f, f2c;
  std::vector<int> nbr10, c2f_read, c2f_write, f2c_write, f2c_read, outlet;
  size_t n() const { return I.size(); }
};

struct DBuild {
  int levels = 1, span = 0;
  float nu_phys = 0.f;
  std::vector<DStl> stls;
  std::vector<DRule> rules;
  std::vector<DLevel> lv;
  std::string err;
};

bool ijk_less(int ia, int ja, int ka, int ib, int jb, int kb) {
  if (ka != kb) return ka < kb;
  if (ja != jb) return ja < jb;
  return ia < ib;
}

// binary search in cells sorted by (k, j, i); -1 when absent
int find_cell(const DLevel& L, int i, int j, int k) {
  size_t lo = 0, hi = L.n();
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (ijk_less(L.I[mid], L.J[mid], L.K[mid], i, j, k)) lo = mid + 1;
    else hi = mid;
  }
  if (lo < L.n() && L.I[lo] == i && L.J[lo] == j && L.K[lo] == k) return (int)lo;
  return -1;
}

template <class T>
void permute(std::vector<T>& a, const std::vector<size_t>& order) {
  if (a.empty()) return;
  std::vector<T> b(order.size());
  for (size_t x = 0; x < order.size(); ++x) b[x] = a[order[x]];
  a.swap(b);
}

// sortIJK (with the enforce markers of finer levels)
void sort_cells(DLevel& L) {
  std::vector<size_t> order(L.n());
  for (size_t x = 0; x < order.size(); ++x) order[x] = x;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return ijk_less(L.I[a], L.J[a], L.K[a], L.I[b], L.J[b], L.K[b]);
  });
  permute(L.I, order); permute(L.J, order); permute(L.K, order);
  permute(L.enfB, order); permute(L.enfF, order);
}

// sortFinestGrid / sortCoarseGrid + resize: the cells are already sorted by
// (k, j, i), so sorting by (keep desc, k, j, i) and truncating keeps the kept
// cells in order.
template <class T>
void keep_only(std::vector<T>& a, const std::vector<uint8_t>& keep) {
  if (a.empty()) return;
  std::vector<T> b;
  for (size_t x = 0; x < keep.size(); ++x)
    if (keep[x]) b.push_back(a[x]);
  a.swap(b);
}

// directions of getDIADNeighbours (include/DIADFunctions.h:561-563)
const int DN_CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1};
const int DN_CY[27] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1};
const int DN_CZ[27] = {0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1};

std::vector<int> neighbours26(const DLevel& L) {
  const size_t n = L.n();
  std::vector<int> nb(26 * n);
  for (int d = 1; d < 27; ++d)
    for (size_t c = 0; c < n; ++c)
      nb[(size_t)(d - 1) * n + c] = find_cell(L, L.I[c] + DN_CX[d], L.J[c] + DN_CY[d], L.K[c] + DN_CZ[d]);
  return nb;
}

// markDIADNeighbours: target |= some existing neighbour is set in source
void mark_neighbours(const std::vector<int>& nb, const std::vector<uint8_t>& src, std::vector<uint8_t>& tgt) {
  const size_t n = src.size();
  for (int d = 0; d < 26; ++d)
    for (size_t c = 0; c < n; ++c) {
      const int m = nb[(size_t)d * n + c];
      if (m >= 0 && src[m]) tgt[c] = 1;
    }
}

// ApplyMarkersInsideSTL on a cell list (include/STLFunctions.h:482-479+)
std::vector<uint8_t> stl_inside(const DLevel& L, const DInfo& I, const DStl& S) {
  const Columns C = stl_columns(S.tri, S.n, S.ox, S.oy, S.oz, I.ox, I.oy, I.oz, I.res, I.nx, I.ny, I.nz);
  std::vector<uint8_t> m(L.n());
  for (size_t c = 0; c < L.n(); ++c) m[c] = column_value(C, L.I[c], L.J[c], L.K[c], 1) ? 1 : 0;
  return m;
}

// apply{Bounceback,Refinement,Outlet}MarkerFromFunction (lines 432-498)
void apply_rules(const DBuild& B, int what, const DLevel& L, const DInfo& I, std::vector<uint8_t>& m) {
  for (size_t c = 0; c < L.n(); ++c) {
    int v = m[c];
    for (const DRule& r : B.rules)
      if (r.what == what && ((r.level_mask >> I.gridID) & 1u) && rule_contains(r, I, L.I[c], L.J[c], L.K[c]))
        v = r.value;
    m[c] = v ? 1 : 0;
  }
}

std::vector<uint8_t> stl_union(const DBuild& B, const DLevel& L, const DInfo& I) {
  std::vector<uint8_t> m(L.n(), 0);
  for (const DStl& S : B.stls) {
    const std::vector<uint8_t> s = stl_inside(L, I, S);
    for (size_t c = 0; c < m.size(); ++c) m[c] |= s[c];
  }
  return m;
}

// getDIADEsotwistNbrArray: +i +j +k +ij +ik +jk +ijk -i -j -k, periodic walk
// until an existing cell is met (the cell itself at the latest)
void esotwist_neighbours(DLevel& L) {
  static const int D[10][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1},
                               {0, 1, 1}, {1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  const size_t n = L.n();
  const DInfo& I = L.info;
  L.nbr10.assign(10 * n, -1);
  for (int d = 0; d < 10; ++d)
    for (size_t c = 0; c < n; ++c) {
      int ti = L.I[c], tj = L.J[c], tk = L.K[c], r = -1;
      while (r < 0) {
        ti = ((ti + D[d][0]) % I.nx + I.nx) % I.nx;
        tj = ((tj + D[d][1]) % I.ny + I.ny) % I.ny;
        tk = ((tk + D[d][2]) % I.nz + I.nz) % I.nz;
        r = find_cell(L, ti, tj, tk);
      }
      L.nbr10[(size_t)d * n + c] = r;
    }
}

std::vector<uint8_t> b_not(const std::vector<uint8_t>& a) {
  std::vector<uint8_t> r(a.size());
  for (size_t c = 0; c < a.size(); ++c) r[c] = !a[c];
  return r;
}
void b_or(std::vector<uint8_t>& a, const std::vector<uint8_t>& b) { for (size_t c = 0; c < a.size(); ++c) a[c] |= b[c]; }
void b_and(std::vector<uint8_t>& a, const std::vector<uint8_t>& b) { for (size_t c = 0; c < a.size(); ++c) a[c] &= b[c]; }

// markFinestFluidBounceback: include/DIADFunctions.h:500-555
void finest_fluid_bounceback(const DBuild& B, const DLevel& L, std::vector<uint8_t>& fluid,
                             std::vector<uint8_t>& finest_bb) {
  const DInfo I0 = L.info;
  fluid.assign(L.n(), 0);
  finest_bb.assign(L.n(), 0);
  for (int tl = I0.gridID; tl < B.levels; ++tl) {
    const int shifts = 1 << (tl - I0.gridID);
    const float target_res = I0.res / shifts;
    const float start = -0.5f * I0.res + 0.5f * target_res;
    for (int sx = 0; sx < shifts; ++sx)
      for (int sy = 0; sy < shifts; ++sy)
        for (int sz = 0; sz < shifts; ++sz) {
          DInfo I = I0;
          I.ox = I0.ox + start + sx * target_res;
          I.oy = I0.oy + start + sy * target_res;
          I.oz = I0.oz + start + sz * target_res;
          std::vector<uint8_t> m = stl_union(B, L, I);
          apply_rules(B, R_BOUNCEBACK, L, I, m);
          b_or(finest_bb, m);
          b_or(fluid, b_not(m));
        }
  }
}

void build_level(DBuild& B, int level) {
  DLevel& L = B.lv[level];
  const DInfo& I = L.info;
  sort_cells(L);
  const bool multi = B.levels > 1;
  if (level == B.levels - 1) {
    const std::vector<int> nb = neighbours26(L);
    std::vector<uint8_t> bb = stl_union(B, L, I);
    apply_rules(B, R_BOUNCEBACK, L, I, bb);
    if (multi) b_or(bb, L.enfB);
    std::vector<uint8_t> fluid = b_not(bb);
    if (multi) b_or(fluid, L.enfF);
    std::vector<uint8_t> keep = fluid;
    mark_neighbours(nb, fluid, keep);
    L.bb = b_not(fluid);
    L.iface.assign(L.n(), 0);
    L.enfB.clear(); L.enfF.clear();
    keep_only(L.I, keep); keep_only(L.J, keep); keep_only(L.K, keep);
    keep_only(L.bb, keep); keep_only(L.iface, keep);
    esotwist_neighbours(L);
    return;
  }
  const std::vector<int> nb = neighbours26(L);
  std::vector<uint8_t> fluid, finest_bb;
  finest_fluid_bounceback(B, L, fluid, finest_bb);
  if (level != 0) b_or(fluid, L.enfF);
  std::vector<uint8_t> border(L.n(), 0);
  mark_neighbours(nb, fluid, border);
  b_and(border, b_not(fluid));
  std::vector<uint8_t> keep = fluid;
  b_or(keep, border);
  std::vector<uint8_t> ref = finest_bb;
  b_or(ref, border);
  std::vector<uint8_t> new_ref(L.n(), 0);
  const int thickness = B.span + (B.levels - level - 1);
  for (int layer = 0; layer < thickness; ++layer) {
    mark_neighbours(nb, ref, new_ref);
    b_or(new_ref, ref);
    ref = new_ref;
    b_and(ref, keep);
  }
  apply_rules(B, R_REFINEMENT, L, I, ref);
  std::vector<uint8_t> f2c(L.n(), 0);
  mark_neighbours(nb, ref, f2c);
  b_and(f2c, keep);
  b_and(f2c, b_not(ref));
  std::vector<uint8_t> c2f(L.n(), 0);
  mark_neighbours(nb, f2c, c2f);
  b_and(c2f, keep);
  b_and(c2f, b_not(f2c));
  b_and(c2f, b_not(ref));
  b_or(ref, f2c);
  b_or(ref, c2f);
  const std::vector<uint8_t> ref_inv = b_not(ref);
  std::vector<uint8_t> coarse_bb(L.n(), 0);
  apply_rules(B, R_BOUNCEBACK, L, I, coarse_bb);
  b_or(coarse_bb, stl_union(B, L, I));
  std::vector<uint8_t> under = c2f;
  b_or(under, f2c);
  std::vector<uint8_t> bb_under = under, fluid_under = under;
  b_and(bb_under, coarse_bb);
  b_and(fluid_under, b_not(coarse_bb));
  keep = fluid;
  b_or(keep, border);
  b_and(keep, ref_inv);
  b_or(keep, c2f);
  b_or(keep, f2c);
  if (level != 0) b_or(keep, L.enfF);

  // children of the refined cells, in coarse order, (i, j, k) offsets nested i-j-k
  DLevel& F = B.lv[level + 1];
  F.I.clear(); F.J.clear(); F.K.clear(); F.enfB.clear(); F.enfF.clear();
  for (size_t c = 0; c < L.n(); ++c) {
    if (!ref[c]) continue;
    for (int ip = 0; ip < 2; ++ip)
      for (int jp = 0; jp < 2; ++jp)
        for (int kp = 0; kp < 2; ++kp) {
          F.I.push_back(2 * L.I[c] + ip);
          F.J.push_back(2 * L.J[c] + jp);
          F.K.push_back(2 * L.K[c] + kp);
          F.enfB.push_back(bb_under[c]);
          F.enfF.push_back(fluid_under[c]);
        }
  }
  keep_only(L.I, keep); keep_only(L.J, keep); keep_only(L.K, keep);
  keep_only(f2c, keep); keep_only(c2f, keep);
  keep_only(L.enfB, keep); keep_only(L.enfF, keep);
  L.iface = f2c;
  b_or(L.iface, c2f);
  std::vector<uint8_t> bb(L.n(), 0);
  apply_rules(B, R_BOUNCEBACK, L, I, bb);
  b_or(bb, stl_union(B, L, I));
  if (level != 0) {
    b_or(bb, L.enfB);
    b_and(bb, b_not(L.enfF));
  }
  L.enfB.clear(); L.enfF.clear();
  L.bb = bb;
  b_and(c2f, b_not(bb));
  b_and(f2c, b_not(bb));
  L.c2f = c2f;
  L.f2c = f2c;
  esotwist_neighbours(L);
  std::vector<uint8_t> out(L.n(), 0);
  apply_rules(B, R_OUTLET, L, I, out);
  L.outlet.clear();
  for (size_t c = 0; c < L.n(); ++c)
    if (out[c]) L.outlet.push_back((int)c);

  DInfo& FI = F.info;
  FI.gridID = I.gridID + 1;
  FI.res = I.res * 0.5f;
  FI.dt = I.dt * 0.5f;
  FI.nu = (FI.dt * B.nu_phys) / ((FI.res / 1000.f) * (FI.res / 1000.f));
  FI.ox = I.ox - FI.res * 0.5f;
  FI.oy = I.oy - FI.res * 0.5f;
  FI.oz = I.oz - FI.res * 0.5f;
  FI.nx = I.nx * 2;
  FI.ny = I.ny * 2;
  FI.nz = I.nz * 2;
}

// mapInterfaceLambda (lines 964-1043): coarse cells of a marker, in order, and
// the fine cell at (2i, 2j, 2k)
bool map_interface(DBuild& B, int level, const std::vector<uint8_t>& m, std::vector<int>& coarse,
                   std::vector<int>& fine) {
  const DLevel& L = B.lv[level];
  const DLevel& F = B.lv[level + 1];
  coarse.clear(); fine.clear();
  for (size_t c = 0; c < m.size(); ++c)
    if (m[c]) {
      coarse.push_back((int)c);
      const int f = find_cell(F, 2 * L.I[c], 2 * L.J[c], 2 * L.K[c]);
      if (f < 0) {
        B.err = "invalid interface connection on level " + std::to_string(level);
        return false;
      }
      fine.push_back(f);
    }
  return true;
}

DBuild* diad_build(const DInfo& base, int levels, int span, float nu_phys, const std::vector<DStl>& stls,
                   const std::vector<DRule>& rules) {
  DBuild* B = new DBuild;
  B->levels = levels; B->span = span; B->nu_phys = nu_phys; B->stls = stls; B->rules = rules;
  B->lv.resize(levels);
  DLevel& L0 = B->lv[0];
  L0.info = base;
  for (int k = 0; k < base.nz; ++k)
    for (int j = 0; j < base.ny; ++j)
      for (int i = 0; i < base.nx; ++i) {
        L0.I.push_back(i); L0.J.push_back(j); L0.K.push_back(k);
      }
  for (int l = 0; l < levels; ++l) build_level(*B, l);
  for (int l = levels - 2; l >= 0; --l) {
    DLevel& L = B->lv[l];
    if (!map_interface(*B, l, L.c2f, L.c2f_read, L.c2f_write)) break;
    if (!map_interface(*B, l, L.f2c, L.f2c_write, L.f2c_read)) break;
  }
  return B;
}

}  // namespace

// =============================================================== C interface
extern "C" {

int oracle_abi_version() { return 1; }

// the velocity set (include/applyStreaming.h:6-8)
void oracle_velocity_set(int* cx, int* cy, int* cz) {
  for (int q = 0; q < 27; ++q) {
    cx[q] = CX[q];
    cy[q] = CY[q];
    cz[q] = CZ[q];
  }
}

#define ORACLE_PER_CELL(SUF, R)                                                                      \
  void oracle_feq_##SUF(const R* rho_u, R* feq, long long n) {                                       \
    for (long long c = 0; c < n; ++c) {                                                              \
      R t[27];                                                                                       \
      get_feq(rho_u[4 * c], rho_u[4 * c + 1], rho_u[4 * c + 2], rho_u[4 * c + 3], t);                \
      for (int q = 0; q < 27; ++q) feq[27 * c + q] = t[q];                                           \
    }                                                                                                \
  }                                                                                                  \
  void oracle_rho_u_##SUF(const R* f, R* rho_u, long long n) {                                       \
    for (long long c = 0; c < n; ++c) {                                                              \
      R t[27];                                                                                       \
      for (int q = 0; q < 27; ++q) t[q] = f[27 * c + q];                                             \
      get_rho_u(t, rho_u[4 * c], rho_u[4 * c + 1], rho_u[4 * c + 2], rho_u[4 * c + 3]);              \
    }                                                                                                \
  }                                                                                                  \
  void oracle_collide_##SUF(R* f, long long n, R nu, R cs) {                                         \
    for (long long c = 0; c < n; ++c) {                                                              \
      R t[27];                                                                                       \
      for (int q = 0; q < 27; ++q) t[q] = f[27 * c + q];                                             \
      collide(t, nu, cs);                                                                            \
      for (int q = 0; q < 27; ++q) f[27 * c + q] = t[q];                                             \
    }                                                                                                \
  }                                                                                                  \
  R oracle_omega_les_##SUF(const R* f, R nu, R cs) {                                                 \
    R t[27], feq[27], fneq[27], rho, ux, uy, uz;                                                     \
    for (int q = 0; q < 27; ++q) t[q] = f[q];                                                        \
    get_rho_u(t, rho, ux, uy, uz);                                                                   \
    get_feq(rho, ux, uy, uz, feq);                                                                   \
    for (int q = 0; q < 27; ++q) fneq[q] = t[q] - feq[q];                                            \
    return get_omega_les(fneq, rho, nu, cs);                                                         \
  }                                                                                                  \
  void oracle_rescale_##SUF(R* f, long long n, int coarse_to_fine) {                                 \
    for (long long c = 0; c < n; ++c) {                                                              \
      R t[27];                                                                                       \
      for (int q = 0; q < 27; ++q) t[q] = f[27 * c + q];                                             \
      rescale_f(t, coarse_to_fine != 0);                                                             \
      for (int q = 0; q < 27; ++q) f[27 * c + q] = t[q];                                             \
    }                                                                                                \
  }                                                                                                  \
  void oracle_bounceback_##SUF(R* f, long long n) {                                                  \
    for (long long c = 0; c < n; ++c) {                                                              \
      R t[27];                                                                                       \
      for (int q = 0; q < 27; ++q) t[q] = f[27 * c + q];                                             \
      bounceback(t);                                                                                 \
      for (int q = 0; q < 27; ++q) f[27 * c + q] = t[q];                                             \
    }                                                                                                \
  }                                                                                                  \
  /* which: 0 mbbc, 1 restore rho+u, 2 restore u (rho given), 3 restore rho (u given) */              \
  int oracle_bc_##SUF(int which, int code, R* f, R* rho_u) {                                         \
    if (!find_bc(code)) return -1;                                                                   \
    R t[27];                                                                                         \
    for (int q = 0; q < 27; ++q) t[q] = f[q];                                                        \
    R rho = rho_u[0], ux = rho_u[1], uy = rho_u[2], uz = rho_u[3];                                   \
    if (which == 0) bc_mbbc(code, t, rho, ux, uy, uz);                                               \
    else if (which == 1) bc_restore_rhou(code, t, rho, ux, uy, uz);                                  \
    else if (which == 2) bc_restore_u(code, t, rho, ux, uy, uz);                                     \
    else if (which == 3) bc_restore_rho(code, t, rho, ux, uy, uz);                                   \
    else return -2;                                                                                  \
    for (int q = 0; q < 27; ++q) f[q] = t[q];                                                        \
    rho_u[0] = rho; rho_u[1] = ux; rho_u[2] = uy; rho_u[3] = uz;                                     \
    return 0;                                                                                        \
  }                                                                                                  \
  void oracle_local_update_##SUF(R* fa, const int* shift, const OracleInfo* I, const uint8_t* flags,  \
                                 const R* bcvals, const R* cs_cell) {                                \
    local_update(fa, shift, *I, flags, bcvals, cs_cell);                                             \
  }                                                                                                  \
  void oracle_periodic_##SUF(R* fa, const int* shift, const OracleInfo* I, int axis) {               \
    periodic_copy(fa, shift, *I, axis);                                                              \
  }                                                                                                  \
  void oracle_fill_equilibrium_##SUF(R* fa, const OracleInfo* I, const R* rho, const R* ux,          \
                                     const R* uy, const R* uz) {                                     \
    fill_equilibrium(fa, *I, rho, ux, uy, uz);                                                       \
  }                                                                                                  \
  void oracle_momentum_exchange_##SUF(const R* fa, const int* shift, const OracleInfo* I,            \
                                      const uint8_t* flags, const int* lo, const int* hi,            \
                                      double* out) {                                                 \
    momentum_exchange(fa, shift, *I, flags, lo, hi, out);                                            \
  }                                                                                                  \
  double oracle_mass_##SUF(const R* fa, const int* shift, const OracleInfo* I, const uint8_t* flags) { \
    return total_mass(fa, shift, *I, flags);                                                         \
  }                                                                                                  \
  void oracle_coarse_fine_##SUF(R* fc, const int* sc, const OracleInfo* Ic, R* ff, const int* sf,    \
                                const OracleInfo* If, const uint8_t* fine_flags) {                   \
    coarse_fine_comm(fc, sc, *Ic, ff, sf, *If, fine_flags);                                          \
  }                                                                                                  \
  void oracle_fill_coarse_from_fine_##SUF(R* fc, const int* sc, const OracleInfo* Ic, const R* ff,   \
                                          const int* sf, const OracleInfo* If) {                     \
    int s[3] = {Ic->i_sub_start + 1, Ic->j_sub_start + 1, Ic->k_sub_start + 1};                      \
    int e[3] = {Ic->i_sub_end - 1, Ic->j_sub_end - 1, Ic->k_sub_end - 1};                            \
    write_to_coarse(fc, sc, *Ic, ff, sf, *If, s, e);                                                 \
  }                                                                                                  \
  void oracle_esotwist_update_##SUF(R* fa, long long n, const int* nbr7, const int* ijk3,            \
                                    const OracleInfo* I, const uint8_t* flags, const R* bcvals,      \
                                    int flipper) {                                                   \
    const int* nb[7];                                                                                \
    for (int t = 0; t < 7; ++t) nb[t] = nbr7 + (size_t)t * n;                                        \
    esotwist_update(fa, n, nb, ijk3, ijk3 + n, ijk3 + 2 * n, *I, flags, bcvals, flipper != 0);       \
  }

ORACLE_PER_CELL(f32, float)
ORACLE_PER_CELL(f64, double)

#define ORACLE_ESO_LEVELS(SUF, R)                                                                              \
  void oracle_eso_to_coarse_##SUF(R* fc, long long nc, const int* nbc7, int flip_c, const R* ff, long long nf,  \
                                  const int* nbf7, int flip_f, const int* write, const int* read, long long m) { \
    const int* a[7];                                                                                           \
    const int* b[7];                                                                                           \
    for (int t = 0; t < 7; ++t) { a[t] = nbc7 + (size_t)t * nc; b[t] = nbf7 + (size_t)t * nf; }              \
    eso_to_coarse(fc, nc, a, flip_c != 0, ff, nf, b, flip_f != 0, write, read, m);                            \
  }                                                                                                            \
  void oracle_eso_to_fine_##SUF(const R* fc, long long nc, const int* nbc7, int flip_c, R* ff, long long nf,    \
                                const int* nbf7, int flip_f, const int* write, const int* read, long long m) {   \
    const int* a[7];                                                                                           \
    const int* b[7];                                                                                           \
    for (int t = 0; t < 7; ++t) { a[t] = nbc7 + (size_t)t * nc; b[t] = nbf7 + (size_t)t * nf; }              \
    eso_to_fine(fc, nc, a, flip_c != 0, ff, nf, b, flip_f != 0, write, read, m);                              \
  }                                                                                                            \
  void oracle_scalar_outlet_x_##SUF(R* Ta, const int* shift, const OracleInfo* I, const uint8_t* flags) {     \
    scalar_outlet_x(Ta, shift, *I, flags);                                                                     \
  }                                                                                                            \
  void oracle_scalar_update_##SUF(R* fa, R* Ta, const int* shift, const OracleInfo* I, const uint8_t* flags,   \
                                  const R* bcvals, const R* givenT, R tauT) {                                  \
    local_update(fa, shift, *I, flags, bcvals, (const R*)nullptr, Ta, givenT, tauT);                         \
  }                                                                                                            \
  void oracle_scalar_collide_##SUF(const R* f, R* T, long long n, R tauT) {                                   \
    for (long long c = 0; c < n; ++c) {                                                                        \
      R a[27], b[27];                                                                                          \
      for (int q = 0; q < 27; ++q) { a[q] = f[27 * c + q]; b[q] = T[27 * c + q]; }                           \
      scalar_collide(a, b, tauT);                                                                              \
      for (int q = 0; q < 27; ++q) T[27 * c + q] = b[q];                                                       \
    }                                                                                                          \
  }                                                                                                            \
  void oracle_eso_outlet_z_##SUF(R* fa, long long n, const int* nbr7, int flipper, const int* outlet,        \
                                 long long m, const int* kmin) {                                               \
    const int* a[7];                                                                                           \
    for (int t = 0; t < 7; ++t) a[t] = nbr7 + (size_t)t * n;                                                   \
    eso_outlet_z(fa, n, a, flipper != 0, outlet, m, kmin);                                                     \
  }                                                                                                            \
  void oracle_eso_moving_bounceback_##SUF(R* fa, long long n, const int* nbr7, int flipper, const int* cells, \
                                         long long m, const float* u3) {                                       \
    const int* a[7];                                                                                           \
    for (int t = 0; t < 7; ++t) a[t] = nbr7 + (size_t)t * n;                                                   \
    eso_moving_bounceback(fa, n, a, flipper != 0, cells, m, u3);                                              \
  }                                                                                                            \
  void oracle_eso_momentum_exchange_##SUF(const R* fa, long long n, const int* nbr7, int flipper,              \
                                          const uint8_t* flags, double* out3) {                                \
    const int* a[7];                                                                                           \
    for (int t = 0; t < 7; ++t) a[t] = nbr7 + (size_t)t * n;                                                   \
    eso_momentum_exchange(fa, n, a, flipper != 0, flags, out3);                                                \
  }                                                                                                            \
  void oracle_eso_forced_velocity_##SUF(R* fa, long long n, const int* nbr7, int flipper, const int* cells,     \
                                        long long m, const float* ut3, R nu, R cs) {                           \
    const int* a[7];                                                                                           \
    for (int t = 0; t < 7; ++t) a[t] = nbr7 + (size_t)t * n;                                                   \
    eso_forced_velocity(fa, n, a, flipper != 0, cells, m, ut3, nu, cs);                                        \
  }                                                                                                            \
  void oracle_eso_forcing_sum_##SUF(const R* fa, long long n, const int* nbr7, int flipper, const int* cells,   \
                                    long long m, const float* ut3, const int* ijk3, float res, float ox,       \
                                    float oy, float oz, float dt, float rho_phys, double* out6) {              \
    const int* a[7];                                                                                           \
    for (int t = 0; t < 7; ++t) a[t] = nbr7 + (size_t)t * n;                                                   \
    eso_forcing_sum(fa, n, a, flipper != 0, cells, m, ut3, ijk3, res, ox, oy, oz, dt, rho_phys, out6);         \
  }                                                                                                            \
  void oracle_collide_forced_##SUF(R* f27, R nu, R cs, const R* g3) {                                          \
    R f[27];                                                                                                   \
    for (int q = 0; q < 27; ++q) f[q] = f27[q];                                                                \
    collide_g(f, nu, cs, g3);                                                                                  \
    for (int q = 0; q < 27; ++q) f27[q] = f[q];                                                                \
  }

ORACLE_ESO_LEVELS(f32, float)
ORACLE_ESO_LEVELS(f64, double)

void oracle_stream_shift(int* shift, int nx, int ny, int nz) { stream_shift(shift, nx, ny, nz); }

int oracle_ray_hit(long long ax, long long ay, long long bx, long long by, long long cx, long long cy) {
  return ray_hit(ax, ay, bx, by, cx, cy) ? 1 : 0;
}

float oracle_ray_hit_z(const float* t9, float rx, float ry) {
  return ray_hit_z(t9[0], t9[1], t9[2], t9[3], t9[4], t9[5], t9[6], t9[7], t9[8], rx, ry);
}

void oracle_markers_inside_stl(const float* tri, int ntri, const float* stl_origin, const float* grid_origin,
                               float res, int nx, int ny, int nz, int inside_value, uint8_t* marker) {
  markers_inside_stl(tri, ntri, stl_origin[0], stl_origin[1], stl_origin[2], grid_origin[0], grid_origin[1],
                     grid_origin[2], res, nx, ny, nz, inside_value, marker);
}

int oracle_check_stl_edges(const float* tri, int ntri, int* counts) { return check_stl_edges(tri, ntri, counts); }

// Compiler flags of this build (the CPU baseline reports them).
const char* oracle_build_flags() { return PSL_ORACLE_FLAGS; }

int oracle_esotwist_tables(int* read_slot, int* write_slot) {
  for (int q = 0; q < 27; ++q) {
    read_slot[q] = ESO_READ[q];
    write_slot[q] = ESO_WRITE[q];
  }
  return 27;
}


// DIAD builder.  base6f = res, ox, oy, oz, nu, dtPhys; base3i = nx, ny, nz.
// stl_tris: triangles of all meshes back to back, stl_counts[s] triangles each,
// stl_origins[3 s]; rules: nrules records laid out as DRule (72 bytes).
void* oracle_diad_build(const float* base6f, const int* base3i, int levels, int span, float nu_phys,
                        const float* stl_tris, const int* stl_counts, const float* stl_origins, int nstl,
                        const void* rules, int nrules) {
  DInfo I{};
  I.gridID = 0;
  I.res = base6f[0]; I.ox = base6f[1]; I.oy = base6f[2]; I.oz = base6f[3]; I.nu = base6f[4]; I.dt = base6f[5];
  I.nx = base3i[0]; I.ny = base3i[1]; I.nz = base3i[2];
  std::vector<DStl> stls;
  size_t off = 0;
  for (int s = 0; s < nstl; ++s) {
    stls.push_back({stl_tris + 9 * off, stl_counts[s], stl_origins[3 * s], stl_origins[3 * s + 1], stl_origins[3 * s + 2]});
    off += (size_t)stl_counts[s];
  }
  std::vector<DRule> rl((const DRule*)rules, (const DRule*)rules + nrules);
  return diad_build(I, levels, span, nu_phys, stls, rl);
}

const char* oracle_diad_error(void* h) { return ((DBuild*)h)->err.c_str(); }

// which: 0 ijk (3n int), 1 bounceback (n u8), 2 interface (n u8), 3 neighbours
// (10n int: +i +j +k +ij +ik +jk +ijk -i -j -k), 4 coarseToFineRead,
// 5 coarseToFineWrite, 6 fineToCoarseWrite, 7 fineToCoarseRead, 8 outlet cells
long long oracle_diad_count(void* h, int level, int which) {
  const DLevel& L = ((DBuild*)h)->lv[level];
  switch (which) {
    case 0: return 3 * (long long)L.n();
    case 1: case 2: return (long long)L.n();
    case 3: return (long long)L.nbr10.size();
    case 4: return (long long)L.c2f_read.size();
    case 5: return (long long)L.c2f_write.size();
    case 6: return (long long)L.f2c_write.size();
    case 7: return (long long)L.f2c_read.size();
    case 8: return (long long)L.outlet.size();
  }
  return -1;
}

void oracle_diad_get(void* h, int level, int which, void* out) {
  const DLevel& L = ((DBuild*)h)->lv[level];
  const size_t n = L.n();
  auto put = [&](const std::vector<int>& v) { std::memcpy(out, v.data(), v.size() * sizeof(int)); };
  switch (which) {
    case 0:
      std::memcpy(out, L.I.data(), n * sizeof(int));
      std::memcpy((int*)out + n, L.J.data(), n * sizeof(int));
      std::memcpy((int*)out + 2 * n, L.K.data(), n * sizeof(int));
      break;
    case 1: std::memcpy(out, L.bb.data(), n); break;
    case 2: std::memcpy(out, L.iface.data(), n); break;
    case 3: put(L.nbr10); break;
    case 4: put(L.c2f_read); break;
    case 5: put(L.c2f_write); break;
    case 6: put(L.f2c_write); break;
    case 7: put(L.f2c_read); break;
    case 8: put(L.outlet); break;
  }
}

// level info: f6 = res, ox, oy, oz, nu, dtPhys; i4 = nx, ny, nz, cellCount
void oracle_diad_info(void* h, int level, float* f6, int* i4) {
  const DLevel& L = ((DBuild*)h)->lv[level];
  f6[0] = L.info.res; f6[1] = L.info.ox; f6[2] = L.info.oy; f6[3] = L.info.oz; f6[4] = L.info.nu; f6[5] = L.info.dt;
  i4[0] = L.info.nx; i4[1] = L.info.ny; i4[2] = L.info.nz; i4[3] = (int)L.n();
}

void oracle_diad_free(void* h) { delete (DBuild*)h; }

}  // extern "C"
