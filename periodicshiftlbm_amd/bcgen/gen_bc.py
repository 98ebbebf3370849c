"""Derive the D3Q27 moment-based boundary-condition tables and emit them as C++.

This is the project's own symbolic derivation of the local boundary conditions
described in paper.pdf section 2 (p.5-14) of the reference project.  The
selection rules (which moments are imposed, which null-space combinations are
used to restore the missing macroscopic quantities) must match the reference
generator exactly, because the reference's shipped headers
include/boundaryConditions/{applyMBBC,restoreRho,restoreUxUyUz,restoreRhoUxUyUz}.h
are the output of that generator and the parity target:

* known / unknown split by the outer normal n: a population q is unknown when
  any component c_qa * n_a < 0
  (symbolicHelperTools/CppCodeGenerator/helperMain.py:155-162);
* MBBC: rows of M_fu are taken greedily in the fixed 23-moment order, keeping a
  row when it raises the rank, until there are as many rows as unknowns;
  f_u = (S M_fu)^-1 (S m_eq - S M_fk f_k)            (helperMain.py:181-194, 272-282);
* restore: C = null space of M_fu^T (sympy's RREF basis), columns stably sorted
  by the summed moment order of their non-zero entries (helperMain.py:197-205);
  rho,u from (Sq C^T Q)^-1 Sq C^T M_fk f_k            (helperMain.py:206-218, 296-301);
  u from rho via Su                                    (helperMain.py:220-233, 315-324);
  rho from u via the first row of C^T Q_rho            (helperMain.py:235-247, 338-348).

Outputs (both generated, both committed):
  periodicshiftlbm_amd/csrc/bc_generated.inc  straight-line device code, one
      function per outer normal and quantity, dispatched by a switch on the
      reference's normal code (n_x+5)*100+(n_y+5)*10+(n_z+5);
  oracle/bc_tables.inc  exact rational tables consumed by the CPU oracle.

Run:  python periodicshiftlbm_amd/bcgen/gen_bc.py
"""
from __future__ import annotations

import json
import os
import sys

import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))

# D3Q27 velocity set in the reference ordering (include/applyStreaming.h:6-8).
CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1]
CY = [0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1]
CZ = [0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1]

# The 23 raw moments m_abc = sum_q cx^a cy^b cz^c f_q, in the order that drives
# the greedy row selection (helperMain.py:24-28).
MOMENTS = ["000", "100", "010", "001", "200", "020", "002", "011", "101", "110",
           "111", "210", "201", "120", "021", "102", "012",
           "022", "202", "220", "211", "121", "112"]
MOMENT_ORDER = [sum(int(ch) for ch in m) for m in MOMENTS]

# Linear parts of the equilibrium moments in (rho, rho ux, rho uy, rho uz):
# m_eq ~ Q [rho, jx, jy, jz]^T.  Q_rho / Q_u are its first column / last three.
def _q_matrix():
    third, ninth = sp.Rational(1, 3), sp.Rational(1, 9)
    rows = []
    for m in MOMENTS:
        a, b, c = (int(ch) for ch in m)
        row = [0, 0, 0, 0]
        if (a, b, c) == (0, 0, 0):
            row[0] = 1
        elif a + b + c == 1:
            row[1 + [a, b, c].index(1)] = 1
        elif sorted((a, b, c)) == [0, 0, 2]:
            row[0] = third
        elif sorted((a, b, c)) == [0, 2, 2]:
            row[0] = ninth
        elif sorted((a, b, c)) == [0, 1, 2]:
            # m_{2,1,0}-type: (1/3) rho u_k for the axis with exponent 1
            row[1 + [a, b, c].index(1)] = third
        rows.append(row)
    return sp.Matrix(rows)


Q = _q_matrix()
Q_RHO = Q[:, 0]
Q_U = Q[:, 1:4]

# Equilibrium moment expressions m_eq(rho, u) as C++ text, same algebraic form
# as the reference generator uses (helperMain.py:30-53); `T` is the scalar type.
def meq_expr(m: str) -> str:
    a, b, c = (int(ch) for ch in m)
    u = {0: "ux", 1: "uy", 2: "uz"}
    e = [a, b, c]
    third = "(T(1)/T(3))"
    ninth = "(T(1)/T(9))"
    if e == [0, 0, 0]:
        return "rho"
    if sum(e) == 1:
        return f"rho * {u[e.index(1)]}"
    if sorted(e) == [0, 0, 2]:
        k = u[e.index(2)]
        return f"{third} * rho + rho * {k} * {k}"
    if sorted(e) == [0, 1, 1]:
        i, j = [d for d in range(3) if e[d] == 1]
        return f"rho * {u[i]} * {u[j]}"
    if e == [1, 1, 1]:
        return "T(0)"
    if sorted(e) == [0, 1, 2]:
        one = e.index(1)
        two = e.index(2)
        # reference text orders the product as ux*ux*uy / ux*uy*uy by axis index
        lo, hi = sorted([one, two])
        prod = " * ".join(u[d] for d in sorted([one, two, two]))
        return f"{third} * rho * {u[one]} + rho * {prod}"
    if sorted(e) == [0, 2, 2]:
        i, j = [d for d in range(3) if e[d] == 2]
        return (f"{ninth} * rho + {third} * rho * {u[i]} * {u[i]} + {third} * rho * {u[j]} * {u[j]}"
                f" + rho * {u[i]} * {u[i]} * {u[j]} * {u[j]}")
    if sorted(e) == [1, 1, 2]:
        i, j = [d for d in range(3) if e[d] == 1]
        return f"{third} * rho * {u[i]} * {u[j]}"
    raise ValueError(m)


# Outer normals in the reference's emission order (helperMain.py:253-257).
ALL_NORMALS = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
               [0, 1, 1], [1, 0, 1], [1, 1, 0], [0, -1, -1], [-1, 0, -1], [-1, -1, 0],
               [0, 1, -1], [1, 0, -1], [1, -1, 0], [0, -1, 1], [-1, 0, 1], [-1, 1, 0],
               [1, 1, 1], [-1, -1, -1],
               [-1, 1, 1], [1, -1, 1], [1, 1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]


def normal_code(n):
    return (n[0] + 5) * 100 + (n[1] + 5) * 10 + (n[2] + 5)


def _greedy_rows(M: sp.Matrix, count: int, start=(0,)):
    """Keep rows of M in order while they raise the rank, until `count` rows."""
    keep = list(start)
    i = max(keep) + 1 if keep else 0
    while len(keep) < count:
        if i >= M.rows:
            raise RuntimeError("ran out of independent rows")
        if M.extract(keep + [i], list(range(M.cols))).rank() > M.extract(keep, list(range(M.cols))).rank():
            keep.append(i)
        i += 1
    return keep


def derive(normal):
    known = [q for q in range(27) if min(CX[q] * normal[0], CY[q] * normal[1], CZ[q] * normal[2]) >= 0]
    unknown = [q for q in range(27) if q not in known]

    def moment_row(m, qs):
        a, b, c = (int(ch) for ch in m)
        return [CX[q] ** a * CY[q] ** b * CZ[q] ** c for q in qs]

    Mfk = sp.Matrix([moment_row(m, known) for m in MOMENTS])
    Mfu = sp.Matrix([moment_row(m, unknown) for m in MOMENTS])

    # --- MBBC: impose len(unknown) equilibrium moments on the unknowns
    srows = _greedy_rows(Mfu, len(unknown))
    S = sp.eye(23).extract(srows, list(range(23)))
    inv = (S * Mfu).inv()
    Smfk = S * Mfk

    # --- restore: combinations of moments that do not involve the unknowns
    null = Mfu.T.nullspace()

    def order_key(v):
        return sum(MOMENT_ORDER[i] for i, val in enumerate(v) if val != 0)

    C = sp.Matrix.hstack(*sorted(null, key=order_key))
    CtQ = C.T * Q
    CtQu = C.T * Q_U
    CtQrho = C.T * Q_RHO
    CtMfk = C.T * Mfk

    sq = _greedy_rows(CtQ, 4)
    Sq = sp.eye(CtQ.rows).extract(sq, list(range(CtQ.rows)))
    full = (Sq * CtQ).inv() * Sq * CtMfk          # 4 x nk

    su = _greedy_rows(CtQu, 3)
    Su = sp.eye(CtQ.rows).extract(su, list(range(CtQ.rows)))
    uinv = (Su * CtQu).inv()                       # 3 x 3
    uSCMfk = Su * CtMfk                            # 3 x nk
    uSCQ = Su * CtQrho                             # 3 x 1

    # rho restoration always uses the first combination (helperMain.py:236-247
    # tests row 0 on every iteration); it must have a non-zero rho coefficient.
    if CtQrho[0] == 0:
        raise RuntimeError(f"normal {normal}: first null-space row has no rho term")
    rSCQrho = CtQrho[0]
    rSCMfk = CtMfk.row(0)
    rSCQu = CtQu.row(0)

    return dict(normal=normal, code=normal_code(normal), known=known, unknown=unknown,
                mbbc_rows=srows, mbbc_inv=inv, mbbc_smfk=Smfk,
                rhou_full=full,
                u_inv=uinv, u_scmfk=uSCMfk, u_scq=uSCQ,
                rho_scqrho=rSCQrho, rho_scmfk=rSCMfk, rho_scqu=rSCQu)


# ----------------------------------------------------------------------------- C++ emission
def _c(x) -> str:
    """Exact rational as a T-typed constant expression."""
    x = sp.Rational(x)
    if x.q == 1:
        return f"T({x.p})"
    return f"(T({x.p}) / T({x.q}))"


def _lincomb(coeffs, names) -> str:
    """sum_j coeffs[j]*names[j], left to right, skipping zeros."""
    terms = []
    for c, n in zip(coeffs, names):
        c = sp.Rational(c)
        if c == 0:
            continue
        if c == 1:
            terms.append(("+", n))
        elif c == -1:
            terms.append(("-", n))
        else:
            terms.append(("+", f"{_c(c)} * {n}"))
    if not terms:
        return "T(0)"
    out = ("-" if terms[0][0] == "-" else "") + terms[0][1]
    for s, t in terms[1:]:
        out += f" {s} {t}"
    return out


def emit_device(data) -> str:
    L = []
    L.append("// Generated by periodicshiftlbm_amd/bcgen/gen_bc.py -- do not edit.")
    L.append("// Moment-based boundary conditions for D3Q27, one function per outer normal.")
    L.append("// Normal code = (nx+5)*100 + (ny+5)*10 + (nz+5); 555 (no normal) is a no-op.")
    L.append("#pragma once")
    L.append("")
    for d in data:
        code = d["code"]
        fk = [f"f[{q}]" for q in d["known"]]
        nu = len(d["unknown"])
        # ---- MBBC
        L.append(f"// normal ({d['normal'][0]},{d['normal'][1]},{d['normal'][2]}): {nu} unknown populations")
        L.append("template <typename T> PSL_HD PSL_INLINE void bc_mbbc_%d(T (&f)[27], T rho, T ux, T uy, T uz)" % code)
        L.append("{")
        for i, row in enumerate(d["mbbc_rows"]):
            coeffs = list(d["mbbc_smfk"].row(i))
            L.append(f"  const T s{i} = ({meq_expr(MOMENTS[row])}) - ({_lincomb(coeffs, fk)});")
        for i, q in enumerate(d["unknown"]):
            coeffs = list(d["mbbc_inv"].row(i))
            L.append(f"  f[{q}] = {_lincomb(coeffs, [f's{j}' for j in range(nu)])};")
        L.append("}")
        # ---- restore rho,u
        L.append("template <typename T> PSL_HD PSL_INLINE void bc_restore_rhou_%d(const T (&f)[27], T& rho, T& ux, T& uy, T& uz)" % code)
        L.append("{")
        for i in range(4):
            L.append(f"  const T q{i} = {_lincomb(list(d['rhou_full'].row(i)), fk)};")
        L.append("  rho = q0;")
        # true divisions, as in the reference: the restored velocities come from
        # cancelling sums and a reciprocal would add a rounding step
        L.append("  ux = q1 / rho; uy = q2 / rho; uz = q3 / rho;")
        L.append("}")
        # ---- restore u given rho
        L.append("template <typename T> PSL_HD PSL_INLINE void bc_restore_u_%d(const T (&f)[27], T rho, T& ux, T& uy, T& uz)" % code)
        L.append("{")
        for i in range(3):
            L.append(f"  const T s{i} = ({_lincomb(list(d['u_scmfk'].row(i)), fk)}) / rho - {_c(d['u_scq'][i])};")
        for a, name in enumerate(["ux", "uy", "uz"]):
            L.append(f"  {name} = {_lincomb(list(d['u_inv'].row(a)), ['s0', 's1', 's2'])};")
        L.append("}")
        # ---- restore rho given u
        L.append("template <typename T> PSL_HD PSL_INLINE void bc_restore_rho_%d(const T (&f)[27], T& rho, T ux, T uy, T uz)" % code)
        L.append("{")
        L.append(f"  const T s = {_c(d['rho_scqrho'])} + ({_lincomb(list(d['rho_scqu']), ['ux', 'uy', 'uz'])});")
        L.append(f"  rho = ({_lincomb(list(d['rho_scmfk']), fk)}) / s;")
        L.append("}")
        L.append("")
    # dispatchers
    for kind, sig, call in [
        ("mbbc", "T (&f)[27], T rho, T ux, T uy, T uz", "f, rho, ux, uy, uz"),
        ("restore_rhou", "const T (&f)[27], T& rho, T& ux, T& uy, T& uz", "f, rho, ux, uy, uz"),
        ("restore_u", "const T (&f)[27], T rho, T& ux, T& uy, T& uz", "f, rho, ux, uy, uz"),
        ("restore_rho", "const T (&f)[27], T& rho, T ux, T uy, T uz", "f, rho, ux, uy, uz"),
    ]:
        L.append(f"template <typename T> PSL_HD PSL_INLINE void bc_{kind}(int normal_code, {sig})")
        L.append("{")
        L.append("  switch (normal_code) {")
        for d in data:
            L.append(f"    case {d['code']}: bc_{kind}_{d['code']}({call}); break;")
        L.append("    default: break;")
        L.append("  }")
        L.append("}")
        L.append("")
    return "\n".join(L) + "\n"


def _frac_pairs(vals):
    out = []
    for v in vals:
        r = sp.Rational(v)
        out.append(f"{{{r.p},{r.q}}}")
    return out


def emit_oracle_tables(data) -> str:
    """Rational tables for the oracle's table-driven boundary conditions."""
    L = ["// Generated by periodicshiftlbm_amd/bcgen/gen_bc.py -- do not edit.",
         "// Exact rational boundary-condition tables, one record per outer normal,",
         "// in the reference generator's normal order.",
         "// Each fraction is {numerator, denominator}.",
         "#pragma once", ""]
    L.append(f"static const int ORACLE_BC_COUNT = {len(data)};")
    L.append("static const OracleBcRecord ORACLE_BC[] = {")
    for d in data:
        nk, nu = len(d["known"]), len(d["unknown"])
        def mat(M, rows, cols):
            return "{" + ",".join("{" + ",".join(_frac_pairs([M[i, j] for j in range(cols)])) + "}" for i in range(rows)) + "}"
        rec = []
        rec.append("{%d,%d,%d}" % tuple(d["normal"]))
        rec.append(str(d["code"]))
        rec.append(str(nk))
        rec.append(str(nu))
        rec.append("{" + ",".join(map(str, d["known"])) + "}")
        rec.append("{" + ",".join(map(str, d["unknown"])) + "}")
        rec.append("{" + ",".join(map(str, d["mbbc_rows"])) + "}")
        rec.append("{" + ",".join("{" + ",".join(str(int(d["mbbc_smfk"][i, j])) for j in range(nk)) + "}" for i in range(nu)) + "}")
        rec.append(mat(d["mbbc_inv"], nu, nu))
        rec.append(mat(d["rhou_full"], 4, nk))
        rec.append(mat(d["u_inv"], 3, 3))
        rec.append(mat(d["u_scmfk"], 3, nk))
        rec.append("{" + ",".join(_frac_pairs([d["u_scq"][i] for i in range(3)])) + "}")
        rec.append("{" + ",".join(_frac_pairs([d["rho_scqrho"]])) + "}")
        rec.append("{" + ",".join(_frac_pairs([d["rho_scmfk"][j] for j in range(nk)])) + "}")
        rec.append("{" + ",".join(_frac_pairs([d["rho_scqu"][j] for j in range(3)])) + "}")
        L.append("  {" + ", ".join(rec) + "},")
    L.append("};")
    return "\n".join(L) + "\n"


def tables_json(data):
    def frac(x):
        r = sp.Rational(x)
        return [int(r.p), int(r.q)]
    out = []
    for d in data:
        out.append(dict(
            normal=d["normal"], code=d["code"], known=d["known"], unknown=d["unknown"],
            mbbc_rows=d["mbbc_rows"],
            mbbc_smfk=[[int(v) for v in d["mbbc_smfk"].row(i)] for i in range(d["mbbc_smfk"].rows)],
            mbbc_inv=[[frac(v) for v in d["mbbc_inv"].row(i)] for i in range(d["mbbc_inv"].rows)],
            rhou_full=[[frac(v) for v in d["rhou_full"].row(i)] for i in range(4)],
            u_inv=[[frac(v) for v in d["u_inv"].row(i)] for i in range(3)],
            u_scmfk=[[frac(v) for v in d["u_scmfk"].row(i)] for i in range(3)],
            u_scq=[frac(d["u_scq"][i]) for i in range(3)],
            rho_scqrho=frac(d["rho_scqrho"]),
            rho_scmfk=[frac(v) for v in d["rho_scmfk"]],
            rho_scqu=[frac(v) for v in d["rho_scqu"]],
        ))
    return out


def main(argv=None):
    data = [derive(n) for n in ALL_NORMALS]
    dev = emit_device(data)
    with open(os.path.join(REPO, "periodicshiftlbm_amd", "csrc", "bc_generated.inc"), "w") as fh:
        fh.write(dev)
    with open(os.path.join(REPO, "oracle", "bc_tables.inc"), "w") as fh:
        fh.write(emit_oracle_tables(data))
    with open(os.path.join(REPO, "periodicshiftlbm_amd", "bcgen", "bc_tables.json"), "w") as fh:
        json.dump(tables_json(data), fh, indent=0)
    print(f"generated {len(data)} normals")


if __name__ == "__main__":
    sys.exit(main())
