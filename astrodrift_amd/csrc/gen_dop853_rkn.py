#!/usr/bin/env python3
"""Generates dop853_rkn.hpp: the DOP853 tableau restated for second-order
systems, as one coefficient table the kernels read with scalar loads.

The orbits integrate y = (x, v) with y' = (v, F(x)).  The velocity half of
every stage derivative k_j is the stage's own velocity V_j = v0 + h sum_m
a_jm F_m, so every linear combination of stage derivatives that the
integrator forms (stage points, the 8th-order update, both error estimates,
the dense-output coefficients) can be written with the forces F_m alone:

    sum_j w_j V_j = (sum_j w_j) v0 + h sum_m (sum_j w_j a_jm) F_m.

The stage positions become X_i = x0 + h c_i v0 + h^2 sum_m (A^2)_im F_m (the
Runge-Kutta-Nystrom form of the same method), and the stage velocities are
never formed.  The products are computed here in exact rational arithmetic
from the decimal tableau in integrators.hpp (namespace hw853, the constants of
src/integrators/galpy/dop853.rs:105-264) and rounded once to double.  Products
that vanish by the method's order conditions come out at the 1e-27 level (the
tableau's decimals are truncated at 30 digits) and are dropped.

Every coefficient the step uses (the derived ones and the tableau's own b, er,
bhh and d weights of the force halves) goes into one table, ordered by use, so
that a kernel fetches each use site's coefficients with one or two 64-byte
scalar loads instead of materialising each 64-bit literal with two scalar
moves.

Stage numbering: 1..12 the attempted step, 13 the FSAL derivative at the new
point (a_13,j = b_j), 14..16 the three dense-output stages.

Run:  python3 astrodrift_amd/csrc/gen_dop853_rkn.py  (rewrites the header;
tests/test_rkn_coefficients.py checks the committed header against it).
"""

from __future__ import annotations

import re
import sys
from fractions import Fraction
from pathlib import Path

HERE = Path(__file__).resolve().parent
ZERO = Fraction(1, 10**20)


def tableau(src: str) -> dict[str, Fraction]:
    ns = src[src.index("namespace hw853 {"):src.index("}  // namespace hw853")]
    return {k: Fraction(v) for k, v in re.findall(r"\b([a-z]+\d+)\s*=\s*([-0-9.e]+)", ns)}


def derive(t: dict[str, Fraction]):
    """Returns the ordered list of (function name, [(stage m, coefficient)])
    sums, plus the stage abscissae."""
    b = {j: t.get(f"b{j}", Fraction(0)) for j in range(1, 13)}

    def A(i: int, j: int) -> Fraction:
        if i == 13:
            return b.get(j, Fraction(0))
        return t.get(f"a{i}{j}", Fraction(0))

    def comb(w: dict[int, Fraction], expect_sum: int):
        """sum_j w_j V_j = alpha v0 + h sum_m beta_m F_m; alpha is checked."""
        alpha = sum(w.values(), Fraction(0))
        if abs(alpha - expect_sum) > ZERO:
            raise SystemExit(f"weights sum to {float(alpha)}, expected {expect_sum}")
        beta = []
        for m in range(1, 17):
            x = sum((w[j] * A(j, m) for j in w if j > m), Fraction(0))
            if abs(x) > ZERO:
                beta.append((m, x))
        return beta

    def direct(w: dict[int, Fraction]):
        return [(m, x) for m, x in sorted(w.items()) if x != 0]

    sums = []
    cs = {}
    for i in list(range(2, 13)) + [14, 15, 16]:
        row = {j: A(i, j) for j in range(1, i)}
        cs[i] = sum(row.values(), Fraction(0))
        beta = comb(row, cs[i]) if i > 2 else []
        sums.append((f"stage{i}", beta, cs[i]))
    # 8th-order update: sum_j b_j k_j; force half directly, position half
    # sum_j b_j V_j = v0 + h bpos(F)
    sums.append(("bvel", direct(b), None))
    sums.append(("bpos", comb(dict(b), 1), None))
    # error estimates
    er = {j: t.get(f"er{j}", Fraction(0)) for j in range(1, 13)}
    sums.append(("errvel", direct(er), None))
    sums.append(("errpos", comb(er, 0), None))
    bhh = {j: b[j] for j in b}
    bhh[1] -= t["bhh1"]
    bhh[9] -= t["bhh2"]
    bhh[12] -= t["bhh3"]
    # force half: sum_j b_j F_j - bhh1 F_1 - bhh2 F_9 - bhh3 F_12, written
    # from the b-weighted sum k4 the step already has
    sums.append(("bhhvel", [(1, t["bhh1"]), (9, t["bhh2"]), (12, t["bhh3"])], None))
    sums.append(("bhhpos", comb(bhh, 0), None))
    # dense output rcont5..8 (without their factor h), split at the attempt's
    # last stage: terms in F_1..F_12 are known before the dense stages run
    for k in (4, 5, 6, 7):
        d = {j: t.get(f"d{k}{j}", Fraction(0)) for j in range(1, 17)}
        dv = direct(d)
        dp = comb({j: x for j, x in d.items() if x != 0}, 0)
        sums.append((f"dense{k}vel_pre", [(m, x) for m, x in dv if m <= 12], None))
        sums.append((f"dense{k}pos_pre", [(m, x) for m, x in dp if m <= 12], None))
    for k in (4, 5, 6, 7):
        d = {j: t.get(f"d{k}{j}", Fraction(0)) for j in range(1, 17)}
        dv = direct(d)
        dp = comb({j: x for j, x in d.items() if x != 0}, 0)
        sums.append((f"dense{k}vel_post", [(m, x) for m, x in dv if m > 12], None))
        sums.append((f"dense{k}pos_post", [(m, x) for m, x in dp if m > 12], None))
    return sums


def lit(x: Fraction) -> str:
    return repr(float(x))


def emit(sums) -> str:
    table = []  # (value, comment)
    funcs = []
    for name, terms, c in sums:
        base = len(table)
        if c is not None:
            table.append((c, f"{name}: c"))
        start = len(table)
        for m, x in terms:
            table.append((x, f"{name}: F{m}"))
        if name.startswith("stage"):
            i = int(name[5:])
            funcs.append(f"constexpr int k_c{i} = {base};")
        if not terms:
            continue
        expr = " + ".join(f"C[{start + q}] * F[{m}][c]" for q, (m, _) in enumerate(terms))
        funcs.append("template <class FA, class CP>")
        funcs.append(f"DRIFT_HD DRIFT_FI double {name}(const FA& F, int c, CP C) {{")
        funcs.append(f"    return {expr};")
        funcs.append("}")
    n = len(table)
    stage_ids = [int(name[5:]) for name, _, c in sums if name.startswith("stage")]
    kc = ["-1"] * 17
    for name, terms, c in sums:
        if name.startswith("stage"):
            kc[int(name[5:])] = f"k_c{int(name[5:])}"
    # stage sums over a local copy of the stage's coefficients (c_i, then the
    # terms in order), for kernels that fetch a stage's coefficients ahead
    kn = ["0"] * 17
    for name, terms, c in sums:
        if name.startswith("stage"):
            i = int(name[5:])
            kn[i] = str(1 + len(terms))
            if terms:
                expr = " + ".join(f"L[{1 + q}] * F[{m}][c]" for q, (m, _) in enumerate(terms))
                funcs.append("template <class FA>")
                funcs.append(f"DRIFT_HD DRIFT_FI double {name}_local(const FA& F, int c, const double* L) {{")
                funcs.append(f"    return {expr};")
                funcs.append("}")
    funcs.append("")
    funcs.append("// number of coefficients of stage i (c_i and its terms), from kStageC[i]")
    funcs.append("constexpr int kStageN[17] = {" + ", ".join(kn) + "};")
    funcs.append("")
    funcs.append("// index of c_i in kCoef for stage i (-1: no such stage)")
    funcs.append("constexpr int kStageC[17] = {" + ", ".join(kc) + "};")
    funcs.append("")
    funcs.append("// sum_m (A^2)_im F_m[c] for stage I (0 for stage 2)")
    funcs.append("template <int I, class FA, class CP>")
    funcs.append("DRIFT_HD DRIFT_FI double stage_sum(const FA& F, int c, CP C) {")
    for i in stage_ids:
        has = any(name == f"stage{i}" and terms for name, terms, _ in sums)
        body = f"return stage{i}(F, c, C);" if has else "return 0.0;"
        funcs.append(f"    if constexpr (I == {i}) {body}")
    funcs.append("}")
    funcs.append("")
    funcs.append("// the same from a local copy L of the stage's coefficients (L[0] = c_I)")
    funcs.append("template <int I, class FA>")
    funcs.append("DRIFT_HD DRIFT_FI double stage_sum_local(const FA& F, int c, const double* L) {")
    for i in stage_ids:
        has = any(name == f"stage{i}" and terms for name, terms, _ in sums)
        body = f"return stage{i}_local(F, c, L);" if has else "return 0.0;"
        funcs.append(f"    if constexpr (I == {i}) {body}")
    funcs.append("}")
    lines = [
        "// dop853_rkn.hpp -- GENERATED by gen_dop853_rkn.py from the DOP853 tableau",
        "// in integrators.hpp (namespace hw853); do not edit by hand.",
        "//",
        "// DOP853 for y = (x, v), y' = (v, F(x)): every combination of stage",
        "// derivatives written with the forces F_m only (stages 1..12, 13 = FSAL",
        "// derivative at the new point, 14..16 the dense-output stages):",
        "//   stage positions  X_i = x0 + h c_i v0 + h^2 stage_i(F)",
        "//   update           sum_j b_j k_j = (v0 + h bpos(F), bvel(F))",
        "//   error estimates  sum_j er_j k_j = (h errpos(F), errvel(F)),",
        "//                    k4 - bhh1 k1 - bhh2 k9 - bhh3 k12",
        "//                      = (h bhhpos(F), bvel(F) - bhhvel(F))",
        "//   dense output     sum_j d_kj k_j = (h dense_k pos(F), dense_k vel(F)),",
        "//                    k = 4..7 (rcont5..8 without their factor h)",
        "// F is indexed F[m][c] (m = stage, c = component); C points at kCoef (or",
        "// a device copy of it), the coefficients in order of use.",
        "#pragma once",
        "",
        '#include "drift_device.hpp"',
        "",
        "namespace drift {",
        "namespace rkn853 {",
        "",
        f"constexpr int kNumCoef = {n};",
        "constexpr double kCoef[kNumCoef] = {",
    ]
    for q, (v, comment) in enumerate(table):
        lines.append(f"    {lit(v)},  // [{q}] {comment}")
    lines.append("};")
    lines.append("")
    lines.append("#if defined(__HIPCC__)")
    lines.append("// the device copy, read with scalar loads (see dop853_coef)")
    lines.append("static __constant__ double kCoefDevice[kNumCoef] = {")
    for v, _ in table:
        lines.append(f"    {lit(v)},")
    lines.append("};")
    lines.append("#endif")
    lines.append("")
    lines += funcs
    lines += ["", "}  // namespace rkn853", "}  // namespace drift", ""]
    return "\n".join(lines)


def main() -> int:
    src = (HERE / "integrators.hpp").read_text()
    text = emit(derive(tableau(src)))
    target = HERE / "dop853_rkn.hpp"
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        return 0 if target.exists() and target.read_text() == text else 1
    target.write_text(text)
    print(f"wrote {target}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
