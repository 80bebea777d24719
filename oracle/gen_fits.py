"""Transcribe the IMRPhenomX (2,2) parameter-space fits into C for the oracle.

This is test infrastructure: the header it writes (``oracle/xas_fits.h``)
is compiled only into the CPU oracle, never into the HIP product.

The fits are LALSimulation's calibration data (LALSimIMRPhenomX_*.c, via
waveformrs), held by the reference as generated Python modules:

* ``src/waveformx/models/xpnr/_generated/utilities.py`` (final state, f_MECO,
  f_ISCO, linb, psi4-to-strain),
* ``.../_generated/qnm.py`` (fring22, fdamp22),
* ``.../_generated/inspiral.py``, ``intermediate.py``, ``ringdown.py``
  (collocation values and ansatzes).

Each function body is a straight-line sequence of assignments of
arithmetic expressions, so the transcription is mechanical: the module's
AST is walked and every expression is written with explicit parentheses in
Python's evaluation order (left to right), ``fma(a, b, c)`` as the unfused
``a * b + c`` the reference uses (``xpnr/_math.py:125-137``), and every
float literal with ``repr`` (which round-trips the exact double).

The IMRPhenomXHM fits (``.../_generated/hm_*.py``) go the same way into
``oracle/xhm_fits.h``; the final-spin fits of the higher modes
(``hm_qnm.py``) are written as their numerator and denominator
coefficients, which the oracle re-expands in powers of ``1 - a_f`` as
``xhm/reference.py:96-99`` does.

Usage: ``python oracle/gen_fits.py [REFERENCE_ROOT]`` (default
``/root/reference``); it rewrites ``oracle/xas_fits.h`` and
``oracle/xhm_fits.h`` next to itself.
"""

from __future__ import annotations

import ast
import sys
from pathlib import Path

MODULES = {
    # module file -> (C prefix, functions to emit)
    "utilities": ("ut", ["sign", "chi_eff", "chi_pn_hat", "s_tot_r", "dchi", "f_isco", "f_meco",
                         "final_mass_2017", "final_spin_2017", "linb", "psi4_to_strain"]),
    "qnm": ("qnm", ["fring22", "fdamp22"]),
    "inspiral": ("ins", ["amp_22_v2", "amp_22_v3", "amp_22_v4", "pow_7_of", "amp_22_rho1", "amp_22_rho2",
                         "amp_22_rho3", "amp_22_ansatz", "amp_22_ansatz_exact", "amp_22_dansatz", "phase_22_v3",
                         "phase_22_d13", "phase_22_d23", "phase_22_d43", "inspiral_dphase_norm",
                         "phase_22_ansatz_exact", "phase_22_ansatz_int_exact"]),
    "intermediate": ("itm", ["amp_22_va", "amp_22_delta0", "amp_22_delta1", "amp_22_delta2", "amp_22_delta3",
                             "amp_22_delta4", "amp_22_ansatz", "phase_22_v2", "phase_22_v2m_rdv4",
                             "phase_22_v3m_rdv4", "phase_22_d43", "phase_22_ansatz_exact",
                             "phase_22_ansatz_int_exact"]),
    "ringdown": ("rd", ["amp_22_gamma2", "amp_22_gamma3", "amp_22_v1", "amp_22_ansatz", "amp_22_ansatz_exact",
                        "amp_22_dansatz", "amp_22_peak_frequency", "phase_22_v4", "phase_22_d12", "phase_22_d24",
                        "phase_22_d34", "phase_22_d54", "phase_22_ansatz_exact", "phase_22_ansatz_int_exact"]),
}

# struct-typed parameters, by name
STRUCT_ARGS = {"powers": "const Powers<R>&", "amp": "const AmpCoeffs<R>&", "phase": "const PhaseCoeffs<R>&",
               "wf": "const WfVars<R>&"}

# statement blocks of hand-written reference functions whose long literal
# expressions are transcribed too: (file, function, C name, outputs)
BLOCKS = [
    ("amp22.py", "amp_coefficients", "amp22_pn_terms",
     ["pn_two_thirds", "pn_three_thirds", "pn_four_thirds", "pn_five_thirds", "pn_six_thirds"]),
    ("phase22.py", "phase_coefficients", "phase22_pn_terms",
     ["phi2_ns", "phi3_ns", "phi3_s", "phi4_ns", "phi4_s", "phi5l_ns", "phi5l_s", "phi6_ns", "phi6_s",
      "phi6l_ns", "phi7_ns", "phi7_s", "phi8_s", "phi8l_s"]),
]
BLOCK_LOCALS = {"phi8_common"}

NAMES = {"PI": "M_PI_", "FRAC_PI_2": "(M_PI_ / 2.0)", "FRAC_PI_4": "(M_PI_ / 4.0)", "EGAMMA": "EGAMMA_",
         "LN_2": "LN2_"}

CALLS = {"fma": "o_fma", "fms": "o_fms", "fnms": "o_fnms", "sqrt_nonneg": "o_sqrt_nonneg"}
JNP = {"cbrt": "o_cbrt", "sqrt": "o_sqrt", "exp": "o_exp", "abs": "o_abs", "arctan": "o_atan", "power": "o_pow",
       "minimum": "o_min", "where": "o_where"}
# calls whose constant arguments are lifted to the scalar type R
LIFT = {"o_min", "o_where"}


class Emitter:
    def __init__(self, prefix, local_functions):
        self.prefix = prefix
        self.local = set(local_functions)

    def expr(self, node):
        if isinstance(node, ast.Constant):
            if isinstance(node.value, bool):
                return "true" if node.value else "false"
            value = float(node.value)
            text = repr(value)
            return text if ("." in text or "e" in text or "n" in text) else text + ".0"
        if isinstance(node, ast.Name):
            return NAMES.get(node.id, node.id)
        if isinstance(node, ast.Attribute):
            base = node.value
            if isinstance(base, ast.Name) and base.id in ("POWERS_OF_LALPI", "pi"):
                return f"lalpi().{node.attr}"
            return f"{self.expr(base)}.{node.attr}"
        if isinstance(node, ast.UnaryOp):
            if isinstance(node.op, ast.USub):
                return f"(-{self.expr(node.operand)})"
            if isinstance(node.op, ast.UAdd):
                return self.expr(node.operand)
            raise NotImplementedError(ast.dump(node))
        if isinstance(node, ast.BinOp):
            ops = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.BitAnd: "&&", ast.BitOr: "||"}
            if isinstance(node.op, ast.Pow):
                return f"o_pow({self.expr(node.left)}, {self.expr(node.right)})"
            return f"({self.expr(node.left)} {ops[type(node.op)]} {self.expr(node.right)})"
        if isinstance(node, ast.Compare):
            assert len(node.ops) == 1
            ops = {ast.Gt: ">", ast.Lt: "<", ast.GtE: ">=", ast.LtE: "<=", ast.Eq: "==", ast.NotEq: "!="}
            return f"({self.expr(node.left)} {ops[type(node.ops[0])]} {self.expr(node.comparators[0])})"
        if isinstance(node, ast.Call):
            func = node.func
            args = node.args
            if isinstance(func, ast.Attribute) and isinstance(func.value, ast.Name) and func.value.id == "jnp":
                name = JNP[func.attr]
            elif isinstance(func, ast.Name) and func.id == "fma_sum":
                # constant + sum_i c_i x_i accumulated left to right with fma
                total = self.expr(args[0])
                for c, x in zip(args[1].elts, args[2].elts):
                    total = f"o_fma({self.expr(c)}, {self.expr(x)}, {total})"
                return total
            elif isinstance(func, ast.Name) and func.id in CALLS:
                name = CALLS[func.id]
            elif isinstance(func, ast.Name) and func.id in self.local:
                name = f"{self.prefix}_{func.id}"
            else:
                raise NotImplementedError(ast.dump(node))
            texts = [self.expr(a) for a in args]
            if name in LIFT:
                texts = [f"R({t})" if isinstance(a, (ast.Constant, ast.UnaryOp)) and not _names(a) else t
                         for a, t in zip(args, texts)]
            return f"{name}({', '.join(texts)})"
        raise NotImplementedError(ast.dump(node))

    def function(self, fn):
        params = []
        for a in fn.args.args:
            params.append(f"{STRUCT_ARGS.get(a.arg, 'R')} {a.arg}")
        if params:
            lines = [f"template <class R> static inline R {self.prefix}_{fn.name}({', '.join(params)}) {{"]
        else:
            lines = [f"static inline double {self.prefix}_{fn.name}() {{"]
        declared = set(a.arg for a in fn.args.args)
        for stmt in fn.body:
            if isinstance(stmt, ast.Expr) and isinstance(stmt.value, ast.Constant):
                continue  # docstring
            if isinstance(stmt, ast.Assign):
                (target,) = stmt.targets
                name = target.id
                value = self.expr(stmt.value)
                if name in declared:
                    lines.append(f"    {name} = {value};")
                else:
                    declared.add(name)
                    lines.append(f"    R {name} = {value};")
            elif isinstance(stmt, ast.AugAssign):
                ops = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}
                name = stmt.target.id
                lines.append(f"    {name} = ({name} {ops[type(stmt.op)]} {self.expr(stmt.value)});")
            elif isinstance(stmt, ast.Return):
                lines.append(f"    return {self.expr(stmt.value)};")
            else:
                raise NotImplementedError(ast.dump(stmt))
        lines.append("}")
        return "\n".join(lines)


def _names(node):
    return {n.id for n in ast.walk(node) if isinstance(n, ast.Name)}


def block(path: Path, function: str, cname: str, outputs):
    """The assignments to ``outputs`` (and their helper locals) of a
    reference function, as a C function with the free names as inputs."""
    tree = ast.parse(path.read_text())
    fn = next(n for n in ast.walk(tree) if isinstance(n, ast.FunctionDef) and n.name == function)
    wanted = set(outputs) | BLOCK_LOCALS
    stmts = [s for s in fn.body if isinstance(s, ast.Assign) and len(s.targets) == 1
             and isinstance(s.targets[0], ast.Name) and s.targets[0].id in wanted]
    free = set()
    for s in stmts:
        free |= _names(s.value)
    free -= wanted | {"fma", "fms", "fnms", "pi", "PI", "EGAMMA", "LN_2", "jnp"}
    inputs = sorted(free)
    emitter = Emitter("", set())
    params = [f"R {n}" for n in inputs] + [f"R& {n}" for n in outputs]
    lines = [f"// {path.name}:{fn.lineno} ({function}): the statements assigning {', '.join(outputs)}",
             f"template <class R> static inline void {cname}({', '.join(params)}) {{"]
    declared = set(outputs)
    for s in stmts:
        name = s.targets[0].id
        value = emitter.expr(s.value)
        if name in declared:
            lines.append(f"    {name} = {value};")
        else:
            declared.add(name)
            lines.append(f"    R {name} = {value};")
    lines.append("}")
    return "\n".join(lines), inputs


def generate(reference_root: Path) -> str:
    base = reference_root / "src/waveformx/models/xpnr/_generated"
    out = [
        "// GENERATED by oracle/gen_fits.py from the reference's",
        "// src/waveformx/models/xpnr/_generated/{utilities,qnm,inspiral,intermediate,ringdown}.py",
        "// (LALSimulation 6.2.0's IMRPhenomX fits via waveformrs; GPL-3.0-or-later data).",
        "// Oracle only: generic over the scalar R (double, or the oracle's forward-mode dual),",
        "// fma(a, b, c) unfused as a * b + c. Do not edit.",
        "#pragma once",
        "",
    ]
    for module, (prefix, wanted) in MODULES.items():
        path = base / f"{module}.py"
        tree = ast.parse(path.read_text())
        functions = {n.name: n for n in tree.body if isinstance(n, ast.FunctionDef)}
        emitter = Emitter(prefix, functions)
        out.append(f"// ---- {module}.py ----")
        for name in wanted:
            fn = functions[name]
            out.append(f"// {module}.py:{fn.lineno}")
            out.append(emitter.function(fn))
            out.append("")
    # the (2,2) QNM fits as coefficient lists in powers of af, for the
    # re-expansion in powers of 1 - af (xas/core.py:183-211)
    qnm = ast.parse((base / "qnm.py").read_text())
    for fn in (n for n in qnm.body if isinstance(n, ast.FunctionDef) and n.name in ("fring22", "fdamp22")):
        for stmt in fn.body:
            if isinstance(stmt, ast.Assign) and isinstance(stmt.value, ast.Call) and \
                    getattr(stmt.value.func, "id", "") == "fma_sum":
                const, coeffs, pows = stmt.value.args
                degree = {"af": 1}
                table = {0: float(const.value)}
                for c, x in zip(coeffs.elts, pows.elts):
                    k = degree.get(x.id) or int(x.id[1:])
                    table[k] = float(c.value) if isinstance(c, ast.Constant) else -float(c.operand.value)
                n = max(table) + 1
                values = ", ".join(repr(table.get(k, 0.0)) for k in range(n))
                out.append(f"// qnm.py:{stmt.lineno}: {fn.name} {stmt.targets[0].id}, coefficients of af^0..af^{n - 1}")
                out.append(f"static const double qnm_{fn.name}_{stmt.targets[0].id}[{n}] = {{{values}}};")
        out.append("")
    for file, function, cname, outputs in BLOCKS:
        text, _ = block(reference_root / "src/waveformx/models/xpnr" / file, function, cname, outputs)
        out.append(text)
        out.append("")
    return "\n".join(out) + "\n"


# the IMRPhenomXHM (release 122022) fits: module -> (C prefix, functions
# not emitted: the dispatchers by mode, which the oracle writes as switches)
HM_MODULES = {
    "hm_amp_fits_inspiral": ("hai", {"amp_fit"}),
    "hm_amp_fits_intermediate": ("ham", {"amp_fit"}),
    "hm_amp_fits_ringdown": ("har", set()),
    "hm_intermediate": ("hpi", {"phase_fit"}),
    "hm_inspiral": ("hpn", {"phase_lambda_fit"}),
    "hm_ringdown": ("hpr", {"phase_32_fit"}),
}
HM_QNM = ["fring21", "fring33", "fring32", "fring44", "fdamp21", "fdamp33", "fdamp32", "fdamp44",
          "re_l2m2lp2", "re_l2m2lp3", "re_l3m2lp2", "re_l3m2lp3", "im_l2m2lp2", "im_l2m2lp3", "im_l3m2lp2",
          "im_l3m2lp3"]
# parameters that are not the scalar R
INT_ARGS = {"mode_tag"}


class HmEmitter(Emitter):
    """The emitter with ``if``/``elif`` on Python ints (``lambda_pn``'s
    choice by mode) and ``math.log`` of constants."""

    def expr(self, node):
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) and \
                isinstance(node.func.value, ast.Name) and node.func.value.id == "math":
            assert node.func.attr == "log"
            return f"std::log({self.expr(node.args[0])})"
        if isinstance(node, ast.Compare) and isinstance(node.left, ast.Name) and node.left.id in INT_ARGS:
            (op,), (right,) = node.ops, node.comparators
            assert isinstance(op, ast.Eq)
            return f"({node.left.id} == {int(right.value)})"
        return super().expr(node)

    def statements(self, body, declared, indent):
        lines = []
        for stmt in body:
            if isinstance(stmt, ast.Expr) and isinstance(stmt.value, ast.Constant):
                continue
            if isinstance(stmt, ast.Assign):
                (target,) = stmt.targets
                value = self.expr(stmt.value)
                if target.id in declared:
                    lines.append(f"{indent}{target.id} = {value};")
                else:
                    declared.add(target.id)
                    lines.append(f"{indent}R {target.id} = {value};")
            elif isinstance(stmt, ast.Return):
                lines.append(f"{indent}return {self.expr(stmt.value)};")
            elif isinstance(stmt, ast.If):
                # (each branch assigns the same names: declare them first)
                for name in sorted(_assigned(stmt) - declared):
                    declared.add(name)
                    lines.append(f"{indent}R {name};")
                lines.append(f"{indent}if {self.expr(stmt.test)} {{")
                lines += self.statements(stmt.body, declared, indent + "    ")
                lines.append(f"{indent}}} else {{")
                lines += self.statements(stmt.orelse, declared, indent + "    ")
                lines.append(f"{indent}}}")
            else:
                raise NotImplementedError(ast.dump(stmt))
        return lines

    def function(self, fn):
        params = [f"{'int' if a.arg in INT_ARGS else STRUCT_ARGS.get(a.arg, 'R')} {a.arg}" for a in fn.args.args]
        lines = [f"template <class R> static inline R {self.prefix}_{fn.name}({', '.join(params)}) {{"]
        lines += self.statements(fn.body, set(a.arg for a in fn.args.args), "    ")
        lines.append("}")
        return "\n".join(lines)


def _assigned(node):
    return {t.id for n in ast.walk(node) if isinstance(n, ast.Assign) for t in n.targets if isinstance(t, ast.Name)}


class _Poly:
    """Polynomial coefficients (ascending) from a fit's AST: ``fma(c, x, rest)``
    chains in the powers ``af``, ``x2`` ... of the final spin."""

    def __init__(self, names):
        self.names = names

    def of(self, node):
        if isinstance(node, ast.Constant):
            return {0: float(node.value)}
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return {k: -v for k, v in self.of(node.operand).items()}
        if isinstance(node, ast.Name):
            return {self.names[node.id]: 1.0}
        if isinstance(node, ast.Call) and node.func.id == "fma":
            a, b, c = (self.of(x) for x in node.args)
            return self.add(self.mul(a, b), c)
        if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
            return self.mul(self.of(node.left), self.of(node.right))
        raise NotImplementedError(ast.dump(node))

    @staticmethod
    def mul(a, b):
        out = {}
        for i, x in a.items():
            for j, y in b.items():
                out[i + j] = out.get(i + j, 0.0) + x * y
        return out

    @staticmethod
    def add(a, b):
        out = dict(a)
        for k, v in b.items():
            out[k] = out.get(k, 0.0) + v
        return out


def qnm_ratio(fn):
    """The numerator and denominator coefficients of a final-spin fit
    ``num / den`` (``hm_qnm.py``), ascending powers of ``af``."""
    names = {"af": 1}
    ret = None
    for stmt in fn.body:
        if isinstance(stmt, ast.Assign):
            name = stmt.targets[0].id
            names[name] = int(name[1:])
        elif isinstance(stmt, ast.Return):
            ret = stmt.value
    assert isinstance(ret, ast.BinOp) and isinstance(ret.op, ast.Div)
    poly = _Poly(names)
    out = []
    for part in (ret.left, ret.right):
        c = poly.of(part)
        out.append([c.get(k, 0.0) for k in range(max(c) + 1)])
    return out


def generate_hm(reference_root: Path) -> str:
    base = reference_root / "src/waveformx/models/xpnr/_generated"
    out = [
        "// GENERATED by oracle/gen_fits.py from the reference's",
        "// src/waveformx/models/xpnr/_generated/hm_*.py (LALSimulation 6.2.0's IMRPhenomXHM fits via",
        "// waveformrs; GPL-3.0-or-later data). Oracle only: generic over the scalar R, fma(a, b, c)",
        "// unfused as a * b + c. Do not edit.",
        "#pragma once",
        "",
    ]
    qnm = ast.parse((base / "hm_qnm.py").read_text())
    functions = {n.name: n for n in qnm.body if isinstance(n, ast.FunctionDef)}
    out.append("// ---- hm_qnm.py: numerator and denominator in ascending powers of af ----")
    for name in HM_QNM:
        num, den = qnm_ratio(functions[name])
        out.append(f"// hm_qnm.py:{functions[name].lineno}")
        for part, c in (("num", num), ("den", den)):
            out.append(f"static const double hmq_{name}_{part}[{len(c)}] = {{{', '.join(repr(x) for x in c)}}};")
    out.append("")
    for module, (prefix, skip) in HM_MODULES.items():
        tree = ast.parse((base / f"{module}.py").read_text())
        functions = {n.name: n for n in tree.body if isinstance(n, ast.FunctionDef)}
        emitter = HmEmitter(prefix, functions)
        out.append(f"// ---- {module}.py ----")
        for name, fn in functions.items():
            if name in skip:
                continue
            out.append(f"// {module}.py:{fn.lineno}")
            out.append(emitter.function(fn))
            out.append("")
    # xhm/reference.py:189: the inverse of the pseudo-PN collocation system
    # of the inspiral amplitude, (x^p) for x = 0.5, 0.75, 1 and p = 7/3, 8/3, 3
    import numpy as np

    inverse = np.linalg.inv(np.array([[x ** q for q in (7.0 / 3.0, 8.0 / 3.0, 3.0)] for x in (0.5, 0.75, 1.0)]))
    rows = ", ".join("{" + ", ".join(repr(float(v)) for v in row) + "}" for row in inverse)
    out.append("// xhm/reference.py:189 (numpy.linalg.inv in float64)")
    out.append(f"static const double hm_inspiral_inverse[3][3] = {{{rows}}};")
    return "\n".join(out) + "\n"


def main():
    root = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    here = Path(__file__).resolve().parent
    (here / "xas_fits.h").write_text(generate(root))
    (here / "xhm_fits.h").write_text(generate_hm(root))


if __name__ == "__main__":
    main()
