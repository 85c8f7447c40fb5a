"""Writes tests/golden/mod_rates_ttpc.npz, mod_rates_mainen.npz,
mod_currents.npz, mod_derivs_ttpc.npz and mod_derivs_mainen.npz.

mod_rates_ttpc.npz: the gating rates of the BBP TTPC
mechanisms evaluated from the NMODL sources' own PROCEDURE rates blocks
(Figures/Figure2_BBP_TTPC_clean/*.mod in the reference tree), on a voltage
grid (a calcium grid for SK_E2), so that the oracle's C restatement of the
rates (oracle/cpu_ref.c, which follows the generated AllModels.cu) can be
checked against the equations it was generated from.

Run here, where the reference tree is present:
    python tests/golden/make_mod_rates.py [/root/reference]

Only the arithmetic of each block is taken: its statements (assignments and
if/else on the gating variables, with NMODL's ^ as a power) are translated
into a Python function of v (or ca) and evaluated in double precision with
the math module.  The fixture holds inputs and outputs only.

It also writes tests/golden/mod_rates_mainen.npz from the Mainen set's
sources (Figures/FigureS4_mainen/{ca,kca,km,kv,na}.mod): their rates
procedures call the FUNCTIONs efun and trap0 of the same file and read the
PARAMETER block's defaults and celsius = 37 (the operating temperature the
files' comments and the model's fit.hoc set); the defaults used are stored
beside the outputs.
"""
import math
import os
import re
import sys

import numpy as np

MECHS = ["Ca_HVA", "Ca_LVAst", "Ih", "Im", "K_Pst", "K_Tst", "Nap_Et2", "NaTa_t", "NaTs2_t", "SK_E2", "SKv3_1"]
OUTS = {"SK_E2": ["zInf"]}


MAINEN = ["ca", "kca", "km", "kv", "na"]
CELSIUS_MAINEN = 37.0


def strip_comments(text):
    text = re.sub(r"COMMENT.*?ENDCOMMENT", "", text, flags=re.S)
    return re.sub(r":[^\n]*", "", text)


def strip_units(e):
    """NMODL unit annotations such as (mV) or (degC) inside expressions."""
    return re.sub(r"\(\s*(?:mV|degC|mM|ms|/ms|pS/um2|um)\s*\)", "", e)


def parameters(text):
    """Defaults of the PARAMETER block: name = number."""
    m = re.search(r"PARAMETER\s*\{(.*?)\}", text, flags=re.S)
    out = {}
    for line in m.group(1).split("\n"):
        mm = re.match(r"\s*([A-Za-z_]\w*)\s*=\s*([-+0-9.eE]+)", line)
        if mm:
            out[mm.group(1)] = float(mm.group(2))
    return out


def function_blocks(text):
    """FUNCTION name(args) { body } -> Python source defining name."""
    out = []
    for m in re.finditer(r"FUNCTION\s+(\w+)\s*\(([^)]*)\)\s*\{", text):
        name, args = m.group(1), [a.split("(")[0].strip() for a in m.group(2).split(",")]
        i, depth, j = m.end(), 1, m.end()
        while depth:
            depth += {"{": 1, "}": -1}.get(text[j], 0)
            j += 1
        src = to_python(text[i:j - 1], ", ".join(args), ret=name)
        out.append(src.replace("def rates(", "def %s(" % name, 1))
    return "\n".join(out)


def rates_block(text):
    """Body and argument name of PROCEDURE rates(...) { ... }."""
    m = re.search(r"PROCEDURE\s+rates\s*\((.*?)\)\s*\{", text)
    if not m:
        raise ValueError("no PROCEDURE rates")
    arg = strip_units(m.group(1)).split("(")[0].strip() or "v"
    i, depth = m.end(), 1
    j = i
    while depth:
        depth += {"{": 1, "}": -1}.get(text[j], 0)
        j += 1
    return text[i:j - 1], arg


def to_python(body, arg, ret=None):
    """The block's statements as the source of a Python function of arg
    (returning the named outputs, or the value assigned to ret)."""
    body = strip_units(re.sub(r":[^\n]*", "", body))        # NMODL comments, units
    body = body.replace("{", "{\n").replace("}", "\n}\n")
    lines = ["def rates(%s):" % arg]
    depth = 1
    for raw in body.split("\n"):
        s = raw.strip()
        if not s or s in ("UNITSOFF", "UNITSON") or s.startswith(("LOCAL", "TABLE", "DEPEND", "FROM")):
            continue
        if s == "}":
            depth -= 1
            continue
        if s.startswith("}"):
            s = s[1:].strip()
            depth -= 1
        m = re.match(r"(else\s*)?if\s*\((.*)\)\s*\{$", s)
        if m:
            kw = "elif" if m.group(1) else "if"
            lines.append("    " * depth + "%s %s:" % (kw, expr(m.group(2))))
            depth += 1
            continue
        if re.match(r"else\s*\{$", s):
            lines.append("    " * depth + "else:")
            depth += 1
            continue
        m = re.match(r"([A-Za-z_]\w*)\s*=\s*(.+)$", s)
        if not m or "__" in m.group(1):
            raise ValueError("statement not understood: %r" % s)
        lines.append("    " * depth + "%s = %s" % (m.group(1), expr(m.group(2))))
    if ret:
        return "\n".join(lines).replace(" %s = " % ret, " __ret = ") + "\n    return __ret"
    lines.append("    return dict(%s)" % ", ".join("%s=locals().get(%r)" % (o, o) for o in
                                               ("mInf", "mTau", "hInf", "hTau", "zInf", "minf", "mtau", "hinf",
                                                "htau", "ninf", "ntau")))
    return "\n".join(lines)


_BUILTINS = {"abs": abs, "dict": dict, "locals": locals}


def expr(e):
    """An NMODL arithmetic expression as Python.  The source files are not
    trusted: only numbers, operators, parentheses, commas and plain names
    pass, and a name may be called only if it is exp, fabs or a FUNCTION of
    the same file (checked when the function is compiled, by the restricted
    builtins and namespace of the exec)."""
    if not re.fullmatch(r"[\w\s.+\-*/^()<>=!,]*", e) or "__" in e or re.search(r"[A-Za-z_]\w*\s*\.", e):
        raise ValueError("expression outside the arithmetic subset: %r" % e)
    e = e.replace("^", "**")
    e = re.sub(r"\bfabs\s*\(", "abs(", e)
    return re.sub(r"\bexp\s*\(", "math.exp(", e)


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(root, "Figures", "Figure2_BBP_TTPC_clean")
    v = np.concatenate([np.arange(-120.0, 60.0, 0.37),
                        [-154.9, -66.0, -64.4, -60.0, -50.0, -38.0, -32.0, -27.0, -17.0, -60.0 - 10, -50.0 - 10]])
    ca = np.concatenate([10.0 ** np.arange(-8.0, -1.0, 0.05), [1e-7, 5e-5, 4.3e-4]])
    out = {"v": v, "ca": ca}
    for name in MECHS:
        body, arg = rates_block(open(os.path.join(src, name + ".mod")).read())
        code = to_python(body, arg)
        ns = {"math": math, "__builtins__": _BUILTINS}
        exec(compile(code, name + ".mod:rates", "exec"), ns)
        xs = ca if name == "SK_E2" else v
        res = [ns["rates"](float(x)) for x in xs]
        keys = OUTS.get(name, [k for k in ("mInf", "mTau", "hInf", "hTau") if res[0][k] is not None])
        out[name] = np.array([[r[k] for k in keys] for r in res])
        out[name + "_keys"] = np.array(keys)
        print(name, arg, keys, out[name].shape)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "mod_rates_ttpc.npz"), **out)

    src = os.path.join(root, "Figures", "FigureS4_mainen")
    cai = np.concatenate([10.0 ** np.arange(-7.0, -1.0, 0.05), [5e-5, 1e-4]])
    vm = np.concatenate([np.arange(-120.0, 60.0, 0.37), [-35.0, -30.0, 25.0, -27.0, -50.0, -75.0, -65.0]])
    out = {"v": vm, "cai": cai}
    for name in MAINEN:
        text = strip_comments(open(os.path.join(src, name + ".mod")).read())
        pars = parameters(text)
        body, arg = rates_block(text)
        code = function_blocks(text) + "\n" + to_python(body, arg)
        ns = {"math": math, "celsius": CELSIUS_MAINEN, "__builtins__": _BUILTINS}
        ns.update(pars)
        exec(compile(code, name + ".mod:rates", "exec"), ns)
        xs = cai if name == "kca" else vm
        res = [ns["rates"](float(x)) for x in xs]
        keys = [k for k in ("minf", "mtau", "hinf", "htau", "ninf", "ntau") if res[0][k] is not None]
        out[name] = np.array([[r[k] for k in keys] for r in res])
        out[name + "_keys"] = np.array(keys)
        out[name + "_pnames"] = np.array(sorted(pars))
        out[name + "_pvalues"] = np.array([pars[k] for k in sorted(pars)])
        print(name, arg, keys, out[name].shape)
    np.savez_compressed(os.path.join(here, "mod_rates_mainen.npz"), **out)
    currents(root, here)
    derivatives(root, here)
    derivatives_mainen(root, here)


# NEURON's faraday constant as the generated code holds it (FARADAY in the
# .mod UNITS block)
FARADAY = 96485.309


def derivatives(root, here, nsamples=200):
    """mod_derivs_ttpc.npz: each TTPC mechanism's states after one cnexp step
    of its DERIVATIVE block.  Every primed equation is linear in its state
    (checked), y' = a + b y with a and b from two evaluations of the
    right-hand side (rates called first, as the block does), so the step's
    exact result y(dt) = y_inf + (y0 - y_inf) exp(b dt), y_inf = -a / b, is
    what cnexp computes."""
    rng = np.random.default_rng(78)
    src = os.path.join(root, "Figures", "Figure2_BBP_TTPC_clean")
    out = {}
    for name in MECHS + ["CaDynamics_E2"]:
        text = strip_comments(open(os.path.join(src, name + ".mod")).read())
        m = re.search(r"DERIVATIVE\s+\w+\s*\{(.*?)\}", text, flags=re.S)
        call, eqs = None, []
        for line in m.group(1).split("\n"):
            line = line.strip()
            if not line:
                continue
            mm = re.match(r"rates\s*\(\s*(\w*)\s*\)$", line)
            if mm:
                call = mm.group(1) or "v"
                continue
            mm = re.match(r"([A-Za-z_]\w*)'\s*=\s*(.+)$", line)
            if not mm or "__" in mm.group(1):
                raise ValueError("statement not understood: %r" % line)
            eqs.append((mm.group(1), compile(expr(mm.group(2)), name + ".mod:DERIVATIVE", "eval")))
        rfun = None
        if call:
            body, arg = rates_block(text)
            ns0 = {"math": math, "__builtins__": _BUILTINS}
            exec(compile(to_python(body, arg), name + ".mod:rates", "exec"), ns0)
            rfun = ns0["rates"]
        pars = parameters(text)
        used = set(re.findall(r"[A-Za-z_]\w*", m.group(1)))
        pnames = sorted(k for k in pars if k in used)
        states = [y for y, _ in eqs]
        inputs = ["v", "dt", "cai", "ica"] + [y for y in states if y != "cai"] + pnames
        X = np.empty((nsamples, len(inputs)))
        Y = np.empty((nsamples, len(states)))
        for j in range(nsamples):
            vals = {"v": rng.uniform(-100.0, 50.0), "dt": float(rng.choice([0.02, 0.025, 0.1])),
                    "cai": 10 ** rng.uniform(-5, -2), "ica": -10 ** rng.uniform(-4, 0)}
            for y in states:
                if y != "cai":
                    vals[y] = rng.uniform(0.0, 1.0)
            for k in pnames:
                vals[k] = {"gamma": rng.uniform(5e-4, 0.05), "decay": rng.uniform(20.0, 1000.0),
                           "depth": rng.uniform(0.05, 1.0), "minCai": 1e-4}.get(k, rng.uniform(1.0, 10.0))
            ns = {"math": math, "__builtins__": _BUILTINS, "FARADAY": FARADAY}
            ns.update(vals)
            if rfun:
                ns.update({k: w for k, w in rfun(vals[call]).items() if w is not None})
            for i, (y, code) in enumerate(eqs):
                y0 = vals[y]
                ns[y] = 0.0
                a = eval(code, ns)
                ns[y] = 1.0
                b = eval(code, ns) - a
                ns[y] = y0
                assert abs(eval(code, ns) - (a + b * y0)) <= 1e-12 * max(abs(a), abs(b * y0), 1e-300)
                yinf = -a / b
                Y[j, i] = yinf + (y0 - yinf) * math.exp(b * vals["dt"])
            X[j] = [vals[k] for k in inputs]
        out[name + "_inputs"] = np.array(inputs)
        out[name + "_states"] = np.array(states)
        out[name + "_x"] = X
        out[name + "_y"] = Y
        print("derivative", name, inputs, states)
    np.savez_compressed(os.path.join(here, "mod_derivs_ttpc.npz"), **out)


# reversal potentials the models set (not part of the .mod files), as the
# oracle holds them: TTPC Ek -85, ENa 50; Mainen Ek -90, ENa 60 (mV)
IONS = {"ttpc": {"ek": -85.0, "ena": 50.0}, "mainen": {"ek": -90.0, "ena": 60.0}}


def currents(root, here, nsamples=200):
    """mod_currents.npz: the BREAKPOINT current of each mechanism (the sum of
    the ionic and nonspecific currents it writes) at random v, states,
    conductances and eca, from the .mod BREAKPOINT block."""
    rng = np.random.default_rng(77)
    out = {}
    sets = [("ttpc", os.path.join(root, "Figures", "Figure2_BBP_TTPC_clean"),
             ["Ca_HVA", "Ca_LVAst", "Ih", "Im", "K_Pst", "K_Tst", "Nap_Et2", "NaTa_t", "NaTs2_t", "SK_E2", "SKv3_1"]),
            ("mainen", os.path.join(root, "Figures", "FigureS4_mainen"), ["ca", "kca", "km", "kv", "na", "pasx"])]
    for tag, src, names in sets:
        for name in names:
            text = strip_comments(open(os.path.join(src, name + ".mod")).read())
            m = re.search(r"BREAKPOINT\s*\{(.*?)\n\}", text, flags=re.S)
            body = "\n".join(l for l in m.group(1).split("\n") if "SOLVE" not in l)
            neuron = re.search(r"NEURON\s*\{(.*?)\}", text, flags=re.S).group(1)
            writes = re.findall(r"WRITE\s+(\w+)", neuron) + re.findall(r"NONSPECIFIC_CURRENT\s+(\w+)", neuron)
            stm = re.search(r"STATE\s*\{(.*?)\}", text, flags=re.S)
            states = [w for w in re.findall(r"[A-Za-z_]\w*", stm.group(1)) if w not in ("FROM", "TO")] if stm else []
            pars = parameters(text)
            code = to_python(body, "v").replace("return dict(", "return dict(__cur=%s, " % " + ".join(writes), 1)
            code = code.replace("__cur=", "cur=")
            used = set(re.findall(r"[A-Za-z_]\w*", body))
            inputs = ["v"] + states + sorted(k for k in pars if k in used) + [k for k in ("eca",) if k in used]
            X = np.empty((nsamples, len(inputs)))
            y = np.empty(nsamples)
            for j in range(nsamples):
                vals = {"v": rng.uniform(-100.0, 50.0), "eca": rng.uniform(50.0, 150.0)}
                for st in states:
                    vals[st] = rng.uniform(0.0, 1.0)
                for k in pars:
                    if k in used:
                        vals[k] = 10 ** rng.uniform(-4, 0) if ("bar" in k or k == "g") else rng.uniform(-90.0, -30.0)
                ns = {"math": math, "__builtins__": _BUILTINS}
                ns.update(IONS[tag])
                if "q10" in pars:
                    ns["tadj"] = pars["q10"] ** ((CELSIUS_MAINEN - pars["temp"]) / 10)
                ns.update({k: vals[k] for k in inputs if k != "v"})
                exec(compile(code, name + ".mod:BREAKPOINT", "exec"), ns)
                X[j] = [vals[k] for k in inputs]
                y[j] = ns["rates"](vals["v"])["cur"]
            out[tag + ":" + name + "_inputs"] = np.array(inputs)
            out[tag + ":" + name + "_x"] = X
            out[tag + ":" + name + "_i"] = y
            print("current", tag, name, inputs, writes)
    np.savez_compressed(os.path.join(here, "mod_currents.npz"), **out)


# model values the Mainen cell sets outside its .mod files: vshift of na
# (fit.hoc: vshift_na = -5) and ca (0), and cad's fitted shell depth, as the
# oracle holds them
MAINEN_MODEL = {"na": {"vshift": -5.0}, "ca": {"vshift": 0.0}, "cad": {"depth": 0.09334562733124982}}


def derivatives_mainen(root, here, nsamples=200):
    """mod_derivs_mainen.npz: as derivatives(), for the Mainen set.  Its
    blocks call trates(x) (the rates procedure at x; the TABLE is not used)
    or rates(cai), and cad's block sets a clamped drive term first."""
    rng = np.random.default_rng(79)
    src = os.path.join(root, "Figures", "FigureS4_mainen")
    out = {}
    for name in ["ca", "cad", "kca", "km", "kv", "na"]:
        text = strip_comments(open(os.path.join(src, name + ".mod")).read())
        m = re.search(r"DERIVATIVE\s+\w+\s*\{(.*?)\n\}", text, flags=re.S)
        pars = parameters(text)
        pars.update(MAINEN_MODEL.get(name, {}))
        call, pre, eqs = None, [], []
        for line in m.group(1).split("\n"):
            line = line.strip()
            if not line:
                continue
            mm = re.match(r"t?rates\s*\(\s*([^)]*)\)$", line)
            if mm:
                call = compile(expr(mm.group(1) or "v"), name + ".mod:call", "eval")
                continue
            mm = re.match(r"([A-Za-z_]\w*)'\s*=\s*(.+)$", line)
            if mm and "__" not in mm.group(1):
                eqs.append((mm.group(1), compile(expr(mm.group(2)), name + ".mod:DERIVATIVE", "eval")))
                continue
            mm = re.match(r"if\s*\((.*)\)\s*\{\s*([A-Za-z_]\w*)\s*=\s*(.+?)\s*\}$", line)
            if mm and "__" not in mm.group(2):
                pre.append((mm.group(2), compile(expr(mm.group(3)), name, "eval"), compile(expr(mm.group(1)), name, "eval")))
                continue
            mm = re.match(r"([A-Za-z_]\w*)\s*=\s*(.+)$", line)
            if mm and "__" not in mm.group(1):
                if eqs:
                    continue                     # assignments after the equations (cai = ca)
                pre.append((mm.group(1), compile(expr(mm.group(2)), name, "eval"), None))
                continue
            raise ValueError("statement not understood: %r" % line)
        rfun = None
        if call:
            body, arg = rates_block(text)
            ns0 = {"math": math, "__builtins__": _BUILTINS, "celsius": CELSIUS_MAINEN}
            ns0.update(pars)
            code = function_blocks(text) + "\n" + to_python(body, arg)
            exec(compile(code, name + ".mod:rates", "exec"), ns0)
            rfun = ns0["rates"]
        states = [y for y, _ in eqs]
        inputs = ["v", "dt", "cai", "ica"] + [y for y in states if y != "cai"]
        X = np.empty((nsamples, len(inputs)))
        Y = np.empty((nsamples, len(states)))
        for j in range(nsamples):
            vals = {"v": rng.uniform(-100.0, 50.0), "dt": float(rng.choice([0.02, 0.025, 0.1])),
                    "cai": 10 ** rng.uniform(-5, -2), "ica": -10 ** rng.uniform(-4, 0) * rng.choice([1.0, -0.1])}
            for y in states:
                vals[y] = rng.uniform(0.0, 1.0) if y != "ca" else 10 ** rng.uniform(-5, -2)
            ns = {"math": math, "__builtins__": _BUILTINS, "FARADAY": FARADAY, "celsius": CELSIUS_MAINEN}
            ns.update(pars)
            ns.update(vals)
            for k, code, cond in pre:
                if cond is None or eval(cond, ns):
                    ns[k] = eval(code, ns)
            if rfun:
                ns.update({k: w for k, w in rfun(eval(call, ns)).items() if w is not None})
            for i, (y, code) in enumerate(eqs):
                y0 = vals[y]
                ns[y] = 0.0
                a = eval(code, ns)
                ns[y] = 1.0
                b = eval(code, ns) - a
                ns[y] = y0
                assert abs(eval(code, ns) - (a + b * y0)) <= 1e-12 * max(abs(a), abs(b * y0), 1e-300)
                yinf = -a / b
                Y[j, i] = yinf + (y0 - yinf) * math.exp(b * vals["dt"])
            X[j] = [vals[k] for k in inputs]
        out[name + "_inputs"] = np.array(inputs)
        out[name + "_states"] = np.array(states)
        out[name + "_x"] = X
        out[name + "_y"] = Y
        out[name + "_pnames"] = np.array(sorted(pars))
        out[name + "_pvalues"] = np.array([pars[k] for k in sorted(pars)])
        print("derivative mainen", name, inputs, states)
    np.savez_compressed(os.path.join(here, "mod_derivs_mainen.npz"), **out)


if __name__ == "__main__":
    main()
