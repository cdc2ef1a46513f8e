"""`caroli run` / `caroli check` for transmission runs, on the MI355X engine.

    python -m caroli_amd.cli run INPUT.toml [--out DIR] [--set NAME=VALUE]... [--only K,...]
                                            [--precision double|mixed]
    python -m caroli_amd.cli check INPUT.toml [--set NAME=VALUE]...

Reads the input format of the reference's command line (crates/caroli-cli/src/input.rs:36-170:
`[parameters]`, `[model]`, `[[disorder]]`, `[system]`, `[[run]]`) and runs its transmission
runs as crates/caroli-cli/src/run.rs:70-190 does: per scan point, the lead self-energies once per
energy on configuration 0, then every configuration's system (disorder seeded `seed + c`,
crates/caroli-cli/src/build.rs:251-310) and ONE batched launch over all (configuration, energy)
points instead of a rayon task per point. Two leads on opposite faces of one axis go to the
two-terminal solver (`<scan>,config,energy,T`); leads on other faces, or `sectors`, to the
multi-terminal one (`T_p_q` or `T_p.a_q.b` columns, run.rs:80-104, :130-185). Each run writes a CSV
with the reference's columns and number formats.

Covered: the engine's presets (`bhz` with the default parameters, `square`, `cubic`, `graphene`
with t = 1 and spacing 1), Anderson and magnetic disorder (isotropic, polarized or cone moments;
same-g or opposite-g spin matrices; `cells` and `sublattices` restrictions), parameters with
expressions, scans, `configurations`, leads on any faces, `sectors`. Everything else of the
reference's input (other run types, continuum models, lattices of one's own, fields) is refused with
an error naming it.
"""
import argparse
import ast
import itertools
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import NestedMultiTerminal, NestedTwoTerminal, Precision, _capi as capi, batch_transmission, multi_batch_transmissions

try:
    import tomllib as _toml
except ModuleNotFoundError:  # Python < 3.11
    import tomli as _toml

PRESETS = {"bhz": 0, "square": 1, "graphene": 2, "cubic": 3}
DIMENSION = {"bhz": 2, "square": 2, "graphene": 2, "cubic": 3}
ORBITALS = {"bhz": 4, "square": 1, "graphene": 1, "cubic": 1}


class InputError(ValueError):
    pass


# -- numbers and parameters (crates/caroli-cli/src/build.rs:11-118) ----------------------------------

_FUNCS = {"sqrt": math.sqrt, "sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log,
          "log": math.log, "abs": abs, "atan": math.atan, "asin": math.asin, "acos": math.acos}


def number(expr, params):
    """A number or an expression of the parameters: + - * / ^ (or **), parentheses, pi, and the
    functions of _FUNCS."""
    if isinstance(expr, bool):
        raise InputError(f"`{expr}` is not a number")
    if isinstance(expr, (int, float)):
        return float(expr)
    if not isinstance(expr, str):
        raise InputError(f"`{expr}` is not a number")
    try:
        tree = ast.parse(expr.replace("^", "**"), mode="eval")
    except SyntaxError as e:
        raise InputError(f"`{expr}`: {e.msg}") from None

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)) and not isinstance(n.value, bool):
            return float(n.value)
        if isinstance(n, ast.Name):
            if n.id in params:
                return params[n.id]
            if n.id == "pi":
                return math.pi
            raise InputError(f"`{expr}`: unknown name `{n.id}`")
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.UAdd, ast.USub)):
            v = ev(n.operand)
            return v if isinstance(n.op, ast.UAdd) else -v
        if isinstance(n, ast.BinOp) and isinstance(n.op, (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow)):
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, ast.Add):
                return a + b
            if isinstance(n.op, ast.Sub):
                return a - b
            if isinstance(n.op, ast.Mult):
                return a * b
            if isinstance(n.op, ast.Pow):
                return a ** b
            if b == 0.0:  # IEEE division, as f64 arithmetic gives
                return math.copysign(math.inf, a) if a else math.nan
            return a / b
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and len(n.args) == 1:
            return float(_FUNCS[n.func.id](ev(n.args[0])))
        raise InputError(f"`{expr}`: unsupported expression")

    return ev(tree)


def resolve(raw, overrides):
    """Params::resolve (build.rs:17-44): the parameters with overrides, expressions in any order."""
    for name in overrides:
        if name not in raw:
            raise InputError(f"unknown parameter `{name}` (not in [parameters])")
    done = dict(overrides)
    pending = {k: v for k, v in raw.items() if k not in done}
    while pending:
        before, last = len(pending), None
        for name in list(pending):
            try:
                done[name] = number(pending[name], done)
                del pending[name]
            except InputError as e:
                last = f"parameter `{name}`: {e}"
        if len(pending) == before:
            raise InputError(last or "unresolved parameters")
    return done


def values(v, params):
    """Params::values (build.rs:67-88): a number, a list, or {from, to, count | step}."""
    if isinstance(v, list):
        return [number(x, params) for x in v]
    if isinstance(v, dict):
        unknown = set(v) - {"from", "to", "count", "step"}
        if unknown:
            raise InputError(f"unknown range key(s) {sorted(unknown)}")
        a, b = number(v["from"], params), number(v["to"], params)
        if "count" in v and "step" not in v:
            n = int(v["count"])
            return [a if n == 1 else a + (b - a) * k / (n - 1) for k in range(n)]
        if "step" in v and "count" not in v:
            step = number(v["step"], params)
            if not step > 0:
                raise InputError("a range step must be positive")
            n = int(math.floor((b - a) / step + 1e-9)) + 1
            return [a + step * k for k in range(n)]
        raise InputError("a range needs either `count` or `step`")
    return [number(v, params)]


def scan_points(raw, base, scan):
    """build.rs:90-118: every combination of the scanned values, the parameters resolved for each."""
    p0 = resolve(raw, base)
    axes = [(k, values(v, p0)) for k, v in sorted(scan.items())]
    out = []
    for combo in itertools.product(*[vals for _, vals in axes]):
        over = dict(base)
        over.update({name: x for (name, _), x in zip(axes, combo)})
        out.append((list(combo), resolve(raw, over)))
    return out


# -- formats of the reference's tables (run.rs:22-35) -------------------------------------------------

def fmt_display(x):
    """Rust's `{}` of an f64: the shortest round-trip digits, never an exponent."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    s = np.format_float_positional(x, trim="-")
    return "-0" if s == "-0" else s


def fmt_exp(x):
    """Rust's `{:e}` of an f64: shortest round-trip mantissa, exponent without sign padding."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    mant, exp = np.format_float_scientific(x, trim="-", exp_digits=1).split("e")
    return f"{mant}e{int(exp)}"


# -- the input -----------------------------------------------------------------------------------

def parse(text):
    """lib.rs:35-43: the input, with at least one [[run]]."""
    data = _toml.loads(text)
    unknown = set(data) - {"title", "parameters", "model", "disorder", "system", "run"}
    if unknown:
        raise InputError(f"unknown table(s) {sorted(unknown)}")
    if not data.get("run"):
        raise InputError("no [[run]] in the input")
    if "model" not in data:
        raise InputError("missing [model]")
    return data


def preset_of(model, params):
    """The engine's preset of [model] (build.rs:10-83 for the presets it covers)."""
    if "preset" not in model:
        raise InputError("the MI355X front end takes [model] presets only (bhz, square, cubic, graphene)")
    name = model["preset"]
    if name not in PRESETS:
        raise InputError(f"preset `{name}` is not one of the engine's (bhz, square, cubic, graphene)")
    defaults = {"spacing": 1.0, "t": 1.0, "a": 1.0, "b": -0.376406, "c": 0.0, "d": -0.280933, "mass": -0.137,
                "lambda": 0.0}
    for key, val in model.items():
        if key == "preset":
            continue
        if key in defaults:
            x = number(val, params)
            if abs(x - defaults[key]) > 1e-12 * max(1.0, abs(defaults[key])):
                raise InputError(f"[model] {key} = {x}: the engine's {name} preset has {key} = {defaults[key]}")
            continue
        if key == "spin" and not val:
            continue
        raise InputError(f"[model] `{key}` is not supported by the MI355X front end")
    return name


def leads_of(system, dim):
    """build.rs:371-413: ("axis", axis) for two leads on opposite faces of one axis (the two-terminal
    solver), else ("faces", [(sign, axis), ...]) (the multi-terminal one)."""
    names = system.get("leads", ["-x", "+x"])
    axes = {"x": 0, "y": 1, "z": 2}
    parsed = []
    for n in names:
        if len(n) != 2 or n[0] not in "+-" or n[1] not in axes:
            raise InputError(f"lead `{n}`: a face such as -x or +y")
        if axes[n[1]] >= dim:
            raise InputError(f"lead `{n}` on a lattice of dimension {dim}")
        parsed.append((1 if n[0] == "+" else -1, axes[n[1]]))
    if len(parsed) < 2:
        raise InputError("transport needs at least two leads")
    if len(parsed) == 2 and parsed[0] == (-1, parsed[0][1]) and parsed[1] == (1, parsed[0][1]):
        return "axis", parsed[0][1]
    return "faces", parsed


def disorders_of(data, params, name):
    """cg_disorder records of the [[disorder]] tables (input.rs:131-170, build.rs:260-310)."""
    out = []
    allowed = {"type", "strength", "concentration", "exchange", "moments", "mz", "theta0", "spin", "cells",
               "sublattices", "seed"}
    for k, d in enumerate(data.get("disorder", [])):
        unknown = set(d) - allowed
        if unknown:
            raise InputError(f"[[disorder]] {k}: unknown key(s) {sorted(unknown)}")
        c = capi.Disorder()
        kind = d.get("type")
        if kind == "anderson":
            c.kind = 0
            if "strength" not in d:
                raise InputError(f"[[disorder]] {k}: anderson disorder needs `strength`")
            c.strength = number(d["strength"], params)
        elif kind == "magnetic":
            c.kind = 1
            for key in ("concentration", "exchange"):
                if key not in d:
                    raise InputError(f"[[disorder]] {k}: magnetic disorder needs `{key}`")
            c.concentration = number(d["concentration"], params)
            c.exchange = number(d["exchange"], params)
            moments = d.get("moments", "isotropic")
            if moments == "isotropic":
                c.moments = 0
            elif moments in ("polarized", "polarised"):
                c.moments, c.moment_param = 2, number(d.get("mz", 1.0), params)
            elif moments == "cone":
                if "theta0" not in d:
                    raise InputError(f"[[disorder]] {k}: cone moments need `theta0`")
                c.moments, c.moment_param = 1, math.radians(number(d["theta0"], params))
            else:
                raise InputError(f"[[disorder]] {k}: unknown moments `{moments}` (isotropic, polarized, cone)")
            spin = d.get("spin")
            orbitals = ORBITALS[name]
            if spin is None and orbitals != 2:
                raise InputError(f"[[disorder]] {k}: give the `spin` matrices of sites with {orbitals} orbitals")
            spins = {"opposite-g": 0, "same-g": 1, "pauli": 2, None: 2}
            if not isinstance(spin, (str, type(None))) or spin not in spins:
                raise InputError(f"[[disorder]] {k}: spin `{spin}` (same-g, opposite-g or pauli here)")
            c.spin = spins[spin]
            if (c.spin == 2) != (orbitals == 2):
                raise InputError(f"[[disorder]] {k}: spin matrices of another size than the {orbitals} orbitals")
        else:
            raise InputError(f"[[disorder]] {k}: unknown disorder `{kind}` (anderson, magnetic)")
        for a, axis in enumerate("xyz"):
            r = d.get("cells", {}).get(axis)
            if r is not None:
                c.has_cells[a] = 1
                c.cells[a][0], c.cells[a][1] = int(r[0]), int(r[1])
        unknown = set(d.get("cells", {})) - set("xyz")
        if unknown:
            raise InputError(f"[[disorder]] {k}: unknown cells key(s) {sorted(unknown)}")
        for sub in d.get("sublattices", []):
            c.sublattices |= 1 << int(sub)
        c.seed = int(d.get("seed", 1 + 1000 * k))
        out.append(c)
    return out


class Setup:
    """The sample of a scan point: preset, size, lead axis, disorders."""

    def __init__(self, data, params):
        self.name = preset_of(data["model"], params)
        system = data.get("system", {})
        unknown = set(system) - {"size", "periodic", "leads"}
        if unknown:
            raise InputError(f"[system] unknown key(s) {sorted(unknown)}")
        if system.get("periodic"):
            raise InputError("a periodic sample has no leads")
        dim = DIMENSION[self.name]
        size = system.get("size")
        if not size or len(size) != dim or any(int(x) <= 0 for x in size):
            raise InputError(f"`size` needs {dim} positive numbers of cells")
        self.extent = [int(x) for x in size] + [0] * (3 - dim)
        self.leads, lead_arg = leads_of(system, dim)
        self.axis = lead_arg if self.leads == "axis" else None
        self.faces = lead_arg if self.leads == "faces" else [(-1, lead_arg), (1, lead_arg)]  # run.rs:131-134
        self.disorders = disorders_of(data, params, self.name)

    def system(self, config):
        """The system of configuration `config` (run.rs:63-68), built by cg_disordered_create."""
        import ctypes as C
        ext = (C.c_uint64 * 3)(*self.extent)
        arr = (capi.Disorder * max(1, len(self.disorders)))(*self.disorders)
        h = C.c_void_p()
        rc = capi.cg_disordered_create(PRESETS[self.name], ext, self.axis, len(self.disorders), arr, config,
                                       C.byref(h))
        if rc != 0:
            raise InputError(f"the engine refused the system (status {rc})")
        return NestedTwoTerminal(h.value)

    def multi(self, config):
        """The multi-terminal system of configuration `config` with a lead on every face
        (run.rs:130-138), built by cg_multi_disordered_create."""
        import ctypes as C
        ext = (C.c_uint64 * 3)(*self.extent)
        arr = (capi.Disorder * max(1, len(self.disorders)))(*self.disorders)
        n = len(self.faces)
        signs = (C.c_int32 * n)(*[f[0] for f in self.faces])
        axes = (C.c_int32 * n)(*[f[1] for f in self.faces])
        h = C.c_void_p()
        rc = capi.cg_multi_disordered_create(PRESETS[self.name], ext, n, signs, axes, len(self.disorders), arr, config,
                                             C.byref(h))
        if rc != 0:
            raise InputError(f"the engine refused the multi-terminal system (status {rc})")
        return NestedMultiTerminal(h.value)


def multi_rows(setup, energies, configs, sectors, prefix, precision, threads):
    """run.rs:130-185: every configuration's multi-terminal system, the self-energies once per
    energy (the leads are clean), one batched launch; per point the transmissions between leads (or
    (lead, sector) labels) p != q in the order of the header. Mixed precision runs these points in
    complex f64 (the refinement traces one lead pair)."""
    if precision == Precision.Mixed:
        precision = Precision.Double
    systems = [setup.multi(c) for c in range(configs)]
    count = len(setup.faces)
    labs = sorted(set(sectors)) if sectors is not None else None

    def sigma(e):
        try:
            return systems[0].self_energies(e, sectors=None if sectors is None else [list(sectors)] * count)
        except Exception:
            return None

    with ThreadPoolExecutor(threads) as ex:
        sig = list(ex.map(sigma, energies))
    pts = [(c, k) for c in range(configs) for k in range(len(energies)) if sig[k] is not None]
    width = count * (count - 1) if labs is None else (count * len(labs)) * (count * len(labs) - 1)
    cells = np.full((configs, len(energies), width), np.nan)
    if pts:
        t, st = multi_batch_transmissions([systems[c] for c, _ in pts], [sig[k] for _, k in pts], precision)
        for (c, k), tm, s in zip(pts, t, st):
            if s != 0:
                continue
            if labs is None:
                cells[c, k] = [tm[p, q] for p in range(count) for q in range(count) if q != p]
            else:
                at = {l: i for i, l in enumerate(sig[k].labels())}
                keys = [(p, a) for p in range(count) for a in labs]
                cells[c, k] = [tm[at[x], at[y]] if x in at and y in at else 0.0 for x in keys for y in keys if x != y]
    rows = []
    for c in range(configs):
        for k, e in enumerate(energies):
            rows.append(f"{prefix}{c},{fmt_display(e)}," + ",".join(fmt_exp(v) for v in cells[c, k]) + "\n")
    return rows


def transmission(data, run, base, precision, threads=16):
    """run.rs:70-190: the configurations x energies of a scan point in one batched launch, on the
    two-terminal solver for two leads along one axis, else on the multi-terminal one."""
    allowed = {"type", "output", "energies", "scan", "configurations", "sectors"}
    unknown = set(run) - allowed
    if unknown:
        raise InputError(f"unknown key(s) {sorted(unknown)}")
    sectors = run.get("sectors")
    if sectors is not None and (not isinstance(sectors, list) or not all(isinstance(x, int) and x >= 0 for x in sectors)):
        raise InputError("`sectors`: a list of non-negative sector numbers, one per orbital of a site")
    if "energies" not in run:
        raise InputError("a transmission run needs `energies`")
    scan = run.get("scan", {})
    points = scan_points(data.get("parameters", {}), base, scan)
    configs = int(run.get("configurations", 1))
    first = Setup(data, points[0][1])
    two = first.leads == "axis" and sectors is None
    count = len(first.faces)
    header = "".join(f"{k}," for k in sorted(scan)) + "config,energy"
    if two:
        header += ",T"
    elif sectors is not None:
        labs = sorted(set(sectors))
        keys = [(p, a) for p in range(count) for a in labs]
        header += "".join(f",T_{p}.{a}_{q}.{b}" for (p, a) in keys for (q, b) in keys if (p, a) != (q, b))
    else:
        header += "".join(f",T_{p}_{q}" for p in range(count) for q in range(count) if p != q)
    rows = [header + "\n"]
    for vals, params in points:
        energies = values(run["energies"], params)
        setup = Setup(data, params)
        prefix = "".join(f"{v}," for v in map(fmt_display, vals))
        if not two:
            rows += multi_rows(setup, energies, configs, sectors, prefix, precision, threads)
            continue
        systems = [setup.system(c) for c in range(configs)]
        reference = systems[0]  # the leads are clean: one set of self-energies serves every configuration

        def sigma(e):
            try:
                return reference.self_energies(e)
            except Exception:
                return None

        with ThreadPoolExecutor(threads) as ex:
            sig = list(ex.map(sigma, energies))
        pts = [(c, k) for c in range(configs) for k in range(len(energies)) if sig[k] is not None]
        t = np.full((configs, len(energies)), np.nan)
        if pts:
            tv, st = batch_transmission([systems[c] for c, _ in pts], [sig[k] for _, k in pts], precision)
            for (c, k), x, s in zip(pts, tv, st):
                t[c, k] = x if s == 0 else np.nan
        for c in range(configs):
            for k, e in enumerate(energies):
                rows.append(f"{prefix}{c},{fmt_display(e)},{fmt_exp(t[c, k])}\n")
    summary = f"transmission: {len(points)} scan points × {configs} configurations, {count} leads"
    return "".join(rows), summary


def output_name(k, run):
    """lib.rs:84-98."""
    return run.get("output") or f"{k:02d}-{run.get('type')}.csv"


def run_all(data, out_dir, overrides, only, precision):
    """lib.rs:100-115: every [[run]] (or those of `only`), one CSV each in out_dir."""
    os.makedirs(out_dir, exist_ok=True)
    lines = []
    for k, r in enumerate(data["run"]):
        if only is not None and k not in only:
            continue
        if r.get("type") != "transmission":
            raise InputError(f"run {k}: `{r.get('type')}` runs are outside the MI355X front end (transmission)")
        t0 = time.perf_counter()
        try:
            csv, summary = transmission(data, r, overrides, precision)
        except InputError as e:
            raise InputError(f"run {k}: {e}") from None
        path = os.path.join(out_dir, output_name(k, r))
        with open(path, "w") as f:
            f.write(csv)
        lines.append(f"{summary} → {path} ({time.perf_counter() - t0:.1f}s)")
    return lines


def describe(data, overrides):
    """lib.rs:46-81: the lattice, parameters, sample, disorder and runs."""
    params = resolve(data.get("parameters", {}), overrides)
    setup = Setup(data, params)
    out = [data["title"]] if data.get("title") else []
    dim = DIMENSION[setup.name]
    subs = 2 if setup.name == "graphene" else 1
    out.append(f"lattice: dimension {dim}, {subs} sublattice(s) with {[ORBITALS[setup.name]] * subs} orbitals")
    out.append(f"parameters: {dict(sorted(params.items()))}")
    s = setup.system(0)
    n, _, nsites = s.sizes()
    out.append(f"sample: {setup.extent[:dim]} cells, {nsites} sites, {n} orbitals")
    out.append(f"disorder: {len(setup.disorders)} term(s)")
    for k, r in enumerate(data["run"]):
        out.append(f"run {k}: {output_name(k, r)}")
    return "\n".join(out) + "\n"


USAGE = """caroli (MI355X front end) — transmission runs of a caroli input file on the GPU engine

usage:
  python -m caroli_amd.cli run INPUT.toml [--out DIR] [--set NAME=VALUE]... [--only K,...] [--precision P]
  python -m caroli_amd.cli check INPUT.toml [--set NAME=VALUE]...
"""


def main(argv=None):
    ap = argparse.ArgumentParser(prog="caroli", usage=USAGE)
    ap.add_argument("command", choices=["run", "check"])
    ap.add_argument("input")
    ap.add_argument("--out")
    ap.add_argument("--set", action="append", default=[])
    ap.add_argument("--only")
    ap.add_argument("--precision", choices=["double", "mixed"], default="double",
                    help="complex f64 fronts (the reference's Precision::Double) or the FP64-refined mixed mode")
    a = ap.parse_args(argv)
    try:
        overrides = {}
        for item in a.set:
            if "=" not in item:
                raise InputError("--set NAME=VALUE")
            name, v = item.split("=", 1)
            try:
                overrides[name] = float(v)
            except ValueError:
                raise InputError(f"--set {item}: not a number") from None
        with open(a.input) as f:
            data = parse(f.read())
        if a.command == "check":
            sys.stdout.write(describe(data, overrides))
            return 0
        out = a.out or os.path.splitext(a.input)[0]
        only = None if a.only is None else [int(k) for k in a.only.split(",")]
        precision = Precision.Double if a.precision == "double" else Precision.Mixed
        for line in run_all(data, out, overrides, only, precision):
            print(line)
        return 0
    except (InputError, _toml.TOMLDecodeError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
