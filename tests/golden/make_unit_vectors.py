"""Collect the numeric golden vectors of the reference's potential unit tests into
reference_unit_vectors.json.

The test files are read as text and walked with `ast`; nothing from the reference is
imported or executed.  For every ``assertAlmostEqual``/``assertEqual`` whose first argument
is a call of a potential method (directly, or through a local name bound to such a call),
the potential's constructor arguments (from setUp), the call's arguments, the expected value
and the number of places are recorded.  Literal arithmetic in the tests (``-3.44 * 3.1``,
``x + 1.e-13``) is evaluated the same way the tests evaluate it.

Usage:  python tests/golden/make_unit_vectors.py [/path/to/reference]
"""
import ast
import json
import math
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_unit_vectors.json")
FILES = ["test_inverse_power_potential.py", "test_lennard_jones_potential.py", "test_hard_sphere_potential.py",
         "test_hard_dipole_potential.py",
         "test_displaced_even_power_potential.py", "test_inverse_power_coulomb_bounding_potential.py",
         "test_merged_image_coulomb_potential.py"]


class Unknown(Exception):
    pass


def ev(node, env):
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, (ast.List, ast.Tuple)):
        return [ev(e, env) for e in node.elts]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -ev(node.operand, env)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.UAdd):
        return +ev(node.operand, env)
    if isinstance(node, ast.BinOp):
        a, b = ev(node.left, env), ev(node.right, env)
        ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b, ast.Div: lambda: a / b,
               ast.Pow: lambda: a ** b}
        if type(node.op) in ops:
            return ops[type(node.op)]()
    if isinstance(node, ast.Name) and node.id in env:
        return env[node.id]
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == "float":
        return float(ev(node.args[0], env))
    if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id == "math":
        return getattr(math, node.attr)
    raise Unknown(ast.dump(node)[:80])


def constructors(fn, env):
    """self._x = Class(...) assignments and HypercubicSetting(system_length=...) in a setup method."""
    out, length = {}, None
    env = dict(env)
    for st in ast.walk(fn):
        if isinstance(st, ast.Call) and isinstance(st.func, ast.Attribute) and st.func.attr == "HypercubicSetting":
            for kw in st.keywords:
                if kw.arg == "system_length":
                    length = ev(kw.value, env)
    for st in fn.body:
        if isinstance(st, ast.Assign) and len(st.targets) == 1 and isinstance(st.targets[0], ast.Name):
            try:
                env[st.targets[0].id] = ev(st.value, env)
            except Unknown:
                pass
            continue
        if isinstance(st, ast.Assign) and isinstance(st.value, ast.Call) and isinstance(st.targets[0], ast.Attribute):
            c = st.value
            name = c.func.id if isinstance(c.func, ast.Name) else getattr(c.func, "attr", None)
            try:
                args = [ev(a, env) for a in c.args]
                kwargs = {kw.arg: ev(kw.value, env) for kw in c.keywords}
            except Unknown:
                continue
            out[st.targets[0].attr] = dict(cls=name, args=args, kwargs=kwargs)
    return out, length


def method_call(node):
    """self._pot.method(args) -> (pot attribute, method, args)"""
    if (isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute)
            and isinstance(node.func.value, ast.Attribute) and isinstance(node.func.value.value, ast.Name)
            and node.func.value.value.id == "self"):
        return node.func.value.attr, node.func.attr, node
    return None


def collect(path):
    tree = ast.parse(open(path).read())
    vectors = []
    for cls in [n for n in tree.body if isinstance(n, ast.ClassDef)]:
        methods = {f.name: f for f in cls.body if isinstance(f, ast.FunctionDef)}
        base_ctor, base_len = constructors(methods["setUp"], {}) if "setUp" in methods else ({}, None)
        for name, fn in methods.items():
            if not name.startswith("test"):
                continue
            env, ctor, length, bound = {}, dict(base_ctor), base_len, {}
            for st in fn.body:
                # self.setUpSystemLengthOne() style set-up calls
                if isinstance(st, ast.Expr) and isinstance(st.value, ast.Call) and \
                        isinstance(st.value.func, ast.Attribute) and st.value.func.attr.startswith("setUp"):
                    c, length = constructors(methods[st.value.func.attr], {})
                    ctor.update(c)
                    continue
                # self._potential._prefactor = ... (tests that change a constructor parameter)
                tgt = st.targets[0] if isinstance(st, ast.Assign) and len(st.targets) == 1 else None
                if (isinstance(tgt, ast.Attribute) and isinstance(tgt.value, ast.Attribute)
                        and isinstance(tgt.value.value, ast.Name) and tgt.value.value.id == "self"
                        and tgt.value.attr in ctor):
                    entry = json.loads(json.dumps(ctor[tgt.value.attr]))
                    entry["kwargs"][tgt.attr.lstrip("_")] = ev(st.value, env)
                    ctor[tgt.value.attr] = entry
                    continue
                if isinstance(st, ast.Assign) and len(st.targets) == 1 and isinstance(st.targets[0], ast.Name):
                    mc = method_call(st.value)
                    if mc:
                        bound[st.targets[0].id] = mc
                        continue
                    try:
                        env[st.targets[0].id] = ev(st.value, env)
                    except Unknown:
                        env.pop(st.targets[0].id, None)
                    continue
                if not (isinstance(st, ast.Expr) and isinstance(st.value, ast.Call)
                        and isinstance(st.value.func, ast.Attribute)
                        and st.value.func.attr in ("assertAlmostEqual", "assertEqual")):
                    continue
                call = st.value
                first = call.args[0]
                mc = method_call(first)
                if mc is None and isinstance(first, ast.Name) and first.id in bound:
                    mc = bound[first.id]
                if mc is None:
                    continue
                pot, meth, node = mc
                if pot not in ctor:
                    continue
                try:
                    args = [ev(a, env) for a in node.args]
                    kwargs = {kw.arg: ev(kw.value, env) for kw in node.keywords}
                    expected = ev(call.args[1], env)
                except Unknown:
                    continue
                places = None
                for kw in call.keywords:
                    if kw.arg == "places":
                        places = ev(kw.value, env)
                vectors.append(dict(file=os.path.basename(path), test=name, line=st.lineno, potential=ctor[pot],
                                    system_length=length, method=meth, args=args, kwargs=kwargs,
                                    expected=expected, places=places,
                                    exact=call.func.attr == "assertEqual"))
    return vectors


def main():
    out = []
    for f in FILES:
        out += collect(os.path.join(REF, "unittests", "test_potential", f))
    for v in out:
        if isinstance(v["expected"], float) and math.isinf(v["expected"]):
            v["expected"] = "inf" if v["expected"] > 0 else "-inf"
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    by = {}
    for v in out:
        by[v["file"]] = by.get(v["file"], 0) + 1
    print(by, len(out))


if __name__ == "__main__":
    main()
