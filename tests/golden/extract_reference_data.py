"""Turn the reference's own test-data modules into ``.npz`` fixtures.

The reference keeps its golden vectors as NumPy literals inside ``*_data.py``
modules under ``test/``.  This script *parses* those files with :mod:`ast`
(nothing from the reference is imported or executed) and stores every literal
array, number and string it finds as a flat ``.npz`` archive here, keyed
``<variable>::<list index>::<dict key>``.  Lambdas (e.g. the ``threshold`` of
``test/mesh/tetrahedron_mesh_data.py:4620``) are not data and are skipped; the
tests restate them.

Run in a container that has the reference checked out at ``/root/reference``::

    python tests/golden/extract_reference_data.py
"""
import ast
import os
import sys

import numpy as np

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))

SOURCES = {
    "ref_tetrahedron_mesh.npz": ("mesh/tetrahedron_mesh_data.py",
                                 ["init_data", "from_one_tetrahedron_data", "entity_measure",
                                  "grad_lambda", "grad_shape_function", "number_of_local_ipoints",
                                  "number_of_global_ipoints", "interpolation_points",
                                  "face_to_ipoint", "cell_to_ipoint", "from_box",
                                  "face_unit_normal", "face_unit_norm"]),
    "ref_triangle_mesh.npz": ("mesh/triangle_mesh_data.py",
                              ["init_data", "from_box_data", "entity_measure_data",
                               "grad_lambda_data", "grad_shape_function_data",
                               "interpolation_point_data"]),
    "ref_lagrange_fe_space.npz": ("functionspace/lagrange_fe_space_data.py",
                                  ["triangle_mesh_one_box"]),
    "ref_scalar_diffusion_integrator.npz": ("fem/scalar_diffusion_integrator_data.py",
                                            ["triangle_mesh_one_box"]),
    "ref_scalar_mass_integrator.npz": ("fem/scalar_mass_integrator_data.py",
                                       ["triangle_mesh_one_box"]),
    "ref_stroud_quadrature.npz": ("quadrature/stroud_quadrature_data.py", ["data"]),
    "ref_bilinear_form.npz": ("fem/bilinear_form_data.py", ["mesh_data"]),
}

_DTYPES = {"int32": np.int32, "int64": np.int64, "float64": np.float64,
           "float32": np.float32, "bool": np.bool_, "bool_": np.bool_, "int_": np.int64}


class Skip(Exception):
    pass


def _eval(node):
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, (ast.List, ast.Tuple)):
        return [_eval(e) for e in node.elts]
    if isinstance(node, ast.Dict):
        out = {}
        for k, v in zip(node.keys, node.values):
            try:
                out[_eval(k)] = _eval(v)
            except Skip:
                pass
        return out
    if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) \
            and node.value.id in ("np", "numpy"):
        if node.attr in _DTYPES:
            return _DTYPES[node.attr]
        raise Skip(node.attr)
    if isinstance(node, ast.Call):
        f = node.func
        if isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name) \
                and f.value.id in ("np", "numpy") and f.attr in ("array", "asarray"):
            data = _eval(node.args[0])
            kw = {k.arg: _eval(k.value) for k in node.keywords}
            dtype = kw.get("dtype", node.args[1] and _eval(node.args[1]) if len(node.args) > 1 else None)
            return np.array(data, dtype=dtype)
        raise Skip(ast.dump(f))
    raise Skip(type(node).__name__)


def _flatten(prefix, value, out):
    if isinstance(value, dict):
        for k, v in value.items():
            _flatten(f"{prefix}::{k}", v, out)
    elif isinstance(value, list) and value and isinstance(value[0], dict):
        for i, v in enumerate(value):
            _flatten(f"{prefix}::{i}", v, out)
    elif value is None:
        return
    else:
        try:
            out[prefix] = np.asarray(value)
        except ValueError:          # ragged literal such as ``([0,1,0,1], 1, 1)``
            out[prefix] = np.asarray(repr(value))


def _parse_prefix(src):
    """Parse the longest run of whole top-level statements that is valid
    Python (``test/mesh/triangle_mesh_data.py`` has a syntax error near its
    end, inside data no test here uses)."""
    lines = src.splitlines(keepends=True)
    while True:
        try:
            return ast.parse("".join(lines))
        except SyntaxError as e:
            cut = e.lineno - 1
            while cut > 0 and (not lines[cut - 1].strip() or lines[cut - 1][0] in " \t#)]}"):
                cut -= 1
            lines = lines[:max(cut - 1, 0)]


def extract(path, names):
    tree = _parse_prefix(open(path).read())
    out = {}
    for st in tree.body:
        if isinstance(st, ast.Assign) and len(st.targets) == 1 \
                and isinstance(st.targets[0], ast.Name) and st.targets[0].id in names:
            try:
                value = _eval(st.value)
            except Skip:
                continue
            _flatten(st.targets[0].id, value, out)
    return out


RULE_SOURCES = {"tri": "fealpy/quadrature/triangle.py", "tet": "fealpy/quadrature/tetrahedron.py"}
ORACLE_RULES = os.path.join(HERE, "..", "..", "oracle", "quadrature_rules.npz")


def extract_rules(path, key):
    """The ``if index == k: A = bm.tensor([...])`` branches of a reference
    rule file's ``make`` method, as ``{f"{key}_{k}": (NQ, TD+2) float64}``."""
    tree = ast.parse(open(path).read())
    out = {}

    def visit_if(node):
        test = node.test
        if isinstance(test, ast.Compare) and isinstance(test.left, ast.Name) \
                and test.left.id == "index" and isinstance(test.comparators[0], ast.Constant):
            k = test.comparators[0].value
            for st in node.body:
                if isinstance(st, ast.Assign) and isinstance(st.value, ast.Call) \
                        and isinstance(st.value.func, ast.Attribute) and st.value.func.attr == "tensor":
                    out[f"{key}_{k}"] = np.array(_eval(st.value.args[0]), dtype=np.float64)
        for st in node.orelse:
            if isinstance(st, ast.If):
                visit_if(st)

    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == "make":
            for st in node.body:
                if isinstance(st, ast.If):
                    visit_if(st)
    return out


def main():
    for fname, (rel, names) in SOURCES.items():
        data = extract(os.path.join(REF, rel), names)
        np.savez_compressed(os.path.join(HERE, fname), **data)
        print(fname, len(data), "arrays", file=sys.stderr)
    rules = {}
    for key, rel in RULE_SOURCES.items():
        rules.update(extract_rules(os.path.join(os.path.dirname(REF), rel), key))
    np.savez_compressed(ORACLE_RULES, **rules)
    print(os.path.basename(ORACLE_RULES), len(rules), "rules", file=sys.stderr)


if __name__ == "__main__":
    main()
