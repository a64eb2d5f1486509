"""Write the golden fixtures under tests/golden/ from the reference checkout.

Run in the build container (needs /root/reference and oracle/_ref built by
``make -C oracle``):

    python tests/golden/make_fixtures.py

Outputs (data only -- inputs and expected outputs):

* ``cc_goldens.json``   -- the expected-value arrays and input arrays of the
  reference's gtest files ``source/lib/tests/test_{env_mat_a,fmt_nlist,
  prod_force_a,prod_virial_a,prod_force_grad_a,prod_virial_grad_a,
  tabulate_se_a}.cc``, parsed from their ``std::vector<...> name = {...};``
  initialisers, keyed by test class.
* ``cc_geometry.json``  -- the extended coordinates / types / mapping and the
  formatted neighbour lists those gtests build with the reference's own
  ``copy_coord_cpu`` / ``build_nlist_cpu`` / ``format_nlist_i_cpu``, produced
  by calling the compiled reference (oracle/_ref).
* ``pt_tabulate_op.json`` -- the tensors of
  ``source/tests/pt/test_tabulate_fusion_se_a.py`` (read with ``ast``, no
  execution).
* ``deeppot_sea.yaml`` / ``deeppot-testcase.yaml`` -- copies of the
  reference's se_a model text and its golden E/F/V cases
  (``source/tests/infer/``).
* ``pair_exclusion.json`` -- the two-frame neighbour-list case of
  ``source/tests/common/test_mixins.py`` (TestCaseSingleFrameWithNlist: the
  literal arrays, read with ``ast``, and its second frame permuted as that
  setUp does) with the expected PairExcludeMask of
  ``source/tests/pt/model/test_exclusion_mask.py`` for exclude_types [[0, 1]].
"""

from __future__ import annotations

import ast
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)

from oracle import refcapi  # noqa: E402

CC_FILES = ["test_env_mat_a.cc", "test_fmt_nlist.cc", "test_prod_force_a.cc",
            "test_prod_virial_a.cc", "test_prod_force_grad_a.cc",
            "test_prod_virial_grad_a.cc", "test_tabulate_se_a.cc"]


def parse_cc(path):
    """{class: {name: list}} for ``std::vector<T> name = {...};`` members."""
    text = open(path).read()
    out = {}
    cls_iter = list(re.finditer(r"class\s+(\w+)\s*:\s*public\s*::testing::Test", text))
    for k, m in enumerate(cls_iter):
        end = cls_iter[k + 1].start() if k + 1 < len(cls_iter) else len(text)
        body = text[m.end():end]
        arrs = {}
        for v in re.finditer(r"std::vector<(double|int)>\s+(\w+)\s*=\s*\{([^}]*)\};", body):
            typ, name, vals = v.groups()
            vals = [x for x in re.split(r"[,\s]+", vals.strip()) if x]
            conv = float if typ == "double" else int
            try:
                arrs[name] = [conv(x) for x in vals]
            except ValueError:  # expression initialisers (e.g. 1 << 24) are not data we use
                continue
        out[m.group(1)] = arrs
    return out


def parse_pt_tabulate(path):
    """self.<name> = torch.tensor([...]) [.reshape(...)] assignments."""
    tree = ast.parse(open(path).read())
    res = {}
    for node in ast.walk(tree):
        if not isinstance(node, ast.Assign) or len(node.targets) != 1:
            continue
        t = node.targets[0]
        if not (isinstance(t, ast.Attribute) and isinstance(t.value, ast.Name)
                and t.value.id == "self"):
            continue
        call, shape = node.value, None
        if isinstance(call, ast.Call) and isinstance(call.func, ast.Attribute) \
                and call.func.attr == "reshape":
            shape = [ast.literal_eval(a) for a in call.args]
            call = call.func.value
        if isinstance(call, ast.Call) and isinstance(call.func, ast.Attribute) \
                and call.func.attr == "tensor" and call.args:
            res[t.attr] = {"value": ast.literal_eval(call.args[0]), "shape": shape}
        elif t.attr in ("last_layer_size", "nloc", "nnei") and isinstance(node.value, ast.Constant):
            res[t.attr] = node.value.value
    return res


def cc_geometry():
    """Geometry of the gtests' 6-atom water fixture (13 A box, rc 6) built by
    the compiled reference: extended coords, types, mapping, formatted lists."""
    posi = np.array([12.83, 2.56, 2.18, 12.09, 2.87, 2.74, 00.25, 3.32, 1.68,
                     3.36, 3.00, 1.81, 3.51, 2.51, 2.60, 4.27, 3.22, 1.56])
    atype = np.array([0, 1, 1, 0, 1, 1], np.int32)
    box = np.array([13., 0., 0., 0., 13., 0., 0., 0., 13.])
    rc = 6.0
    ext, ext_t, mapping = refcapi.copy_coord(posi, atype, rc, box)
    nloc = 6
    raw, nn = refcapi.build_nlist(ext, nloc, rc)
    out = {"posi": posi.tolist(), "atype": atype.tolist(), "box": box.tolist(), "rc": rc,
           "posi_cpy": ext.ravel().tolist(), "atype_cpy": ext_t.tolist(),
           "mapping": mapping.tolist(), "raw_nlist": [raw[i, :nn[i]].tolist() for i in range(nloc)],
           "fmt": {}}
    for sel in ([10, 10], [5, 5], [2, 2]):
        f = [refcapi.format_nlist_i(ext, ext_t, i, raw[i, :nn[i]], rc, sel).tolist()
             for i in range(nloc)]
        out["fmt"]["%d_%d" % tuple(sel)] = f
    return out


def _ast_arrays(path, cls, names):
    """Literal ``self.<name> = np.array([...])`` / list assignments inside class
    ``cls`` of a Python file, read with ``ast`` (no execution)."""
    tree = ast.parse(open(path).read())
    out = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.ClassDef) and node.name == cls:
            for sub in ast.walk(node):
                if not isinstance(sub, ast.Assign) or len(sub.targets) != 1:
                    continue
                t = sub.targets[0]
                name = t.attr if isinstance(t, ast.Attribute) else getattr(t, "id", None)
                if name not in names or name in out:
                    continue
                v = sub.value
                if isinstance(v, ast.Call) and v.args:  # np.array([...], ...) (.reshape ignored)
                    f = v.func
                    while isinstance(f, ast.Attribute) and isinstance(f.value, ast.Call):
                        v = f.value
                        f = v.func
                    v = v.args[0]
                try:
                    out[name] = ast.literal_eval(v)
                except ValueError:
                    pass
    return out


def pair_exclusion_fixture():
    mix = _ast_arrays(os.path.join(REF, "source/tests/common/test_mixins.py"),
                      "TestCaseSingleFrameWithNlist", {"atype_ext", "nlist", "sel", "perm"})
    exp = _ast_arrays(os.path.join(REF, "source/tests/pt/model/test_exclusion_mask.py"),
                      "TestPairExcludeMask", {"expected_mask", "exclude_types"})
    sel = mix["sel"]
    nloc = 3
    atype = np.array(mix["atype_ext"]).reshape(1, -1)
    nlist = np.array(mix["nlist"]).reshape(1, nloc, sum(sel))
    perm = np.array(mix["perm"])
    inv = np.argsort(perm)
    # second frame: the first one permuted (TestCaseSingleFrameWithNlist.setUp)
    atype = np.concatenate([atype, atype[:, perm]], axis=0)
    n1 = nlist[:, perm[:nloc], :]
    n1 = np.where(n1 == -1, -1, inv[n1])
    nlist = np.concatenate([nlist, n1], axis=0)
    return {"ntypes": 2, "sel": sel, "atype_ext": atype.tolist(), "nlist": nlist.tolist(),
            "exclude_types": exp["exclude_types"],
            "expected_mask": np.array(exp["expected_mask"]).reshape(2, nloc, -1).tolist()}


def main():
    if refcapi.load() is None:
        raise SystemExit("build oracle/_ref first: make -C oracle")
    gold = {}
    for f in CC_FILES:
        gold[f] = parse_cc(os.path.join(REF, "source/lib/tests", f))
    with open(os.path.join(HERE, "cc_goldens.json"), "w") as fh:
        json.dump(gold, fh)
    with open(os.path.join(HERE, "cc_geometry.json"), "w") as fh:
        json.dump(cc_geometry(), fh)
    with open(os.path.join(HERE, "pt_tabulate_op.json"), "w") as fh:
        json.dump(parse_pt_tabulate(os.path.join(REF, "source/tests/pt/test_tabulate_fusion_se_a.py")), fh)
    with open(os.path.join(HERE, "pair_exclusion.json"), "w") as fh:
        json.dump(pair_exclusion_fixture(), fh)
    for f in ("deeppot_sea.yaml", "deeppot-testcase.yaml"):
        shutil.copyfile(os.path.join(REF, "source/tests/infer", f), os.path.join(HERE, f))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
