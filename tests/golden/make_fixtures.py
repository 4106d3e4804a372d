"""Writes the golden fixtures of tests/golden/ from the reference checkout.

Run in a container that has the reference at /root/reference:

    python tests/golden/make_fixtures.py [/root/reference]

It reads the reference's own test data as TEXT/YAML (no reference code is
imported or executed) and stores plain numbers:

* lib_goldens.json -- the expected arrays and their literal inputs from
  source/lib/tests/test_{env_mat_a,prod_force_a,prod_virial_a,
  prod_force_grad_a,prod_virial_grad_a,coord}.cc (5-decimal goldens);
* pt_tabulate_se_a.json -- table / table_info / em_x / em and the FP64
  expected forward and backward of source/tests/pt/test_tabulate_fusion_se_a.py;
* deeppot_sea_model.npz + deeppot_sea_cases.json -- the se_e2_a model weights
  of source/tests/infer/deeppot_sea.yaml and the expected energies, forces,
  virials and descriptors of source/tests/infer/deeppot-testcase.yaml.
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def cc_arrays(path: str) -> dict:
    """name -> list of numbers for every `std::vector<T> name = {...};`
    literal of a gtest source (first occurrence of each name wins, a second
    class in the same file gets the suffix @2, @3...)."""
    text = open(path).read()
    out = {}
    for m in re.finditer(r"std::vector<\s*(double|int|float)\s*>\s+(\w+)\s*=\s*\{(.*?)\};", text, re.S):
        name = m.group(2)
        vals = [float(v) for v in re.findall(NUM, m.group(3))]
        if m.group(1) == "int":
            vals = [int(v) for v in vals]
        key, k = name, 2
        while key in out:
            key = f"{name}@{k}"
            k += 1
        out[key] = vals
    return out


def py_tensors(path: str) -> dict:
    """self.NAME = torch.tensor([...]) literals of a unittest source."""
    text = open(path).read()
    out = {}
    for m in re.finditer(r"self\.(\w+)\s*=\s*torch\.tensor\(\s*\[(.*?)\]", text, re.S):
        out[m.group(1)] = [float(v) for v in re.findall(NUM, m.group(2))]
    return out


def main(ref: str) -> None:
    lt = os.path.join(ref, "source", "lib", "tests")
    lib = {}
    for name in ("test_env_mat_a", "test_prod_force_a", "test_prod_virial_a", "test_prod_force_grad_a",
                 "test_prod_virial_grad_a", "test_coord"):
        lib[name] = cc_arrays(os.path.join(lt, name + ".cc"))
    with open(os.path.join(HERE, "lib_goldens.json"), "w") as f:
        json.dump(lib, f)

    pt = py_tensors(os.path.join(ref, "source", "tests", "pt", "test_tabulate_fusion_se_a.py"))
    keep = ["table_tensor", "table_info_tensor", "em_x_tensor", "em_tensor", "expected_descriptor_tensor",
            "expected_dy_dem_x", "expected_dy_dem"]
    pt = {k: pt[k] for k in keep}
    pt["shapes"] = {"table_tensor": [8, 132], "em_x_tensor": [4, 4], "em_tensor": [4, 4, 4],
                    "expected_descriptor_tensor": [4, 4, 8], "expected_dy_dem_x": [4, 4],
                    "expected_dy_dem": [4, 4, 4], "last_layer_size": 8}
    with open(os.path.join(HERE, "pt_tabulate_se_a.json"), "w") as f:
        json.dump(pt, f)

    inf = os.path.join(ref, "source", "tests", "infer")
    with open(os.path.join(inf, "deeppot_sea.yaml")) as f:
        mdl = yaml.safe_load(f)["model"]
    desc, fit = mdl["descriptor"], mdl["fitting"]
    arrs = {
        "davg": np.array(desc["@variables"]["davg"]["value"], dtype=np.float64),
        "dstd": np.array(desc["@variables"]["dstd"]["value"], dtype=np.float64),
        "bias_atom_e": np.array(fit["@variables"]["bias_atom_e"]["value"], dtype=np.float64).reshape(-1),
        "out_bias": np.array(mdl["@variables"]["out_bias"]["value"], dtype=np.float64).reshape(-1),
        "out_std": np.array(mdl["@variables"]["out_std"]["value"], dtype=np.float64).reshape(-1),
    }
    meta = {"sel": desc["sel"], "rcut": desc["rcut"], "rcut_smth": desc["rcut_smth"], "neuron": desc["neuron"],
            "axis_neuron": desc["axis_neuron"], "type_one_side": desc["type_one_side"],
            "fit_neuron": fit["neuron"], "ntypes": fit["ntypes"], "embed": [], "fitting": []}
    for kind, coll in (("embed", desc["embeddings"]["networks"]), ("fitting", fit["nets"]["networks"])):
        for ni, net in enumerate(coll):
            layers = []
            for li, layer in enumerate(net["layers"]):
                v = layer["@variables"]
                key = f"{kind}{ni}_l{li}"
                arrs[key + "_w"] = np.array(v["w"]["value"], dtype=np.float64)
                arrs[key + "_b"] = np.array(v["b"]["value"], dtype=np.float64)
                if v.get("idt") is not None:
                    arrs[key + "_idt"] = np.array(v["idt"]["value"], dtype=np.float64)
                layers.append({"key": key, "act": layer["activation_function"] != "none",
                               "resnet": bool(layer["resnet"]), "idt": v.get("idt") is not None})
            meta[kind].append(layers)
    np.savez_compressed(os.path.join(HERE, "deeppot_sea_model.npz"), meta=json.dumps(meta), **arrs)

    with open(os.path.join(inf, "deeppot-testcase.yaml")) as f:
        tc = yaml.safe_load(f)
    cases = []
    for blk in tc["results"]:
        cases.append({k: blk[k] for k in ("coord", "atype", "box", "atomic_energy", "force", "atomic_virial",
                                          "descriptor") if k in blk})
    with open(os.path.join(HERE, "deeppot_sea_cases.json"), "w") as f:
        json.dump(cases, f)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
