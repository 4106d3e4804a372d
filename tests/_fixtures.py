"""Loaders for the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lib_goldens():
    with open(os.path.join(GOLD, "lib_goldens.json")) as f:
        return json.load(f)


def pt_tabulate():
    with open(os.path.join(GOLD, "pt_tabulate_se_a.json")) as f:
        d = json.load(f)
    sh = d["shapes"]
    out = {k: np.array(v, dtype=np.float64).reshape(sh[k]) if k in sh else np.array(v, dtype=np.float64)
           for k, v in d.items() if k != "shapes"}
    out["M"] = sh["last_layer_size"]
    return out


def deeppot_sea_model():
    """The reference's se_e2_a test model as an SeAModel (uncompressed)."""
    from dpmi.model import Layer, SeAModel

    z = np.load(os.path.join(GOLD, "deeppot_sea_model.npz"))
    meta = json.loads(str(z["meta"]))

    def layers(spec):
        return [Layer(z[s["key"] + "_w"], z[s["key"] + "_b"], z[s["key"] + "_idt"] if s["idt"] else None,
                      s["act"], s["resnet"]) for s in spec]

    nt = meta["ntypes"]
    nnei = sum(meta["sel"])
    m = SeAModel(ntypes=nt, sel=meta["sel"], rcut=meta["rcut"], rcut_smth=meta["rcut_smth"],
                 neuron=meta["neuron"], axis_neuron=meta["axis_neuron"], fit_neuron=meta["fit_neuron"],
                 type_one_side=meta["type_one_side"], davg=z["davg"].reshape(nt, nnei, 4),
                 dstd=z["dstd"].reshape(nt, nnei, 4), embed=[layers(s) for s in meta["embed"]],
                 fitting=[layers(s) for s in meta["fitting"]],
                 bias_atom_e=z["bias_atom_e"].copy())
    assert np.all(z["out_bias"] == 0) and np.all(z["out_std"] == 1)
    return m


def deeppot_sea_cases():
    with open(os.path.join(GOLD, "deeppot_sea_cases.json")) as f:
        return json.load(f)
