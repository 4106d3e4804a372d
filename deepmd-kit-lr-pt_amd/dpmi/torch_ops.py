"""Registers torch.ops.deepmd.* from the in-tree op library.

Mirrors deepmd/pt/cxx_op.py:56,130 (torch.ops.load_library of the op
library); raises when the library is missing instead of stubbing the op.
"""
from __future__ import annotations

import os

import torch

from . import lib as _lib

HERE = os.path.dirname(os.path.abspath(__file__))
OP_LIB = os.path.join(HERE, "libdeepmd_op_pt.so")
_loaded = False


def load() -> None:
    global _loaded
    if _loaded:
        return
    if not os.path.exists(OP_LIB):
        raise RuntimeError(f"{OP_LIB} is missing: build it with `python -m dpmi.torch_op_build`")
    _lib.load()
    torch.ops.load_library(OP_LIB)
    _loaded = True


def tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size: int):
    load()
    return torch.ops.deepmd.tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size)
