"""CPU checkers for the se_e2_a path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package.  It holds:

* ``liboracle.so`` (sea_oracle.cc): the repo's C++ restatement of the
  reference CPU library functions on the path;
* ``_ref/libdeepmd_ref.so``: the reference CPU library itself, compiled from
  its sources by oracle/Makefile where the reference checkout exists;
* ``sea.py``: the Python glue of the reference path (descriptor contraction,
  fitting net, net_deriv, force/virial assembly) over either library, plus a
  direct (uncompressed) embedding-net evaluation.
"""
from .cpu import CpuLib, oracle_lib, ref_lib, ref_available  # noqa: F401
