"""TorchScript route of the TIP3P nonbonded model without Python registrations (no GPU needed).

The reference traces its TIP3P model for OpenMM's TorchForce (examples/tip3p/md.py:451-520),
which runs it in a libtorch process: every torch.ops.torchff operator has to be registered in
C++.  Here the model is scripted and saved, then loaded in a fresh interpreter that loads only
libtorchff_ops.so (the torchff Python package is never imported there), and the graph's
operators resolve; without the library the load fails.
"""

import os
import subprocess
import sys
import textwrap

import torch

from torchff import _lib
from torchff.tip3p import for_workload
from torchff.workloads import make_workload

OPS = ["build_neighbor_list_nsquared", "compute_lennard_jones_energy", "compute_multipolar_energy_from_atom_pairs",
       "pme_long_range"]


def _run(code):
    return subprocess.run([sys.executable, "-c", textwrap.dedent(code)], capture_output=True, text=True, timeout=300)


def test_scripted_tip3p_model_loads_through_the_operator_library_alone(tmp_path):
    w = make_workload("water900")
    path = str(tmp_path / "tip3p_nonbonded.pt")
    torch.jit.script(for_workload(w)).save(path)
    out = _run(f"""
        import sys, torch
        torch.ops.load_library({_lib.OPS_PATH!r})
        m = torch.jit.load({path!r})
        kinds = set()
        def walk(block):
            for n in block.nodes():
                kinds.add(n.kind())
                for b in n.blocks():
                    walk(b)
        walk(m.energy_components.graph)
        assert "torchff" not in sys.modules
        print(sorted(k for k in kinds if k.startswith("torchff::")))
        for k in kinds:
            if k.startswith("torchff::"):
                torch._C._jit_get_operation(k)
        print("ok")
    """)
    assert out.returncode == 0, out.stderr
    for op in OPS:
        assert f"torchff::{op}" in out.stdout, out.stdout
    assert out.stdout.strip().endswith("ok")


def test_scripted_model_needs_the_operator_library(tmp_path):
    w = make_workload("water900")
    path = str(tmp_path / "tip3p_nonbonded.pt")
    torch.jit.script(for_workload(w)).save(path)
    out = _run(f"""
        import torch
        torch.jit.load({path!r})
    """)
    assert out.returncode != 0 and "torchff" in out.stderr


def test_operator_library_is_built_next_to_the_kernels():
    assert os.path.exists(_lib.OPS_PATH) and os.path.dirname(_lib.OPS_PATH) == os.path.dirname(_lib.LIB_PATH)
