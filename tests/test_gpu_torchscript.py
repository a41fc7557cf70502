"""The traced TIP3P nonbonded model in a process with only the operator library loaded (GPU).

As the reference does for OpenMM's TorchForce (examples/tip3p/md.py:451-472), the model is
traced with torch.jit.trace on the device and saved; a fresh interpreter that never imports the
torchff package loads libtorchff_ops.so and the saved module, evaluates the energy and the
forces by autograd, and the result is compared with the float64 oracle on the 900-atom box
(order-6 PME on 24^3): float64 to 1e-9 of the summed |terms| and force RMS 1e-9, float32 (the
reference traces in float32) to 1e-5 and 1e-4.
"""

import subprocess
import sys
import textwrap

import numpy as np
import pytest
import torch

from oracle import composite
from torchff import _lib
from torchff.tip3p import for_workload
from torchff.workloads import make_workload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,te,tf", [("float64", 1e-9, 1e-9), ("float32", 1e-5, 1e-4)])
def test_traced_model_matches_oracle_without_python_ops(tmp_path, dtype, te, tf):
    w = make_workload("water900")
    E_ref, F_ref, terms, _ = composite.energy_forces(w)
    dt = getattr(torch, dtype)
    x = torch.tensor(w["coords"], dtype=dt, device="cuda")
    box = torch.tensor(w["box"], dtype=dt, device="cuda")
    model = for_workload(w).to("cuda")
    traced = torch.jit.trace(model, (x, box), check_trace=False)
    path = str(tmp_path / "tip3p_traced.pt")
    traced.save(path)
    np.save(tmp_path / "x.npy", w["coords"])
    np.save(tmp_path / "box.npy", w["box"])
    out_path = str(tmp_path / "out.npz")
    code = f"""
        import sys
        import numpy as np
        import torch
        torch.ops.load_library({_lib.OPS_PATH!r})
        m = torch.jit.load({path!r}, map_location="cuda")
        x = torch.tensor(np.load({str(tmp_path / 'x.npy')!r}), dtype=torch.{dtype}, device="cuda",
                         requires_grad=True)
        box = torch.tensor(np.load({str(tmp_path / 'box.npy')!r}), dtype=torch.{dtype}, device="cuda")
        E = m(x, box)
        E.backward()
        assert "torchff" not in sys.modules
        np.savez({out_path!r}, E=float(E), F=-x.grad.double().cpu().numpy())
        print("ok", float(E))
    """
    print(f"traced model {dtype}: running in a fresh interpreter", flush=True)
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(code)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    got = np.load(out_path)
    scale = sum(abs(v) for v in terms.values())
    rms = float(np.sqrt(np.mean((got["F"] - F_ref) ** 2)) / np.sqrt(np.mean(F_ref ** 2)))
    print(f"traced {dtype}: E={float(got['E']):.8f} ref={E_ref:.8f} ({abs(float(got['E']) - E_ref) / scale:.2e} of "
          f"sum|terms|), force rms rel {rms:.2e}")
    assert abs(float(got["E"]) - E_ref) < te * scale
    assert rms < tf
