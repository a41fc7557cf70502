"""The allocation-free fused entry point (static device scratch) against the workspace entry point
(GPU), both called through the C ABI on the same inputs: energies, forces and status bitwise
equal (the fused path accumulates in int64 fixed point, so it is deterministic), on two calls in
a row, the second with a smaller system, in both precisions; and a system above the scratch
size is refused."""

import ctypes

import numpy as np
import pytest
import torch

from torchff import _lib as L
from torchff.fused import for_workload
from torchff.workloads import make_workload

pytestmark = pytest.mark.gpu


def _call(w, precision, noalloc):
    lib = L.load()
    m = for_workload(w, precision)
    dev = "cuda"
    c, box, q, s, e = (torch.tensor(w[k], dtype=torch.float64, device=dev)
                       for k in ("coords", "box", "charges", "sigma", "epsilon"))
    mod = [x.to(device=dev, dtype=m.compute_dtype).contiguous() for x in (m.xmoduli, m.ymoduli, m.zmoduli)]
    rp, col = m.excl_row_ptr.to(dev), m.excl_col_indices.to(dev)
    bh = L.box_host(box)
    cd = L.F32 if precision == "mixed" else L.F64
    out = torch.full((5,), float("nan"), dtype=torch.float64, device=dev)
    forces = torch.full_like(c, float("nan"))
    status = torch.full((2,), -1, dtype=torch.int32, device=dev)
    args = [L.F64, cd, L.ptr(c), L.ptr(q), L.ptr(s), L.ptr(e), m.natoms, bh, L.ptr(rp), L.ptr(col), m.cutoff,
            m.list_cutoff, m.alpha, m.coul, *m.K, m.order, *(L.ptr(x) for x in mod), L.ptr(out), L.ptr(forces),
            L.ptr(status), L.ptr(box), None]
    if noalloc:
        rc = lib.tff_fused_energy_forces_noalloc(*args, L.stream_of(c))
    else:
        nbytes = lib.tff_fused_workspace_bytes(cd, m.natoms, bh, m.list_cutoff, *m.K)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        rc = lib.tff_fused_energy_forces(*args, L.ptr(ws), nbytes, L.stream_of(c))
    L.check(rc, "fused")
    torch.cuda.synchronize()
    return out.cpu().numpy(), forces.cpu().numpy(), status.cpu().numpy()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_matches_workspace_entry_point(precision):
    for name in ("water96k", "water900"):  # the second call is the smaller system
        print(f"noalloc {precision} {name}", flush=True)
        w = make_workload(name)
        a = _call(w, precision, noalloc=True)
        b = _call(w, precision, noalloc=False)
        assert a[2][0] == 0 and np.array_equal(a[2], b[2]), (a[2], b[2])
        assert np.array_equal(a[0], b[0]), (a[0], b[0])
        assert np.array_equal(a[1], b[1])
        assert np.isfinite(a[1]).all()


def test_noalloc_refuses_a_system_above_the_scratch_size():
    lib = L.load()
    w = make_workload("water999k")
    box = (ctypes.c_double * 9)(*np.asarray(w["box"], dtype=np.float64).reshape(9).tolist())
    need = lib.tff_fused_workspace_bytes(0, w["natoms"], box, w["cutoff"], w["K"], w["K"], w["K"])
    assert need > (64 << 20)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = L.ptr(dummy)
    rc = lib.tff_fused_energy_forces_noalloc(L.F64, L.F32, p, p, p, p, w["natoms"], box, None, None, w["cutoff"],
                                             w["cutoff"], w["alpha"], 138.935456, w["K"], w["K"], w["K"], 5, p, p, p,
                                             p, p, p, None, None, L.stream_of(dummy))
    assert rc != 0 and b"static array" in lib.tff_last_error()
