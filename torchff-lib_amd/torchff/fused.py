"""Fused energy + forces of the fixed-charge water composition in one device call.

``FusedNonbondedPME`` evaluates what the reference's TIP3P model composes from
``NeighborList`` + ``Vdw(LennardJones)`` + ``MultipolarInteraction(rank=0, pairs_excl)``
+ ``PME(rank=0)`` (examples/tip3p/model.py:100-123): Lennard-Jones with
Lorentz-Berthelot mixing and erfc-screened Coulomb over all pairs within the
cutoff, the (erfc - 1)/r correction on excluded pairs, and smooth PME, returning
the energy terms and F = -dE/dr.  No pair list is materialised: atoms are
sorted into 32-atom spatial clusters and the pair kernel runs over cluster-pair
tiles (libtorchff_hip.so, ``tff_fused_energy_forces``).

precision="mixed": float32 arithmetic, int64 fixed-point accumulation of forces
and energies, float64 k-space and self sums.  precision="fp64": float64 throughout.

Differentiation: when an input requires grad, (E, F) are differentiable with respect to
coords, charges, sigma and epsilon through a hand-written vector-Jacobian product
(``tff_fused_energy_forces_backward``): a loss of the forces (force matching) gets
gE dE/dtheta - d(sum_i v_i . dE/dr_i)/dtheta with v = dL/dF, i.e. the double backward the
reference computes with create_graph on its pure-PyTorch path.  Such a forward call runs in
the backward's (larger) workspace, and its backward takes the cluster arrays and pair list the
forward left there (``tff_fused_energy_forces_backward_reuse``) when nothing has written that
workspace since, on the same stream; otherwise it rebuilds them.  The backward is not itself
differentiable.
"""

import math
import time
from collections import deque

import torch
import torch.nn as nn

from . import _lib as L
from . import bsplines
from .nblist import NeighborList
from .workloads import COULOMB_KJMOL

PAIR_OVERFLOW = 1
EXCL_OVERFLOW = 2
COLUMN_OVERFLOW = 4
BRICK_OVERFLOW = 8
BOX_STALE = 16
# every flag means the call's energies and forces are wrong
INVALID = PAIR_OVERFLOW | EXCL_OVERFLOW | COLUMN_OVERFLOW | BRICK_OVERFLOW | BOX_STALE
_MESSAGES = (
    (PAIR_OVERFLOW, "cluster-pair capacity exceeded (neighbour count per cluster far above the box average)"),
    (EXCL_OVERFLOW, "an atom has more than 8 exclusion partners"),
    (BRICK_OVERFLOW, "a charge-spreading brick is reached by more than 1024 clusters, or holds the stencil "
                     "bases of more than four times the mean number of atoms per brick plus 64 (strongly "
                     "inhomogeneous system or a grid much coarser than the atom spacing)"),
    (BOX_STALE, "the box tensor was changed without a version bump (through .data, a view or a graph "
                "replay) after its host copy was cached; call torchff.invalidate_box_cache(box) or "
                "torchff.set_box_cache(False)"),
    (COLUMN_OVERFLOW, "an x-y column holds more than 8192 atoms (strongly inhomogeneous system): it is left "
                      "unsorted in z, which the charge spreading's cluster search relies on"),
)
_RING = 8  # calls whose status may be outstanding at once


def status_message(flags):
    return "; ".join(m for bit, m in _MESSAGES if flags & bit)


class FusedNonbondedPME(nn.Module):
    def __init__(self, natoms: int, cutoff: float, ewald_alpha: float, max_hkl, order: int = 5,
                 coul_constant: float = COULOMB_KJMOL, skin: float = 0.0, exclusions=None,
                 precision: str = "mixed"):
        super().__init__()
        if precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        self.natoms = int(natoms)
        self.cutoff = float(cutoff)
        self.list_cutoff = float(cutoff) + float(skin)
        self.alpha = float(ewald_alpha)
        self.K = tuple(int(k) for k in max_hkl) if isinstance(max_hkl, (tuple, list)) else (int(max_hkl),) * 3
        self.order = int(order)
        self.coul = float(coul_constant)
        self.precision = precision
        self.compute_dtype = torch.float32 if precision == "mixed" else torch.float64
        if exclusions is not None:
            rp, col = NeighborList.convert_pairs_coo_to_csr(torch.as_tensor(exclusions), self.natoms)
            self.register_buffer("excl_row_ptr", rp)
            self.register_buffer("excl_col_indices", col)
        else:
            self.excl_row_ptr = None
            self.excl_col_indices = None
        K0, K1, K2 = self.K
        self.register_buffer("xmoduli", bsplines.compute_bspline_moduli_1d(K0, order=self.order))
        self.register_buffer("ymoduli", bsplines.compute_bspline_moduli_1d(K1, order=self.order))
        self.register_buffer("zmoduli", bsplines.compute_bspline_moduli_z(K2, order=self.order))
        self._ws = None
        self._ws_key = None
        self._mod = None
        self.status = None
        # status words of recent calls, mirrored into a row of pinned host memory by each call's last
        # kernel (word 1, a fence, then word 0 over the -1 the host stored before the call):
        # (row, what) in call order; read once word 0 has arrived, so that an overflow surfaces on a
        # later call (or at check()) without synchronising the call itself or recording an event
        self._ring = None
        self._ring_np = None
        self._ring_next = 0
        self._ring_dev = None
        self._ring_dstat = {}  # device -> (_RING + 1, 2) int32: each row the status words of its call
        self._pending = deque()
        self._captured = False  # a call was captured into a CUDA graph (its replays write ring row _RING)
        # the backward-sized workspace: writes to it are counted, so that a backward can tell whether
        # the cluster arrays and pair list of its forward call are still there
        self._bws = None
        self._bws_key = None
        self._bws_gen = 0
        self._last_token = None

    def _workspace(self, box_host, device):
        lib = L.load()
        cd = L.F32 if self.compute_dtype == torch.float32 else L.F64
        key = (tuple(box_host), device)
        if self._ws is None or self._ws_key != key:
            nbytes = lib.tff_fused_workspace_bytes(cd, self.natoms, box_host, self.list_cutoff, *self.K)
            if nbytes == 0:
                raise RuntimeError("FusedNonbondedPME needs an orthorhombic periodic box")
            self._ws = L.workspace(nbytes, device)
            self._ws_key = key
        return self._ws

    def _backward_workspace(self, box_host, device):
        lib = L.load()
        cd = L.F32 if self.compute_dtype == torch.float32 else L.F64
        key = (tuple(box_host), device)
        if self._bws is None or self._bws_key != key:
            nbytes = lib.tff_fused_backward_workspace_bytes(cd, self.natoms, box_host, self.list_cutoff, *self.K)
            if nbytes == 0:
                raise RuntimeError("FusedNonbondedPME needs an orthorhombic periodic box")
            self._bws = L.workspace(nbytes, device)
            self._bws_key = key
            self._bws_gen += 1
        return self._bws

    def _status_slot(self, device):
        """(device status tensor, pinned host mirror) for a new call, after reading the status of
        every earlier call that has completed (raising if one of them was invalid).  A call made
        while the stream is being captured into a CUDA graph takes the ring's last row, which its
        replays rewrite and check() reads (events cannot be polled under capture)."""
        capturing = torch.cuda.is_current_stream_capturing()
        if not capturing:
            self._poll(block=len(self._pending) >= _RING)
        if self._ring is None:
            # pinned memory is allocated outside any capture: the first call of a module is eager
            self._ring = torch.full((_RING + 1, 2), -1, dtype=torch.int32, pin_memory=True)
            self._ring_np = self._ring.numpy()
            self._ring_dev = device
        if capturing:
            row = _RING
            self._captured = True
        else:
            row = self._ring_next
            self._ring_next = (row + 1) % _RING
        self._ring_np[row] = -1
        dstat = self._ring_dstat.get(device)
        if dstat is None:
            if capturing:  # allocated outside any capture only
                return torch.empty(2, dtype=torch.int32, device=device), self._ring[row], row
            dstat = self._ring_dstat[device] = torch.empty((_RING + 1, 2), dtype=torch.int32, device=device)
        # a row is reused only after its earlier call has completed (at most _RING calls are
        # outstanding, and the oldest is waited for first)
        return dstat[row], self._ring[row], row

    def _record(self, row, what):
        if not torch.cuda.is_current_stream_capturing():
            self._pending.append((row, what))

    def _wait_row(self, row):
        """Flags of ring row ``row`` once its call's last kernel has written them: a short spin,
        then sleeps; after a second, a device synchronisation (which raises on a device error)."""
        t0 = time.perf_counter()
        spins = 0
        while True:
            flags = int(self._ring_np[row, 0])
            if flags >= 0:
                return flags
            spins += 1
            if spins > 64:
                time.sleep(2e-5)
            if time.perf_counter() - t0 > 1.0:
                torch.cuda.synchronize(self._ring_dev)
                flags = int(self._ring_np[row, 0])
                if flags < 0:
                    raise RuntimeError("FusedNonbondedPME: a call's status words never arrived")
                return flags

    def _poll(self, block=False, all_=False):
        """Read the status of completed calls in call order; raise on the first invalid one.
        block: wait for the oldest call; all_: wait for every call."""
        while self._pending:
            row, what = self._pending[0]
            flags = int(self._ring_np[row, 0])
            if flags < 0:
                if not (all_ or (block and len(self._pending) >= _RING)):
                    break
                flags = self._wait_row(row)
            self._pending.popleft()
            if flags & INVALID:
                self._pending.clear()
                raise RuntimeError(f"FusedNonbondedPME: an earlier {what} returned invalid results: "
                                   f"{status_message(flags)}")

    def _box_device(self, box, io):
        b = box.detach()
        if b.dtype != io or not b.is_contiguous():
            b = b.to(io).contiguous()
        return b

    @torch._dynamo.disable
    def forward(self, coords, box, charges, sigma, epsilon, check: bool = False):
        """Returns (energy float64 scalar, forces shaped like coords, terms float64 (4,) =
        [Lennard-Jones, real-space Coulomb incl. exclusion correction, reciprocal, self]).

        Capacity overflows and a stale box cache are reported: at once with check=True (one
        synchronisation), otherwise by raising on a later call as soon as this call's status has
        arrived on the host (no synchronisation), and by check()."""
        if torch.is_grad_enabled() and any(t.requires_grad for t in (coords, charges, sigma, epsilon)):
            E, F, terms = _FusedEnergyForces.apply(self, coords, charges, sigma, epsilon, box)
            if check:
                self.check()
            return E, F, terms
        return self._forward_raw(coords, box, charges, sigma, epsilon, check)

    def _forward_raw(self, coords, box, charges, sigma, epsilon, check=False, for_grad=False):
        """for_grad: run in the backward's workspace and leave in _last_token what its backward
        needs to take the cluster arrays and pair list from there (None under graph capture)."""
        L.require_cuda(coords, box, charges, sigma, epsilon)
        lib = L.load()
        dev = coords.device
        io = coords.dtype
        c = coords.detach().contiguous()
        if c.shape != (self.natoms, 3):
            raise RuntimeError(f"coords must be ({self.natoms}, 3)")
        q, s, e = (t.detach().to(io).contiguous() for t in (charges, sigma, epsilon))
        bh = L.box_host(box)
        ws = self._backward_workspace(bh, dev) if for_grad else self._workspace(bh, dev)
        if self._mod is None or self._mod[0].device != dev or self._mod[0].dtype != self.compute_dtype:
            self._mod = tuple(m.to(device=dev, dtype=self.compute_dtype).contiguous()
                              for m in (self.xmoduli, self.ymoduli, self.zmoduli))
        if self.excl_row_ptr is not None and self.excl_row_ptr.device != dev:
            # keep the exclusion CSR resident on the device (moved once, not per call)
            self.excl_row_ptr = self.excl_row_ptr.to(dev)
            self.excl_col_indices = self.excl_col_indices.to(dev)
        rp, col = self.excl_row_ptr, self.excl_col_indices
        out = torch.empty(5, dtype=torch.float64, device=dev)
        forces = torch.empty_like(c)
        status, host, row = self._status_slot(dev)
        bd = self._box_device(box, io)
        cd = L.F32 if self.compute_dtype == torch.float32 else L.F64
        if for_grad:
            # counted before the call: a call that fails part-way has still written the workspace
            self._bws_gen += 1
            self._last_token = None
        rc = lib.tff_fused_energy_forces(L.dtype_code(c), cd, L.ptr(c), L.ptr(q), L.ptr(s), L.ptr(e), self.natoms,
                                         bh, L.ptr(rp), L.ptr(col), self.cutoff, self.list_cutoff, self.alpha,
                                         self.coul, *self.K, self.order, *(L.ptr(m) for m in self._mod),
                                         L.ptr(out), L.ptr(forces), L.ptr(status), L.ptr(bd), L.ptr(host),
                                         L.ptr(ws), ws.numel(), L.stream_of(c))
        L.check(rc, "FusedNonbondedPME")
        if for_grad:
            capturing = torch.cuda.is_current_stream_capturing()
            self._last_token = None if capturing else (self._bws_gen, L.stream_of(c).value)
        self._record(row, "forward call")
        self.status = status
        if check:
            self.check()
        return out[4], forces, out[:4]

    def _backward_raw(self, coords, box, charges, sigma, epsilon, grad_energy, grad_forces, needs, token=None):
        """dL/d(coords, charges, sigma, epsilon) for upstream grad_energy (0-d, or None) and
        grad_forces ((N, 3), or None); entries of ``needs`` False give None.  token: the forward
        call's (from _last_token): while no call has written the backward workspace since, on the
        same stream and with no call of this module captured into a graph (whose replays write it
        unseen), the forward's cluster arrays and pair list are reused."""
        lib = L.load()
        dev = coords.device
        io = coords.dtype
        c = coords.detach().contiguous()
        q, s, e = (t.detach().to(io).contiguous() for t in (charges, sigma, epsilon))
        bh = L.box_host(box)
        cd = L.F32 if self.compute_dtype == torch.float32 else L.F64
        ws = self._backward_workspace(bh, dev)
        stream = L.stream_of(c)
        reuse = (token is not None and not self._captured and token == (self._bws_gen, stream.value)
                 and not torch.cuda.is_current_stream_capturing())
        ge = None if grad_energy is None else grad_energy.detach().to(device=dev, dtype=torch.float64).reshape(1)
        gf = None if grad_forces is None else grad_forces.detach().to(device=dev, dtype=io).contiguous()
        # the requested gradients as views of one allocation
        sizes = [3 * self.natoms if needs[0] else 0] + [self.natoms if n else 0 for n in needs[1:]]
        buf = torch.empty(sum(sizes), dtype=io, device=dev)
        outs, o = [], 0
        for k, n in enumerate(sizes):
            outs.append(None if n == 0 else (buf[o:o + n].view(self.natoms, 3) if k == 0 else buf[o:o + n]))
            o += n
        status, host, row = self._status_slot(dev)
        bd = self._box_device(box, io)
        rp, col = self.excl_row_ptr, self.excl_col_indices
        entry = lib.tff_fused_energy_forces_backward_reuse if reuse else lib.tff_fused_energy_forces_backward
        if not reuse:
            self._bws_gen += 1  # rebuilt in place (counted before the call, as in the forward)
        rc = entry(
            L.dtype_code(c), cd, L.ptr(c), L.ptr(q), L.ptr(s), L.ptr(e), self.natoms, bh, L.ptr(rp), L.ptr(col),
            self.cutoff, self.list_cutoff, self.alpha, self.coul, *self.K, self.order, *(L.ptr(m) for m in self._mod),
            L.ptr(ge), L.ptr(gf), *(L.ptr(o) for o in outs), L.ptr(status), L.ptr(bd), L.ptr(host),
            L.ptr(ws), ws.numel(), stream)
        L.check(rc, "FusedNonbondedPME backward")
        self._record(row, "backward call")
        self.backward_status = status
        return outs

    def tiles(self):
        """Number of 32 x 32 cluster-pair tiles within the list cutoff in the last call (synchronises)."""
        return None if self.status is None else int(self.status[1].item())

    def check(self):
        """Wait for every outstanding call (forward and backward) and raise if one of them
        overflowed an internal capacity or met a stale box cache.  Calls captured into a CUDA
        graph report the status of their most recent replay (after a device synchronisation)."""
        while self._pending:
            row, what = self._pending.popleft()
            flags = self._wait_row(row)
            if flags:
                self._pending.clear()
                raise RuntimeError(f"FusedNonbondedPME: {what}: {status_message(flags)}")
        if self._captured:
            torch.cuda.synchronize(self._ring_dev)
            flags = int(self._ring_np[_RING, 0])
            if flags > 0:  # -1: not replayed yet
                raise RuntimeError(f"FusedNonbondedPME: a replay of a captured call: {status_message(flags)}")


class _FusedEnergyForces(torch.autograd.Function):
    """(E, F, terms) with a hand-written vector-Jacobian product (see the module docstring)."""

    @staticmethod
    def forward(ctx, module, coords, charges, sigma, epsilon, box):
        E, F, terms = module._forward_raw(coords, box, charges, sigma, epsilon, for_grad=True)
        ctx.token = module._last_token
        ctx.module = module
        ctx.save_for_backward(coords, charges, sigma, epsilon, box)
        # E and terms are views of the call's 5-value output (no copies)
        ctx.mark_non_differentiable(terms)
        return E, F, terms

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_energy, grad_forces, grad_terms):
        coords, charges, sigma, epsilon, box = ctx.saved_tensors
        needs = ctx.needs_input_grad[1:5]
        if not any(needs):
            return None, None, None, None, None, None
        d = ctx.module._backward_raw(coords, box, charges, sigma, epsilon, grad_energy, grad_forces, needs,
                                     ctx.token)
        d = [None if x is None else x.to(t.dtype) for x, t in zip(d, (coords, charges, sigma, epsilon))]
        return (None, *d, None)


def for_workload(w, precision="mixed"):
    """A FusedNonbondedPME configured for a water workload dict (see torchff.workloads)."""
    return FusedNonbondedPME(w["natoms"], w["cutoff"], w["alpha"], w["K"], order=w["order"],
                             skin=w.get("skin", 0.0), exclusions=w["exclusions"], precision=precision)


def ewald_alpha_from_tolerance(cutoff, tol=5e-4):
    return math.sqrt(-math.log(2.0 * tol)) / cutoff
