"""ctypes binding of the C ABI in include/torchff_hip.h (libtorchff_hip.so, gfx950).

The library is loaded on first use.  There is no fallback: if the shared
object is missing, or a tensor is not on a ROCm device, the call raises.
"""

import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtorchff_hip.so")
OPS_PATH = os.path.join(_HERE, "libtorchff_ops.so")

F32, F64 = 0, 1
NBLIST_NSQUARED, NBLIST_CELL_LIST = 0, 1
SERVE_FULL = 5  # tff_serve_table_add: the system does not fit the capacities
PAIR_NONBONDED, PAIR_EWALD_LJ, PAIR_COUL_EWALD, PAIR_LJ_TYPES, PAIR_LJ_PAIRS, PAIR_COUL_EWALD_SYM = 0, 1, 2, 3, 4, 5

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t

_SIGNATURES = {
    "tff_last_error": (ctypes.c_char_p, []),
    "tff_abi_version": (_i, []),
    "tff_prof_enable": (None, [_i]),
    "tff_prof_collect": (_i, [ctypes.c_char_p, ctypes.POINTER(_d), ctypes.POINTER(_ll)]),
    "tff_fused_set_overlap": (None, [_i]),
    "tff_fused_set_graphs": (None, [_i]),
    "tff_nblist_workspace_bytes": (_sz, [_ll, _vp, _d, _i]),
    "tff_build_neighbor_list": (_i, [_i, _vp, _ll, _vp, _d, _i, _i, _vp, _vp, _vp, _ll, _vp, _vp, _sz, _vp]),
    "tff_pair_workspace_bytes": (_sz, [_i, _ll, _ll, _ll]),
    "tff_pair_energy": (_i, [_i, _i, _vp, _ll, _vp, _vp, _ll, _vp, _ll, _vp, _vp, _vp, _vp, _ll, _d, _d, _d, _i,
                             _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "tff_pme_workspace_bytes": (_sz, [_i, _ll, _i, _i, _i]),
    "tff_pme_energy": (_i, [_i, _vp, _vp, _ll, _vp, _i, _i, _i, _i, _d, _vp, _vp, _vp, _i, _d, _vp, _vp, _vp,
                            _sz, _vp]),
    "tff_pme_gather": (_i, [_i, _vp, _vp, _vp, _ll, _vp, _i, _i, _i, _i, _d, _d, _vp, _vp, _vp]),
    "tff_pme_energy_multipole": (_i, [_i, _vp, _vp, _vp, _vp, _ll, _vp, _i, _i, _i, _i, _d, _vp, _vp, _vp, _i, _d,
                                      _vp, _vp, _vp, _sz, _vp]),
    "tff_pme_gather_multipole": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _ll, _vp, _i, _i, _i, _i, _d, _d, _vp, _vp, _vp,
                                      _vp, _vp]),
    "tff_cluster_pairs_workspace_bytes": (_sz, [_ll, _vp, _d, _d]),
    "tff_cluster_pairs_build": (_i, [_i, _vp, _ll, _vp, _d, _d, _vp, _vp, _vp, _vp, _sz, _vp]),
    "tff_cluster_pairs_emit": (_i, [_i, _vp, _ll, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _ll, _vp, _vp, _ll, _vp, _sz,
                                    _vp]),
    "tff_cluster_pairs_traverse_workspace_bytes": (_sz, [_ll, _ll, _ll]),
    "tff_cluster_pairs_decode": (_i, [_i, _vp, _ll, _vp, _d, _vp, _vp, _vp, _ll, _vp, _vp, _ll, _vp, _ll, _vp, _vp,
                                      _sz, _vp]),
    "tff_cluster_pairs_nonbonded": (_i, [_i, _vp, _ll, _vp, _vp, _vp, _vp, _d, _d, _i, _vp, _vp, _vp, _ll, _vp, _vp,
                                         _ll, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "tff_ewald_workspace_bytes": (_sz, [_i, _ll, _i]),
    "tff_ewald_multipole": (_i, [_i, _vp, _ll, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _sz, _vp]),
    "tff_ewald": (_i, [_i, _vp, _vp, _ll, _vp, _i, _d, _vp, _vp, _vp, _vp, _sz, _vp]),
    "tff_fused_workspace_bytes": (_sz, [_i, _ll, _vp, _d, _i, _i, _i]),
    "tff_fused_energy_forces": (_i, [_i, _i, _vp, _vp, _vp, _vp, _ll, _vp, _vp, _vp, _d, _d, _d, _d, _i, _i, _i, _i,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "tff_fused_energy_forces_noalloc": (_i, [_i, _i, _vp, _vp, _vp, _vp, _ll, _vp, _vp, _vp, _d, _d, _d, _d, _i, _i,
                                             _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tff_fused_backward_workspace_bytes": (_sz, [_i, _ll, _vp, _d, _i, _i, _i]),
    "tff_fused_energy_forces_backward": (_i, [_i, _i, _vp, _vp, _vp, _vp, _ll, _vp, _vp, _vp, _d, _d, _d, _d, _i, _i,
                                              _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                              _vp, _sz, _vp]),
    "tff_fused_energy_forces_backward_reuse": (_i, [_i, _i, _vp, _vp, _vp, _vp, _ll, _vp, _vp, _vp, _d, _d, _d, _d,
                                                    _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                    _vp, _vp, _vp, _vp, _sz, _vp]),
    "tff_serve_table_bytes": (_sz, [_i, _i]),
    "tff_serve_atoms_offset": (_sz, [_i]),
    "tff_serve_table_reset": (None, [_vp]),
    "tff_serve_table_totals": (_i, [_vp, _vp]),
    "tff_serve_table_add": (_i, [_vp, _i, _i, _i, _i, _ll, _i, _i, _vp, _vp, _vp, _d, _d]),
    "tff_serve_workspace_bytes": (_sz, [_i, _i, _i, _i, _ll]),
    "tff_serve_workspace_init": (_i, [_i, _i, _i, _i, _ll, _vp, _sz, _vp]),
    "tff_serve_upload": (_i, [_vp, _i, _i, _i, _i, _ll, _vp, _sz, _vp]),
    "tff_serve_energy_forces": (_i, [_i, _d, _i, _vp, _i, _i, _i, _ll, _vp, _vp, _sz, _vp]),
}

_lib = None


def exported_symbols():
    """Names the C header declares (checked against the shared object by the tests)."""
    return sorted(_SIGNATURES)


def load(require_gpu=True):
    """Load libtorchff_hip.so (raises if it is missing, or if no ROCm device is visible)."""
    global _lib
    if require_gpu and not torch.cuda.is_available():
        raise RuntimeError("torchff (gfx950 build) needs a ROCm GPU; none is visible")
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python torchff-lib_amd/build_lib.py` "
                               "(or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc, what):
    if rc != 0:
        msg = load(require_gpu=False).tff_last_error().decode()
        raise RuntimeError(f"{what}: {msg} (status {rc})")


def dtype_code(t):
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise TypeError(f"torchff: unsupported dtype {t.dtype} (float32 or float64 expected)")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("torchff (gfx950 build): tensors must be on the ROCm device; "
                               f"got one on {t.device}")


_box_cache = True


def set_box_cache(enabled: bool):
    """Switch the host copy cache of device boxes on (default) or off.  Off: every call copies the
    box to the host (one device-to-host synchronisation per call, not graph-capturable)."""
    global _box_cache
    _box_cache = bool(enabled)


def invalidate_box_cache(box):
    """Forget the cached host copy of ``box`` (after writing it in a way autograd does not
    track: through ``.data``, a DLPack or numpy view, or a replayed graph)."""
    try:
        del box._tff_host_box
    except AttributeError:
        pass


def box_host(box):
    """float64 host copy (9,) of a (3,3) box.

    A host (CPU) box is read on every call.  A device box is copied to the host
    once and cached on the tensor object, keyed by its data pointer and its
    in-place version counter, so repeated calls with the same box stay free of
    device-to-host synchronisation (and can be captured into a graph).  Writes
    that bypass the version counter (``.data``, DLPack / numpy views, graph
    replays) are not seen by the cache: call :func:`invalidate_box_cache`, or
    :func:`set_box_cache` (False).  The fused path also compares the device box
    with the cached copy on the device and reports a mismatch (TFF_FUSED_BOX_STALE).
    """
    if box is None:
        return None
    cached = _box_cache and box.is_cuda
    if cached:
        key = (box.data_ptr(), box._version)
        cache = getattr(box, "_tff_host_box", None)
        if cache is not None and cache[0] == key:
            return cache[1]
    b = box.detach().to("cpu", torch.float64).contiguous().reshape(9).numpy().copy()
    arr = (ctypes.c_double * 9)(*b.tolist())
    if cached:
        try:
            box._tff_host_box = (key, arr)
        except (AttributeError, RuntimeError):
            pass
    return arr


def box_numpy(box):
    return np.frombuffer(box_host(box), dtype=np.float64).reshape(3, 3).copy()


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def prof_enable(on=True):
    load().tff_prof_enable(int(bool(on)))


def prof_collect(name=None):
    """(total_ms, launches) of the timed launches named ``name`` (all when None), then forget them."""
    lib = load()
    tot = ctypes.c_double(0.0)
    cnt = ctypes.c_longlong(0)
    check(lib.tff_prof_collect(None if name is None else name.encode(), ctypes.byref(tot), ctypes.byref(cnt)),
          "tff_prof_collect")
    return tot.value, cnt.value
