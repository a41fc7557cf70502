"""torchff nonbonded hot path for AMD Instinct MI355X (gfx950).

Drop-in for the ``use_customized_ops=True`` path of THGLab/torchff-lib's
neighbor list, fused nonbonded / rank-0 Ewald real-space / Lennard-Jones pair
terms and smooth PME: the same module classes, functional wrappers and
``torch.ops.torchff.*`` operators, computed by hand-written HIP kernels in
``libtorchff_hip.so`` (C ABI declared in include/torchff_hip.h).
"""

from . import _ops  # noqa: F401  (registers torch.ops.torchff.*)
from ._lib import set_box_cache, invalidate_box_cache  # noqa: F401
from .pbc import PBC, applyPBC  # noqa: F401
from .nblist import (  # noqa: F401
    NeighborList, build_neighbor_list_nsquared, build_neighbor_list_cell_list, build_cluster_pairs,
    decode_cluster_pairs, build_neighbor_list_cluster_pairs,
)
from .nonbonded import (  # noqa: F401
    Nonbonded, compute_nonbonded_energy_from_atom_pairs, compute_nonbonded_forces_from_atom_pairs,
    compute_nonbonded_ewald_energy_from_atom_pairs, compute_nonbonded_energy_from_cluster_pairs,
    compute_nonbonded_forces_from_cluster_pairs,
)
from .multipoles import MultipolarInteraction  # noqa: F401
from .vdw import Vdw, compute_lennard_jones_energy  # noqa: F401
from .pme import PME  # noqa: F401
from .ewald import Ewald, ewald_long_range, ewald_long_range_with_fields  # noqa: F401
from . import bsplines  # noqa: F401
from .md import BatchedSimModel  # noqa: F401
