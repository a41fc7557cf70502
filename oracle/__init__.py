"""CPU FP64 oracle for the nonbonded hot path of THGLab/torchff-lib.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain PyTorch (CPU,
float64) and numpy, the algorithms that the reference's pure-PyTorch path
(``use_customized_ops=False``) and its CUDA kernels compute for

* the periodic neighbor list            (torchff/nblist.py, csrc/nblist/*),
* real-space erfc-Coulomb + Lennard-Jones with exclusions
                                         (torchff/multipoles.py rank 0,
                                          torchff/coulomb.py, torchff/vdw.py,
                                          csrc/nonbonded/nonbonded_atom_pairs_cuda.cu),
* the smooth-PME reciprocal pipeline    (torchff/pme_ref.py, torchff/bsplines.py,
                                          csrc/pme/*),

each function citing the reference file:line it follows.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
it, and only as the checker.  The shipped library never calls into it.

How each part is pinned (beyond restating the reference) is recorded in the
docstring of the function concerned; see ``tests/test_oracle_*.py``.
"""

from . import pbc, bspline, nblist, realspace, pme, ewald, water, composite  # noqa: F401
