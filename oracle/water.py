"""Re-export of the seeded water-box generator used as input by every check (test infrastructure only)."""

from torchff.workloads import (  # noqa: F401
    COULOMB_KJMOL, WORKLOADS, box_length_for_density, ewald_alpha, make_workload, water_box,
)
