"""MI355X-native multigrid Monte Carlo (MGMC) sampling for Gaussian posteriors
of shifted-Laplace priors, built on hand-written gfx950 HIP kernels.

The compute path is libmgmc_amd.so (C ABI: include/mgmc_amd.h); this package
is the host-side mirror of the reference's operator and sampler interfaces.
"""
from ._lib import MGMCError, lib  # noqa: F401
from .operators import (Lattice, Lattice2d, Lattice3d, MeasuredOperator, ShiftedLaplaceFDOperator,  # noqa: F401
                        ShiftedLaplaceFEMOperator, measurement_vector, seeded_ball_measurements)
from .sampler import (LoopSolver, MultigridMCSampler, MultigridParameters, MultigridPreconditioner,  # noqa: F401
                      SSORSampler)
from .stats import tau_int_chains, tau_int_sokal, tau_int_window  # noqa: F401
from . import measurements  # noqa: F401,E402
