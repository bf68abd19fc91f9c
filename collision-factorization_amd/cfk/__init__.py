"""cfk -- MI355X (gfx950) Wigner-Eckart-factorized collision operator.

Python host mirror of the reference's MEX interface over the C-ABI library libcfk.so
(include/cfk.h).  Every compute call runs on the GPU through that library; there is no CPU
fallback.
"""
from ._lib import CfkError, LIB_PATH, lib  # noqa: F401
from .basis import MOMENT_NAMES, basis_evaluate, moment_weights, moments  # noqa: F401
from .geometry import (build_rtensor, gauss_gen_laguerre, gauss_legendre, real_gaunt,  # noqa: F401
                       set_sumfac_chunk_nodes, set_sumfac_kernel, sort_t_q1, sumfac_grid_sizes,
                       sumfac_last_chunks, sumfac_last_timing)
from .matfile import load_operator, load_precalc, read_mat_dataset, save_mat, save_operator  # noqa: F401
from .mex import (MexError, angular_first_collision_kernel_mex, compute_rtensor_sumfac_mex,  # noqa: F401
                  dense_tensor_kernel_mex, naive_collision_kernel_mex, radial_first_collision_kernel_mex)
from .plan import CollisionPlan, MultiPlan, conservation_project, dense_assemble, dense_execute  # noqa: F401
from .surrogate import (Mlp, Surrogate, make_mlp, maxwellian_mixture_params, project_maxwellians,  # noqa: F401
                        split_params)
