"""mbirtorch_amd: the projector / VCD hot path of model-based iterative CT
reconstruction, built for AMD MI355X (gfx950) with hand-written HIP kernels.

It keeps the reference mbirtorch's ParallelBeamModel / ConeBeamModel API
(forward_project, back_project, recon_direct, recon, sparse_*,
compute_hessian_diagonal, set_params / get_params) so code written against
that API runs unchanged on one MI355X.
"""

__version__ = '0.1.0'

from ._lib import load as load_library  # noqa: F401
from .autograd import (TorchProjector, back_project_differentiable,  # noqa: F401
                       forward_project_differentiable)
from .cone_beam import ConeBeamModel, recon_simple_cone  # noqa: F401
from .parallel_beam import ParallelBeamModel, recon_simple_parallel  # noqa: F401
from .tomography_model import TomographyModel, dot_f64  # noqa: F401
from .vcd_utils import (gen_full_indices, gen_partition_sequence,  # noqa: F401
                        gen_pixel_partition, gen_set_of_pixel_partitions, gen_weights,
                        get_2d_ror_mask)
