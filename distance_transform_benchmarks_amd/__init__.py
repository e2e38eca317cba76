"""MI355X-native generalised geodesic distance transform (GGDT).

A from-scratch HIP/gfx950 implementation of the GGDT of
baptisteesteban/distance_transform_benchmarks, behind the same API:

  - C ABI:      include/dtk.h            (lib/libdtk.so)
  - C++ API:    include/dt/*.hpp         (lib/libdt.so, namespace dt)
  - Python API: this package             (torch device tensors)

The Python names mirror the reference's C++ functions; see ``api``.
"""
from .api import (  # noqa: F401
    BatchPlan,
    autotune,
    backward,
    backward_batched,
    predecessors_batched,
    block_priorities,
    colorize,
    generalised_distance_transform,
    generalised_distance_transform_blocks,
    generalised_distance_transform_task,
    ggdt,
    ggdt_batched,
    heatmap,
    inferno,
    minmax,
    normalize_u8,
    predecessors,
    round_per_blocks_buffer,
    selftest,
    synchronize,
    tile_count,
    tile_grid,
    tile_image,
    tile_size,
    visits_buffer,
)
from .autograd import GGDTFunction, ggdt_differentiable  # noqa: F401
from .imageio import image_info, imread, imsave  # noqa: F401

__version__ = "0.1.0"
