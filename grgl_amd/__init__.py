"""grgl_amd — MI355X-native GRG graph matrix multiplication.

A drop-in for the matmul path of GRGL (pygrgl.matmul / pygrgl.dot_product /
GRG::matrixMultiplication, reference src/python/_grgl.cpp:100-241 and
include/grgl/grg.h:1342-1517): the graph is uploaded once to the GPU as level-bucketed
gather lists, and every call runs hand-written gfx950 HIP kernels through the C ABI of
include/grgl_hip.h.  There is no CPU fallback: if the native extension is missing the
import fails, and if no GPU is usable matmul raises RuntimeError.

    import grgl_amd as pygrgl
    g = pygrgl.grg_from_coalescent(1000, 10000)
    counts = pygrgl.matmul(g, np.ones((1, g.num_samples), np.int32), pygrgl.TraversalDirection.UP)
"""

import os as _os

from ._runtime import share_hip_runtime_with_torch as _share

_share()

try:
    from ._grgl_amd import (  # noqa: F401
        COAL_COUNT_NOT_SET,
        GRG,
        INVALID_NODE,
        NO_UP_EDGES,
        MutableGRG,
        Mutation,
        MutationMappingStats,
        TraversalDirection,
        device_count,
        GraphBuildStats,
        dot_product,
        fast_grg_from_samples,
        get_device,
        grg_from_arrays,
        grg_from_coalescent,
        load_immutable_grg,
        map_mutations,
        matmul,
        save_grg,
        set_device,
    )
except ImportError as exc:  # pragma: no cover - exercised only when the build is missing
    raise ImportError(
        "grgl_amd native extension is not built (run `python grgl_amd/build.py`): " + str(exc)
    ) from exc

DOWN = TraversalDirection.DOWN
UP = TraversalDirection.UP

__all__ = [
    "COAL_COUNT_NOT_SET",
    "GRG",
    "INVALID_NODE",
    "NO_UP_EDGES",
    "MutableGRG",
    "Mutation",
    "MutationMappingStats",
    "TraversalDirection",
    "UP",
    "DOWN",
    "device_count",
    "GraphBuildStats",
    "dot_product",
    "fast_grg_from_samples",
    "get_device",
    "grg_from_arrays",
    "grg_from_coalescent",
    "load_immutable_grg",
    "map_mutations",
    "matmul",
    "save_grg",
    "set_device",
]
