"""sparse-voxel-octrees_amd: MI355X-native sparse voxel octree raymarch and construction.

The product is libsvo_hip.so (C ABI in include/svo.h); `svo` is its Python
binding and `scenes` generates seeded synthetic inputs.
"""
from . import scenes  # noqa: F401
from . import svo  # noqa: F401
from .svo import VoxelOctree, SvoError  # noqa: F401
