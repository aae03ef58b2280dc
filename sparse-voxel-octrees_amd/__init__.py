"""sparse-voxel-octrees_amd: MI355X-native sparse voxel octree raymarch and construction."""
from . import scenes  # noqa: F401
