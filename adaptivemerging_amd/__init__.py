"""MI355X (gfx950) implementation of AdaptiveMerging's 3D per-step contact path.

The compute path is libam.so (HIP kernels behind the C ABI of include/am.h);
this package holds its ctypes binding, scene builders and a Python mirror of
the reference's RigidBodySystem / CollisionProcessor / PGS call surface.
"""
from ._lib import AMError, Params, default_params_host, load  # noqa: F401
from .scene import Scene, load_sph, random_rotation, rot_axis_angle  # noqa: F401

__all__ = ["AMError", "Params", "Scene", "load", "default_params_host", "load_sph", "random_rotation",
           "rot_axis_angle"]
