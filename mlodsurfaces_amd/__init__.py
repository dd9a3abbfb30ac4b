"""mlodsurfaces_amd — MI355X-native MLoD / APSS projection (HIP kernels for gfx950 behind a C ABI).

The compute lives in ``libmlod.so`` (mlodsurfaces_amd/csrc, C ABI in include/mlod.h);
``mlodsurfaces_amd.apss`` mirrors the reference's projection interface on top of it.
"""
from .apss import APSS, APSSOctree, Fastapss, GlobalAPSS, Kernel, MlodError  # noqa: F401

__all__ = ["APSS", "APSSOctree", "Fastapss", "GlobalAPSS", "Kernel", "MlodError"]
