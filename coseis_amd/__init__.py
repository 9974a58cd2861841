"""
MI355X-native SORD (Support Operator Rupture Dynamics) solver step.

``coseis_amd.sord`` mirrors the reference's ``cst.sord`` interface; the
solver itself is a HIP shared library (``coseis_amd/libsord_amd.so``) driven
through the C ABI declared in ``include/sord_amd.h``.
"""
from . import sord  # noqa: F401
