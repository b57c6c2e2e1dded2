"""MI355X TV/TGV primal-dual toolbox, backed by libpdt_hip.so.

As in the reference package (primaldualtoolbox/__init__.py:21-22) the two
binding modules are ``pymrireconstruction`` and ``pydenoising``, exposed as
``primaldualtoolbox.mri`` and ``primaldualtoolbox.denoising``.
"""
from . import pymrireconstruction as mri  # noqa: F401
from . import pydenoising as denoising  # noqa: F401
