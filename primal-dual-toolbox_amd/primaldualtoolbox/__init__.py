"""MI355X TV/TGV primal-dual toolbox: ``primaldualtoolbox.denoising`` and
``primaldualtoolbox.mri`` as in the reference package
(primaldualtoolbox/__init__.py:21-22), backed by libpdt_hip.so."""
from . import denoising, mri  # noqa: F401
