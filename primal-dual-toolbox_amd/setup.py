"""Installs the ``primaldualtoolbox`` package under the reference's
distribution name (setup.py there, which packages its compiled modules), so
that ``import primaldualtoolbox`` and its ``pymrireconstruction`` /
``pydenoising`` modules work without a path edit.

From a checkout of this repository (the library's build reads the C ABI
header in ``include/`` at the repository root):

    pip install ./primal-dual-toolbox_amd        # or: pip install -e ...
    pip wheel --no-deps ./primal-dual-toolbox_amd

The HIP library (``libpdt_hip.so``, gfx950) is built with ``make -C csrc`` when
it is not there yet and shipped as package data, together with the trained
warm-start network in ``primaldualtoolbox/data``.
"""
import os
import subprocess

from setuptools import Distribution, setup
from setuptools.command.build_py import build_py

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "primaldualtoolbox", "libpdt_hip.so")


class BuildWithLibrary(build_py):
    """Builds libpdt_hip.so (hipcc --offload-arch=gfx950) before the package
    files are collected, if no build is present."""

    def run(self):
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-C", os.path.join(HERE, "csrc")])
        super().run()


class PlatformDistribution(Distribution):
    """The package carries a native library, so its wheels are platform wheels."""

    def has_ext_modules(self):
        return True


setup(
    name="PrimalDualToolbox",
    version="2.0.0",
    description="TV / TGV primal-dual solvers and Cartesian MRI reconstruction on AMD Instinct MI355X (HIP, gfx950).",
    packages=["primaldualtoolbox"],
    package_data={"primaldualtoolbox": ["libpdt_hip.so", "data/*.npz", "data/*.json"]},
    python_requires=">=3.8",
    install_requires=["numpy"],
    cmdclass={"build_py": BuildWithLibrary},
    distclass=PlatformDistribution,
    zip_safe=False,
)
