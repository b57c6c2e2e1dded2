"""The package installs as the reference's does: a wheel built from
primal-dual-toolbox_amd/setup.py holds the ``primaldualtoolbox`` package with
the reference's binding module names, the HIP library and the trained
network, so ``import primaldualtoolbox`` needs no path edit once installed
(ref: setup.py, primaldualtoolbox/__init__.py:21-22)."""
import glob
import os
import shutil
import subprocess
import sys
import zipfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "primal-dual-toolbox_amd")


@pytest.mark.skipif(not os.path.exists(os.path.join(PKG, "primaldualtoolbox", "libpdt_hip.so")),
                    reason="libpdt_hip.so is not built")
def test_wheel_holds_package_library_and_network(tmp_path):
    # a copy, so that the build's own folders stay out of the source tree
    src = tmp_path / "src" / "primal-dual-toolbox_amd"
    shutil.copytree(PKG, src, ignore=shutil.ignore_patterns("__pycache__", "_build", "build", "*.egg-info"))
    shutil.copytree(os.path.join(ROOT, "include"), tmp_path / "src" / "include")
    out = tmp_path / "wheels"
    r = subprocess.run([sys.executable, "-m", "pip", "wheel", "--no-deps", "--no-build-isolation", "--no-index",
                        "-w", str(out), str(src)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    wheels = glob.glob(str(out / "PrimalDualToolbox-*.whl"))
    assert len(wheels) == 1 and "linux_x86_64" in wheels[0]  # a platform wheel (native library inside)
    names = {n.split("/", 2)[-1] if n.startswith("PrimalDualToolbox-") and ".data/" in n else n
             for n in zipfile.ZipFile(wheels[0]).namelist()}
    for f in ("primaldualtoolbox/__init__.py", "primaldualtoolbox/pymrireconstruction.py",
              "primaldualtoolbox/pydenoising.py", "primaldualtoolbox/_capi.py", "primaldualtoolbox/libpdt_hip.so",
              "primaldualtoolbox/data/warmstart_tgv_mri_320.npz", "primaldualtoolbox/data/warmstart_tgv_mri_320.json"):
        assert any(n.endswith(f) for n in names), f
