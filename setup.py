"""Linux wheel of the engine: the `nexusmd_amd` package with its in-tree
built libraries (libnexusmd_hip.so for gfx950, libnexusmd_host.so) and the
`NexusMD` Python module at the top level, importable as the reference's
SWIG-built module is (`import NexusMD`; the reference's setup.py packages a
Windows .pyd, setup.py:19-72).

    python -m nexusmd_amd.build          # build the libraries in-tree first
    python setup.py bdist_wheel          # -> dist/nexusmd_amd-*-linux_x86_64.whl

The wheel carries no sources to compile and no test infrastructure (oracle/,
tests/ stay out).
"""
import glob
import os
import shutil
import sysconfig

from setuptools import setup
from setuptools.command.build_py import build_py
from setuptools.dist import Distribution

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(HERE, "nexusmd_amd")
EXT = "NexusMD" + sysconfig.get_config_var("EXT_SUFFIX")
LIBS = ["libnexusmd_hip.so", "libnexusmd_host.so"]


class BinaryDistribution(Distribution):
    """Platform wheel: the package holds compiled libraries."""

    def has_ext_modules(self):
        return True


class BuildPy(build_py):
    """Checks that the libraries are built, and puts the NexusMD module at the
    top level of the wheel (its rpath finds the libraries in nexusmd_amd/)."""

    def run(self):
        missing = [f for f in LIBS + [EXT] if not os.path.exists(os.path.join(PKG, f))]
        if missing:
            raise SystemExit("not built: %s (run `python -m nexusmd_amd.build` first)" % ", ".join(missing))
        super().run()
        shutil.copy2(os.path.join(PKG, EXT), os.path.join(self.build_lib, EXT))


setup(
    name="nexusmd_amd",
    version="0.2.0",
    description="NexusMD force-and-integrate engine for MI355X (gfx950): HIP kernels behind a C-ABI, "
                "C++ dispatcher and the NexusMD Python module",
    packages=["nexusmd_amd"],
    package_data={"nexusmd_amd": LIBS},
    exclude_package_data={"nexusmd_amd": [EXT]},
    data_files=[("include", [os.path.join("include", "nexusmd.h")])],
    python_requires=">=3.8",
    install_requires=["numpy"],
    cmdclass={"build_py": BuildPy},
    distclass=BinaryDistribution,
    zip_safe=False,
)
