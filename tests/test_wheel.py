"""The Linux wheel (setup.py): built from the in-tree libraries, installed
into an empty directory, and imported from there in a fresh interpreter the
way the reference's wheel is used (`import NexusMD`), with the C-ABI library
of the installed package loadable.  CPU only: nothing touches a GPU."""
import glob
import os
import subprocess
import sys
import zipfile

import pytest

from nexusmd_amd import build as B

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = [os.path.join(REPO, "nexusmd_amd", f) for f in ("libnexusmd_hip.so", "libnexusmd_host.so")]


@pytest.mark.skipif(not all(os.path.exists(p) for p in LIBS + [B.module_path()]),
                    reason="the libraries are not built (python -m nexusmd_amd.build)")
def test_wheel_installs_and_imports(tmp_path):
    out, work = tmp_path / "dist", tmp_path / "work"
    work.mkdir()
    # every intermediate file under tmp_path: nothing is written into the tree
    r = subprocess.run([sys.executable, "setup.py", "-q", "egg_info", "--egg-base", str(work), "build",
                        "--build-base", str(work / "build"), "bdist_wheel", "-d", str(out), "-b",
                        str(work / "bdist")], cwd=REPO, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    wheels = glob.glob(str(out / "nexusmd_amd-*-linux_x86_64.whl"))
    assert len(wheels) == 1
    names = zipfile.ZipFile(wheels[0]).namelist()
    ext = os.path.basename(B.module_path())
    assert any(n.endswith("/" + ext) and n.count("/") == 2 for n in names)          # NexusMD at the top level
    for lib in ("nexusmd_amd/libnexusmd_hip.so", "nexusmd_amd/libnexusmd_host.so", "nexusmd_amd/engine.py"):
        assert any(n.endswith(lib) for n in names), lib
    assert not any("/oracle/" in n or "/tests/" in n for n in names)                # no test infrastructure
    target = tmp_path / "site"
    r = subprocess.run([sys.executable, "-m", "pip", "install", "-q", "--no-deps", "--no-index", "--target",
                        str(target), wheels[0]], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    probe = ("import NexusMD, nexusmd_amd; from nexusmd_amd import _lib as L; "
             "assert NexusMD.__file__.startswith(%r); assert L.lib()._name.startswith(%r); "
             "d = NexusMD.TaskDispatcher.CreateDispatcher('GPU_MIXED'); d.EnableNonbondedForce(); "
             "print(NexusMD.Coords3D(1, 2, 3))" % (str(target), str(target)))
    env = dict(os.environ, PYTHONPATH=str(target))
    r = subprocess.run([sys.executable, "-c", probe], cwd=str(tmp_path), env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "Coords3D(1" in r.stdout
