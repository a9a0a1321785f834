"""The ISO_C_BINDING module a GenASiS host would use compiles against the
C-ABI and runs its host-only checks (no device calls)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang"
                                  if os.path.exists("/opt/rocm/bin/amdflang") else None)


@pytest.mark.skipif(FC is None, reason="no Fortran compiler (amdflang) in this image")
def test_fortran_binding_compiles_links_and_runs(tmp_path):
    src = os.path.join(ROOT, "genasis_amd", "fortran")
    lib = os.path.join(ROOT, "genasis_amd", "lib")
    subprocess.run([FC, "-c", os.path.join(src, "GX_Fluid_Binding.f90")],
                   cwd=tmp_path, check=True)
    subprocess.run([FC, os.path.join(src, "GX_Binding_Check.f90"), "GX_Fluid_Binding.o",
                    "-L" + lib, "-lgx_fluid", "-Wl,-rpath," + lib, "-o", "check"],
                   cwd=tmp_path, check=True)
    r = subprocess.run([str(tmp_path / "check")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "nCells" in r.stdout
