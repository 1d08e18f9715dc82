"""The Fortran side of the boundary: cosa_amd/fortran/binding_demo (built by
__graft_entry__.build() with amdflang against libcosahip.so through the
iso_c_binding module cosahip_binding) drives one smooth() on the GPU and
checks that a uniform freestream is left unchanged."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cosa_amd", "fortran", "binding_demo")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(EXE), reason="binding_demo not built (needs amdflang)")
def test_fortran_driver_through_iso_c_binding():
    print("running", EXE, flush=True)
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "binding_demo: ok" in out.stdout
