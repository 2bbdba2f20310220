"""The Fortran ISO_C_BINDING host (fortran/pnamp_driver, the PNAMP_Main pair
loop on the library) writes post_general .elem files equal, to their 8
printed digits, to the kernels the Python front end gets from the same
library; plain and factored forms."""
import os
import subprocess

import numpy as np
import pytest

from sccm_amd import configs
from sccm_amd.driver import chop_kernels, read_elem, write_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortran", "pnamp_driver")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["plain", "factored"])
def test_fortran_host_matches_library(tmp_path, form):
    if not os.path.exists(EXE):
        pytest.skip("fortran/pnamp_driver not built (amdflang absent)")
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN, PNAMP
    w = configs.pf_gcm(nstates=3, nalp=2, nbet=3, nphi=3, jmax=4)
    case = str(tmp_path / "case.bin")
    out = tmp_path / "out"
    out.mkdir()
    write_case(case, w.basis, w.inter, w.mesh, w.U, w.V, w.pairs, w.coords, w.zval, w.nval)
    r = subprocess.run([EXE, case, str(out), form], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ("factored form" in r.stdout) == (form == "factored")
    ctx = PNAMP(w.basis, w.inter, w.mesh)
    ctx.set_states(w.U, w.V)
    ctx.set_form(FORM_FACTORED if form == "factored" else FORM_PLAIN)
    kern = ctx.kernels_batched(w.pairs, w.zval, w.nval)
    ctx.close()
    jm = w.mesh.jmax
    for i, (q0, q1) in enumerate(w.pairs):
        e = read_elem(str(out / f"kern.{q0:04d}.{q1:04d}.elem"))
        ch = chop_kernels(kern[i])
        assert set(e) == {J for J in range(jm + 1) if J != 1}
        scale = max(np.abs(ch["njkk"]).max(), 1e-300)
        hscale = max(np.abs(ch["hjkk"]).max(), 1e-300)
        for J, blk in e.items():
            for (k1, k2), (n, h, nn, zz) in blk.items():
                a, b = k1 + jm, k2 + jm
                assert abs(n - ch["njkk"][J, a, b]) <= 1e-7 * abs(ch["njkk"][J, a, b]) + 1e-9 * scale
                assert abs(h - ch["hjkk"][J, a, b]) <= 1e-7 * abs(ch["hjkk"][J, a, b]) + 1e-9 * hscale
                if ch["njkk"][J, a, b] != 0:
                    assert abs(nn - ch["nn"][J, a, b]) <= 1e-6 * max(1.0, abs(ch["nn"][J, a, b]))
                    assert abs(zz - ch["zz"][J, a, b]) <= 1e-6 * max(1.0, abs(ch["zz"][J, a, b]))
