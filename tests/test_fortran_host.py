"""The Fortran ISO_C_BINDING host (fortran/pnamp_driver, the PNAMP_Main pair
loop on the library) writes post_general .elem files (Codes/PNAMP/PNAMP_IO.f90:39-94)
whose every line follows the record's format statements ((10i5), (8e15.8);
sccm_amd/driver.py:read_elem rejects any other layout) and whose numbers
equal, to their 8 printed digits, the kernels the Python front end gets from
the same library: plain and factored forms, and the E2 fields (qpred, qcpred
for dJ = 0, 1, 2 with effective charges) when the case carries Q2 matrices."""
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
        for J, rec in e.items():
            assert rec["coords"] == pytest.approx((*w.coords[q0], *w.coords[q1]), rel=1e-7)
            for (k1, k2), (n, h, nn, zz, j2, *_q) in rec["k"].items():
                a, b = k1 + jm, k2 + jm
                assert abs(n - ch["njkk"][J, a, b]) <= 1e-7 * abs(ch["njkk"][J, a, b]) + 1e-9 * scale
                assert abs(h - ch["hjkk"][J, a, b]) <= 1e-7 * abs(ch["hjkk"][J, a, b]) + 1e-9 * hscale
                if ch["njkk"][J, a, b] != 0:
                    assert abs(nn - ch["nn"][J, a, b]) <= 1e-6 * max(1.0, abs(ch["nn"][J, a, b]))
                    assert abs(zz - ch["zz"][J, a, b]) <= 1e-6 * max(1.0, abs(ch["zz"][J, a, b]))


def test_fortran_host_writes_e2_fields(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("fortran/pnamp_driver not built (amdflang absent)")
    from sccm_amd.operators import q2_matrices
    from sccm_amd.pnamp import PNAMP
    w = configs.pf_gcm(nstates=3, nalp=2, nbet=3, nphi=3, jmax=4)
    q2 = q2_matrices(w.basis, b_osc=2.0)
    ep, en = 1.5, 0.5
    case = str(tmp_path / "case.bin")
    out = tmp_path / "out"
    out.mkdir()
    write_case(case, w.basis, w.inter, w.mesh, w.U, w.V, w.pairs, w.coords, w.zval, w.nval, q2=q2, echarge=(ep, en))
    r = subprocess.run([EXE, case, str(out), "factored"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ctx = PNAMP(w.basis, w.inter, w.mesh)
    ctx.set_states(w.U, w.V)
    ctx.set_q2(q2)
    _, kern, _, qred, _ = ctx.kernels_e2(w.pairs, w.zval, w.nval)
    ctx.close()
    jm = w.mesh.jmax

    def qp(i, J, k1, k2, dJ):
        if J < 0 or J > jm or J + dJ < 0 or abs(k1) > J + dJ or abs(k2) > J or abs(dJ) > 2:
            return 0.0
        return ep * qred[i, 0, dJ + 2, J, k1 + jm + 2, k2 + jm] + en * qred[i, 1, dJ + 2, J, k1 + jm + 2, k2 + jm]

    def qcp(i, J, k1, k2, dJ):
        return 0.0 if J + dJ > jm else (-1) ** abs(dJ) * np.conj(qp(i, J + dJ, k2, k1, -dJ))

    scale = np.abs(qred).max() * (ep + en)
    nonzero = 0
    for i, (q0, q1) in enumerate(w.pairs):
        e = read_elem(str(out / f"kern.{q0:04d}.{q1:04d}.elem"))
        for J, rec in e.items():
            fields = [(0, rec["k"], 5), (1, rec["dJ1"], 0), (2, rec["dJ2"], 0)]
            for dJ, blk, off in fields:
                for (k1, k2), vals in blk.items():
                    a, b = vals[off], vals[off + 1]
                    ra, rb = qp(i, J, k1, k2, dJ), qcp(i, J, k1, k2, dJ)
                    assert abs(a - ra) <= 1e-7 * abs(ra) + 1e-9 * scale, (q0, q1, J, dJ, k1, k2)
                    assert abs(b - rb) <= 1e-7 * abs(rb) + 1e-9 * scale, (q0, q1, J, dJ, k1, k2)
                    nonzero += abs(ra) > 1e-6 * scale
            for vals in rec["k"].values():
                assert vals[7] == 0 and vals[8] == 0  # q0p, q0pc: not set by the reference's kernel text
    assert nonzero > 500
