"""pnamp_kernels_batched_noalloc (plain form, static scratch array, no device
allocation) against pnamp_kernels_batched in the plain form, through the
C-ABI: two calls in a row, the second with fewer pairs, and a pair list long
enough that the first call runs in several chunks of the scratch array."""
import numpy as np
import pytest

from sccm_amd import configs
from sccm_amd.basis import make_basis
from sccm_amd.interaction import synthetic_interaction
from sccm_amd.mesh import make_mesh

pytestmark = pytest.mark.gpu

PNAMP_E_SIZE = -2


def test_noalloc_matches_batched_plain():
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN, PNAMP
    b = make_basis("sd")
    inter = synthetic_interaction(b, 1.5, 1001, 1002)
    # J <= 12: 0.78 MB of projected kernels per pair, so the 105 pairs of 14
    # states take two chunks (85 + 20) of the 64 MiB array
    mesh = make_mesh(2, 2, 2, 2, 12)
    U, V = configs.make_states(b, 2, 2, configs.gcm_grid(14))
    g = PNAMP(b, inter, mesh)
    g.set_states(U, V)
    pairs_all = np.array([(q, q1) for q in range(14) for q1 in range(q, 14)], dtype=np.int32)
    calls = [pairs_all, pairs_all[40:50]]
    for k, pairs in enumerate(calls):
        g.set_form(FORM_PLAIN)
        zr_ref, kn_ref = g.kernels_batched(pairs, 2, 2, want_zres=True)
        # the no-alloc entry point runs the plain form whatever form is set
        g.set_form(FORM_FACTORED)
        zr, kn = g.kernels_batched(pairs, 2, 2, want_zres=True, noalloc=True)
        print(f"call {k}: {len(pairs)} pairs, max |dkern| {np.abs(kn - kn_ref).max():.3e}")
        assert np.array_equal(kn, kn_ref), k
        assert np.array_equal(zr, zr_ref), k
    g.close()


def test_noalloc_rejects_pairs_beyond_scratch():
    from sccm_amd.pnamp import PNAMP, PnampError
    b = make_basis("sd")
    # 100^3 Euler points: ZRES alone is 96 MB per pair, above the 64 MiB array
    g = PNAMP(b, synthetic_interaction(b, 1.5, 1001, 1002), make_mesh(100, 100, 100, 1, 2))
    U, V = configs.make_states(b, 2, 2, configs.gcm_grid(1))
    g.set_states(U, V)
    with pytest.raises(PnampError) as e:
        g.kernels_batched([(0, 0)], 2, 2, noalloc=True)
    assert e.value.code == PNAMP_E_SIZE
    g.close()
