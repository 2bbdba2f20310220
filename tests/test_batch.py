"""Uneven GCM batch across model spaces: the seeded instance set, and on the
GPU a small batch whose factored kernels equal the plain FP64 ones."""
import numpy as np
import pytest

from sccm_amd import batch


def test_instances_are_seeded_and_in_range():
    a, b = batch.draw_instances(), batch.draw_instances()
    assert a == b and len(a) == 64
    assert {i.space for i in a} == set(batch.SPACES)
    for i in a:
        assert i.zval in (2, 4, 6) and i.nval in (2, 4, 6) and 5 <= i.nconf <= 30


@pytest.mark.gpu
def test_small_uneven_batch_matches_plain():
    r = batch.run(n=4, nalp=2, nbet=3, nphi=3, jmax=4, plain_pairs=3, log=lambda *a: None)
    assert r["max_rel_kernel_diff_vs_plain_on_sampled_pairs"] < 1e-10
    assert r["mesh_points"] > 0


@pytest.mark.gpu
def test_per_pair_valence_numbers_equal_per_instance_calls():
    """States of two instances with different Z, N set together: one call with
    per-pair valence numbers gives the kernels of the per-instance calls."""
    from sccm_amd import configs
    from sccm_amd.basis import make_basis
    from sccm_amd.interaction import synthetic_interaction
    from sccm_amd.mesh import make_mesh
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN, PNAMP
    b = make_basis("sd")
    mesh = make_mesh(2, 3, 2, 3, 4)
    g = PNAMP(b, synthetic_interaction(b, 1.5, *batch.SEEDS["sd"]), mesh)
    Ua, Va = configs.make_states(b, 2, 4, configs.gcm_grid(3), seed=11)
    Ub, Vb = configs.make_states(b, 4, 2, configs.gcm_grid(2), seed=12)
    pa = [(0, 1), (1, 2), (2, 2)]
    pb = [(0, 1), (1, 1)]
    for form in (FORM_PLAIN, FORM_FACTORED):
        g.set_form(form)
        g.set_states(Ua, Va)
        ka = g.kernels_batched(pa, 2, 4)
        g.set_states(Ub, Vb)
        kb = g.kernels_batched(pb, 4, 2)
        g.set_states(np.concatenate([Ua, Ub]), np.concatenate([Va, Vb]))
        pairs = pa + [(p + 3, q + 3) for p, q in pb]
        k = g.kernels_batched_zn(pairs, [2, 2, 2, 4, 4], [4, 4, 4, 2, 2])
        assert np.abs(k - np.concatenate([ka, kb])).max() <= 1e-13 * np.abs(k).max(), form
    g.close()
