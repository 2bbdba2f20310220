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
