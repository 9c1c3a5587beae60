"""GPU template QA against the literal restatement: the same test units, pairs
and shifted template sums (bit-identical FP64 residuals decide the minima)."""
import numpy as np
import pytest

from oracle import qa_ref
from spikesorting_fullpursuit_amd.c_cython import sort_cython

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,C,L,U,pre,post", [(1, 2, 10, 6, 3, 3), (2, 3, 16, 7, 4, 2), (3, 1, 40, 5, 6, 6),
                                                  (4, 4, 9, 8, 2, 5)])
def test_find_overlap_templates_matches_oracle(seed, C, L, U, pre, post):
    rng = np.random.default_rng(seed)
    T = (rng.normal(size=(U, C * L)) * 3).astype(np.float32)
    # planted sums of shifted templates plus a little noise, and ties from repeated templates
    T[U - 1] = (qa_ref._shift(T[0], 1, C, L) + qa_ref._shift(T[1], -2, C, L) + 0.01 * rng.normal(size=C * L)
                ).astype(np.float32)
    T[U - 2] = T[2]
    spikes = rng.integers(1, 200, size=U)
    spikes[:3] = 1000
    print(f"[qa] seed={seed} C={C} L={L} U={U}", flush=True)
    got = sort_cython.find_overlap_templates(T, L, C, pre, post, spikes)
    want = qa_ref.find_overlap_templates(T, L, C, pre, post, spikes)
    assert len(got) == len(want) and len(want) > 0
    for g, w in zip(got, want):
        assert g[0] == w[0] and g[2] == w[2]
        assert np.array_equal(g[1], w[1])


def test_fewer_than_three_templates():
    assert sort_cython.find_overlap_templates(np.ones((2, 8), np.float32), 4, 2, 1, 1, np.array([5, 5])) == []


def test_check_template_pair_matches_oracle():
    from spikesorting_fullpursuit_amd.overlap import neuron_separability
    rng = np.random.default_rng(7)
    C, L = 3, 12
    A = rng.normal(size=(C, L, L))
    covs = [a @ a.T / L + np.eye(L) for a in A]
    for _ in range(5):
        t1 = rng.normal(size=C * L) * 2
        t2 = t1 + rng.normal(size=C * L) * 0.5
        p = neuron_separability.check_template_pair(t1, t2, covs, C, L)
        q = qa_ref.check_template_pair(t1, t2, covs, C, L)
        assert abs(p - q) <= 1e-12
    assert neuron_separability.check_template_pair(t1, t1, covs, C, L) == 1.0
