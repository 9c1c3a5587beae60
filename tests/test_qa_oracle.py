"""Pins the template-QA restatement (oracle/qa_ref.py) by a planted answer: a
test template that is exactly the sum of two shifted templates must come back
with that pair, those shifts and a zero residual."""
import numpy as np

from oracle import qa_ref


def test_planted_sum_is_found():
    rng = np.random.default_rng(1)
    C, L = 2, 10
    # integer-valued templates: every sum below is exact in float32 and FP64
    T = rng.integers(-6, 7, size=(4, C * L)).astype(np.float32)
    T[3] = (qa_ref._shift(T[0], 2, C, L) + qa_ref._shift(T[1], -1, C, L)).astype(np.float32)
    spikes = np.array([100, 100, 1, 10])
    out = qa_ref.find_overlap_templates(T, L, C, 3, 3, spikes)
    got = {r[0]: r for r in out}
    assert 3 in got and got[3][2] == [0, 1]
    assert np.array_equal(got[3][1], qa_ref._shift(T[0], 2, C, L) + qa_ref._shift(T[1], -1, C, L))
    assert np.sum((T[3] - got[3][1]) ** 2) == 0.0


def test_shift_zero_fills_and_pair_confusion_is_symmetric():
    t = np.arange(1, 7, dtype=np.float32)  # C = 2, L = 3
    assert list(qa_ref._shift(t, 1, 2, 3)) == [0, 1, 2, 0, 4, 5]
    assert list(qa_ref._shift(t, -2, 2, 3)) == [3, 0, 0, 6, 0, 0]
    rng = np.random.default_rng(2)
    covs = [np.eye(3), 2 * np.eye(3)]
    a, b = rng.normal(size=6), rng.normal(size=6)
    assert abs(qa_ref.check_template_pair(a, b, covs, 2, 3) - qa_ref.check_template_pair(b, a, covs, 2, 3)) < 1e-12
