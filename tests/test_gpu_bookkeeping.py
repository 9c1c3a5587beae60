"""GPU window bookkeeping: the HIP path's per-pass trace (spikes added,
windows listed, windows flagged) through the C ABI (fbp_get_pass_log).

  * against the hand-derived known answers of tests/test_oracle_bookkeeping.py
    (every screen path and both likelihood precisions);
  * against the oracle's pass log on seeded noisy cases and workload slices,
    where the trace is the work the throughput figure counts."""
import numpy as np
import pytest

from oracle import fbp_oracle
from spikesorting_fullpursuit_amd import _native, synthetic
from tests import cases
from tests.test_gpu_parity import MODES, gpu_run
from tests.test_oracle_bookkeeping import CASES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def handle():
    h = _native.Handle(0)
    yield h
    h.close()


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("literal", MODES)
@pytest.mark.parametrize("name", sorted(CASES))
def test_known_answer_traces(handle, name, literal, mixed):
    c = CASES[name]()
    idx, lab, R = gpu_run(handle, c["V"], c["T"], c["bias"], c["thr"], c["n"], literal=literal, mixed=mixed)
    log = handle.pass_log()
    assert [tuple(r) for r in log.tolist()] == c["trace"]
    assert list(zip(idx.tolist(), lab.tolist())) == c["spikes"]
    np.testing.assert_array_equal(R, c["residual"])


@pytest.mark.parametrize("literal", ["literal", "split"])
@pytest.mark.parametrize("seed", range(6))
def test_pass_log_matches_oracle_noisy(handle, seed, literal):
    V, T, bias, thr, on, lab = cases.noisy_case(seed + 40, U=4, nsp=40)
    n = T.shape[2] - 1
    gpu_run(handle, V, T, bias, thr, n, literal=literal)
    _, _, log = fbp_oracle.pursuit(V, T, bias, thr, n, return_log=True)
    np.testing.assert_array_equal(handle.pass_log(), log[:, :3])


@pytest.mark.parametrize("mixed", [False, "auto"])
@pytest.mark.parametrize("name,seconds", [("small", 2.0), ("probe16", 0.5), ("probe64", 0.3)])
def test_pass_log_matches_oracle_workload_slice(handle, name, seconds, mixed):
    d = synthetic.generate(name, duration_s=seconds)
    T, bias, thr = synthetic.kernel_inputs(d)
    V, n = d["voltage"], d["n_max_shift_inds"]
    gpu_run(handle, V, T, bias, thr, n, literal="split", mixed=mixed)
    _, _, log = fbp_oracle.pursuit(V, T, bias, thr, n, return_log=True)
    np.testing.assert_array_equal(handle.pass_log(), log[:, :3])
    st = handle.stats()
    assert st["windows_listed"] == int(log[:, 1].sum())
    assert st["windows_flagged"] == int(log[:, 2].sum())
