"""The product's synthetic-input generators agree with the oracle's closed-form copies and the
batched-workload draw is deterministic."""
import numpy as np

import refcpu as R
import workloads as W


def test_vortex_and_noise_match_oracle_generators():
    rng = np.random.default_rng(0)
    xy = rng.uniform(0, 10, size=(50, 7, 2))
    pert = W.density_noise(11, 50, 7)
    assert np.array_equal(pert[:, 0], 0.01 * R.uniform_pm1(11, np.arange(50)))
    for t in (0.0, 3.7, 12.0):
        assert np.allclose(W.vortex(xy, t=t, rho_pert=pert), R.vortex_state(xy, t=t, rho_pert=pert),
                           rtol=1e-14, atol=1e-14)
    assert np.allclose(W.taylor_green(xy), R.taylor_green_state(xy), rtol=1e-14)


def test_vortex_closed_form_properties():
    # far from the centre the state is the free stream; at the centre the velocity is u_inf
    U = W.vortex(np.array([[5.0 + 9.0, 5.0], [5.0, 5.0]]), length=100.0)
    assert np.allclose(U[0], [1.0, 1.0, 0.0, 1 / 0.4 + 0.5], atol=1e-12)
    assert U[1, 0] < 1.0 and abs(U[1, 1] / U[1, 0] - 1.0) < 1e-15 and abs(U[1, 2]) < 1e-15


def test_batch_workload_generator_is_deterministic():
    a = W.batch_instances(64, seed=5)
    assert a == W.batch_instances(64, seed=5)
    assert {x["n"] for x in a} <= {16, 32, 64, 128} and {x["p"] for x in a} <= {1, 2, 3, 4}
    assert len({x["n"] for x in a}) == 4 and len({x["p"] for x in a}) == 4
    assert all(3 <= x["beta"] <= 6 and 3 <= x["xc"] <= 7 and 3 <= x["yc"] <= 7 for x in a)
