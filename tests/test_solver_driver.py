"""The solve() driver of the Python mirror (dgx.CompressibleDGSolver) against a restatement of
the reference's run_time_integration loop (S/solver/compressible_solver_base.cpp:242-306):
step count from `while (t < T_final) { ...; t += dt; }`, frames at step 0 and at the top of
each iteration whose step is a positive multiple of save_every, the NaN check after every
step (which ends the loop), and a final frame when the step count is not a multiple of
save_every.  The context is a stand-in whose state records the step it belongs to, so these
run without a GPU; the GPU run of the same driver is in tests/test_gpu_nan_abort.py."""
import numpy as np
import pytest

import dgx


def reference_loop(T_final, dt, save_every, nan_step=None):
    """(steps of the frames, steps taken) as the reference's loop produces them."""
    frames = [0]
    t, step = 0.0, 0
    while t < T_final:
        if step % save_every == 0 and step > 0:
            frames.append(step)
        step += 1  # time_step_ssp_rk3
        t += dt
        if nan_step is not None and step >= nan_step:  # a NaN persists once it appears
            break
    if step % save_every != 0:
        frames.append(step)
    return frames, step


class FakeContext:
    """Holds `step` as its state; a NaN appears at step nan_step and stays."""
    K, Np = 2, 3

    def __init__(self, nan_step=None):
        self.step = 0
        self.nan_step = nan_step
        self.calls = []

    def volume_points(self):
        return np.zeros((self.K, 3, 2))

    def project(self, Uq):
        self.step = 0

    def get_state(self):
        return np.full((self.K, self.Np, 4), float(self.step))

    def rk3_run(self, dt, n):
        self.calls.append(n)
        for k in range(1, n + 1):
            self.step += 1
            if self.nan_step is not None and self.step >= self.nan_step:
                return k, True
        return n, False


CASES = [(1.0, 0.1, 5, None), (1.0, 0.1, 11, None), (1.0, 0.1, 3, None), (0.5, 0.1, 5, None),
         (5e-4, 1e-5, 50, None), (2.0, 0.25, 4, None), (1.0, 0.1, 5, 7), (1.0, 0.1, 5, 5),
         (1.0, 0.1, 5, 10), (1.0, 0.1, 5, 1), (1.0, 0.1, 5, 11), (0.3, 0.1, 1, 2)]


@pytest.mark.parametrize("T_final,dt,save_every,nan_step", CASES)
def test_solve_follows_the_reference_loop(T_final, dt, save_every, nan_step):
    ctx = FakeContext(nan_step)
    solver = dgx.CompressibleDGSolver(ctx)
    frames = solver.solve(lambda xy: np.zeros(xy.shape[:-1] + (4,)), T_final, dt, save_every)
    want, steps = reference_loop(T_final, dt, save_every, nan_step)
    assert [int(f[0, 0]) for f in frames] == want
    assert solver.steps_taken == steps
    assert all(f.shape == (ctx.K, 4 * ctx.Np) for f in frames)
    # the device runs whole chunks between frames, not one step per call
    assert len(ctx.calls) <= steps // save_every + 1


def test_reference_step_count_accumulates_like_the_reference():
    assert dgx.reference_step_count(1.0, 0.1) == 11    # 0.1 * 10 accumulates to 0.9999999999999999
    assert dgx.reference_step_count(0.5, 0.1) == 5
    assert dgx.reference_step_count(5e-4, 1e-5) == 50
    assert dgx.reference_step_count(0.0, 0.1) == 0
    with pytest.raises(ValueError):
        dgx.reference_step_count(1.0, 0.0)


def test_frame_rule_drops_the_aligned_final_state():
    # n_steps a multiple of save_every: the final state is not a frame (the reference adds a
    # frame at the top of an iteration, and after the loop only for an unaligned count)
    assert reference_loop(0.5, 0.1, 5) == ([0], 5)
    assert reference_loop(1.0, 0.1, 5) == ([0, 5, 10, 11], 11)
