"""Launch plan of the headline benchmark's timed region (CPU only)."""

import pytest

import bench


@pytest.mark.parametrize("steps,warmup,chunk", [(20, 5, 200), (4000, 200, 200), (1, 0, 200), (250, 0, 200),
                                                (997, 3, 200), (200, 200, 200), (7, 0, 1)])
def test_step_plan_times_exactly_the_requested_steps(steps, warmup, chunk):
    c, n_full, tail, w = bench.step_plan(steps, warmup, chunk)
    assert 1 <= c <= min(chunk, steps)
    assert n_full >= 1 and 0 <= tail < c
    assert n_full * c + tail == steps
    assert w == warmup


def test_step_plan_driver_invocation():
    # the driver runs `bench.py --gpus 1 --steps 20 --warmup 5`
    assert bench.step_plan(20, 5, 200) == (20, 1, 0, 5)
    assert bench.step_plan(4000, 200, 200) == (200, 20, 0, 200)


@pytest.mark.parametrize("steps,warmup,chunk", [(0, 0, 200), (10, -1, 200), (10, 0, 0)])
def test_step_plan_rejects_bad_arguments(steps, warmup, chunk):
    with pytest.raises(ValueError):
        bench.step_plan(steps, warmup, chunk)
