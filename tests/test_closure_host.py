"""Host logic of the closure's run-time stabiliser (closure.SafeMode): the
hysteresis between the two CFL thresholds, the per-run state and the
reported fraction; no GPU needed."""
import pytest

from rot2d3c_amd import closure as L


class _Net:
    def __init__(self):
        self.strength = None

    def set_sign_clip(self, s):
        self.strength = s


def test_safe_mode_hysteresis():
    sm = L.SafeMode(2.0, 1.5, base_sign_clip=0.1)
    seq = [1.2, 1.9, 2.1, 1.8, 1.6, 1.49, 1.7, 2.5]
    want = [False, False, True, True, True, False, False, True]
    changes = []
    for c, w in zip(seq, want):
        changes.append(sm.check(0, c))
        assert sm.active[0] is w
    assert changes == [False, False, True, False, False, True, False, True]
    # fraction counts the mode in force during each checked interval
    assert sm.fraction == pytest.approx(3 / 8)
    net = _Net()
    sm.apply(net, True)
    assert net.strength == 1.0
    sm.apply(net, False)
    assert net.strength == 0.1


def test_safe_mode_runs_are_independent_and_nan_is_not_recovery():
    sm = L.SafeMode(2.0, 1.5)
    assert sm.check(1, 2.5) and sm.active[1]
    assert not sm.check(0, 1.0) and not sm.active[0]
    # a non-finite CFL number never counts as recovered
    assert not sm.check(1, float("nan")) and sm.active[1]
    with pytest.raises(ValueError):
        L.SafeMode(1.0, 2.0)
