"""Host logic of the level-by-level ground-state driver (no device work: the
batch class is replaced by a recorder)."""
import numpy as np
import pytest

from gp5_amd import capi, ground_state


class _FakeBatch:
    made = []

    def __init__(self, n, L, N):
        self.n, self.closed, self.uploads = list(n), False, 0
        _FakeBatch.made.append(self)

    def upload(self, i, psi):
        self.uploads += 1

    def resample_from(self, i, src, j):
        assert not src.closed

    def tune(self, *a):
        pass

    def ground_state(self, f, dt_factor, tol, check_every, max_steps):
        if len(_FakeBatch.made) > 1:
            raise RuntimeError("device error on the fine level")
        return 200, [0.0] * len(self.n), 1.0

    def close(self):
        self.closed = True


def test_level_sizes_halve_while_even():
    assert ground_state.level_sizes(192, 2) == [96, 192]
    assert ground_state.level_sizes(96, 3) == [24, 48, 96]
    assert ground_state.level_sizes(30, 2) == [30]          # half below the minimum
    assert ground_state.level_sizes(64, 2, L=22.4, max_spacing=0.5) == [64]


def test_batches_are_released_when_a_level_fails(monkeypatch):
    _FakeBatch.made = []
    monkeypatch.setattr(capi, "Batch", _FakeBatch)
    with pytest.raises(RuntimeError, match="fine level"):
        ground_state.two_level([64, 96], [22.4, 33.6], [500.0, 2000.0], None, tune=False,
                               seed=lambda i, g, L, N: np.zeros((g, g, g), np.complex128))
    assert [b.n for b in _FakeBatch.made] == [[32, 48], [64, 96]]
    assert all(b.closed for b in _FakeBatch.made)
    assert _FakeBatch.made[0].uploads == 2
