"""The library integrator's and adjoint's work schedules, replayed on the host
by the library's own schedule arithmetic (`smx_integrator_schedule_check`,
`smx_adjoint_schedule_check`; no GPU).

The kernel runs as many waves as are resident at once and spreads each wave's
share of integrations over its 64 lanes (McNaughton's wrap-around rule); an
integration cut at a lane boundary makes its first attempts on one lane and
the rest on the next, with the state handed over in LDS.  The replay checks
that every integration makes attempts 0..L-1 exactly once and in order, that
a second part starts only after its first part has handed over, and that the
loop length is max(L, ceil(m L / 64)) for a share of m integrations -- the
point of the schedule: plain-20k's 80,000 integrations on 1,024 waves (78 or
79 per wave) take 2,438 or 2,469 steps per lane, not the 4,000 of two rounds
of whole integrations.
"""
import ctypes

import pytest

from schwarmax_amd import _lib


def _check(n_int, waves, L):
    lib = _lib.load_library()
    steps = ctypes.c_int64(0)
    rc = lib.smx_integrator_schedule_check(n_int, waves, L, ctypes.byref(steps))
    return rc, steps.value


@pytest.mark.parametrize("n_int,waves,L", [
    (80_000, 1_024, 2_000),    # plain-20k
    (80, 1, 10), (65, 1, 9), (64, 1, 9), (127, 2, 3), (300, 4, 1), (1_000, 3, 13),
    (7_000, 100, 11), (5_000, 64, 20), (0, 8, 5), (10, 16, 4),
])
def test_schedule_replay(n_int, waves, L):
    rc, steps = _check(n_int, waves, L)
    assert rc == 0, rc
    m = n_int // waves
    if m > 0:
        assert steps == max(L, -(-m * L // 64))


def test_plain20k_loop_length():
    rc, steps = _check(80_000, 1_024, 2_000)     # wave 0 holds 78 integrations
    assert rc == 0 and steps == 2_438
    rc, steps = _check(79 * 1_024, 1_024, 2_000)  # 79 per wave
    assert rc == 0 and steps == 2_469


def test_bad_arguments_are_rejected():
    rc, _ = _check(10, 0, 5)
    assert rc != 0


def _check_adjoint(n_int, waves, U):
    lib = _lib.load_library()
    steps = ctypes.c_int64(0)
    rc = lib.smx_adjoint_schedule_check(n_int, waves, U, ctypes.byref(steps))
    return rc, steps.value


@pytest.mark.parametrize("n_int,waves,U", [
    (80_000, 1_024, 32),       # grad-20k: 2,000 attempts in segments of 64
    (80, 1, 10), (65, 1, 9), (64, 1, 9), (127, 2, 3), (300, 4, 1), (1_000, 3, 13),
    (7_000, 100, 11), (0, 8, 5), (10, 16, 4),
])
def test_adjoint_schedule_replay(n_int, waves, U):
    """The balanced adjoint's step loop: units swept once and in order,
    hand-overs read after they are written, lane 0 on the shadowed unit."""
    rc, steps = _check_adjoint(n_int, waves, U)
    assert rc == 0, rc
    m = n_int // waves
    if m > 0:
        assert steps == max(U, -(-m * U // 64))
