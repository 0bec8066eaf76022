"""Checkpointing schedules for the adjoint of a window of steps
(wam_checkpoint_schedule, host code): every schedule is executed here on a
model of states (integers), which checks that each adjoint step finds the
state it needs, that the adjoint steps come in reverse order, and that no
more slots are used than given; the recomputed steps are the minimum of the
binomial theory, r*l - C(s + r + 1, r - 1) with r the least repetition
number for which C(s + 1 + r, r) >= l (Griewank and Walther, ACM TOMS 26,
2000, Algorithm 799), s the slots besides the initial state."""
from math import comb

import pytest

from wam_amd.model import (CK_ADJOINT, CK_ADVANCE, CK_RESTORE, CK_STORE, WamError, checkpoint_recomputed_steps,
                           checkpoint_schedule)


def run(nsteps, nsnap):
    acts = checkpoint_schedule(nsteps, nsnap)
    assert acts[0] == (CK_STORE, 0, 0)
    slots, cur, order, advanced = {}, 0, [], 0      # the window's initial state is current
    for op, a, b in acts:
        if op == CK_STORE:
            assert cur == a and 0 <= b <= nsnap
            slots[b] = a
        elif op == CK_RESTORE:
            assert slots.get(b) == a
            cur = a
        elif op == CK_ADVANCE:
            assert cur == a and a < b <= nsteps
            advanced += b - a
            cur = b
        else:
            assert op == CK_ADJOINT and cur == a
            order.append(a)
            cur = None
    assert order == list(range(nsteps - 1, -1, -1))
    return advanced


def binomial_minimum(l, s):
    c = s + 1                      # checkpoints including the initial state
    r = 0
    while comb(c + r, c) < l:
        r += 1
    return r * l - comb(c + r, r - 1) if r > 0 else 0


@pytest.mark.parametrize("nsnap", [0, 1, 2, 3, 5, 8])
def test_schedules_execute_and_reach_the_binomial_minimum(nsnap):
    for nsteps in list(range(1, 41)) + [97, 144]:
        adv = run(nsteps, nsnap)
        assert adv == checkpoint_recomputed_steps(nsteps, nsnap)
        if nsnap == 0:
            assert adv == nsteps * (nsteps - 1) // 2
        else:
            assert adv == binomial_minimum(nsteps, nsnap), (nsteps, nsnap)


def test_enough_slots_store_every_state():
    assert run(144, 143) == 143
    assert run(144, 500) == 143


def test_bad_requests_are_refused():
    with pytest.raises(WamError):
        checkpoint_schedule(0, 3)
    with pytest.raises(WamError):
        checkpoint_schedule(10, -1)
