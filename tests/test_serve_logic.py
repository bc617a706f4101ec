"""Admission policy of the serving loop on the CPU (tr_serve_admit through the
C ABI, no device): first fit in arrival order within the HBM budget and the
member cap, the starvation guard, and, over many rounds of a model of the
loop that retires requests after a number of rounds, a budget that is never
exceeded and requests that all finish without starving."""
import numpy as np

from triton_amd.serve import admit_round


def test_first_fit_in_arrival_order():
    skipped = np.zeros(3, np.int32)
    take = admit_round([70, 50, 20], skipped, held=0, budget=100, running=0)
    assert list(take) == [True, False, True]       # 1 does not fit next to 0; 2 does
    assert list(skipped) == [0, 1, 0]


def test_held_bytes_and_member_cap_count():
    skipped = np.zeros(3, np.int32)
    assert list(admit_round([30, 30, 30], skipped, held=50, budget=100, running=0)) == [True, False, False]
    skipped = np.zeros(3, np.int32)
    assert list(admit_round([1, 1, 1], skipped, held=0, budget=0, running=2, max_members=4)) == [True, True, False]
    assert list(skipped) == [0, 0, 1]


def test_no_budget_admits_everything():
    skipped = np.zeros(4, np.int32)
    assert admit_round([10 ** 12] * 4, skipped, held=0, budget=0, running=0).all()


def test_starved_request_holds_back_later_ones():
    skipped = np.array([2, 0], np.int32)           # request 0 already passed over twice
    take = admit_round([60, 10], skipped, held=90, budget=100, running=1, starve_rounds=2)
    assert list(take) == [False, False]            # 10 B would fit, but 0 is now starving
    assert list(skipped) == [3, 0]
    skipped = np.array([1, 0], np.int32)           # not starving yet: 1 goes ahead
    take = admit_round([60, 10], skipped, held=90, budget=100, running=1, starve_rounds=2)
    assert list(take) == [False, True]


def test_rounds_keep_the_budget_and_finish_everything():
    """The loop's rounds with the policy: requests hold their bytes for a
    number of rounds, then retire; the held bytes never exceed the budget and
    every request finishes, the large one too (the guard stops smaller later
    arrivals from starving it)."""
    rng = np.random.default_rng(7)
    budget = 100
    arrivals = [(int(b), int(n)) for b, n in zip(rng.integers(5, 60, 40), rng.integers(1, 6, 40))]
    arrivals.insert(3, (95, 2))                    # needs nearly the whole budget
    waiting, running, done, held_max = [], [], [], 0
    k = 0
    for _ in range(10000):
        if k < len(arrivals):                      # one arrival per round
            waiting.append([k, *arrivals[k], 0])
            k += 1
        if not waiting and not running and k == len(arrivals):
            break
        skipped = np.array([w[3] for w in waiting], np.int32)
        held = sum(r[1] for r in running)
        take = admit_round([w[1] for w in waiting], skipped, held=held, budget=budget, running=len(running),
                           starve_rounds=4)
        keep = []
        for w, t, s in zip(waiting, take, skipped):
            w[3] = int(s)
            (running if t else keep).append(w)
        waiting = keep
        held_max = max(held_max, sum(r[1] for r in running))
        for r in running:
            r[2] -= 1
        done += [r[0] for r in running if r[2] <= 0]
        running = [r for r in running if r[2] > 0]
    assert held_max <= budget
    assert sorted(done) == list(range(len(arrivals)))
