"""Host replay of the dynamic work queues of the persistent kernels (CPU only).

Two queues are replayed with random interleavings of the waves' atomic takes:
* the J kernels' queue (`wave_take` in device_common.hpp: lane 0 adds 1 to the counter and the
  old value goes to every lane; a wave takes tasks until the value reaches ntask);
* the queue of the persistent exchange kernel that was switched off (8 partitions of the unit
  range, one counter each; a wave sweeps the partitions starting with its XCD's, blockIdx & 7,
  and takes units of a partition until its counter passes the partition's end), together with
  that kernel's per-wave compaction queue (128 slots: ballot survivors appended at qn, evaluated
  64 at a time whenever qn >= 64).

Checked: every task or unit is taken exactly once whatever the order of the takes, every wave
leaves after a bounded number of takes (no wave ever waits for another), and the compaction queue
never writes past its 128 slots.  The kernels' arithmetic is restated here, so this pins the
queue logic, not the device code's compilation.
"""
import numpy as np


def replay_j_queue(ntask, nwaves, rng):
    ctr = 0
    taken = []
    alive = list(range(nwaves))
    takes = 0
    while alive:
        w = alive[rng.integers(len(alive))]
        t = ctr  # atomicAdd returns the old value
        ctr += 1
        takes += 1
        if t >= ntask:
            alive.remove(w)
        else:
            taken.append(t)
    return taken, takes


def replay_partitioned_queue(nunits, nwaves, waves_per_group, rng):
    ctr = [0] * 8
    state = {}
    for w in range(nwaves):
        xcd = (w // waves_per_group) & 7  # blockIdx & 7
        state[w] = [xcd, 0]  # (xcd, sweep)
    taken = []
    takes = 0
    alive = list(range(nwaves))
    while alive:
        w = alive[rng.integers(len(alive))]
        xcd, sweep = state[w]
        part = (xcd + sweep) & 7
        lo, hi = nunits * part // 8, nunits * (part + 1) // 8
        t = ctr[part]
        ctr[part] += 1
        takes += 1
        unit = lo + t
        if unit >= hi:
            if sweep == 7:
                alive.remove(w)
            else:
                state[w][1] = sweep + 1
        else:
            taken.append(unit)
    return taken, takes


def test_j_queue_takes_every_task_once_and_ends():
    rng = np.random.default_rng(7)
    for ntask, nwaves in ((0, 5), (1, 64), (1000, 37), (4096, 1024), (5, 2048)):
        taken, takes = replay_j_queue(ntask, nwaves, rng)
        assert sorted(taken) == list(range(ntask))
        assert takes == ntask + nwaves  # one failing take per wave ends it


def test_partitioned_queue_takes_every_unit_once_and_ends():
    rng = np.random.default_rng(11)
    for nunits, nwaves, wpg in ((0, 8, 4), (7, 16, 8), (1000, 96, 4), (12345, 512, 8), (3, 1, 1)):
        taken, takes = replay_partitioned_queue(nunits, nwaves, wpg, rng)
        assert sorted(taken) == list(range(nunits))
        # each wave makes one failing take per partition (8) on top of the units it took
        assert takes == nunits + 8 * nwaves


def test_compaction_queue_stays_within_its_slots():
    rng = np.random.default_rng(13)
    kq = 128
    worst = 0
    for _ in range(20000):
        qn = 0
        for _ in range(int(rng.integers(1, 40))):
            survivors = int(rng.integers(0, 65))  # popcount of one ballot
            # lanes write s_e2[qn + rank] for rank < survivors
            worst = max(worst, qn + survivors - 1)
            assert qn + survivors - 1 < kq
            qn += survivors
            if qn >= 64:  # evaluate 64, move the rest to the front
                qn -= 64
            assert qn < 64
    assert worst <= 126
