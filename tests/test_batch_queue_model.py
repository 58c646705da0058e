"""CPU model of the batch work queue's protocol: k_batch_queue
(opencfd-sc_amd/csrc/ocfd_batch_queue.h) as batch_queue_plan
(opencfd-sc_amd/csrc/ocfd_api.hip) lays out its items.

The model keeps the planner's segment order (per stage: halo fill of every
box; x, y, z sweeps per scheme group; viscous flux of the viscous boxes;
stage update; repair + halo copy) and the kernel's dependency rule
(q_dependency: an item waits for every item of its own box in the phase
before; a stage's halo fill for the previous stage's repair; an inviscid
box's update for its z sweep).  For random batches and item counts it
checks that every item a ticket waits for has a smaller ticket, and that
W workgroups taking tickets in order, with random item times, always drain
the queue (W = 1, 7, 256) with each box's phases completing in order.

This pins the protocol, not the kernel: the persistent launch mode stays
switched off in the library, and its GPU tests are skipped."""
import heapq
import random

import pytest

Q_HALO, Q_X, Q_Y, Q_Z, Q_VFLUX, Q_UPDATE, Q_REPAIR = range(7)


def random_batch(rng):
    nbox = rng.randint(1, 9)
    boxes = []
    for _ in range(nbox):
        visc = rng.random() < 0.7
        need = {Q_HALO: rng.randint(1, 5), Q_X: rng.randint(1, 6), Q_Y: rng.randint(1, 6), Q_Z: rng.randint(1, 6),
                Q_VFLUX: rng.randint(1, 4) if visc else 0, Q_UPDATE: rng.randint(1, 4), Q_REPAIR: rng.randint(1, 3)}
        boxes.append({"group": rng.randint(0, 2), "visc": visc, "need": need})
    return boxes


def plan(boxes):
    """Tickets in the planner's order: (kind, stage, box) per item."""
    groups = sorted({b["group"] for b in boxes})
    tickets = []

    def add(kind, stage, ids):
        for i in ids:
            tickets.extend([(kind, stage, i)] * boxes[i]["need"][kind])

    allb = range(len(boxes))
    for s in (1, 2, 3):
        add(Q_HALO, s, allb)
        for a in range(3):
            for g in groups:
                add(Q_X + a, s, [i for i in allb if boxes[i]["group"] == g])
        add(Q_VFLUX, s, [i for i in allb if boxes[i]["visc"]])
        for v in (True, False):
            add(Q_UPDATE, s, [i for i in allb if boxes[i]["visc"] == v])
        add(Q_REPAIR, s, allb)
    return tickets


def dependency(boxes, box, kind, stage):
    """(stage, kind) whose items of `box` an item waits for, or None."""
    if kind == Q_HALO:
        return None if stage == 1 else (stage - 1, Q_REPAIR)
    if kind == Q_UPDATE and boxes[box]["need"][Q_VFLUX] == 0:
        return (stage, Q_Z)
    return (stage, kind - 1)


def simulate(boxes, tickets, nwg, rng):
    """Event simulation of nwg workgroups; returns the completion order of
    (box, stage, kind) phases, or raises on a deadlock."""
    done = {}
    finished_phase = []
    head = 0
    clock = 0.0
    running = []   # (end time, wg, ticket)
    waiting = {}   # wg -> ticket
    idle = list(range(nwg))

    def ready(tk):
        kind, stage, box = tickets[tk]
        d = dependency(boxes, box, kind, stage)
        if d is None:
            return True
        ds, dk = d
        return done.get((box, ds, dk), 0) >= boxes[box]["need"][dk]

    while True:
        # idle workgroups take the next tickets
        while idle and head < len(tickets):
            waiting[idle.pop()] = head
            head += 1
        # waiting workgroups whose dependency has completed start
        for wg in sorted(waiting):
            tk = waiting[wg]
            if ready(tk):
                del waiting[wg]
                heapq.heappush(running, (clock + rng.random(), wg, tk))
        if not running:
            if waiting:
                raise AssertionError("deadlock: %d workgroups wait, none runs" % len(waiting))
            break
        clock, wg, tk = heapq.heappop(running)
        kind, stage, box = tickets[tk]
        key = (box, stage, kind)
        done[key] = done.get(key, 0) + 1
        if done[key] == boxes[box]["need"][kind]:
            finished_phase.append(key)
        idle.append(wg)
    assert head == len(tickets)
    return finished_phase


@pytest.mark.parametrize("seed", range(40))
def test_dependencies_precede_their_waiters(seed):
    rng = random.Random(seed)
    boxes = random_batch(rng)
    tickets = plan(boxes)
    last = {}
    for tk, (kind, stage, box) in enumerate(tickets):
        last[(box, stage, kind)] = tk
    for tk, (kind, stage, box) in enumerate(tickets):
        d = dependency(boxes, box, kind, stage)
        if d is not None:
            assert last[(box,) + d] < tk, (tk, kind, stage, box)
    # every box has each of its phases once per stage, with its item count
    for i, b in enumerate(boxes):
        for s in (1, 2, 3):
            for k, n in b["need"].items():
                assert sum(1 for t in tickets if t == (k, s, i)) == n


@pytest.mark.parametrize("nwg", [1, 7, 256])
@pytest.mark.parametrize("seed", range(15))
def test_queue_drains_with_phases_in_order(nwg, seed):
    rng = random.Random(1000 * nwg + seed)
    boxes = random_batch(rng)
    tickets = plan(boxes)
    order = simulate(boxes, tickets, nwg, rng)
    for i, b in enumerate(boxes):
        phases = [(s, k) for (box, s, k) in order if box == i]
        expect = []
        for s in (1, 2, 3):
            for k in (Q_HALO, Q_X, Q_Y, Q_Z, Q_VFLUX, Q_UPDATE, Q_REPAIR):
                if b["need"][k]:
                    expect.append((s, k))
        assert phases == expect, (i, phases)


def test_model_detects_a_broken_rule():
    """The checks can fail: an update that waits for the repair of its own
    stage (a later ticket) is flagged by the ordering check and deadlocks a
    single workgroup."""
    rng = random.Random(5)
    boxes = random_batch(rng)
    tickets = plan(boxes)
    global dependency
    saved = dependency

    def broken(bx, box, kind, stage):
        return (stage, Q_REPAIR) if kind == Q_UPDATE else saved(bx, box, kind, stage)

    dependency = broken
    try:
        with pytest.raises(AssertionError):
            test_dependencies_precede_their_waiters(5)
        with pytest.raises(AssertionError, match="deadlock"):
            simulate(boxes, tickets, 1, rng)
    finally:
        dependency = saved
