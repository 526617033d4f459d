"""CPU model of the batched scans' work-queue walk (batched_scan_kernel and
batched_scan_i8_kernel, csrc/batched.hip): persistent waves claim work
items from per-XCD-group queue heads with an atomic add, take their own
group's items first, then the other groups' leftovers, and leave after one
empty claim in each of the 8 groups.

The model runs the walk of every wave as a generator, one atomic claim per
step, under adversarial interleavings (random, round-robin, one wave
starved until the end), and checks what the kernel's correctness and
termination rest on:
  * every item is processed exactly once;
  * every wave leaves after at most (its items + 8) claims;
  * the same holds for a walk that claims one item ahead (the claim for the
    next item issued before the current one is processed), as long as a
    claim is compared with the bounds of the group it was taken from.  A
    walk that issues its claim ahead before checking the current one, and so
    compares a claim held across a group switch with the next group's
    bounds, still terminates, but processes items twice or never: a
    coverage bug, not a hang.
"""
import random
import zlib

import pytest

KXCDS = 8


def make_queues(rng, max_items=None):
    sizes = [rng.randint(0, 30) for _ in range(KXCDS)]
    off = [0]
    for s in sizes:
        off.append(off[-1] + s)
    if max_items is None:
        max_items = off[-1]
    return off, max_items


def walk_shipped(wave, off, max_items, heads, done):
    """The shipped loop: claim, compare with the claimed group's end, switch
    group after an empty claim, leave after KXCDS empty claims."""
    grp, switches = wave % KXCDS, 0
    start, end = off[grp], min(off[grp + 1], max_items)
    claims = 0
    while True:
        it = start + heads[grp]
        heads[grp] += 1  # atomicAdd: the step boundary below is the interleaving point
        claims += 1
        yield
        if it >= end:
            switches += 1
            if switches == KXCDS:
                return claims
            grp = (grp + 1) % KXCDS
            start, end = off[grp], min(off[grp + 1], max_items)
            continue
        done.append(it)


def walk_one_ahead(wave, off, max_items, heads, done, claim_before_check=False):
    """A walk that claims the next item before processing the current one.
    claim_before_check: the claim ahead is issued before the current claim
    is checked, so after a group switch the claim held from the old group is
    compared with the new group's bounds."""
    grp, switches = wave % KXCDS, 0
    end = min(off[grp + 1], max_items)

    def claim(g):
        v = off[g] + heads[g]
        heads[g] += 1
        return v

    it = claim(grp)
    claims = 1
    yield
    while True:
        if claim_before_check:
            nxt = claim(grp)
            claims += 1
            yield
        if it >= end:
            switches += 1
            if switches == KXCDS:
                return claims
            grp = (grp + 1) % KXCDS
            end = min(off[grp + 1], max_items)
            if claim_before_check:
                it = nxt  # taken from the previous group
            else:
                it = claim(grp)
                claims += 1
                yield
            continue
        if not claim_before_check:
            nxt = claim(grp)  # one step ahead
            claims += 1
            yield
        done.append(it)
        it = nxt


def run(walk, off, max_items, waves, order, rng, **kw):
    heads = [0] * KXCDS
    done = []
    gens = {w: walk(w, off, max_items, heads, done, **kw) for w in range(waves)}
    claims = {}
    live = list(range(waves))
    starved = live[-1] if order == "starve" else None
    step = 0
    while live:
        if order == "random":
            w = rng.choice(live)
        elif order == "roundrobin":
            w = live[step % len(live)]
        else:  # starve: the last wave runs only when it is the only one left
            cands = [x for x in live if x != starved] or live
            w = rng.choice(cands)
        step += 1
        try:
            next(gens[w])
        except StopIteration as e:
            claims[w] = e.value
            live.remove(w)
        assert step < 10_000_000, "the walk did not terminate"
    return done, claims


@pytest.mark.parametrize("order", ["random", "roundrobin", "starve"])
@pytest.mark.parametrize("walk", [walk_shipped, walk_one_ahead])
def test_walk_terminates_and_covers_every_item_once(walk, order):
    rng = random.Random(zlib.crc32(f"{walk.__name__}/{order}".encode()))
    for trial in range(200):
        off, max_items = make_queues(rng)
        if trial % 5 == 0:  # the planning's item bound below the planned items
            max_items = rng.randint(0, off[-1])
        waves = rng.choice([1, 2, 3, 8, 13, 64])
        done, claims = run(walk, off, max_items, waves, order, rng)
        want = sorted(i for g in range(KXCDS) for i in range(off[g], min(off[g + 1], max_items)))
        assert sorted(done) == want, (trial, off, max_items)
        items_by_wave = len(done)
        # every wave: at most its own items + one empty claim per group (+1 for
        # the claim held ahead)
        assert sum(claims.values()) <= items_by_wave + waves * (KXCDS + 1)


def test_claim_checked_against_the_wrong_group_breaks_coverage_not_termination():
    rng = random.Random(7)
    bad = 0
    for trial in range(300):
        off, max_items = make_queues(rng)
        waves = rng.choice([2, 8, 13])
        done, claims = run(walk_one_ahead, off, max_items, waves, "random", rng, claim_before_check=True)
        want = sorted(i for g in range(KXCDS) for i in range(off[g], min(off[g + 1], max_items)))
        bad += sorted(done) != want
    assert bad > 0  # the mistake shows as items processed twice or never
