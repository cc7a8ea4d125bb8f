"""The batched clamp-at-level rule used when a workgroup applies summed
decrements (flush_agg in kc_peel.hip), simulated on the CPU under random
interleavings of the atomic subtractions and their restores:

  old = atomicSub(deg, c); eff = min(c, max(0, old - level));
  atomicAdd(deg, c - eff); the op with old > level and old - eff == level
  claims the vertex.

Whatever the interleaving, the degree must settle at max(level, d - sum(c))
and exactly one op must claim iff d > level >= d - sum(c).
"""
import random


def simulate(d, level, deltas, rng):
    x = d
    pending = []  # restores not yet applied
    order = list(range(len(deltas)))
    rng.shuffle(order)
    claims = 0
    events = [("sub", i) for i in order]
    while events or pending:
        if pending and (not events or rng.random() < 0.5):
            x += pending.pop(rng.randrange(len(pending)))
            continue
        _, i = events.pop(0)
        c = deltas[i]
        old = x
        x -= c
        eff = min(c, max(0, old - level))
        if c > eff:
            pending.append(c - eff)
        if old > level and old - eff == level:
            claims += 1
    return x, claims


def test_batched_clamp_random_interleavings():
    rng = random.Random(7)
    for _ in range(20000):
        level = rng.randrange(0, 20)
        d = level + rng.randrange(0, 12)
        deltas = [rng.randrange(1, 6) for _ in range(rng.randrange(1, 6))]
        x, claims = simulate(d, level, deltas, rng)
        assert x == max(level, d - sum(deltas))
        assert claims == (1 if d > level >= d - sum(deltas) else 0)
