"""A CPU model of the guard's concurrent union-find (optflow_soil_amd/csrc/guard.hip: find_root,
unite, k_cc_union). Each model thread is a generator that yields after every shared-memory access,
so a scheduler can interleave the threads at any access. The kernel's path compression walks only
while the node is above the root it read (`x > r`). The model shows why:

- with the earlier rule (walk until the node equals that root), a scripted interleaving, and
  some random ones, close a parent cycle; a find_root on that component then never ends;
- with the kernel's rule, no interleaving makes a cycle, and the final partition is the
  sequential union-find's.

The model restates the kernel's algorithm; it does not run the kernel. The GPU tests of the
guard (large components, validity) check the kernel's results against the oracle."""
import random

import pytest


def find_root(p, x, rule):
    r = x
    while True:
        q = p[r]
        yield
        if q == r:
            break
        r = q
    while (x != r) if rule == "earlier" else (x > r):
        q = p[x]
        yield
        if q == x:
            break
        if q != r:
            p[x] = r
            yield
        x = q
    return r


def unite(p, a, c, rule):
    while True:
        a = yield from find_root(p, a, rule)
        c = yield from find_root(p, c, rule)
        if a == c:
            return
        if a < c:
            a, c = c, a
        if p[a] == a:  # atomicCAS(&parent[a], a, c) succeeds
            p[a] = c
            yield
            return
        yield


def has_cycle(p):
    for s in range(len(p)):
        x, k = s, 0
        while p[x] != x:
            x = p[x]
            k += 1
            if k > len(p):
                return True
    return False


def run_to_end(gen):
    for _ in gen:
        pass


def scripted(rule):
    """Thread A reads root 5 for node 9 (path 9 -> 7 -> 5). Other threads link 5 under 3 and 3
    under 1; thread C compresses 7 and 5 straight to 1; 1 is linked under 0. Then A compresses."""
    p = list(range(10))
    p[9], p[7] = 7, 5
    a = find_root(p, 9, rule)
    for _ in range(3):  # A reads p[9], p[7], p[5]: r = 5
        next(a)
    p[5] = 3
    p[3] = 1
    run_to_end(find_root(p, 7, rule))
    p[1] = 0
    run_to_end(a)
    return p


@pytest.mark.parametrize("rule,cycle", [("earlier", True), ("kernel", False)])
def test_scripted_interleaving(rule, cycle):
    p = scripted(rule)
    assert has_cycle(p) == cycle
    if not cycle:
        assert all(p[x] <= x for x in range(len(p)))


def random_unions(rule, rng, n, n_edges):
    p = list(range(n))
    edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(n_edges)]
    threads = [unite(p, a, c, rule) for a, c in edges]
    live = list(range(len(threads)))
    for _ in range(200000):
        if not live:
            break
        k = rng.choice(live)
        try:
            next(threads[k])
        except StopIteration:
            live.remove(k)
        if has_cycle(p):
            return p, edges, True
    return p, edges, False


def test_random_interleavings_make_cycles_only_with_the_earlier_rule():
    rng = random.Random(0)
    cycles = sum(random_unions("earlier", rng, 40, 40)[2] for _ in range(150))
    assert cycles > 0
    rng = random.Random(0)
    assert not any(random_unions("kernel", rng, 40, 40)[2] for _ in range(150))


def test_kernel_rule_gives_the_sequential_partition():
    rng = random.Random(1)
    for _ in range(200):
        p, edges, cycle = random_unions("kernel", rng, 40, 20)
        assert not cycle
        q = list(range(40))

        def root(t, x):
            while t[x] != x:
                x = t[x]
            return x
        for a, c in edges:
            ra, rc = root(q, a), root(q, c)
            if ra != rc:
                q[max(ra, rc)] = min(ra, rc)
        assert [root(p, x) for x in range(40)] == [root(q, x) for x in range(40)]
        assert all(p[x] <= x for x in range(40))
