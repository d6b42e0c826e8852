"""The property the block-parallel re-binning relies on (k_step.hip, b_bin_apply).

A cell's replay of the reference's swap-remove scan with the early arrivals appended
(misc.c:279-303) moves atoms one after another.  The kernel lists those moves and has all
threads of the block carry them out in two unordered stages: the moves into the surviving
prefix first, then the late arrivals appended after it.  That equals the sequential replay
only if, in the first stage, no move reads a slot another move writes, and only late
arrivals land on slots beyond the prefix.  Random cells, from empty to full, with every mix
of leavers, early and late arrivals, check both against the sequential replay."""
import random


def replay(nc, leaves, A, na, cap, act):
    """the kernel's walk: act(destination, staged, source) per move; returns (prefix, count)"""
    nb = min(na, cap)
    i, n, cur = 0, nc + A, 0
    while i < n:
        if cur < nc and leaves[cur]:
            n -= 1
            if n != i:
                cur = n
        else:
            if cur >= nc:
                q = cur - nc
                if q < nb and i < cap:
                    act(i, True, q)
            elif cur != i:
                act(i, False, cur)
            i += 1
            cur = i
    prefix, t = n, n
    for q in range(A, na):
        if q < nb and t < cap:
            act(t, True, q)
        t += 1
    return prefix, min(t, cap)


def test_two_stage_moves_equal_the_sequential_replay():
    rng = random.Random(7)
    checked = 0
    for _ in range(20000):
        cap = rng.choice([32, 64])
        nc = rng.randint(0, cap)
        p = rng.choice([0.0, 0.05, 0.3, 0.7, 1.0])
        leaves = [rng.random() < p for _ in range(nc)]
        na = rng.randint(0, 6)
        A = rng.randint(0, na)
        staged = ["arrival%d" % q for q in range(na)]
        seq = ["atom%d" % k for k in range(nc)] + [None] * (cap - nc)

        def apply(dst, st, src):
            seq[dst] = staged[src] if st else seq[src]
        _, count = replay(nc, leaves, A, na, cap, apply)

        moves = []
        prefix, count2 = replay(nc, leaves, A, na, cap, lambda d, st, src: moves.append((d, st, src)))
        assert count2 == count
        first = [m for m in moves if m[0] < prefix]
        late = [m for m in moves if m[0] >= prefix]
        assert all(st for _, st, _ in late)
        assert not ({src for _, st, src in first if not st} & {d for d, _, _ in first})
        par = ["atom%d" % k for k in range(nc)] + [None] * (cap - nc)
        orig = list(par)
        rng.shuffle(first)
        rng.shuffle(late)
        for d, st, src in first:
            par[d] = staged[src] if st else orig[src]
        for d, st, src in late:
            par[d] = staged[src]
        assert par[:count] == seq[:count]
        checked += bool(moves)
    assert checked > 10000
