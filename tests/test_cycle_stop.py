"""The jump rounds' early stop for cycle-bound vertices, checked by brute force
on the CPU (no GPU, no oracle involved): on a functional graph where every
vertex leads into a cycle, let x_l(v) be v's 2^l-th successor and lambda_l(v)
the parity of the largest priority in the window of 2^l vertices that starts
at x_l(v) (the label a valuation ending at level l gives v).  The device stops
jumping once lambda_l(v) == lambda_l(x_l(v)) for every such vertex (both
windows are read by the gathers a two-level launch makes anyway); the claim
is that from then on lambda_F, for every F >= l, is the parity of the largest
priority on v's cycle (cpu_parity.cpp:35-52), tails included.  The same holds
with the immediate successor in place of x_l(v)."""
import numpy as np
import pytest


def _levels(f, pri, L):
    """x[l][v] = f^(2^l)(v), y[l][v] = max priority over v .. f^(2^l - 1)(v)."""
    xs, ys = [f.copy()], [pri.copy()]
    for _ in range(L):
        x, y = xs[-1], ys[-1]
        xs.append(x[x])
        ys.append(np.maximum(y, y[x]))
    return xs, ys


def _cycle_parity(f, pri):
    """Parity of the largest priority on the cycle each vertex leads into."""
    n = len(f)
    out = np.full(n, -1)
    for v in range(n):
        if out[v] >= 0:
            continue
        seen, path, u = {}, [], v
        while u not in seen and out[u] < 0:
            seen[u] = len(path)
            path.append(u)
            u = f[u]
        if out[u] >= 0:
            p = out[u]
        else:
            cyc = path[seen[u]:]
            p = int(max(pri[c] for c in cyc)) & 1
        for w in path:
            out[w] = p
    return out


def _graph(rng, n, n_cycles, max_cycle, chain_bias, pri_bound):
    """Random functional graph: cycles over the first vertices of a random
    order, every later vertex linked to an earlier one (to the one just
    before it with probability chain_bias, which makes long tails)."""
    order = rng.permutation(n)
    f = np.empty(n, np.int64)
    pos = 0
    for _ in range(n_cycles):
        L = int(rng.integers(1, max_cycle + 1))
        L = min(L, n - pos)
        if L <= 0:
            break
        cyc = order[pos:pos + L]
        f[cyc] = np.roll(cyc, -1)
        pos += L
    if pos == 0:
        f[order[0]] = order[0]
        pos = 1
    for i in range(pos, n):
        j = i - 1 if rng.random() < chain_bias else int(rng.integers(0, i))
        f[order[i]] = order[j]
    pri = rng.integers(0, pri_bound, n)
    return f, pri


CASES = [(s, n, c, m, b, p) for s, (n, c, m, b, p) in enumerate([
    (50, 3, 8, 0.5, 10),
    (300, 5, 40, 0.9, 1000),
    (300, 1, 300, 0.0, 4),         # one long cycle, few priorities (repeated maxima)
    (1000, 20, 60, 0.97, 100000),  # long tails with large priorities
    (1000, 2, 200, 0.99, 3),
    (2000, 50, 5, 0.995, 50),      # short cycles, very long tails
    (2000, 1, 1, 0.5, 2),          # self-loop
    (4096, 30, 500, 0.9, 7),
])]


@pytest.mark.parametrize("along", ["jump", "succ"])
@pytest.mark.parametrize("seed,n,cycles,max_cycle,bias,pb", CASES)
def test_consistent_bits_give_cycle_labels(seed, n, cycles, max_cycle, bias, pb, along):
    rng = np.random.default_rng(1234 + seed)
    f, pri = _graph(rng, n, cycles, max_cycle, bias, pb)
    truth = _cycle_parity(f, pri)
    L = int(np.ceil(np.log2(n))) + 2
    xs, ys = _levels(f, pri, L)
    lam = [ys[l][xs[l]] & 1 for l in range(L + 1)]
    first = None
    for l in range(L + 1):
        nxt = xs[l] if along == "jump" else f
        if np.array_equal(lam[l], lam[l][nxt]):
            first = l
            break
    # the reference's full window (2^l >= n) always passes
    assert first is not None
    for F in range(first, L + 1):
        assert np.array_equal(lam[F], truth), (first, F)


def test_check_fails_while_windows_miss_the_maximum():
    """A cycle 0 -> 1 -> ... -> 7 -> 0 whose only even priority is at vertex 0:
    windows of one or two vertices disagree along the cycle, so the check
    keeps the rounds going until the window covers the whole cycle."""
    f = np.array([1, 2, 3, 4, 5, 6, 7, 0])
    pri = np.array([10, 3, 5, 7, 9, 1, 3, 5])
    xs, ys = _levels(f, pri, 4)
    lam = [ys[l][xs[l]] & 1 for l in range(5)]
    passed = [bool(np.array_equal(lam[l], lam[l][f])) for l in range(5)]
    assert passed == [False, False, False, True, True]
    passed = [bool(np.array_equal(lam[l], lam[l][xs[l]])) for l in range(5)]
    assert passed == [False, False, False, True, True]
    assert (lam[3] == 0).all()


def test_check_fails_on_a_tail_with_a_larger_odd_priority():
    """Tail 3 -> 2 -> 1 -> 0 into the self-loop 0 (priority 2, even); the tail
    holds 9 (odd).  Short windows at tail vertices still see 9: inconsistent
    until every window that decides a label lies on the cycle."""
    f = np.array([0, 0, 1, 2])
    pri = np.array([2, 9, 4, 6])
    xs, ys = _levels(f, pri, 3)
    lam = [ys[l][xs[l]] & 1 for l in range(4)]
    for nxt in (lambda l: f, lambda l: xs[l]):
        assert not np.array_equal(lam[0], lam[0][nxt(0)])
        ok = [l for l in range(4) if np.array_equal(lam[l], lam[l][nxt(l)])]
        assert ok and all((lam[l] == 0).all() for l in range(ok[0], 4))


@pytest.mark.parametrize("along", ["jump", "succ"])
def test_alternating_parities_are_not_accepted(along):
    """Cycle 0 -> 1 -> 2 -> 3 -> 0 with priorities 9, 2, 7, 4 (largest 9, odd):
    one-vertex windows alternate in parity, so a check that compared each
    vertex with the vertex two steps on (windows with gaps) would accept
    level 0 with wrong labels; the device's check does not."""
    f = np.array([1, 2, 3, 0])
    pri = np.array([9, 2, 7, 4])
    xs, ys = _levels(f, pri, 3)
    lam = [ys[l][xs[l]] & 1 for l in range(4)]
    truth = _cycle_parity(f, pri)
    assert (truth == 1).all()
    for l in range(4):
        nxt = xs[l] if along == "jump" else f
        if np.array_equal(lam[l], lam[l][nxt]):
            assert np.array_equal(lam[l], truth), l
    # the gapped check (two steps on) would pass at level 0, with wrong labels
    assert np.array_equal(lam[0], lam[0][f[f]]) and not np.array_equal(lam[0], truth)
