"""Large uploads and result downloads go through pinned chunks on a host
worker pool; small ones are direct copies.  A batch big enough for the staged
paths gives, game by game, exactly what each game gives alone (direct paths),
through both the bulk results call and the per-game getters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_staged_batch_equals_alone(gpu):
    games = [gpu.gen_random_game(900 + k, 900000 + 1000 * k, 5000, 1, 4) for k in range(10)]
    total = sum(g.n for g in games)
    assert total * 4 >= 32 << 20 and sum(3 * g.n + g.m for g in games) * 4 >= 64 << 20
    print(f"staged batch n={total}", flush=True)
    b = gpu.Solver(games)
    b.solve()
    bulk = b.results()
    for k, g in enumerate(games):
        s = gpu.Solver(g)
        s.solve()
        st, strat, lab = s.stats(), s.strategy(), s.labels()
        s.close()
        assert bulk[k][0] == st, k
        assert np.array_equal(bulk[k][1], strat), k
        assert np.array_equal(bulk[k][2], lab), k
        assert np.array_equal(b.strategy(k), strat) and np.array_equal(b.labels(k), lab), k
    b.close()
