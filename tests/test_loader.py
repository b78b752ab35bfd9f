"""The library's PGSolver loader (Game::parse_pgsolver restated in
parallel-si_amd/host/game.cpp) against the oracle's restatement and the
reference's error messages (game.cpp:51-107).  Host only."""
import numpy as np
import pytest

import oracle
import psi
from conftest import fixture_game, load_fixtures
from test_oracle import PARSE_ERRORS


@pytest.mark.parametrize("text,msg", PARSE_ERRORS)
def test_errors_match_reference_messages(text, msg):
    with pytest.raises(psi.PsiError, match=msg):
        psi.game_from_gm(text)


def test_rejects_what_the_reference_leaves_undefined():
    with pytest.raises(psi.PsiError, match="successor out of range"):
        psi.game_from_gm("parity 1;\n0 1 0 7;\n")
    with pytest.raises(psi.PsiError, match="priority"):
        psi.game_from_gm("parity 1;\n0 -3 0 0;\n")


@pytest.mark.parametrize("seed", range(5))
def test_same_csr_as_oracle(seed):
    g = psi.gen_random_game(seed, 500 + 100 * seed, 37, 1, 5)
    text = g.to_gm()
    a = psi.game_from_gm(text)
    owner, pri, off, tgt, _ = oracle.parse_gm(text)
    for x, y in ((a.owner, owner), (a.priority, pri), (a.offsets, off), (a.targets, tgt)):
        assert np.array_equal(x, y)


def test_reference_fixture_text():
    fx = next(f for f in load_fixtures() if f["name"] == "test2")
    g = fixture_game(fx)
    text = ("parity 9;\n" + "".join(
        f'{v} {v % 2} {v % 2} {(v + 1) % 10},{(v + 2) % 10} "{v}";\n' for v in range(10)))
    a = psi.game_from_gm(text)
    assert np.array_equal(a.targets, g.targets) and np.array_equal(a.priority, g.priority)
