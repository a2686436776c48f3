"""Scene tiles at the two solve shapes: 256 threads per scene (a scene of
more than 256 rows reads its rows from memory for every colour) and 512
threads (rows of scenes up to 512 rows stay in registers).  A scene's rows
are updated in the same order either way, so the states are bit-identical;
one step with large scenes also matches the oracle."""
import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd import scenes
from oracle.step import Snapshot, gpu_order, oracle_step
import oracle as O

pytestmark = pytest.mark.gpu


def piles():
    return scenes.batched_piles(n_scenes=20, lo=16, hi=512, pool8=32, iterations=6)


def settled(block, steps):
    w = scenes.build_world(piles(), max_pairs_per_body=64)
    w.set_tuning(solve_block=block)
    w.use_graph(True)
    w.step(steps)
    return w


def tile_rows(w):
    ph = w.read(L.BUF_PHASES, np.int32, (4,))
    return ph[:, 1] - ph[:, 0]


def test_shapes_agree():
    a = settled(256, 110)
    b = settled(512, 110)
    rows = tile_rows(a)
    assert rows.max() > 256 and (rows[rows > 0] <= 256).any()
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ta, tb = a.stats(), b.stats()
    for k in ("num_contacts", "normal_iterations", "friction_iterations", "normal_residual",
              "friction_residual", "overflow_flags"):
        assert ta[k] == tb[k], k
    assert ta["overflow_flags"] == 0
    a.close()
    b.close()


def test_large_scenes_one_step_matches_oracle():
    w = settled(256, 110)
    w.use_graph(False)
    assert tile_rows(w).max() > 256
    s = Snapshot(w)
    w.step_partial(L.STAGE_COLOR)
    order = gpu_order(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    ref = oracle_step(s, order, contacts)
    for k in ("linvel", "angvel", "pos", "quat"):
        n = 4 if k == "quat" else 3
        assert np.allclose(g[k][:, :n], ref[k][:, :n], rtol=1e-4, atol=1e-4), k
    w.close()
