"""The trained warm-start network shipped with the package
(experiments_amd/warmstart_weights.npz) on held-out batched piles (another
seed than its training scenes): every solve runs to a 1e-4 residual, and
the normal-phase sweeps per scene are compared for the zero start (the
plain solver), the previous-step start and the learned start.  The learned
start must need far fewer normal sweeps than the plain solver and no more
than the previous-step start; the friction phase, whose sweeps the
reference's spin damping sets, is reported beside them."""
import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd import scenes
from experiments_amd.warmstart import load_weights, run_iterations

pytestmark = pytest.mark.gpu


def test_trained_network_against_the_plain_starts():
    weights = load_weights()
    assert weights is not None, "experiments_amd/warmstart_weights.npz is missing"
    scene = scenes.batched_piles(n_scenes=512, seed=scenes.SEED + 1, iterations=500)
    res = {}
    for name, mode in (("zero", L.WARM_ZERO), ("previous", L.WARM_PREVIOUS),
                       ("learned", L.WARM_LEARNED)):
        w = scenes.build_world(scene, solver_iterations=10)
        w.use_graph(True)
        w.step(240)  # the training protocol's settling steps
        if mode == L.WARM_LEARNED:
            w.set_warmstart_weights(*weights)
        w.set_solver(500, 1e-4, mode)
        w.use_graph(True)
        # (the first solves to the tolerance start from the impulses of the
        # fixed-sweep settling steps: left out)
        w.step(5)
        r = run_iterations(w, 10)
        res[name] = (r["mean_iters_normal"], r["mean_iters_friction"])
        assert w.stats()["overflow_flags"] == 0
        w.close()
    print("normal / friction sweeps per scene: " +
          ", ".join("%s %.1f / %.1f" % (k, *v) for k, v in res.items()))
    zero, prev, learned = res["zero"][0], res["previous"][0], res["learned"][0]
    assert learned < zero / 1.3        # the network's gain over the plain solver
    assert learned <= prev * 1.03      # and no loss against the previous-step start
