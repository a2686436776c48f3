"""The trained warm-start network shipped with the package
(experiments_amd/warmstart_weights.npz) on held-out batched piles (another
seed than its training scenes): every solve runs to a 1e-4 residual, and the
sweeps per scene and step are compared for the zero start (the plain
solver), the previous-step start and the learned start over the scene-steps
that every start solved to the tolerance.  The learned start must need far
fewer normal sweeps than the plain solver and no more than the previous-step
start; the friction phase, whose sweeps the reference's spin damping sets, is
reported beside them."""
import pytest

from experiments_amd import scenes
from experiments_amd.warmstart import compare_starts, load_weights

pytestmark = pytest.mark.gpu


def test_trained_network_against_the_plain_starts():
    weights = load_weights()
    assert weights is not None, "experiments_amd/warmstart_weights.npz is missing"
    scene = scenes.batched_piles(n_scenes=512, seed=scenes.SEED + 1, iterations=10)
    # the training protocol's 240 settling steps, then 5 + 10 solves to 1e-4
    res = compare_starts(scene, weights, settle=240, steps=10, tol=1e-4, max_sweeps=4096)
    assert all(res[k]["overflow_flags"] == 0 for k in ("zero", "previous", "learned"))
    conv = res["converged_in_every_start"]
    print("scene-steps converged in every start %d of %d; normal / all sweeps: %s" %
          (conv["scene_steps"], conv["of"],
           ", ".join("%s %.1f / %.1f" % (k, conv[k]["normal_sweeps"], conv[k]["sweeps"])
                     for k in ("zero", "previous", "learned"))))
    assert conv["scene_steps"] >= 0.95 * conv["of"]
    zero, prev, learned = (conv[k]["normal_sweeps"] for k in ("zero", "previous", "learned"))
    assert learned < zero / 1.3        # the network's gain over the plain solver
    assert learned <= prev * 1.03      # and no loss against the previous-step start
