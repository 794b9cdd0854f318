"""Seeded mixed scenes (scenes.mixed: 94-316 tape records; spheres, boxes, slabs, products of two
registers, divisions with a literal numerator, MIN/MAX against literals, a square root of |u v|,
a sine whose operand has two uses) against the oracle, bit for bit, every frame of a temporal
sequence: the compiled engine's unpruned modules (subtrees skipped where their sensitivity is
zero across a wavefront, operands recomputed instead of saved), its pruned modules and the
interpreter.  For the full programs the work counts must equal the oracle's too: a skipped
subtree counts exactly what its records would have counted."""
import pytest

from helpers import assert_same_nodes, run_pair, tape_of
from temporalsetinversion_amd import scenes

pytestmark = pytest.mark.gpu

SEQ = [k / 60.0 for k in range(0, 6)] + [0.5, 0.75]

ENGINES = {"compiled-full": dict(engine="compiled", prune=False),
           "compiled-pruned": dict(engine="compiled", prune=True, prune_live_max=100),
           "interp-full": dict(engine="interpreter", prune=False)}


@pytest.mark.parametrize("engine", sorted(ENGINES))
@pytest.mark.parametrize("seed,n", scenes.mixed_scenes())
def test_mixed_scenes_match_oracle(seed, n, engine):
    tape = tape_of("mixed", seed=seed, n=n)
    swept = seed % 2 == 0          # every cell takes the reverse pass
    for k, g, o, gs, os_, _, _ in run_pair(tape, SEQ, localsubdiv=5, swept_volume=swept, **ENGINES[engine]):
        print(f"mixed seed={seed} n={n} engine={engine} swept={swept} frame={k} nodes={len(o)} "
              f"cells={os_.cell_evals} rev_ops={gs['rev_ops']}/{os_.rev_ops}", flush=True)
        assert_same_nodes(g, o, f"mixed {seed} frame {k}")
        assert gs["cell_evals"] == os_.cell_evals
        if engine != "compiled-pruned":
            assert gs["rev_evals"] == os_.rev_evals
            assert gs["fwd_ops"] == os_.fwd_ops
            assert gs["rev_ops"] == os_.rev_ops
