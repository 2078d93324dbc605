"""The GPU's validate mode (rayna_set_validate / rayna_validate_counts): the engine's debug invariants
(rayna_engine/src/shared/validate.rs:13-198) counted by the kernels.  Parity scenes break none and render the
same image as without validation; scenes that break them on purpose give the oracle's counts exactly."""
import numpy as np
import pytest

import oracle
from tests.test_oracle_validate import CASES, ZERO, light_wall, negative_ball
from rayna_amd import Renderer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(CASES))
def test_parity_scenes_break_no_invariant(gpu, name):
    sc, cam, opts = CASES[name]()
    print("case validate", name, flush=True)
    ref, _, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    with Renderer(sc, cam, opts) as r:
        r.set_validate(True)
        img = r.render(layout="rowmajor").img
        counts = r.validate_counts()
    assert counts == ZERO, counts
    assert np.array_equal(img, ref)


@pytest.mark.parametrize("case", ["negative_emission", "negative_albedo"])
def test_planted_violations_count_as_the_oracle(gpu, case):
    sc, cam, opts = light_wall((-1.0, 0.5, 0.5)) if case == "negative_emission" else negative_ball()
    print("case validate", case, flush=True)
    _, want = oracle.TwinScene(sc).render_checked(cam, opts, 32)
    with Renderer(sc, cam, opts) as r:
        r.render()  # validation off: nothing counted
        assert r.validate_counts() == ZERO
        r.set_validate(True)
        r.render()
        assert r.validate_counts(reset=True) == want
        assert r.validate_counts() == ZERO
        r.set_validate(False)
        r.render()
        assert r.validate_counts() == ZERO
    assert want["sample"] == opts.width * opts.height * opts.samples
