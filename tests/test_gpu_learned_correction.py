"""The learned correction end to end: training data made by the method
itself (PM, TreePM = long range + exact short range), training on the GPU,
evaluation on held-out seeds against TreePM."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_learned_correction_beats_plain_pm_on_held_out_boxes(pm):
    from gadget_ng_pm_hip import learned
    print("case learned correction 32^3 boxes", flush=True)
    lc = learned.LearnedCorrection(cells=32, batch=8)
    info = lc.build_training_set(range(1, 25), workers=8)
    print(info, flush=True)
    hist = lc.train(steps=600, lr=2e-3, log_every=100, log=lambda s: print(s, flush=True))
    assert hist[-1] < hist[0]
    for seed in (1001, 1002):
        ev = lc.evaluate(seed, workers=8)
        print(seed, "PM", ev["pm"], flush=True)
        print(seed, "PM+net", ev["pm_net"], flush=True)
        assert ev["pm_net"]["field_mean_rel"] < ev["pm"]["field_mean_rel"]
        assert ev["pm_net"]["field_median_rel"] < ev["pm"]["field_median_rel"]
