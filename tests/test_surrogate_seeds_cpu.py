"""Training and evaluation crops of the warm start never overlap (CPU).

Crop i of `surrogate.make_crops(n, seed0)` uses seed seed0 + i, so a training
set of n crops from TRAIN_SEED0 covers seeds TRAIN_SEED0 .. TRAIN_SEED0 + n - 1
and must stop short of EVAL_SEED0, where every evaluation (the pipeline, the
CLI, the benchmark's warm-start leg, the GPU warm-start test) draws its crops.
The pipeline refuses a training set that would reach them, before it
generates a crop or touches the GPU.
"""
import numpy as np
import pytest

from lbflow_amd import surrogate as S


def test_seed_ranges_are_disjoint_for_the_shipped_training_sets():
    for n_train in (4096, 16384, 32768):
        assert S.TRAIN_SEED0 + n_train <= S.EVAL_SEED0


def test_pipeline_refuses_a_training_set_reaching_the_evaluation_seeds():
    n = S.EVAL_SEED0 - S.TRAIN_SEED0 + 1
    with pytest.raises(ValueError, match='evaluation seeds'):
        S.pipeline(n_train=n, n_eval=1, side=16)


def test_crop_seed_is_seed0_plus_index():
    a = S.make_crops(3, 7, side=16, workers=1)
    b = S.make_crops(1, 9, side=16, workers=1)
    assert np.array_equal(a[2], b[0])
    assert not np.array_equal(a[0], a[1])
