"""Host-side checks of the batched-crop entry points (no GPU needed).

lbf_batch_create validates the description, the labels and every crop before
it touches the device, with the reference's messages where one exists
(solver/src/file/fileprm.cpp:136-295; an undefined label is the std::map::at
failure of Simulator's constructor, simulator.cpp:30).
"""
import numpy as np
import pytest

from lbflow_amd.engine import Batch, EngineError

LABELS = {75: ('GV', 1e-5, 1, 0), 255: ('SD', 90.0, 0, 0)}


def crop(shape=(4, 5, 6)):
    g = np.full(shape, 75, np.uint8)
    g[:, 0, :] = 255
    return g


def test_rejects_unsupported_lattice():
    with pytest.raises(EngineError, match='Only D2Q9 and D3Q19'):
        Batch(3, 15, 'MRT', 0.8, 1.0, LABELS, [crop()])


def test_rejects_relaxation_time():
    with pytest.raises(EngineError, match='relaxation times must exceed 0.5'):
        Batch(3, 19, 'MRT', 0.5, 1.0, LABELS, [crop()])


def test_rejects_undefined_label():
    g = crop()
    g[1, 2, 3] = 7
    with pytest.raises(EngineError, match='map::at'):
        Batch(3, 19, 'MRT', 0.8, 1.0, LABELS, [crop(), g])


def test_rejects_boundary_labels():
    labels = {**LABELS, 9: ('PB', 1.0, 1, 0)}
    g = crop()
    g[0, 1, 1] = 9
    with pytest.raises(EngineError, match='VD/GV fluid and BB/SD solids'):
        Batch(3, 19, 'SRT', 0.8, 1.0, labels, [g])


def test_rejects_dimension_mismatch():
    with pytest.raises(ValueError):
        Batch(2, 9, 'SRT', 0.8, 1.0, LABELS, [crop()])


def test_empty_batch():
    with pytest.raises(EngineError, match='no crops'):
        Batch(3, 19, 'SRT', 0.8, 1.0, LABELS, [])
