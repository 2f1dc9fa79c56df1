"""Batched stepping: every scene of a batch advances exactly as it does
alone (same kernels per scene, one launch per pass over all scenes)."""
import numpy as np
import pytest

import naga_amd
from tests.scenes import gpu_engine, grid_nodes, rel_l2

pytestmark = pytest.mark.gpu


def _scene(oracle64, D, h, k, layer_cells, centre):
    nodes = grid_nodes(oracle64, D, h, layer=layer_cells * h)
    e = gpu_engine(nodes, 9 if D == 2 else 27, k, h)
    e.set_initial_pulse(centre, 1e-3, 0.1)
    return e


def _specs():
    return [(2, 0.02, 16, 5, [0.5, 0.5]), (2, 0.025, 12, 4, [0.3, 0.6]),
            (2, 0.03125, 20, 3, [0.6, 0.4]), (2, 0.05, 16, 0, [0.5, 0.5]),
            (3, 0.1, 32, 2, [0.5, 0.4, 0.5]), (3, 0.125, 32, 2, [0.4, 0.5, 0.5])]


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_alone(oracle64, graph):
    alone = [_scene(oracle64, *s) for s in _specs()]
    mixed = [_scene(oracle64, *s) for s in _specs()]
    batch = naga_amd.SceneBatch(mixed, graph=graph)
    batch.step_forward(25)
    for e in alone:
        e.step_forward(25)
    for a, b in zip(alone, mixed):
        ma, mb = a.macroscopic(fresh=True), b.macroscopic(fresh=True)
        assert abs(a.time() - b.time()) < 1e-12
        assert rel_l2(mb["drho"], ma["drho"]) < 1e-6
        assert rel_l2(mb["u"], ma["u"]) < 1e-6


def test_batch_membership_changes(oracle64):
    """Scenes join and leave between steps (as a serving loop does)."""
    specs = _specs()
    ref = [_scene(oracle64, *s) for s in specs[:3]]
    for e in ref:
        e.step_forward(30)
    run = [_scene(oracle64, *s) for s in specs[:3]]
    batch = naga_amd.SceneBatch([run[0]])
    batch.step_forward(10)           # scene 0: 10
    batch.add(run[1])
    batch.step_forward(10)           # scene 0: 20, scene 1: 10
    batch.add(run[2])
    batch.remove(run[0])
    run[0].step_forward(10)          # scene 0 alone: 30
    batch.step_forward(20)           # scene 1: 30, scene 2: 20
    batch.remove(run[2])
    run[2].step_forward(10)          # scene 2: 30
    for a, b in zip(ref, run):
        assert rel_l2(b.macroscopic(fresh=True)["drho"], a.macroscopic(fresh=True)["drho"]) < 1e-6


def test_batch_with_density_sources(oracle64):
    specs = _specs()[:2]
    alone = [_scene(oracle64, *s) for s in specs]
    mixed = [_scene(oracle64, *s) for s in specs]
    for e in alone + mixed:
        e.register_density_source(naga_amd.DensitySource([0.5, 0.5], 0.1, 1e-3, 5.0))
    batch = naga_amd.SceneBatch(mixed)
    batch.step_forward(15)
    for e in alone:
        e.step_forward(15)
    for a, b in zip(alone, mixed):
        assert rel_l2(b.macroscopic(fresh=True)["drho"], a.macroscopic(fresh=True)["drho"]) < 1e-6


def test_state_save_load_roundtrip(oracle64):
    """save_state/load_state: resuming from a snapshot repeats the run
    exactly, also in a second engine built from the same nodes."""
    s = _specs()[0]
    a = _scene(oracle64, *s)
    a.step_forward(10)
    snap = a.save_state()
    a.step_forward(10)
    r1 = a.macroscopic(fresh=True)
    a.load_state(snap)
    assert abs(a.time() - 10 * 0.4 * s[1]) < 1e-12
    a.step_forward(10)
    assert np.array_equal(a.macroscopic(fresh=True)["drho"], r1["drho"])
    b = _scene(oracle64, *s)
    b.load_state(snap)
    b.step_forward(10)
    assert np.array_equal(b.macroscopic(fresh=True)["drho"], r1["drho"])
    c = _scene(oracle64, *_specs()[1])
    with pytest.raises(Exception):
        c.load_state(snap)


def test_autotune_leaves_scenes_unchanged(oracle64):
    alone = [_scene(oracle64, *s) for s in _specs()]
    mixed = [_scene(oracle64, *s) for s in _specs()]
    batch = naga_amd.SceneBatch(mixed)
    gb = batch.autotune(3)
    assert gb in (8, 16, 32)
    t = batch.tuning()
    assert t["gather_batch"] == gb and (t["q9_lanes"], t["q9_unions"]) in ((3, True), (3, False),
                                                                            (9, False))
    batch.step_forward(12)
    for e in alone:
        e.step_forward(12)
    for a, b in zip(alone, mixed):
        assert a.time() == pytest.approx(b.time(), abs=1e-12)
        assert rel_l2(b.macroscopic(fresh=True)["drho"], a.macroscopic(fresh=True)["drho"]) < 1e-6


@pytest.mark.parametrize("form", [(3, False), (9, False)])
def test_q9_kernel_forms(oracle64, monkeypatch, form):
    """The autotuner's D2Q9 kernel forms advance every scene alike: three
    distributions per lane reading the tile unions staged in LDS (scenes
    set up with NAGA_Q9_UNIONS=1), the same gathering directly, one
    distribution per lane."""
    monkeypatch.setenv("NAGA_Q9_UNIONS", "1")
    specs = [s for s in _specs() if s[0] == 2]
    runs = {}
    for lanes, unions in ((3, True), form):
        scenes = [_scene(oracle64, *s) for s in specs]
        batch = naga_amd.SceneBatch(scenes)
        batch.set_q9_kernels(lanes, unions)
        t = batch.tuning()
        assert (t["q9_lanes"], t["q9_unions"]) == (lanes, unions)
        batch.step_forward(20)
        runs[(lanes, unions)] = [e.macroscopic(fresh=True) for e in scenes]
    with pytest.raises(ValueError):
        batch.set_q9_kernels(4)
    with pytest.raises(ValueError):
        batch.set_q9_kernels(9, True)
    for a, b in zip(runs[(3, True)], runs[form]):
        assert rel_l2(a["drho"], b["drho"]) < 1e-6
        assert rel_l2(a["u"], b["u"]) < 1e-6


def test_batch_mixes_union_and_direct_scenes(oracle64, monkeypatch):
    """D3Q27 scenes built with and without tile unions (different slab
    layouts) step in separate groups of one batch, each as it does alone."""
    specs = [(3, 0.1, 32, 2, [0.5, 0.4, 0.5]), (3, 0.125, 32, 2, [0.4, 0.5, 0.5])]
    scenes, alone = [], []
    for flag, s in zip(("1", "0"), specs):
        monkeypatch.setenv("NAGA_UNIONS", flag)
        scenes.append(_scene(oracle64, *s))
        alone.append(_scene(oracle64, *s))
    batch = naga_amd.SceneBatch(scenes)
    batch.step_forward(15)
    for e in alone:
        e.step_forward(15)
    for a, b in zip(alone, scenes):
        ma, mb = a.macroscopic(fresh=True), b.macroscopic(fresh=True)
        assert rel_l2(mb["drho"], ma["drho"]) < 1e-6
    batch.close()


def test_batch_sources_cleared_and_added_in_a_batch(oracle64):
    """Density sources added or cleared while the engine is in a batch take
    effect at the next batched step (the batch rebuilds its scene table):
    a scene whose source is cleared after 5 steps ends as the same scene
    run alone with the same history, and differs from one that keeps it."""
    spec = _specs()[0]
    src = naga_amd.DensitySource([0.5, 0.5], 0.1, 1e-3, 5.0)
    alone, kept = _scene(oracle64, *spec), _scene(oracle64, *spec)
    mixed = _scene(oracle64, *spec)
    other = _scene(oracle64, *_specs()[1])
    batch = naga_amd.SceneBatch([mixed, other])
    batch.step_forward(2)                  # no source yet: graph replay
    mixed.register_density_source(src)     # added while in the batch
    batch.step_forward(5)
    mixed.clear_density_sources()          # cleared while in the batch
    batch.step_forward(6)
    for e in (alone, kept):
        e.step_forward(2)
        e.register_density_source(src)
        e.step_forward(5)
    alone.clear_density_sources()
    alone.step_forward(6)
    kept.step_forward(6)
    got = mixed.macroscopic(fresh=True)["drho"]
    assert rel_l2(got, alone.macroscopic(fresh=True)["drho"]) < 1e-6
    assert rel_l2(got, kept.macroscopic(fresh=True)["drho"]) > 1e-3
    batch.close()
