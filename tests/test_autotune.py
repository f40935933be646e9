"""Autotuner selection logic on synthetic timing tables (CPU)."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def at():
    return importlib.import_module("1dsweptcuda_amd.autotune")


def test_size_classes(at):
    assert at.size_class(1024) == 10
    assert at.size_class(1025) == 10
    assert at.size_class(2047) == 10
    assert at.size_class(65535) == 15
    assert at.size_class(65536) == 16


def test_choose_picks_fastest(at):
    t = {10: {128: 3.0e-12, 256: 2.0e-12, 512: 2.5e-12},
         14: {128: 9e-12, 256: 5e-12, 512: 1e-12}}
    assert at.choose(t) == {10: 256, 14: 512}


def test_choose_tie_goes_to_wider_tile(at):
    t = {12: {128: 1.00e-12, 256: 1.01e-12, 512: 1.5e-12}}
    assert at.choose(t) == {12: 256}
    assert at.choose(t, tie=0.0) == {12: 128}


def test_choose_is_deterministic_and_ignores_missing(at):
    t = {11: {"128": 2e-12, "256": None, "512": 0}}
    assert at.choose(t) == at.choose(dict(t)) == {11: 128}
    assert at.choose({13: {}}) == {}


def test_tile_points_for_fills_untimed_classes(at):
    table = {10: 128, 13: 512}
    nx = np.array([1024, 2000, 9000, 70000, 500])
    assert list(at.tile_points_for(nx, table)) == [128, 128, 512, 512, 128]
    assert list(at.tile_points_for(nx, {})) == [256] * 5


def test_save_load_roundtrip(at, tmp_path):
    p = tmp_path / "t.json"
    at.save({10: 128, 11: 256}, {10: {128: 1e-12}}, str(p), meta={"device": "test"})
    assert at.load(str(p)) == {10: 128, 11: 256}
    assert at.load(str(tmp_path / "missing.json")) == {}


def test_workload_generator_is_seeded(sw):
    wl = importlib.import_module("1dsweptcuda_amd.workloads")
    a = wl.batch_problems(64, seed=3)
    b = wl.batch_problems(64, seed=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    eq, nx, dt, dx = a
    assert ((nx >= 1024) & (nx < 65536)).all()
    # 0.8x the stability limits: heat Fo = 0.4, KS z_min = -1.6
    heat = eq == 0
    assert np.allclose(8.418e-5 * dt[heat] / dx[heat] ** 2, 0.4)
    ks = ~heat
    assert np.allclose(dt[ks] * (16 / dx[ks] ** 4 - 4 / dx[ks] ** 2), 1.6)


def test_choose_waves(at):
    # the fastest wins; the default 4 wins a near tie; missing / bad entries are ignored
    assert at.choose_waves({2: 3e-12, 4: 2e-12, 8: 1e-12}) == 8
    assert at.choose_waves({2: 1.00e-12, 4: 1.01e-12, 8: 1.5e-12}) == 4
    assert at.choose_waves({2: 1.00e-12, 4: 1.01e-12}, tie=0.0) == 2
    assert at.choose_waves({"2": 1e-12, "8": 1e-12}) == 2  # exact tie without the default: fewer waves
    assert at.choose_waves({}) == at.DEFAULT_WAVES == 4
    assert at.choose_waves({2: None, 8: 0}) == 4


def test_save_load_waves(at, tmp_path):
    p = tmp_path / "t.json"
    at.save({10: 128}, {10: {128: 1e-12}}, str(p), waves=8, wave_timings={2: 3e-12, 4: 2e-12, 8: 1e-12})
    assert at.load_waves(str(p)) == 8 and at.load(str(p)) == {10: 128}
    assert at.load_waves(str(tmp_path / "missing.json")) == 4


def test_node_width_lookup_and_round_trip(at, tmp_path):
    """Single-plan node widths: exact size class, nearest class below, else
    above; no entry -> 0 (the caller's tpb); the JSON table round-trips."""
    table = {(1, 0): {14: 128, 18: 256, 20: 1024}, (2, 0): {16: 256}}
    assert at.node_width_for(1, 1 << 14, 0, table) == 128
    assert at.node_width_for(1, 1 << 19, 0, table) == 256   # class 19: nearest below is 18
    assert at.node_width_for(1, 3 << 18, 0, table) == 256   # 786432 points: class 19
    assert at.node_width_for(1, 3 << 19, 0, table) == 1024  # 1572864 points: class 20
    assert at.node_width_for(1, 1 << 12, 0, table) == 128   # below every class -> lowest
    assert at.node_width_for(2, 1 << 20, 0, table) == 256
    assert at.node_width_for(0, 1 << 20, 0, table) == 0
    assert at.node_width_for(1, 1 << 20, 1, table) == 0
    timings = {k: {c: {w: 1e-12} for c, w in r.items()} for k, r in table.items()}
    path = str(tmp_path / "nw.json")
    at.save_node_widths(table, timings, path, meta={"test": True})
    assert at.load_node_widths(path) == table
    assert at.load_node_widths(str(tmp_path / "missing.json")) == {}


def test_node_width_choice_from_timings(at):
    """The per-class choice is `choose`: fastest, ties within 2% to the wider node."""
    rows = {14: {128: 1.00e-12, 256: 1.015e-12, 512: 1.3e-12},
            20: {512: 1.2e-12, 1024: 1.1e-12}}
    assert at.choose(rows) == {14: 256, 20: 1024}
