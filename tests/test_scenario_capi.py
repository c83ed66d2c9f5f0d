"""The library's own scenario readers (nbg_scenario_load_json / _tipsy, C
ABI, n-body_amd/csrc/scenario.cpp) against the Python reader of the same
formats (Simulation.cpp:154-228 JSON with its passive rule, :273-395
tipsy), on the reference's scenarios/threebody.json and test.json, on files
exercising the passive rule and number formats, on tipsy images in both byte
orders with and without header padding, and on malformed input.  Host code
only: runs without a GPU."""

import importlib
import json
import os

import numpy as np
import pytest

from test_scenario import HERE, _records, _tipsy


@pytest.fixture(scope="module")
def scn():
    return importlib.import_module("n-body_amd.scenario")


def _same(nbg, scn, path, kind):
    x, y, z, vx, vy, vz, m, flags = nbg.load_scenario(str(path), kind)
    s = scn.load(str(path)) if kind == "json" else scn.load_tipsy(str(path))
    assert len(m) == s.n
    for a, b in ((x, s.position[:, 0]), (y, s.position[:, 1]), (z, s.position[:, 2]), (vx, s.velocity[:, 0]),
                 (vy, s.velocity[:, 1]), (vz, s.velocity[:, 2]), (m, s.mass)):
        assert np.array_equal(a.view(np.uint32), np.asarray(b, np.float32).view(np.uint32))
    assert np.array_equal((flags & nbg.SCENARIO_POSITION) != 0, s.has_position)
    assert np.array_equal((flags & nbg.SCENARIO_VELOCITY) != 0, s.has_velocity)
    assert np.array_equal((flags & nbg.SCENARIO_ACTIVE) != 0, s.active)
    assert np.array_equal((flags & nbg.SCENARIO_PASSIVE) != 0, s.passive)
    return flags


@pytest.mark.parametrize("name", ["threebody.json", "test.json"])
def test_reference_scenarios(nbg, scn, name):
    flags = _same(nbg, scn, os.path.join(HERE, name), "json")
    if name == "test.json":
        assert not np.any(flags & nbg.SCENARIO_ACTIVE)  # no masses: passive only


def test_passive_rule_and_numbers(nbg, scn, tmp_path):
    doc = {"particles": [
        {"position": {"x": 1, "y": -2.5e-3, "z": 3E2}, "mass": 2.0},
        {"position": {"x": 0, "y": 0, "z": 1}, "mass": 1.0, "passive": False},
        {"position": {"x": 5, "y": 0, "z": 0}, "passive": True},
        {"position": {"x": 0, "y": 5, "z": 0}, "acceleration": {"x": 1, "y": 0, "z": 0}, "passive": False},
        {"velocity": {"x": 0.1, "y": 1e-45, "z": -0.0}, "color": {"r": 1, "g": 0, "b": 0.5}, "name": 'a "q" \u00e9'},
        {"mass": 0.30000001192092896, "sphere": {"radius": 1.0}, "tags": [1, [2, 3], {"k": None}]},
    ]}
    p = tmp_path / "s.json"
    p.write_text(json.dumps(doc, indent=2, ensure_ascii=False))
    flags = _same(nbg, scn, p, "json")
    assert [bool(f & nbg.SCENARIO_PASSIVE) for f in flags] == [True, False, True, True, True, True]


@pytest.mark.parametrize("text", ['{"particles": [{"mass": }]}', '{"particles": 3}', '{"bodies": []}',
                                  '{"particles": [{"position": {"x": 1, "y": 2}}]}', '{"particles": []} x',
                                  '{"particles": [{"mass": "heavy"}]}', '[' * 200])
def test_malformed_json_is_refused(nbg, tmp_path, text):
    p = tmp_path / "bad.json"
    p.write_text(text)
    with pytest.raises(nbg.NbgError):
        nbg.load_scenario(str(p), "json")


def test_missing_file(nbg, tmp_path):
    with pytest.raises(nbg.NbgError):
        nbg.load_scenario(str(tmp_path / "none.json"), "json")


@pytest.mark.parametrize("big,pad", [(False, False), (True, False), (False, True), (True, True)])
def test_tipsy_images(nbg, scn, tmp_path, big, pad):
    rng = np.random.default_rng(5)
    p = tmp_path / "a.bin"
    p.write_bytes(_tipsy(_records(rng, 3, 12), _records(rng, 4, 9), _records(rng, 5, 11), big=big, pad=pad))
    _same(nbg, scn, p, "tipsy")          # stars only
    p.write_bytes(_tipsy(_records(rng, 6, 12), _records(rng, 2, 9), [], big=big, pad=pad))
    _same(nbg, scn, p, "tipsy")          # gas without stars
    # 8 bytes short matches neither the padded nor the unpadded layout (4
    # bytes short of a padded image is a valid unpadded one)
    p.write_bytes(_tipsy(_records(rng, 6, 12), [], [], big=big, pad=pad)[:-8])
    with pytest.raises(nbg.NbgError):
        nbg.load_scenario(str(p), "tipsy")  # size does not match the header
