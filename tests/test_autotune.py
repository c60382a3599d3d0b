"""Autotuner host logic (CPU): candidate selection, cache keys and the JSON
cache (round trip, per-device entries, version check, atomic replace)."""
import json
import types

import __graft_entry__

__graft_entry__.load_package()
from jgcm_amd import autotune  # noqa: E402


def test_select_fastest_per_direction_and_ties():
    fwd = {1: 30.0, 2: 27.5, 4: 26.0, 8: 26.0, 16: 29.0}
    inv = {1: 31.0, 2: 28.0, 4: 27.9, 8: 27.1, 16: 27.4}
    assert autotune.select(fwd, inv) == (4, 8)      # tie 4/8 -> the smaller shape
    assert autotune.select({2: 1.0}, {16: 2.0}) == (2, 16)


def test_tuning_key_names_truncation_grid_precision_and_batch():
    m = types.SimpleNamespace(num_fourier=85, nθ=128, nλ=256, precision="split32")
    assert autotune.tuning_key(m, 61) == "T85|128x256|split32|61"


def test_cache_round_trip_and_devices(tmp_path):
    path = str(tmp_path / "sub" / "tuning.json")
    c = autotune.Tuning_Cache(path)
    assert c.get("devA", "k") is None
    c.put("devA", "k", {"fold_ppw": 4, "unfold_ppw": 8})
    c.put("devB", "k", {"fold_ppw": 2, "unfold_ppw": 2})
    d = autotune.Tuning_Cache(path)
    assert d.get("devA", "k") == {"fold_ppw": 4, "unfold_ppw": 8}
    assert d.get("devB", "k")["fold_ppw"] == 2
    raw = json.load(open(path))
    assert raw["version"] == autotune.VERSION and set(raw["entries"]) == {"devA", "devB"}
    assert not (tmp_path / "sub" / "tuning.json.tmp").exists()


def test_cache_of_another_version_is_ignored(tmp_path):
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"version": 999, "entries": {"d": {"k": {"fold_ppw": 1}}}}))
    c = autotune.Tuning_Cache(str(path))
    assert c.get("d", "k") is None


def test_default_path_follows_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("JGCM_TUNING_FILE", str(tmp_path / "x.json"))
    assert autotune.default_cache_path() == str(tmp_path / "x.json")
    monkeypatch.delenv("JGCM_TUNING_FILE")
    assert autotune.default_cache_path().endswith("jgcm_amd/tuning.json")


def test_tile_candidates_are_the_library_shapes():
    # jgcm_mesh_set_legendre_shape accepts 0 (built-in), 1, 2 and 4
    assert autotune.TILE_CANDIDATES == (1, 2, 4)
    assert autotune.select({1: 24.0, 2: 23.0, 4: 23.0}, {1: 20.0, 2: 21.0, 4: 22.0}) == (2, 1)


def test_cache_of_the_previous_format_is_ignored(tmp_path):
    # entries of the FFT-only format carry no Legendre tiles; they are not read
    path = tmp_path / "t.json"
    old = {"version": 1, "entries": {"d": {"k": {"fold_ppw": 4, "unfold_ppw": 8}}}}
    path.write_text(json.dumps(old))
    assert autotune.VERSION != 1
    assert autotune.Tuning_Cache(str(path)).get("d", "k") is None
