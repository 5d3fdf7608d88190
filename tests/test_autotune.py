"""The launch-shape autotuner's host side (autotune.py), CPU only: cache
keys, the JSON cache round trip and its tolerance of a damaged file, and the
candidate sets a device's compute-unit count allows."""
import json

from multicellular_phase_field_active_gel_organoids_amd import autotune


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "sub" / "tune.json")
    assert autotune.load_cache(p) == {}
    k = autotune.cache_key("AMD Instinct MI355X|gfx950|cu=256", 1024, 64, 65)
    assert k == "AMD Instinct MI355X|gfx950|cu=256|N=1024|cells=64|zone=65"
    autotune.save_cache(p, {k: {"tuning": {"lum_blocks": 48, "roles": 128, "graph_steps": 16}}})
    assert autotune.load_cache(p)[k]["tuning"]["roles"] == 128
    open(p, "w").write("{not json")
    assert autotune.load_cache(p) == {}
    open(p, "w").write(json.dumps([1, 2]))
    assert autotune.load_cache(p) == {}


def test_cache_path_from_environment(monkeypatch, tmp_path):
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "x.json"))
    assert autotune.default_cache_path() == str(tmp_path / "x.json")


def test_candidates_respect_the_device():
    c = autotune.candidates(80)
    assert max(c["lum_blocks"]) <= 80 and 64 in c["lum_blocks"]
    assert max(autotune.candidates(256)["lum_blocks"]) == 128
    assert all(1 <= v <= 64 for v in c["graph_steps"]) and all(1 <= v <= 512 for v in c["roles"])
