"""The autotuner's choices persist in a JSON file that later processes read (host logic only)."""
import json
import types

from hotspice_amd import tune


def _fresh(monkeypatch, path):
    monkeypatch.setenv("HOTSPICE_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(tune, "_CACHE", {})
    monkeypatch.setattr(tune, "_LOADED", False)


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    _fresh(monkeypatch, path)
    assert tune.cache_path() == str(path)
    tune._load()
    tune._store("neel|64|64", 8)
    tune._store("metropolis|2048|2048", 4096)
    assert json.loads(path.read_text()) == {"neel|64|64": 8, "metropolis|2048|2048": 4096}
    _fresh(monkeypatch, path)  # a new process
    tune._load()
    assert tune._CACHE == {"neel|64|64": 8, "metropolis|2048|2048": 4096}


def test_unreadable_cache_is_ignored(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    _fresh(monkeypatch, path)
    tune._load()
    assert tune._CACHE == {}
    tune._store("k", 1)
    assert json.loads(path.read_text()) == {"k": 1}


def test_candidates():
    mm = types.SimpleNamespace(nx=2048, ny=2048, n=2048 * 2048)
    assert tune.metro_candidates(mm) == [1024, 2048, 4096, 8192, 16384]
    small = types.SimpleNamespace(nx=200, ny=200, n=40000)
    assert tune.metro_candidates(small) == [175]  # at most one workgroup per 32 x 8 tile
    g = tune.neel_candidates(types.SimpleNamespace(n=65536), 256)
    assert g[0] == 256 and all(1 <= x <= 256 for x in g) and len(set(g)) == len(g)
