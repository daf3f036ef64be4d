"""The autotuner's cache (host logic): keys by kernel configuration and mesh size
class, JSON persistence."""
import types

from atum_hip.tuning import TuneCache, _size_class


def _fake(key, K):
    return types.SimpleNamespace(kernel_key=key, grid=types.SimpleNamespace(K=K))


def test_size_classes():
    assert _size_class(1) == 0 and _size_class(64) == 6 and _size_class(90) == 6 and _size_class(100) == 7


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "t.json")
    c = TuneCache(path)
    a = _fake(("euler", 2, 4, 1, 2, 1, 3, "fast"), 64)
    c.put(a, 5)
    assert c.get(a) == 5
    assert c.get(_fake(("euler", 2, 4, 1, 2, 1, 3, "fast"), 70)) == 5      # same size class
    assert c.get(_fake(("euler", 2, 4, 1, 2, 1, 3, "fast"), 4096)) is None  # other size class
    assert c.get(_fake(("euler", 2, 5, 1, 2, 1, 3, "fast"), 64)) is None    # other N
    assert TuneCache(path).get(a) == 5


def test_banded_entry_round_trip(tmp_path):
    """A banded solve choice is stored with the MFMA solve's waves per workgroup."""
    path = str(tmp_path / "b.json")
    c = TuneCache(path)
    c.store("band|x", {"kernel": "mfma", "waves": 4})
    assert TuneCache(path).table["band|x"] == {"kernel": "mfma", "waves": 4}
