"""Host logic of the serving layer (CPU, no device): the order in which free
slots take ready requests, and admission under the HBM budget (dry-run
footprints before any device allocation, leases counted until released,
LRU eviction of unleased meshes, waiting and refusal)."""
import importlib.util
import os
import sys
import threading
import time
from types import SimpleNamespace

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _serve_mod():
    if "dkt_serve" in sys.modules:
        return sys.modules["dkt_serve"]
    spec = importlib.util.spec_from_file_location("dkt_serve", os.path.join(REPO, "dendro-kt_amd", "serve.py"))
    m = importlib.util.module_from_spec(spec)
    sys.modules["dkt_serve"] = m
    spec.loader.exec_module(m)
    return m


def _item(nodes, submitted):
    return SimpleNamespace(req=SimpleNamespace(nodes=nodes, submitted=submitted))


def test_select_next_smallest_mesh_then_oldest():
    sv = _serve_mod()
    assert sv.select_next([], 0.0, 0.5) is None
    ready = [_item(500, 1.0), _item(100, 1.2), _item(100, 1.1), _item(900, 0.9)]
    assert sv.select_next(ready, 1.3, 0.5) == 2  # smallest mesh, oldest of the two


def test_select_next_aged_requests_first_in_arrival_order():
    sv = _serve_mod()
    ready = [_item(500, 1.0), _item(100, 1.6), _item(900, 0.2), _item(800, 0.4)]
    # at t = 1.7 the requests of t = 0.2 and 0.4 (and 1.0) waited longer than 0.5 s
    assert sv.select_next(ready, 1.7, 0.5) == 2
    # no starvation: repeatedly taking the choice empties the list, aged first
    order = []
    while ready:
        i = sv.select_next(ready, 1.7, 0.5)
        order.append(ready.pop(i).req.submitted)
    assert order == [0.2, 0.4, 1.0, 1.6]


class _FakeBuild:
    """build(spec) for PlanCache: spec = (nodes, map_bytes); records device builds."""

    def __init__(self):
        self.dry, self.device = [], []

    def __call__(self, spec):
        nodes, map_bytes = spec
        self.dry.append(spec)
        da = SimpleNamespace(n_nodes=nodes)

        def make_plans():
            self.device.append(spec)
            return ("stiffness", "mass")
        return da, map_bytes, make_plans


def test_admission_refuses_before_device_allocation():
    sv = _serve_mod()
    fb = _FakeBuild()
    cache = sv.PlanCache(None, 1000.0, build=fb, wait_timeout=0.2)
    vec = sv.VEC_BYTES_PER_NODE
    with pytest.raises(sv.AdmissionError):
        cache.acquire((10, 1000.0 - 5 * vec))  # maps fit, maps + vectors do not
    assert fb.dry == [(10, 1000.0 - 5 * vec)] and fb.device == []  # footprint from the dry run only
    assert cache.used() == 0.0


def test_leases_count_until_released_and_lru_eviction():
    sv = _serve_mod()
    vec = sv.VEC_BYTES_PER_NODE
    fb = _FakeBuild()
    budget = 3 * (100.0 + vec)
    cache = sv.PlanCache(None, budget, build=fb, wait_timeout=0.3)
    a, built = cache.acquire((1, 100.0))
    assert built and cache.used() == 100.0 + vec
    b, _ = cache.acquire((1, 100.0 + 1e-9))
    a2, built2 = cache.acquire((1, 100.0))  # hit: a second lease on one mesh adds its vectors only
    assert not built2 and cache.used() == 2 * (100.0 + vec) + 1e-9 + vec
    # a third mesh does not fit while every entry is leased: the request waits, then is refused
    t0 = time.monotonic()
    with pytest.raises(sv.AdmissionError, match="timed out"):
        cache.acquire((1, 2 * vec + 50.0))
    assert time.monotonic() - t0 >= 0.25
    assert fb.device.count((1, 2 * vec + 50.0)) == 0
    # released leases make entries evictable (least recently used first)
    a.release()
    a2.release()
    a.release()  # idempotent
    c, built = cache.acquire((1, 2 * vec + 50.0))
    assert built and (1, 100.0) not in cache.entries and (1, 100.0 + 1e-9) in cache.entries
    assert cache.used() <= budget
    b.release()
    c.release()
    assert cache.used() == sum(e["map_bytes"] for e in cache.entries.values())


def test_waiting_request_admitted_when_a_lease_ends():
    sv = _serve_mod()
    vec = sv.VEC_BYTES_PER_NODE
    fb = _FakeBuild()
    cache = sv.PlanCache(None, 2 * (102.0 + vec), build=fb, wait_timeout=5.0)
    a, _ = cache.acquire((1, 100.0))
    b, _ = cache.acquire((1, 101.0))
    got = {}

    def third():
        got["lease"] = cache.acquire((1, 102.0))[0]
    th = threading.Thread(target=third)
    th.start()
    time.sleep(0.2)
    assert "lease" not in got  # waiting: both entries are leased
    a.release()
    th.join(5.0)
    assert "lease" in got and (1, 100.0) not in cache.entries
    b.release()
    got["lease"].release()


def test_concurrent_misses_build_once():
    sv = _serve_mod()
    fb = _FakeBuild()
    cache = sv.PlanCache(None, 1e9, build=fb)
    leases = []
    ths = [threading.Thread(target=lambda: leases.append(cache.acquire((7, 10.0)))) for _ in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert fb.device == [(7, 10.0)] and sum(1 for _, built in leases if built) == 1
    for lease, _ in leases:
        lease.release()


def _autotune_mod():
    spec = importlib.util.spec_from_file_location("dkt_autotune_host", os.path.join(REPO, "dendro-kt_amd",
                                                                                   "autotune.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_autotune_cache_key_and_file_round_trip(tmp_path):
    """The autotuner's cache is keyed by (dim, order, columns, operator, size
    class), the size class being the element count's power of four, and a
    saved cache loads back to the same entries."""
    at = _autotune_mod()
    assert [at.size_class(n) for n in (1, 3, 4, 15, 16, 1000, 1 << 20)] == [0, 0, 1, 1, 2, 4, 10]
    entry = {"best": {"f64": 128, "f32": 256}, "columns": {"f64": 0, "f32": 1},
             "ms": {(128, 0): {"f64": 1.0, "f32": 0.7}, (256, 1): {"f64": 1.2, "f32": 0.6}}}
    at._cache[(4, 1, 8, 0, 9)] = entry
    at._cache[(3, 2, 1, 1, 5)] = dict(entry, columns={"f64": 0, "f32": 0})
    path = str(tmp_path / "tune.json")
    at.save(path)
    saved = dict(at._cache)
    at._cache.clear()
    at.load(path)
    assert at._cache == saved
