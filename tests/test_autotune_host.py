"""CPU tests of the autotuner's selection and caching logic, with a stand-in
plan whose launch timings are a fixed table (no library or GPU calls): the
two-step shape and x tiling that time fastest are applied, recorded in the
cache, and restored from it."""
import importlib
import json

# the package exports the autotune() function under the module's name
at = importlib.import_module("awcpml.autotune")


class FakePlan:
    """Duck-typed stand-in for awcpml.Plan: times come from `table`."""

    def __init__(self, ndim=3, dtype=1, two_ms=None, one_ms=1.0):
        self.ndim, self.dtype, self.nshots, self.halo = ndim, dtype, 8, 20
        self.n = [298] * ndim
        self.two_ms = two_ms or {}
        self.one_ms = one_ms
        self.align = dtype == 1
        self.tb = (1, 0, 0, 0)
        self.tile = (4, 64, 3)
        self.group = 1
        self.calls = []

    # launch shapes
    def set_tiling(self, by, ml, pf):
        self.tile = (by, ml, pf)

    def set_shot_group(self, g):
        self.group = g

    def set_time_block(self, steps, rows=0, march=0, prefetch=0):
        self.tb = (steps, rows, march, prefetch)

    def set_two_step_align(self, on):
        self.align = bool(on)
        self.calls.append(("align", bool(on)))

    def two_step_align(self):
        return self.align

    def reset(self):
        pass

    # timings (ms for n launches)
    def time_stencil(self, n):
        return self.one_ms * n

    def time_stencil2(self, n):
        key = (self.tb[1], self.tb[2], self.tb[3], self.align)
        return self.two_ms.get(key, 10.0) * n


def test_autotune_picks_fastest_tiling_and_caches(tmp_path):
    cache = tmp_path / "tuning.json"
    # two-step launch times per launch (two steps): aligned tiles faster for
    # one shape, 62-node tiles faster for another; the overall best is aligned
    ms = {(16, 144, 3, True): 1.49, (16, 144, 3, False): 1.62,
          (8, 24, 2, True): 1.70, (8, 24, 2, False): 1.55}
    P = FakePlan(two_ms=ms, one_ms=0.95)
    best = at.autotune(P, cache=str(cache))
    assert not best["cached"]
    assert best["time_block"] == {"block_rows": 16, "march_len": 144, "prefetch": 3, "align": True}
    assert abs(best["ms_per_step"] - 0.745) < 1e-12
    assert P.tb == (2, 16, 144, 3) and P.align is True
    rec = json.loads(cache.read_text())
    assert list(rec.values())[0]["time_block"]["align"] is True

    # a later plan of the same key takes the cached choice, tiling included
    Q = FakePlan(two_ms=ms, one_ms=0.95)
    Q.align = False
    again = at.autotune(Q, cache=str(cache))
    assert again["cached"]
    assert Q.tb == (2, 16, 144, 3) and Q.align is True


def test_autotune_keeps_62_node_tiles_when_faster(tmp_path):
    ms = {(8, 24, 2, True): 0.52, (8, 24, 2, False): 0.44}
    P = FakePlan(dtype=0, two_ms=ms, one_ms=0.33)
    best = at.autotune(P, cache=str(tmp_path / "t.json"))
    assert best["time_block"]["align"] is False and P.align is False
    assert P.tb == (2, 8, 24, 2)


def test_autotune_keeps_single_steps_when_faster(tmp_path):
    ms = {(16, 144, 3, True): 3.0}
    P = FakePlan(two_ms=ms, one_ms=0.9)
    align0 = P.align
    best = at.autotune(P, cache=str(tmp_path / "t.json"))
    assert best["time_block"] is None
    assert P.tb[0] == 1
    assert P.align == align0  # the plan's own tiling setting is left as it was


def test_cached_entry_without_tiling_leaves_it(tmp_path):
    # cache entries written before the tiling choice existed have no "align"
    P = FakePlan()
    key = at._key(P)
    cache = tmp_path / "old.json"
    cache.write_text(json.dumps({key: {"block_y": 4, "march_len": 64, "prefetch": 3, "shot_group": 1,
                                       "ms_per_launch": 1.0, "ms_per_step": 0.8, "tried": 3,
                                       "time_block": {"block_rows": 16, "march_len": 96,
                                                      "prefetch": 3}}}))
    best = at.autotune(P, cache=str(cache))
    assert best["cached"] and P.tb == (2, 16, 96, 3)
    assert not any(c[0] == "align" for c in P.calls)
