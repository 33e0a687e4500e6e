"""The autotuner's search (galaxy_amd/autotune.py) on a synthetic cost: no GPU.  The
coordinate descent repeats its pass while a pass improves, so parameters whose best
values interact (the traversal block and wave count, the group count) still reach
the best choice; every measured candidate is recorded and the cache round-trips."""
import itertools

import pytest

from galaxy_amd import Config, autotune


class FakeGpu:
    def info(self):
        return {"driver_info": "synthetic"}


def synthetic_cost(choice):
    # best at (4, 2048, 256, 16, 2).  Block and waves interact: with 32 waves per
    # CU the 128-thread block is better, so a single pass that tunes the block first
    # (at the default 32 waves) settles on 128 and misses the optimum.  Waves and
    # groups interact too: one group prefers 32 waves, two groups only pay off at 16
    # or fewer, so neither coordinate improves alone from (32, 1).
    c = 1.0
    c += {4: 0.0, 8: 0.3, 16: 0.9}[choice.bucket]
    c += {2048: 0.0, 1024: 0.02}[choice.sort_tile]
    if choice.traverse_waves_per_cu == 32:
        c += {256: 0.05, 128: 0.0, 64: 0.08}[choice.traverse_block]
    else:
        c += {256: 0.0, 128: 0.04, 64: 0.08}[choice.traverse_block]
    if choice.streams == 1:
        c += {32: 0.0, 24: 0.05, 20: 0.08, 16: 0.1, 12: 0.15}[choice.traverse_waves_per_cu]
    else:
        c += {32: 0.3, 24: 0.1, 20: 0.0, 16: -0.1, 12: -0.05}[choice.traverse_waves_per_cu]
    return c


@pytest.fixture
def fake_measure(monkeypatch):
    calls = []

    def measure(gpu, states, config, dt, choice, steps=5, warmup=2):
        calls.append(choice)
        return synthetic_cost(choice)

    monkeypatch.setattr(autotune, "measure", measure)
    return calls


def test_repeated_passes_reach_the_best_choice(fake_measure, tmp_path):
    states = [[0] * 100, [0] * 300]
    cache = str(tmp_path / "tune.json")
    r = autotune.tune(FakeGpu(), states, Config(), 1e-3, cache=cache)
    space = [autotune.Choice(*v) for v in itertools.product(*autotune.SPACE.values())]
    best = min(space, key=synthetic_cost)
    assert r["choice"] == best
    assert not r["cached"]
    # every candidate was measured once (timings are kept between passes)
    assert len(fake_measure) == len(set(fake_measure))
    assert {autotune.Choice(**c) for c, _ in r["measured"]} == set(fake_measure)
    again = autotune.tune(FakeGpu(), states, Config(), 1e-3, cache=cache)
    assert again["cached"] and again["choice"] == best
