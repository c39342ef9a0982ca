"""The autotuner's search (coordinate descent over a discrete space) on synthetic cost functions."""
from vkpathtracer_amd.autotune import PATH_STATE_BYTES, fitting_capacity, tune


def test_finds_the_minimum_of_a_separable_cost():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30], "c": ["x", "y"]}
    cost = lambda p: (p["a"] - 3) ** 2 + abs(p["b"] - 20) / 10 + (0 if p["c"] == "y" else 5) + 1
    r = tune(cost, space, {"a": 1, "b": 10, "c": "x"})
    assert r.params == {"a": 3, "b": 20, "c": "y"}
    assert r.ms == 1 and r.default_ms == cost({"a": 1, "b": 10, "c": "x"})
    assert r.speedup > 1


def test_coupled_cost_needs_a_second_round():
    """With a coupled cost the first round's choice of `a` is revised after `b` moves."""
    space = {"a": [0, 1, 2], "b": [0, 1, 2]}
    cost = lambda p: 10 - 3 * (p["a"] == p["b"]) - 2 * p["b"] - 0.5 * p["a"]
    r = tune(cost, space, {"a": 0, "b": 0})
    assert r.params == {"a": 2, "b": 2}


def test_each_point_is_measured_once_and_noise_does_not_move():
    calls = []

    def cost(p):
        calls.append(tuple(sorted(p.items())))
        return 100.0 - 0.1 * p["a"]  # 0.1 % better: inside the noise margin

    r = tune(cost, {"a": [0, 1, 2]}, {"a": 0})
    assert r.params == {"a": 0}
    assert len(calls) == len(set(calls)) == 3


def test_tuned_capacity_is_limited_to_free_device_memory():
    """A whole-frame capacity is kept when its path state fits in 80 % of the free memory, cut to what
    fits otherwise, and never below 1 Mi paths (or the capacity itself when smaller)."""
    cap = 874971136
    assert fitting_capacity(cap, 288 << 30) == cap
    small = fitting_capacity(cap, 64 << 30)
    assert small < cap and small * PATH_STATE_BYTES <= 0.8 * (64 << 30)
    assert small == int(0.8 * (64 << 30)) // PATH_STATE_BYTES
    assert fitting_capacity(cap, 1 << 20) == 1 << 20
    assert fitting_capacity(5000, 1 << 20) == 5000
