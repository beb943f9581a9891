"""Serving admission and completion logic without a GPU: a fake engine
stands in for the native batch and a fake clock for time."""
import numpy as np
import pytest
import torch

from diffbmp_amd import serve as SV


class FakeState:
    def __init__(self, spec):
        self.spec = spec
        self.steps_done = 0
        self.nbytes = SV.job_bytes(spec, torch.float16)
        self.loss = np.linspace(1.0, 0.5, spec.steps)
        self.t_admit = self.t_done = None
        self.done_event = None

    def unpack(self):
        return {}


class FakeEngine:
    store_dtype = torch.float16

    def __init__(self, log):
        self.jobs = []
        self.log = log

    def add(self, j):
        self.jobs.append(j)
        self.log.append(("add", j.spec.name))

    def remove(self, j):
        self.jobs.remove(j)
        self.log.append(("remove", j.spec.name))

    def step(self):
        for j in self.jobs:
            j.steps_done += 1
        self.log.append(("step", tuple(j.spec.name for j in self.jobs)))

    def grow_if_needed(self):
        return False


class Marker:
    def query(self):
        return True

    def synchronize(self):
        pass


class Clock:
    """Advances 1 ms per reading."""
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t


def spec(name, side, N, T, steps, arrival):
    return SV.JobSpec(side, side, N, tuple(range(T)), steps=steps, arrival=arrival, name=name,
                      template_size=64)


def run(specs, budget, max_jobs=8):
    log = []
    eng = FakeEngine(log)
    srv = SV.Server(eng, SV.HbmBudget(budget), max_jobs=max_jobs, clock=Clock(),
                    make_state=FakeState, make_marker=Marker)
    stats = srv.run(specs)
    return stats, log, srv


def test_job_bytes_counts_every_buffer():
    s = spec("a", 256, 1000, 4, 200, 0.0)
    b16 = SV.job_bytes(s, torch.float16)
    # RGBA records: Ht + 1 rows of Wt 16-byte entries per template
    want = 1000 * 8 * 4 * 3 + 1000 * 8 * 2 + 1000 * 4 + 4 * 64 * 64 * (2 + 6) + 4 * 65 * 64 * 16 \
        + 4 * 16 + 256 * 256 * 9 * 4 + 200 * 4
    assert b16 == want
    assert SV.job_bytes(s, torch.float32) > b16


def test_budget_accounting():
    b = SV.HbmBudget(100)
    assert b.reserve(60) and not b.reserve(41) and b.reserve(40)
    b.release(100)
    assert b.used == 0
    with pytest.raises(RuntimeError):
        b.release(1)


def test_admission_respects_budget_and_fifo():
    a, b, c = spec("a", 256, 800, 2, 5, 0.0), spec("b", 256, 800, 2, 5, 0.0), spec("c", 256, 800, 2, 3, 0.0)
    one = SV.job_bytes(a, torch.float16)
    stats, log, srv = run([a, b, c], budget=2 * one + 10)
    # a and b run together; c is admitted only after one of them retired
    steps = [e[1] for e in log if e[0] == "step"]
    assert steps[0] == ("a", "b")
    assert max(len(s) for s in steps) == 2
    first_c = next(i for i, e in enumerate(log) if e == ("add", "c"))
    first_rm = next(i for i, e in enumerate(log) if e[0] == "remove")
    assert first_rm < first_c
    assert sorted(r.spec.name for r in stats.results) == ["a", "b", "c"]
    assert srv.budget.used == 0
    for r in stats.results:
        assert r.latency >= r.queue >= 0.0
    # every job got exactly its number of steps
    for nm, n in (("a", 5), ("b", 5), ("c", 3)):
        assert sum(nm in s for s in steps) == n


def test_slot_limit_and_arrival_order():
    specs = [spec(f"j{i}", 256, 600, 1, 4, 0.002 * i) for i in range(6)]
    stats, log, _ = run(specs, budget=1 << 40, max_jobs=2)
    steps = [e[1] for e in log if e[0] == "step"]
    assert max(len(s) for s in steps) <= 2
    adds = [e[1] for e in log if e[0] == "add"]
    assert adds == [f"j{i}" for i in range(6)]
    assert stats.summary()["jobs"] == 6


def test_job_larger_than_budget_is_refused():
    big = spec("big", 2048, 32768, 16, 200, 0.0)
    with pytest.raises(ValueError, match="more than the whole budget"):
        run([big], budget=SV.job_bytes(big, torch.float16) - 1)


def test_poisson_arrivals_rate():
    t = SV.poisson_arrivals(50.0, 20000, np.random.default_rng(0))
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 50.0) < 1.5


def test_random_spec_ranges():
    rng = np.random.default_rng(1)
    specs = [SV.random_spec(rng) for _ in range(400)]
    assert {s.height for s in specs} == {256, 512, 1024, 2048}
    ns = np.array([s.num_primitives for s in specs])
    assert ns.min() >= 512 and ns.max() <= 32768
    # log-uniform: the median is near the geometric mean of the range
    assert 2500 < np.median(ns) < 6000
    ts = [len(s.template_ids) for s in specs]
    assert min(ts) == 1 and max(ts) == 16
    assert all(len(set(s.template_ids)) == len(s.template_ids) for s in specs)
