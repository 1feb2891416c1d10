"""Device scheduler (cytom_amd/scheduler.py) against the scheduling rules of
scripts/run.py:415-681, with fake tasks on a fake clock (no device needed).

The expected launch sequences are worked out by hand from run.py's loops:
sort a dataset's algorithms by profiled usage (descending, stable), then at
level L try windows of n - L consecutive algorithms, launch a window when
current + predicted usage <= 100, slide by one otherwise, stop a pass once the
current usage reaches 90.
"""
import json
import os

import numpy as np
import pytest

from cytom_amd import scheduler as S


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeHandle:
    def __init__(self, clock, duration, busy_fraction, task, rc=0):
        self.clock, self.end, self.rc = clock, clock() + duration, rc
        self.busy_ms = busy_fraction * duration * 1e3
        self.task = task

    def poll(self):
        return self.rc if self.clock() >= self.end else None

    def wait(self):
        self.clock.t = max(self.clock.t, self.end)
        return self.rc

    def summaries(self):
        if self.rc:
            return []
        n = len(self.task.algos)
        return [{"job": i, "store": 0, "compute_ms": self.busy_ms / n, "ingest_ms": 0.0}
                for i in range(n)]

    def tail(self, n=5):
        return ["error: fake failure"]


class FakeSpawn:
    """Every task runs `duration` seconds; its measured device usage is
    usage(task) percent; algorithms named in `fail` exit with status 1."""

    def __init__(self, clock, duration=10.0, usage=None, fail=()):
        self.clock, self.duration = clock, duration
        self.usage = usage or (lambda task: 40.0)
        self.fail = set(fail)
        self.launched = []
        self.max_alive = 0
        self.handles = []

    def __call__(self, task):
        self.launched.append(task.names)
        rc = 1 if self.fail & set(task.names) else 0
        h = FakeHandle(self.clock, self.duration, self.usage(task) / 100.0, task, rc)
        self.handles.append(h)
        alive = sum(1 for x in self.handles if x.poll() is None)
        self.max_alive = max(self.max_alive, alive)
        return h


def _job(graph="rmat-16", algorithm="bfs", V=1 << 16, E=1 << 20):
    return {"max-vertices": V, "vertices-per-tile": 16384, "batch-insertion-count": 1 << 16,
            "count-edges": E, "deletion-percentage": 0.0, "algorithm-iterations": 100,
            "delta": 0.01, "start-algorithm-count": 0, "enable-interactive-algorithm": "1",
            "enable-reading-edges": "0", "path-to-edges-input": "/x",
            "enable-writing-edges": "0", "path-to-edges-output": "/y",
            "enable-rewrite-ids": "0", "enable-in-memory-ingestion": "0",
            "algorithm": algorithm}


def _algos(graph, names, **kw):
    return [S.Algo(n, graph, ["--algorithm", n], _job(graph, n, **kw)) for n in names]


def _sched(tmp_path, profiles=None, spawn=None, memory=1 << 40, model=None):
    st = S.SchedulerState(str(tmp_path / "state.json"))
    st.profile_data = profiles or {}
    if model:
        st.model_X, st.model_y = model
    clock = Clock()
    spawn = spawn or FakeSpawn(clock)
    spawn.clock = clock
    out = []
    sch = S.Scheduler(state=st, spawn=spawn, busy=lambda: 0.0, memory_bytes=memory, clock=clock,
                      sleep=clock.sleep, out=out.append)
    return sch, spawn, out


def _prof(graph, **usage):
    return {graph: {k.replace("_", "-"): {"gpu_usage": v, "running_time": 1.0}
                    for k, v in usage.items()}}


# ------------------------------------------------------------ linear model
def test_linear_model_least_squares():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 200, 40)
    y = 0.7 * x + 5 + rng.normal(0, 3, 40)
    m = S.LinearModel().fit([[v] for v in x], list(y))
    b, a = np.polyfit(x, y, 1)
    assert m.slope == pytest.approx(b, rel=1e-12) and m.intercept == pytest.approx(a, rel=1e-9)
    assert m.predict([[10.0]])[0] == pytest.approx(a + 10 * b)
    # one point or equal x: slope 0, intercept the mean (minimum-norm solution)
    m = S.LinearModel().fit([[50.0]], [37.0])
    assert (m.slope, m.intercept) == (0.0, 37.0)
    m = S.LinearModel().fit([[50.0], [50.0]], [30.0, 40.0])
    assert m.predict([[0.0]]) == [35.0]


def test_linear_model_matches_scikit_learn():
    lm = pytest.importorskip("sklearn.linear_model")
    rng = np.random.default_rng(4)
    X = rng.uniform(0, 300, (25, 1))
    y = rng.uniform(0, 100, 25)
    ref = lm.LinearRegression().fit(X, y)
    m = S.LinearModel().fit(X.tolist(), y.tolist())
    assert m.slope == pytest.approx(ref.coef_[0], rel=1e-10)
    assert m.intercept == pytest.approx(ref.intercept_, rel=1e-10)


# ------------------------------------------------------------ scheduling rules
def test_window_slides_past_an_overfull_group(tmp_path):
    # four algorithms at 30 %: the window of four (120 %) does not fit, the
    # window [1:4] (90 %) does; the pass then stops at 90 % and `a` runs once
    # the first task is done.
    sch, sp, out = _sched(tmp_path, _prof("g", a=30, b=30, c=30, d=30))
    assert sch.run_all(_algos("g", "abcd")) == 0
    assert sp.launched == [["b", "c", "d"], ["a"]]
    assert sp.max_alive == 1
    assert out[0] == "-------Now launch task 0 : ['b', 'c', 'd'] on g"


def test_algorithms_sorted_by_profile(tmp_path):
    sch, sp, _ = _sched(tmp_path, _prof("g", a=10, b=60, c=20))
    sch.run_all(_algos("g", "abc"))
    assert sp.launched == [["b", "c", "a"]]


def test_levels_split_a_dataset(tmp_path):
    # 60/45/45/40 sorted a, b, c, d.  Level 0: [0:4] 190 no, [1:4] 130 no,
    # [2:4] 85 yes.  Level 1 at 85 %: [a] and [b] do not fit; level 2 has no
    # window.  Once [c, d] is done (measured 40 %), the model fitted on
    # (85, 40) predicts 40 for [a, b], which then runs as one task.
    sch, sp, _ = _sched(tmp_path, _prof("g", a=60, b=45, c=45, d=40))
    sch.run_all(_algos("g", "abcd"))
    assert sp.launched == [["c", "d"], ["a", "b"]]
    assert sch.state.model_X == [[85.0], [105.0]]


def test_two_datasets_share_the_device(tmp_path):
    sch, sp, _ = _sched(tmp_path, {**_prof("g1", bfs=40, cc=20), **_prof("g2", bfs=30)})
    sch.run_all(_algos("g1", ["bfs", "cc"]) + _algos("g2", ["bfs"]))
    assert sp.launched == [["bfs", "cc"], ["bfs"]]
    assert sp.max_alive == 2  # 60 % + 30 % run together


def test_stop_at_ninety_percent(tmp_path):
    sch, sp, _ = _sched(tmp_path, {**_prof("g1", bfs=92), **_prof("g2", bfs=5)})
    sch.run_all(_algos("g1", ["bfs"]) + _algos("g2", ["bfs"]))
    assert sp.launched == [["bfs"], ["bfs"]] and sp.max_alive == 1


def test_device_memory_admission(tmp_path):
    one = S.estimate_device_bytes([_job()])
    sch, sp, _ = _sched(tmp_path, {**_prof("g1", bfs=10), **_prof("g2", bfs=10)},
                        memory=int(1.5 * one / S.MEMORY_FRACTION))
    sch.run_all(_algos("g1", ["bfs"]) + _algos("g2", ["bfs"]))
    assert len(sp.launched) == 2 and sp.max_alive == 1  # usage fits, memory does not


def test_algorithm_above_the_limit_runs_alone(tmp_path):
    # Before any model fit a 150 % prediction never fits; with nothing else
    # running it is started by itself instead of waiting forever.
    sch, sp, _ = _sched(tmp_path, _prof("g", big=150))
    assert sch.run_all(_algos("g", ["big"])) == 0
    assert sp.launched == [["big"]]
    # With a second algorithm, `small` runs first (the window of both and
    # `big` alone do not fit); `big` follows once `small` is done.
    sch, sp, _ = _sched(tmp_path / "second", _prof("g", big=150, small=10))
    assert sch.run_all(_algos("g", ["big", "small"])) == 0
    assert sp.launched == [["small"], ["big"]]


def test_unprofiled_algorithms_take_the_graph_mean_or_default(tmp_path):
    sch, _, _ = _sched(tmp_path, _prof("g", bfs=20, cc=40))
    assert sch.profile_usage(["bfs", "sssp"], "g") == [20, 30.0]
    assert sch.profile_usage(["bfs"], "unknown") == [S.DEFAULT_USAGE]


# ------------------------------------------------------------ measurement, model, state
def test_model_learns_from_measured_groups(tmp_path):
    # a group of profiled 30 + 30 measures 45 % on the device: the model is
    # fitted on (60, 45) and predicts 45 for any sum until more data comes.
    clock = Clock()
    sp = FakeSpawn(clock, usage=lambda t: 45.0)
    sch, sp, _ = _sched(tmp_path, _prof("g", a=30, b=30), spawn=sp)
    sch.run_all(_algos("g", "ab"))
    assert sch.state.model_X == [[60.0]] and sch.state.model_y == [pytest.approx(45.0)]
    saved = json.load(open(tmp_path / "state.json"))
    assert saved["model_X"] == [[60.0]]
    sch2, _, _ = _sched(tmp_path, _prof("g", a=30, b=30), model=(saved["model_X"],
                                                                 saved["model_y"]))
    assert not sch2.unfit and sch2.predict(200.0) == pytest.approx(45.0)


def test_profile_all_records_single_runs(tmp_path):
    clock = Clock()
    sp = FakeSpawn(clock, duration=8.0, usage=lambda t: {"bfs": 70.0, "cc": 35.0}[t.names[0]])
    sch, sp, out = _sched(tmp_path, spawn=sp)
    assert sch.profile_all(_algos("rmat-16", ["bfs", "cc"])) == 0
    assert sp.launched == [["bfs"], ["cc"]] and sp.max_alive == 1
    prof = json.load(open(tmp_path / "state.json"))["profile_data"]["rmat-16"]
    assert prof["bfs"]["gpu_usage"] == pytest.approx(70.0, rel=0.02)
    assert prof["cc"]["gpu_usage"] == pytest.approx(35.0, rel=0.02)
    assert prof["bfs"]["running_time"] == pytest.approx(8.0, abs=0.6)
    assert any(o.startswith("Task 1 has finished:  ['cc'] on rmat-16") for o in out)


def test_failed_task_is_counted_and_not_learned(tmp_path):
    clock = Clock()
    sp = FakeSpawn(clock, fail={"cc"})
    sch, sp, out = _sched(tmp_path, _prof("g", bfs=60, cc=60), spawn=sp)
    assert sch.run_all(_algos("g", ["bfs", "cc"])) == 1
    assert sp.launched == [["cc"], ["bfs"]]
    assert sch.state.model_X == [[60.0]]  # only the task that succeeded
    assert any("fake failure" in o for o in out)


def test_busy_time_counts_each_store_ingest_once():
    s = [{"job": 0, "store": 0, "compute_ms": 10.0, "ingest_ms": 5.0},
         {"job": 1, "store": 0, "compute_ms": 20.0, "ingest_ms": 5.0},
         {"job": 2, "store": 1, "compute_ms": 1.0, "ingest_ms": 2.0}]
    assert S.busy_ms(s) == 38.0


def test_device_bytes_estimate():
    a = S.estimate_device_bytes([_job(algorithm="bfs")])
    two = S.estimate_device_bytes([_job(algorithm="bfs"), _job(algorithm="cc")])
    assert a < two < 2 * a  # the two programs share one store
    V = 1 << 16
    assert two - a == V * 21 + 2 * 4
    w = S.estimate_device_bytes([_job(algorithm="sssp")])
    cap = max(1 << 22, (1 << 20) + (1 << 18))
    assert w - a == 2 * cap * 4  # weights double the slot bytes
    other = S.estimate_device_bytes([_job(algorithm="bfs"), _job(algorithm="bfs", E=1 << 22)])
    assert other > 2 * a - 1  # different stores add up


def test_state_round_trip(tmp_path):
    p = str(tmp_path / "s.json")
    st = S.SchedulerState(p)
    st.profile_data = _prof("g", bfs=12.5)
    st.model_X, st.model_y = [[1.0], [2.0]], [3.0, 4.0]
    st.save()
    st2 = S.SchedulerState(p)
    assert (st2.profile_data, st2.model_X, st2.model_y) == (st.profile_data, st.model_X,
                                                            st.model_y)
    assert not os.path.exists(p + ".tmp")


def test_interrupt_stops_running_tasks(tmp_path):
    # An exception inside the scheduling loop (say KeyboardInterrupt while
    # waiting) terminates every running child before it propagates.
    clock = Clock()

    class Handle(FakeHandle):
        stopped = []

        def terminate(self):
            Handle.stopped.append(self.task.names)

    class Spawn(FakeSpawn):
        def __call__(self, task):
            self.launched.append(task.names)
            return Handle(self.clock, self.duration, 0.4, task)

    def boom(dt):
        raise KeyboardInterrupt

    sp = Spawn(clock)
    sch, sp, _ = _sched(tmp_path, {**_prof("g1", bfs=40), **_prof("g2", bfs=40)}, spawn=sp)
    sch.sleep = boom
    with pytest.raises(KeyboardInterrupt):
        sch.run_all(_algos("g1", ["bfs"]) + _algos("g2", ["bfs"]))
    assert sorted(map(tuple, Handle.stopped)) == [("bfs",), ("bfs",)] and not sch.alive


def test_alive_tasks_capped_when_the_model_predicts_nothing(tmp_path):
    # A model fitted on windows that measured ~0 % predicts zero (or less)
    # for every window, so usage never stops admission; the alive cap does.
    # Forty single-algorithm datasets at 0 % profiled usage each.
    profiles = {}
    algos = []
    for i in range(40):
        g = "g%d" % i
        profiles.update(_prof(g, bfs=0.0))
        algos += _algos(g, ["bfs"])
    sch, sp, _ = _sched(tmp_path, profiles, model=([[10.0], [20.0]], [-10.0, -20.0]))
    assert sch.predict(0.0) == S.MIN_PREDICT  # floored, never <= 0
    assert sch.run_all(algos) == 0
    assert len(sp.launched) == 40
    assert sp.max_alive == S.MAX_ALIVE <= 16


def test_alive_cap_is_bounded_by_sixteen(tmp_path):
    st = S.SchedulerState(str(tmp_path / "s.json"))
    sch = S.Scheduler(state=st, spawn=FakeSpawn(Clock()), busy=lambda: 0.0, memory_bytes=1 << 40,
                      max_alive=100)
    assert sch.max_alive == 16
