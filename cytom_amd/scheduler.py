"""Concurrent task scheduling on the device (scripts/run.py:415-681 Scheduler,
run.py:779-864 execute_task, run.py:866-929 CytomManager).

The reference schedules algorithm tasks on a CPU host.  Per dataset it sorts
the pending algorithms by their profiled CPU usage, then tries groups of
decreasing size: at ``level`` L, windows of n - L consecutive algorithms
(run.py:485-560).  A window is launched as one process when its predicted
usage plus the current usage stays within 100 % (check_cpu_usage,
run.py:562-582), and scheduling stops while the current usage is at 90 % or
more.  A window's usage is predicted by a linear model fitted on (sum of the
members' single-run usage, measured usage of the window) pairs (fitTrueData,
run.py:623-629), or by the plain sum before the first fit.  Single runs
(``--profile`` / ``--sequential``) record each algorithm's usage and running
time per graph, and the profiles and the model's data persist between
launcher runs (exit, run.py:631-670).

Here the shared resource is the one GPU:

* a task's usage is the percentage of its running time that the device
  spends on its work: the compute and ingest time the engine reports per job
  (``python -m cytom_amd.execute`` prints them), over the task's wall time;
* the current usage is the larger of the running tasks' predictions and the
  busy percentage the amdgpu driver reports (in place of psutil.cpu_percent);
* a window must also fit the device memory that the running tasks leave
  (MEMORY_FRACTION of the card), predicted from the jobs' sizes the way the
  engine allocates them (estimate_device_bytes);
* tasks are child processes of ``python -m cytom_amd.execute`` that share the
  GPU; the jobs of one task that agree on the graph settings share one tile
  store and run as programs of it.

Deliberate differences from the reference:

* an algorithm with no profile takes the mean of its graph's profiles, and a
  graph with none takes DEFAULT_USAGE (the reference indexes
  profile_data[graph] and fails for an unprofiled graph, run.py:600-622);
* when nothing runs and a whole scheduling pass launched nothing, the first
  pending algorithm runs alone (the reference waits forever when a single
  algorithm's prediction is above 100 %);
* the state is a JSON file (CYTOM_SCHEDULER_STATE) rather than a rewrite of
  config.py.
"""
import glob
import json
import os
import subprocess
import sys
import time
from collections import OrderedDict

import numpy as np

from .execute import GROUP_KEYS

USAGE_LIMIT = 100.0   # run.py:577
STOP_USAGE = 90.0     # run.py:505
# Tasks alive at once, whatever the usage model predicts: every task is a
# process with the GPU open, and a card serves only so many such processes.
MAX_ALIVE = 8
# Floor of a predicted usage (percent): a fitted line can go to zero or below
# for small windows, which would admit any number of them.
MIN_PREDICT = 1.0
DEFAULT_USAGE = 50.0  # usage of an algorithm on a graph with no profile
MEMORY_FRACTION = 0.9
DEVICE_MEMORY_DEFAULT = 288 << 30  # MI355X HBM3E


def state_path():
    return os.environ.get("CYTOM_SCHEDULER_STATE",
                          os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                       "scheduler-state.json"))


def task_command(job_lists):
    """Command line of one process running the given jobs."""
    cmd = [sys.executable, "-m", "cytom_amd.execute"]
    for j in job_lists:
        cmd += list(j)
    return cmd


# ------------------------------------------------------------ linear model
class LinearModel:
    """Least squares y = a + b x on one feature (run.py:433 LinearRegression
    with the single feature sum_profile_cpu_usage).  With every x equal the
    slope is 0 and the intercept the mean of y, the minimum-norm solution."""

    def __init__(self):
        self.intercept = 0.0
        self.slope = 0.0

    def fit(self, X, y):
        x = np.asarray(X, dtype=np.float64).reshape(len(X), -1)[:, 0]
        yy = np.asarray(y, dtype=np.float64)
        dx = x - x.mean()
        sxx = float(dx @ dx)
        self.slope = float(dx @ (yy - yy.mean())) / sxx if sxx > 0.0 else 0.0
        self.intercept = float(yy.mean() - self.slope * x.mean())
        return self

    def predict(self, X):
        return [self.intercept + self.slope * float(np.ravel(r)[0]) for r in X]


# ------------------------------------------------------------ persistent state
class SchedulerState:
    """PROFILE_DATA, model_X and model_y of scripts/config.py:66-68."""

    def __init__(self, path=None):
        self.path = path or state_path()
        self.profile_data = {}
        self.model_X = []
        self.model_y = []
        if os.path.exists(self.path):
            with open(self.path) as f:
                d = json.load(f)
            self.profile_data = d.get("profile_data", {})
            self.model_X = d.get("model_X", [])
            self.model_y = d.get("model_y", [])

    def save(self):
        d = os.path.dirname(os.path.abspath(self.path))
        os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({"profile_data": self.profile_data, "model_X": self.model_X,
                       "model_y": self.model_y}, f, indent=1)
        os.replace(tmp, self.path)


# ------------------------------------------------------------ device probes
def _drm_values(name):
    vals = []
    for p in glob.glob("/sys/class/drm/card*/device/" + name):
        try:
            with open(p) as f:
                vals.append(float(f.read().strip()))
        except (OSError, ValueError):
            pass
    return vals


def device_busy_percent():
    """Busy percentage the amdgpu driver reports (0 where it is not readable)."""
    v = _drm_values("gpu_busy_percent")
    return max(v) if v else 0.0


def device_memory_bytes():
    v = _drm_values("mem_info_vram_total")
    return int(max(v)) if v else DEVICE_MEMORY_DEFAULT


# ------------------------------------------------------------ memory estimate
def estimate_device_bytes(jobs):
    """Device bytes of one process running the given jobs (dicts of the 40 job
    keys): per shared store a slot pool of 1.25x count-edges (4 B a slot, 8 B
    when a job is sssp), doubled for the repack that may grow it, tile arrays,
    degrees and per-batch scratch; per job the program's vertex arrays."""
    stores = OrderedDict()
    for j in jobs:
        stores.setdefault(tuple(j[k] for k in GROUP_KEYS), []).append(j)
    total = 0
    for group in stores.values():
        j0 = group[0]
        V = int(j0["max-vertices"])
        E = int(j0["count-edges"])
        batch = int(j0["batch-insertion-count"])
        weighted = any(j["algorithm"] == "sssp" for j in group)
        cap = max(1 << 22, E + E // 4)
        vpt = 16384
        T = max(1, -(-V // vpt))
        total += 2 * cap * (8 if weighted else 4)    # pool (+ weights), repack headroom
        total += T * T * 24 + V * 8                  # tile arrays, degrees
        total += max(batch, 1) * 96 + (64 << 20)     # per-batch scratch, fixed buffers
        total += len(group) * (V * 21 + 2 * T)       # program state (PageRank-delta size)
    return total


# ------------------------------------------------------------ tasks
class Algo:
    """One configuration: an algorithm on a graph as one 80-argument job."""

    def __init__(self, algorithm, graph, argv, job):
        self.algorithm = algorithm
        self.graph = graph
        self.argv = list(argv)
        self.job = job


class Task:
    def __init__(self, task_id, algos, graph, sum_profile, predict, device_bytes, single):
        self.id = task_id
        self.algos = algos
        self.graph = graph
        self.sum_profile = sum_profile
        self.predict = predict
        self.device_bytes = device_bytes
        self.single = single
        self.handle = None
        self.start = 0.0
        self.rc = None

    @property
    def names(self):
        return [a.algorithm for a in self.algos]


def busy_ms(summaries):
    """Device time of a task from its job summaries: every job's compute time
    plus each shared store's ingest time once."""
    stores = {}
    ms = 0.0
    for s in summaries:
        ms += float(s.get("compute_ms", 0.0))
        stores[s.get("store", s.get("job"))] = float(s.get("ingest_ms", 0.0))
    return ms + sum(stores.values())


class ProcessHandle:
    def __init__(self, popen, log, path):
        self.popen = popen
        self.log = log
        self.path = path

    def terminate(self, grace=5.0):
        """Stops the child (SIGTERM, then SIGKILL after `grace` seconds)."""
        if self.popen.poll() is None:
            self.popen.terminate()
            try:
                self.popen.wait(timeout=grace)
            except subprocess.TimeoutExpired:
                self.popen.kill()
                self.popen.wait()
        self.poll()

    def poll(self):
        rc = self.popen.poll()
        if rc is not None and not self.log.closed:
            self.log.close()
        return rc

    def wait(self):
        self.popen.wait()
        return self.poll()

    def summaries(self):
        out = []
        with open(self.path) as f:
            for line in f:
                if line.startswith("{"):
                    try:
                        out.append(json.loads(line))
                    except ValueError:
                        pass
        return out

    def tail(self, n=5):
        with open(self.path) as f:
            return f.read().splitlines()[-n:]


class ProcessRunner:
    """Starts a task as ``python -m cytom_amd.execute`` with its jobs; the
    output goes to task_<id>.log in log_root."""

    def __init__(self, log_root):
        self.log_root = log_root

    def __call__(self, task):
        os.makedirs(self.log_root, exist_ok=True)
        path = os.path.join(self.log_root, "task_%d.log" % task.id)
        log = open(path, "w")
        env = dict(os.environ)
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH")
                                    else "")
        p = subprocess.Popen(task_command([a.argv for a in task.algos]), stdout=log,
                             stderr=subprocess.STDOUT, env=env)
        return ProcessHandle(p, log, path)


# ------------------------------------------------------------ scheduler
class Scheduler:
    def __init__(self, state=None, spawn=None, busy=device_busy_percent, memory_bytes=None,
                 clock=time.monotonic, sleep=time.sleep, out=print, poll_interval=0.5,
                 log_root=None, max_alive=MAX_ALIVE):
        self.state = state if state is not None else SchedulerState()
        self.max_alive = max(1, min(int(max_alive), 16))
        self.spawn = spawn or ProcessRunner(log_root or os.path.join(os.getcwd(), "logs"))
        self.busy = busy
        self.memory_limit = MEMORY_FRACTION * (memory_bytes if memory_bytes is not None
                                               else device_memory_bytes())
        self.clock = clock
        self.sleep = sleep
        self.out = out
        self.poll_interval = poll_interval
        self.model = LinearModel()
        self.unfit = True
        if self.state.model_X and self.state.model_y:
            self.model.fit(self.state.model_X, self.state.model_y)
            self.unfit = False
        self.pending = OrderedDict()  # graph -> [Algo] (dataset_algo_map)
        self.n_pending = 0
        self.alive = OrderedDict()    # task id -> Task
        self.finished = []
        self.predicted_current = 0.0
        self.bytes_current = 0
        self.task_id = 0
        self.stats = []               # the reference's stat_q
        self.failed = 0

    # -- profiles and model
    def profile_usage(self, names, graph):
        """get_profile_cpu_usage (run.py:600-611)."""
        data = self.state.profile_data.get(graph, {})
        mean = (sum(v["gpu_usage"] for v in data.values()) / len(data)) if data else DEFAULT_USAGE
        return [data[n]["gpu_usage"] if n in data else mean for n in names]

    def save_profile(self, algorithm, graph, usage, running_time):
        """saveProfileData (run.py:631-641)."""
        self.state.profile_data.setdefault(graph, {})[algorithm] = {
            "gpu_usage": usage, "running_time": running_time}

    def fit_true_data(self, X, y):
        """fitTrueData (run.py:623-629)."""
        self.state.model_X = self.state.model_X + X
        self.state.model_y = self.state.model_y + y
        self.model.fit(self.state.model_X, self.state.model_y)
        self.unfit = False

    def check_stats(self):
        """check_stat_q (run.py:450-481)."""
        X, y = [], []
        for single, task, usage, running_time in self.stats:
            if single:
                self.save_profile(task.algos[0].algorithm, task.graph, usage, running_time)
            else:
                X.append([task.sum_profile])
                y.append(usage)
        self.stats = []
        if X:
            self.fit_true_data(X, y)

    def predict(self, sum_profile):
        p = sum_profile if self.unfit else self.model.predict([[sum_profile]])[0]
        return max(MIN_PREDICT, float(p))

    # -- running tasks
    def add_algos(self, algos):
        for a in algos:
            self.pending.setdefault(a.graph, []).append(a)
            self.n_pending += 1

    def launch(self, algos, graph, sum_profile, predict, device_bytes, single=False):
        """Starts a task; unless it is a single profiling run its algorithms
        leave the pending count."""
        t = Task(self.task_id, algos, graph, sum_profile, predict, device_bytes, single)
        self.task_id += 1
        self.out("-------Now launch task %d : %s on %s" % (t.id, t.names, graph))
        t.start = self.clock()
        t.handle = self.spawn(t)
        self.alive[t.id] = t
        self.predicted_current += predict
        self.bytes_current += device_bytes
        if not single:
            self.n_pending -= len(algos)
        return t

    def _finish(self, t, rc):
        elapsed = max(self.clock() - t.start, 1e-9)
        t.rc = rc
        self.predicted_current -= t.predict
        self.bytes_current -= t.device_bytes
        del self.alive[t.id]
        self.finished.append(t)
        self.out("------------------------------------------")
        self.out("Task %d has finished:  %s on %s (%.1f s, exit %d)"
                 % (t.id, t.names, t.graph, elapsed, rc))
        summaries = t.handle.summaries()
        for s in summaries:
            self.out(json.dumps(s))
        if rc != 0 or not summaries:
            self.failed += 1
            if hasattr(t.handle, "tail"):
                for line in t.handle.tail():
                    self.out("  | " + line)
            return
        usage = min(100.0, 100.0 * busy_ms(summaries) / (elapsed * 1e3))
        self.stats.append((t.single, t, usage, elapsed))

    def update_alive(self):
        """update_predict_cur_cpu_usage (run.py:584-598)."""
        for t in list(self.alive.values()):
            rc = t.handle.poll()
            if rc is not None:
                self._finish(t, rc)

    def current_usage(self):
        self.update_alive()
        return max(self.predicted_current, self.busy())

    def check_usage(self, algos, graph):
        """check_cpu_usage (run.py:562-582) plus the device-memory test."""
        usage = self.profile_usage([a.algorithm for a in algos], graph)
        sum_profile = sum(usage)
        predict = self.predict(sum_profile)
        device_bytes = estimate_device_bytes([a.job for a in algos])
        current = self.current_usage()
        avail = (current + predict <= USAGE_LIMIT
                 and self.bytes_current + device_bytes <= self.memory_limit
                 and len(self.alive) < self.max_alive)
        return avail, sum_profile, predict, device_bytes

    def schedule(self):
        """One scheduling pass (run.py:483-560)."""
        self.check_stats()
        for graph, algos in self.pending.items():
            usage = self.profile_usage([a.algorithm for a in algos], graph)
            order = sorted(range(len(algos)), key=lambda i: usage[i], reverse=True)
            self.pending[graph] = [algos[i] for i in order]
        launched = 0
        level = 0  # windows of n - level algorithms of a dataset
        while True:
            untouched = 0
            stop = False
            for graph in self.pending:
                if self.n_pending == 0 or self.current_usage() >= STOP_USAGE:
                    stop = True
                    break
                algos = self.pending[graph]
                n = len(algos)
                stride = n - level
                if n == 0 or stride <= 0:
                    untouched += 1
                    continue
                taken = set()
                start = 0
                while start < n:
                    end = min(start + stride, n)
                    avail, sum_profile, predict, device_bytes = self.check_usage(
                        algos[start:end], graph)
                    if avail:
                        self.launch(algos[start:end], graph, sum_profile, predict, device_bytes)
                        launched += 1
                        taken.update(range(start, end))
                        start += stride
                    else:
                        start += 1
                self.pending[graph] = [a for i, a in enumerate(algos) if i not in taken]
            if stop or untouched == len(self.pending):
                break
            level += 1
        if launched == 0 and not self.alive and self.n_pending:
            for graph, algos in self.pending.items():
                if algos:
                    _, sum_profile, predict, device_bytes = self.check_usage(algos[:1], graph)
                    self.launch(algos[:1], graph, sum_profile, predict, device_bytes)
                    self.pending[graph] = algos[1:]
                    break

    def wait_any(self):
        n = len(self.finished)
        while self.alive:
            self.update_alive()
            if len(self.finished) > n:
                return
            self.sleep(self.poll_interval)

    def stop_all(self):
        """Terminates the running tasks (used when the scheduler itself is
        interrupted, so no child keeps the GPU)."""
        for t in list(self.alive.values()):
            if hasattr(t.handle, "terminate"):
                t.handle.terminate()
        self.alive.clear()

    def run_all(self, algos):
        """Schedules every configuration, waits for all tasks and saves the
        profiles and the model's data (sched_func and exit, run.py:387-412,
        631-670).  Returns the number of failed tasks."""
        self.add_algos(algos)
        try:
            while self.n_pending:
                self.schedule()
                if self.n_pending:
                    self.wait_any()
            while self.alive:
                self.wait_any()
        finally:
            self.stop_all()
        self.check_stats()
        self.state.save()
        return self.failed

    def profile_all(self, algos):
        """One task per configuration, one after the other, each recording its
        single-run profile (CytomManager.profile_all, run.py:905-916)."""
        try:
            for a in algos:
                usage = self.profile_usage([a.algorithm], a.graph)
                self.launch([a], a.graph, usage[0], usage[0], estimate_device_bytes([a.job]),
                            single=True)
                while self.alive:
                    self.wait_any()
        finally:
            self.stop_all()
        self.check_stats()
        self.state.save()
        return self.failed
