"""Host logic of stride_amd.workers.ShotWorkers without a GPU: items are dealt round-robin, each
worker runs its items in order, results come back in item order, a worker's exception is
raised on the caller, and close() releases every solver and stream (the solver and stream
calls are replaced by recording stubs)."""

import threading

import pytest

from stride_amd import workers
from stride_amd._native import Solver


class _Stub(Solver):
    def __init__(self, log):  # no library calls
        self.log = log
        self.stream = None
        self.closed = False

    def set_stream(self, stream):
        self.stream = stream

    def synchronize(self):
        self.log.append(('sync', self.stream))

    def close(self):
        self.closed = True


@pytest.fixture
def streams(monkeypatch):
    made, destroyed = [], []
    counter = iter(range(100, 200))
    monkeypatch.setattr(Solver, 'stream_create', staticmethod(lambda device=0: made.append(next(counter)) or made[-1]))
    monkeypatch.setattr(Solver, 'stream_destroy', staticmethod(lambda s: destroyed.append(s)))
    return made, destroyed


def test_map_deals_round_robin_and_keeps_order(streams):
    made, destroyed = streams
    log = []
    lock = threading.Lock()
    seen = {}

    def fn(solver, item):
        with lock:
            seen.setdefault(solver.stream, []).append(item)
        return item * item

    with workers.ShotWorkers(lambda: _Stub(log), n=3) as pool:
        assert len(pool) == 3
        assert [s.stream for s in pool.solvers] == made
        out = pool.map(fn, range(8))
        solvers = list(pool.solvers)
    assert out == [i * i for i in range(8)]
    assert seen == {made[0]: [0, 3, 6], made[1]: [1, 4, 7], made[2]: [2, 5]}
    assert sorted(e[1] for e in log if e[0] == 'sync') == sorted(made)
    assert all(s.closed for s in solvers)
    assert destroyed == made


def test_map_raises_a_worker_exception(streams):
    def fn(solver, item):
        if item == 2:
            raise RuntimeError('shot 2 failed')
        return item

    pool = workers.ShotWorkers(lambda: _Stub([]), n=2)
    with pytest.raises(RuntimeError, match='shot 2 failed'):
        pool.map(fn, range(4))
    pool.close()


def test_rejects_bad_arguments(streams):
    with pytest.raises(ValueError):
        workers.ShotWorkers(lambda: _Stub([]), n=0)
    with pytest.raises(TypeError):
        workers.ShotWorkers(lambda: object(), n=1)


def test_partial_construction_failure_releases_built_workers(streams):
    made, destroyed = streams
    log = []
    built = []

    def make():
        if len(built) == 2:
            raise RuntimeError('out of device memory')
        built.append(_Stub(log))
        return built[-1]

    with pytest.raises(RuntimeError, match='out of device memory'):
        workers.ShotWorkers(make, n=3)
    assert all(s.closed for s in built) and len(built) == 2
    assert sorted(destroyed) == sorted(made) and len(made) == 2

    class NotASolver:
        closed = False

        def close(self):
            self.closed = True

    stray = NotASolver()
    with pytest.raises(TypeError):
        workers.ShotWorkers(lambda: stray, n=2)
    assert stray.closed
