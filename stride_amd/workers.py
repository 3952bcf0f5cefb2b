"""Several solvers of one grid on one GPU, each on its own HIP stream and driven by its own host
thread, so that independent shots overlap on the device.

This is what running several Mosaic workers per GPU gives Stride (each worker holds its own
IsoAcousticDevito instance and runs whole shots, stride/__init__.py:110-141). On the 2D grids a
time step is one short launch that leaves much of the GPU idle, and two workers give about 1.4x
the shot throughput of one; on the 3D grids one shot fills the GPU and a second worker gains
nothing. The streams are the library's own (non-blocking, default priority).

A solver's setters (medium, boundary, taps) first wait for the work queued on that solver's
stream. Setting new taps also frees the old tap buffers, and a device free waits for the whole
device, so a worker that changes taps between shots briefly waits for the other workers' queued
work too.
"""

import threading

from ._native import Solver

__all__ = ['ShotWorkers']


class ShotWorkers:
    """``n`` solvers built by ``make_solver()`` (which returns a configured :class:`Solver`), each
    given its own stream. ``map(fn, items)`` calls ``fn(solver, item)`` for every item: items are
    dealt round-robin to the workers, each worker runs its items in order on its own thread and
    synchronises its stream at the end; results come back in item order, and the first exception
    raised by a worker is raised again. ``close()`` (or leaving a ``with`` block) releases the
    solvers and streams."""

    def __init__(self, make_solver, n=2, device=0):
        if n < 1:
            raise ValueError('n must be at least 1')
        self.streams = []
        self.solvers = []
        try:
            for _ in range(n):
                s = make_solver()
                if not isinstance(s, Solver):
                    if hasattr(s, 'close'):
                        s.close()
                    raise TypeError('make_solver must return a Solver')
                self.solvers.append(s)
                st = Solver.stream_create(device)
                self.streams.append(st)
                s.set_stream(st)
        except BaseException:
            self.close()
            raise

    def __len__(self):
        return len(self.solvers)

    def map(self, fn, items):
        items = list(items)
        n = len(self.solvers)
        results = [None] * len(items)
        errors = []

        def run(k):
            try:
                s = self.solvers[k]
                for i in range(k, len(items), n):
                    results[i] = fn(s, items[i])
                s.synchronize()
            except BaseException as e:  # re-raised on the calling thread
                errors.append(e)

        threads = [threading.Thread(target=run, args=(k,)) for k in range(n)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return results

    def close(self):
        solvers, streams = self.solvers, self.streams
        self.solvers = []
        self.streams = []
        for s in solvers:
            s.close()
        for st in streams:
            Solver.stream_destroy(st)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
