"""The Stride-side binding of INTEGRATION.md, on the CPU.

``class IsoAcousticHIP(_HIP, ProblemTypeBase)`` with Stride's ProblemTypeBase /
Gridded / Operator (the stand-in of tests/stride_contract.py, built on
stride/core.py:639-829, stride/problem/base.py:11-73 and
stride/physics/problem_type.py:12-56) must:

- run Stride's __init__ chain (operator name and uid, the Gridded grid);
- keep Stride's asynchronous Operator.__call__: ``await pde(...)`` records the adjoint
  graph node on the traces, and ``__call_adjoint__`` runs the adjoint with the inputs of
  the forward call;
- return the gradients as ``vp.alike(name='vp_grad', ...)`` with ``.prec``
  (stride/physics/iso_acoustic/devito.py:1088-1095) and honour dump_local_grad /
  dump_local_prec (:1097-1111).

The solver itself is replaced by a recording stand-in (no GPU here); the GPU test of the
same binding is tests/test_gpu_operator.py::test_stride_operator_contract.
"""

import asyncio
import inspect
import os

import numpy as np
import pytest

import stride_contract as sc
from stride_amd import IsoAcousticHIP, ScalarField, Shot, SubProblem, Traces
from stride_amd.configs import config1
from stride_amd.physics import iso_acoustic as iso_mod


class _FakeSolver:
    """Records the calls the operator makes; returns fixed arrays of the right shapes."""

    def __init__(self, dim, shape, spacing, absorbing, dt, nt, device=0):
        self.key = (dim, tuple(shape), tuple(spacing), tuple(absorbing), float(dt), int(nt), int(device))
        self.shape = tuple(shape)
        self.nt = nt
        self.calls = []

    def __getattr__(self, name):
        if name.startswith('set_') or name in ('free_saved', 'close'):
            return lambda *a, **k: self.calls.append(name)
        raise AttributeError(name)

    def forward(self, src):
        self.calls.append('forward')
        self.nrec = 128
        return np.full((self.nt, self.nrec), 0.5, np.float32)

    def adjoint(self, res, want_vp=True, want_rho=False, want_src=False):
        self.calls.append('adjoint')
        g = np.full(self.shape, 2.0, np.float32)
        return dict(grad_vp=g, prec_vp=g * 3, grad_rho=None, prec_rho=None, grad_src=None)


class Bound(IsoAcousticHIP, sc.ProblemTypeBase):
    """The mixin of INTEGRATION.md, exactly as a Stride maintainer writes it."""
    pass


def _problem(tmp_path=None):
    cfg = config1(scale=0.2, nt=120)
    space, time = cfg['space'], cfg['time']
    vp = ScalarField('vp', space=space, data=cfg['vp'], needs_grad=True)
    observed = sc.GraphTraces('observed', data=np.zeros((128, time.num), np.float32))
    shot = Shot(7, cfg['sources'], cfg['receivers'], wavelets=cfg['wavelet'][None].astype(np.float32),
                observed=observed)
    problem = SubProblem(shot, space, time, name='bind', output_folder=str(tmp_path or '.'))
    return space, time, vp, shot, problem


def test_mixin_init_chain_and_grid(monkeypatch):
    space, time, vp, shot, problem = _problem()
    pde = Bound(space=space, time=time)
    # Stride's Operator.__init__ ran: operator-style name and uid; Gridded built the grid
    assert pde.name.startswith('head:bound_') and pde.uname.startswith('bound-')
    assert pde.inputs is None and pde.num_outputs is None
    assert pde.space is space and pde.time is time and pde.grid.space is space
    named = Bound(name='pde', grid=problem.grid)
    assert named.name == 'pde' and named.space is space
    # standing alone the operator still builds its own grid and name
    alone = IsoAcousticHIP(space=space, time=time)
    assert alone.name == 'isoacoustichip' and alone.space is space


def test_mixin_call_is_stride_async_call(monkeypatch, tmp_path):
    monkeypatch.setattr(iso_mod, 'Solver', _FakeSolver)
    space, time, vp, shot, problem = _problem(tmp_path)
    pde = Bound(space=space, time=time)
    coro = pde(shot.wavelets, vp, problem=problem, autotune=False)
    assert inspect.iscoroutine(coro)
    traces = asyncio.run(coro)
    assert pde.calls == 1                       # through Operator.__call__
    assert isinstance(traces, sc.GraphTraces) and traces.name == 'modelled'
    assert traces.data.shape == (128, time.num) and np.all(traces.data == 0.5)
    # the adjoint graph node points back at this operator
    assert traces.needs_grad and traces.prev_op.op is pde and traces.prev_op.method == '__call_adjoint__'
    assert traces.graph.nodes == [traces.prev_op]
    assert [n.method for n in traces.prev_op.next_ops] == ['__noop__', '__call_adjoint__']
    assert pde._solver.calls[-1] == 'forward'

    class It:
        abs_id = 4

    residual = Traces('residual', data=np.ones((128, time.num), np.float32))
    grads = asyncio.run(pde.__call_adjoint__(residual, iteration=It(), dump_local_grad=True,
                                             dump_local_prec=True))
    assert pde.inputs is None
    # after_adjoint returns get_grad(wavelets, vp, rho, alpha) (devito.py:981): Stride maps
    # the entries to the forward call's positional inputs (wavelets, vp) in order
    assert isinstance(grads, tuple) and len(grads) == 4 and grads[0] is None and grads[2:] == (None, None)
    g = grads[1]
    assert isinstance(g, ScalarField) and g.name == 'vp_grad' and g.prec.name == 'vp_prec'
    assert g.data.shape == space.shape and g.extended_data.shape == space.shape
    np.testing.assert_allclose(g.data, 2.0 / vp.data, rtol=1e-6)          # x 1/vp (slowness)
    np.testing.assert_allclose(g.prec.data, 6.0 / vp.data ** 2, rtol=1e-6)
    # dump_local_grad / dump_local_prec: version = iteration.abs_id + 1, file names as Mosaic
    # builds them (camel-cased parameter, five-digit version)
    names = sorted(os.listdir(tmp_path))
    assert names == ['bind-vpLocalGradShot00007-00005.h5', 'bind-vpLocalSrcPrecShot00007-00005.h5']
    back = ScalarField('back', space=space, shape=space.shape, extended_shape=space.shape).load(
        path=str(tmp_path), project_name='bind', parameter='vp_local_grad-Shot00007', version=5)
    np.testing.assert_array_equal(back.data, g.data)


def test_standalone_call_runs_forward(monkeypatch):
    monkeypatch.setattr(iso_mod, 'Solver', _FakeSolver)
    space, time, vp, shot, problem = _problem()
    pde = IsoAcousticHIP(space=space, time=time)
    traces = pde(shot.wavelets, vp, problem=problem, autotune=False)
    assert not inspect.iscoroutine(traces) and traces.data.shape == (128, time.num)
