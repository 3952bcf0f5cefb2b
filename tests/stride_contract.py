"""A stand-in for the parts of Stride's class contract that a plugin operator meets.

Stride itself cannot be imported here (it needs mosaic, zmq, devito and more), so the
binding test builds the same class structure in a few lines and drives the mixin a
Stride maintainer writes (``class IsoAcousticHIP(_HIP, ProblemTypeBase): pass``)
through it:

- ``Operator`` (stride/core.py:639-793): ``__init__`` pops ``name`` and numbers the
  instance; ``__call__`` is a coroutine that records which inputs need gradients, keeps
  the inputs, calls ``forward`` (a coroutine or not) and sets ``prev_op``, ``graph`` and
  ``needs_grad`` on every output; ``__call_adjoint__`` (:795-829) calls
  ``adjoint(*output_grads, *inputs, **kwargs)``;
- ``Gridded`` (stride/problem/base.py:11-73): pops grid / space / time into a grid
  object and exposes them as read-only properties;
- ``ProblemTypeBase`` (stride/physics/problem_type.py:12-56): ``ABC, Gridded, Operator``
  whose ``__init__`` initialises both bases and whose ``forward`` / ``adjoint`` run the
  before / run / after template.

Outputs are marked through the stand-in ``Variable`` attributes (``graph``, ``prev_op``),
which Stride's ``Traces`` carry and this package's ``Traces`` do not: the stand-in
``shot.observed`` is a ``GraphTraces``.
"""

import inspect
import uuid
from abc import ABC

from stride_amd.problem import Grid, Traces

__all__ = ['Node', 'Operator', 'Gridded', 'ProblemTypeBase', 'GraphTraces', 'tessera']


def tessera(cls):
    """mosaic.tessera stands for remote placement; locally the class is unchanged."""
    return cls


class Node:
    def __init__(self, op, method, idx=0, next_ops=None):
        self.op = op
        self.method = method
        self.idx = idx
        self.next_ops = next_ops or []


class _Graph:
    def __init__(self):
        self.nodes = []

    def add(self, node):
        self.nodes.append(node)


class GraphTraces(Traces):
    """Traces with the graph attributes of a Stride Variable."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.graph = _Graph()
        self.prev_op = None

    def alike(self, name=None, data=None, **kwargs):
        return GraphTraces(name or self.name, data=data if data is not None else self._data * 0,
                           transducer_ids=self.transducer_ids)


class Operator:
    _count = 0

    def __init__(self, *args, **kwargs):
        cls = self.__class__
        name = kwargs.pop('name', None)
        uname = 'head:%s_%d' % (cls.__name__.lower(), cls._count)
        self.name = name or uname
        self.uname = '%s-%d-%s' % (name or cls.__name__.lower(), cls._count,
                                   uuid.uuid5(uuid.NAMESPACE_OID, uname).hex)
        cls._count += 1
        self.inputs = None
        self.num_outputs = None
        self.calls = 0

    async def __call__(self, *args, **kwargs):
        self.calls += 1
        needs_grad = False
        next_ops = []
        for arg in args:
            if hasattr(arg, 'needs_grad'):
                needs_grad |= bool(arg.needs_grad)
                next_ops.append(Node(arg, '__call_adjoint__' if arg.needs_grad else '__noop__', 0))
        self.inputs = (args, kwargs)
        if inspect.iscoroutinefunction(self.forward):
            outputs = await self.forward(*args, **kwargs)
        else:
            outputs = self.forward(*args, **kwargs)
        outputs = (outputs,) if not isinstance(outputs, tuple) else outputs
        for idx, output in enumerate(outputs):
            if needs_grad:
                prev_op = Node(self, '__call_adjoint__', idx, next_ops)
                output.graph.add(prev_op)
                output.prev_op = prev_op
            output.needs_grad = needs_grad
        self.num_outputs = len(outputs)
        return outputs if len(outputs) > 1 else outputs[0]

    async def __call_adjoint__(self, *output_grads, **kwargs):
        input_args = self.inputs[0]
        input_kwargs = {**self.inputs[1], **kwargs}
        if inspect.iscoroutinefunction(self.adjoint):
            input_grads = await self.adjoint(*output_grads, *input_args, **input_kwargs)
        else:
            input_grads = self.adjoint(*output_grads, *input_args, **input_kwargs)
        self.inputs = None
        self.num_outputs = None
        return input_grads


class Gridded:
    def __init__(self, **kwargs):
        grid = kwargs.pop('grid', None)
        space = kwargs.pop('space', None)
        time = kwargs.pop('time', None)
        self._grid = Grid(space, time) if grid is None else Grid(grid.space, grid.time)

    @property
    def grid(self):
        return self._grid

    @property
    def space(self):
        return self._grid.space

    @property
    def time(self):
        return self._grid.time


@tessera
class ProblemTypeBase(ABC, Gridded, Operator):
    def __init__(self, **kwargs):
        Gridded.__init__(self, **kwargs)
        Operator.__init__(self, **kwargs)

    def forward(self, *args, **kwargs):
        self.before_forward(*args, **kwargs)
        fw_output = self.run_forward(*args, **kwargs)
        output = self.after_forward(*args, **kwargs)
        return output if output is not None else fw_output

    def adjoint(self, *args, **kwargs):
        self.before_adjoint(*args, **kwargs)
        ad_output = self.run_adjoint(*args, **kwargs)
        output = self.after_adjoint(*args, **kwargs)
        return output if output is not None else ad_output
