"""Minimal problem-model types consumed by the iso-acoustic operator.

These mirror the attribute surface that Stride's ``IsoAcousticDevito`` reads from
its inputs and writes to its outputs, and nothing else (no plotting, h5 or
resampling):

- ``Space`` / ``Time``           stride/problem/domain.py:9-74, 285-330
- ``ScalarField``                stride/problem/data.py:791-862 (extended-shape float32,
                                 edge padding data.py:513-557, ``vp`` slowness transform
                                 data.py:31-33, 131-132)
- ``Traces``                     stride/problem/data.py:1384-1430 ``(n_transducers, nt)``
- ``Shot`` / ``SubProblem``      stride/problem/acquisitions.py:236-290,
                                 stride/problem/problem.py:252-294

Layouts are identical to the reference: float32, C order, last axis fastest,
extended shape ``shape + 2*extra``.
"""

import numpy as np

__all__ = ['Space', 'Time', 'Grid', 'ScalarField', 'Traces', 'Shot', 'SubProblem',
           'inv_transform']


class Space:
    """Spatial grid: inner ``shape`` plus ``extra`` cells each side, of which
    ``absorbing`` are the absorbing layer (stride/problem/domain.py:48-74)."""

    def __init__(self, shape, spacing, extra=None, absorbing=None):
        if isinstance(spacing, (float, int)):
            spacing = (float(spacing),) * len(shape)
        extra = extra or (0,) * len(shape)
        absorbing = absorbing or (0,) * len(shape)
        self.dim = len(shape)
        self.shape = tuple(int(s) for s in shape)
        self.spacing = tuple(float(s) for s in spacing)
        self.extra = tuple(int(e) for e in extra)
        self.absorbing = tuple(int(a) for a in absorbing)
        self.origin = (0.,) * self.dim
        self.pml_origin = tuple(-h * e for h, e in zip(self.spacing, self.extra))
        self.extended_shape = tuple(s + 2 * e for s, e in zip(self.shape, self.extra))
        self.limit = tuple(h * (s - 1) for h, s in zip(self.spacing, self.shape))
        self.extended_limit = tuple(h * (s + e - 1)
                                    for h, s, e in zip(self.spacing, self.shape, self.extra))

    @property
    def size(self):
        return self.limit

    @property
    def inner(self):
        """Slices of the inner domain inside the extended domain."""
        return tuple(slice(e, e + s) for e, s in zip(self.extra, self.shape))


class Time:
    """Time grid fully defined by three of start/step/num/stop
    (stride/problem/domain.py:285-330)."""

    def __init__(self, start=None, step=None, num=None, stop=None):
        if start is None:
            start = stop - step * (num - 1)
        elif step is None:
            step = (stop - start) / (num - 1)
        elif num is None:
            num = int(np.ceil((stop - start) / step + 1))
            stop = step * (num - 1) + start
        elif stop is None:
            stop = start + step * (num - 1)
        if not isinstance(num, int):
            raise TypeError('"input" argument must be of type int')
        self.start = start
        self.stop = stop
        self.step = step
        self.num = num
        self.extra = 0
        self.extended_num = num

    @property
    def grid(self):
        return np.linspace(self.start, self.stop, self.num, endpoint=True, dtype=np.float32)


class Grid:
    def __init__(self, space=None, time=None):
        self.space = space
        self.time = time


def inv_transform(x):
    """Slowness transform attached by default to fields named ``vp``
    (stride/problem/data.py:31-33, 131-132)."""
    return 1 / x


def _inner_desc(inner):
    return [[str(sl.start), str(sl.stop), str(sl.step)] for sl in inner]


def _inner_from_desc(desc):
    out = []
    for each in desc:
        each = [e.decode() if isinstance(e, bytes) else e for e in each]
        out.append(slice(*[int(e) if e != 'None' else None for e in each]))
    return tuple(out)


def _structured_desc(obj):
    """StructuredData.__get_desc__ (stride/problem/data.py:725-752), without compression."""
    return {'shape': tuple(obj.shape), 'extended_shape': tuple(obj.extended_shape),
            'inner': _inner_desc(obj._inner), 'dtype': 'float32', 'data': np.ascontiguousarray(obj.data),
            'compression': False, 'step_size': None}


def _dump(obj, description, args, kwargs):
    """GriddedSaved.dump (stride/problem/base.py:186-204): grid description + the object's."""
    from . import h5
    desc = {}
    space = getattr(obj, 'space', None)
    if space is not None:
        desc['space'] = {'shape': tuple(space.shape), 'spacing': tuple(space.spacing), 'extra': tuple(space.extra),
                         'absorbing': tuple(space.absorbing)}
    desc.update(description)
    kwargs['parameter'] = kwargs.get('parameter', obj.name)
    with h5.HDF5(*args, mode='w', **kwargs) as f:
        f.dump(desc)
        return f.filename


def _load(obj, args, kwargs):
    from . import h5
    kwargs['parameter'] = kwargs.get('parameter', obj.name)
    with h5.HDF5(*args, mode='r', **kwargs) as f:
        desc = f.load()
    if desc.get('compression'):
        raise ValueError('compressed data (%s) is not supported' % desc['compression'])
    return desc


class ScalarField:
    """Extended-shape float32 field with an inner view.

    ``data`` is the inner domain, ``extended_data`` the whole padded array
    (stride/problem/data.py:280-293). ``pad`` edge-pads inner data into the
    extended shape (data.py:513-557, without the optional smoothing).
    """

    def __init__(self, name, grid=None, space=None, data=None, needs_grad=False,
                 transform=None, shape=None, extended_shape=None, inner=None):
        self.name = name
        self.space = space if space is not None else (grid.space if grid is not None else None)
        self.needs_grad = needs_grad
        if transform is None and 'vp' in name:
            transform = inv_transform
        self.transform = transform
        if extended_shape is None:
            extended_shape = self.space.extended_shape if self.space is not None else np.shape(data)
        if shape is None:
            shape = self.space.shape if self.space is not None else np.shape(data)
        self._extended_shape = tuple(extended_shape)
        self._shape = tuple(shape)
        if inner is None:
            inner = tuple(slice((e - s) // 2, (e - s) // 2 + s)
                          for e, s in zip(self._extended_shape, self._shape))
        self._inner = inner
        self._data = np.zeros(self._extended_shape, dtype=np.float32)
        self.grad = None
        self.prec = None
        if data is not None:
            data = np.asarray(data, dtype=np.float32)
            if data.shape == self._extended_shape:
                self._data[...] = data
            elif data.shape == self._shape:
                self._data[...] = self.pad_data(data)
            else:
                raise ValueError('data of shape %s fits neither %s nor %s'
                                 % (data.shape, self._shape, self._extended_shape))

    @property
    def data(self):
        return self._data[self._inner]

    @property
    def extended_data(self):
        return self._data

    @property
    def shape(self):
        return self._shape

    @property
    def extended_shape(self):
        return self._extended_shape

    def fill(self, value):
        self._data.fill(value)

    def pad_data(self, data):
        widths = [[(e - s) // 2] * 2 for e, s in zip(self._extended_shape, data.shape)]
        return np.pad(data, widths, mode='edge')

    def pad(self):
        self._data[...] = self.pad_data(np.array(self.data))

    def dump(self, *args, **kwargs):
        """Write as Stride does (Mosaic HDF5 protocol 0.1; base.py:186-204, data.py:725-752,
        1163-1167): ``path``, ``project_name``, ``parameter`` (default: the name), ``version``
        or ``filename``. Returns the file name."""
        desc = _structured_desc(self)
        desc.update(time_dependent=False, slow_time_dependent=False)
        return _dump(self, desc, args, kwargs)

    def load(self, *args, **kwargs):
        """Read a file written by ``dump`` or by Stride (base.py:232-276, data.py:754-781): the
        inner data is edge-padded into the extended shape; the space is taken from the file
        when this field has none."""
        desc = _load(self, args, kwargs)
        if self.space is None and 'space' in desc:
            sp = desc['space']
            self.space = Space(shape=sp['shape'], spacing=sp['spacing'], extra=sp['extra'],
                               absorbing=sp['absorbing'])
        self._shape = tuple(desc['shape'])
        self._extended_shape = tuple(desc['extended_shape'])
        self._inner = _inner_from_desc(desc['inner'])
        data = np.asarray(desc['data'], dtype=np.float32)
        self._data = np.zeros(self._extended_shape, dtype=np.float32)
        self._data[...] = data if data.shape == self._extended_shape else self.pad_data(data)
        return self

    def alike(self, name=None, data=None, shape=None, extended_shape=None, inner=None, **kwargs):
        shape = shape if shape is not None else self._shape
        extended_shape = extended_shape if extended_shape is not None else (
            self._extended_shape if shape == self._shape else shape)
        return ScalarField(name or self.name, space=self.space, data=data,
                           shape=shape, extended_shape=extended_shape, inner=inner,
                           transform=self.transform if name is None else kwargs.get('transform'))


class Traces:
    """Time traces ``(n_transducers, nt)`` float32 (stride/problem/data.py:1384-1430)."""

    def __init__(self, name, data=None, shape=None, needs_grad=False, transducer_ids=None):
        self.name = name
        if data is None:
            data = np.zeros(shape if shape is not None else (0, 0), dtype=np.float32)
        self._data = np.ascontiguousarray(data, dtype=np.float32)
        self.needs_grad = needs_grad
        self.transducer_ids = transducer_ids
        self.grad = None

    @property
    def data(self):
        return self._data

    @property
    def extended_data(self):
        return self._data

    @property
    def shape(self):
        return self._data.shape

    @property
    def extended_shape(self):
        return self._data.shape

    @property
    def num_transducers(self):
        return self._data.shape[0]

    _inner = property(lambda self: tuple(slice(0, n, None) for n in self._data.shape))

    def dump(self, *args, **kwargs):
        """As ``ScalarField.dump`` (the Traces description, data.py:1631-1636)."""
        desc = _structured_desc(self)
        desc.update(num_transducers=self.num_transducers,
                    transducer_ids=(np.asarray(self.transducer_ids) if self.transducer_ids is not None
                                    else np.arange(self.num_transducers)))
        return _dump(self, desc, args, kwargs)

    def load(self, *args, **kwargs):
        desc = _load(self, args, kwargs)
        self._data = np.ascontiguousarray(desc['data'], dtype=np.float32)
        if desc.get('transducer_ids') is not None:
            self.transducer_ids = np.asarray(desc['transducer_ids'])
        return self

    def alike(self, name=None, data=None, **kwargs):
        return Traces(name or self.name, data=data if data is not None else np.zeros_like(self._data),
                      transducer_ids=self.transducer_ids)

    def deallocate(self):
        pass


class Shot:
    """One shot: source and receiver point coordinates (metres, inner frame)
    plus wavelets and observed traces (stride/problem/acquisitions.py:57-290)."""

    def __init__(self, id, source_coordinates, receiver_coordinates, wavelets=None, observed=None,
                 num_t=None):
        self.id = int(id)
        self.source_coordinates = np.asarray(source_coordinates, dtype=np.float32).reshape(
            len(source_coordinates), -1)
        self.receiver_coordinates = np.asarray(receiver_coordinates, dtype=np.float32).reshape(
            len(receiver_coordinates), -1)
        if wavelets is None:
            wavelets = Traces('wavelets', shape=(self.num_points_sources, num_t))
        elif not isinstance(wavelets, Traces):
            wavelets = Traces('wavelets', data=wavelets)
        self.wavelets = wavelets
        if observed is None:
            observed = Traces('observed', shape=(self.num_points_receivers, self.wavelets.shape[-1]))
        elif not isinstance(observed, Traces):
            observed = Traces('observed', data=observed)
        self.observed = observed

    @property
    def num_points_sources(self):
        return self.source_coordinates.shape[0]

    @property
    def num_points_receivers(self):
        return self.receiver_coordinates.shape[0]

    @property
    def num_sources(self):
        return self.num_points_sources

    @property
    def num_receivers(self):
        return self.num_points_receivers


class SubProblem:
    """The per-shot problem handed to the operator (stride/problem/problem.py:252-294)."""

    def __init__(self, shot, space, time, name='problem', output_folder='.'):
        self.shot = shot
        self.space = space
        self.time = time
        self.grid = Grid(space, time)
        self.name = name
        self.output_folder = output_folder

    @property
    def shot_id(self):
        return self.shot.id
