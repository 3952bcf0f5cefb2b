"""Drop-in iso-acoustic operator backed by the gfx950 HIP library.

``IsoAcousticHIP`` keeps the plugin surface of Stride's ``IsoAcousticDevito``
(stride/physics/iso_acoustic/devito.py:23-1707) on top of ``ProblemTypeBase``
(stride/physics/problem_type.py:12-321):

- ``forward(wavelets, vp, rho=None, alpha=None, problem=..., **kwargs) -> Traces``
  = before_forward / run_forward / after_forward;
- ``adjoint(adjoint_source, wavelets, vp, rho=None, alpha=None, problem=..., **kwargs)``
  returns one entry per positional input (gradient or None), vp / rho gradients
  on the inner grid with a ``.prec`` attribute;
- ``prepare_grad_*`` / ``init_grad_*`` / ``get_grad_*`` per variable,
  ``clear_operators``, ``deallocate_wavefield``.

The same keyword options are accepted (kernel, boundary_type, interpolation_type,
attenuation_power, save_wavefield, save_undersampling, time_bounds,
diff_source, platform, deviceid, deallocate). ``adjoint_mode='exact'`` selects
the exact discrete adjoint (every forward level stored) instead of the
reference's as-written adjoint; the default, ``'compat'``, reproduces the
reference. ``autotune`` (True / False; default: on for 3D grids of at least 2^24
points) times the candidate tiles of the 3D update the scheme runs on, once per grid
and scheme, and uses the fastest, as Devito's runtime autotuning picks block sizes
(stride/physics/common/devito.py:1061-1078).

The forward state (solver, saved wavefield, plan) is kept on the instance
between ``forward`` and ``adjoint`` of a shot, as the reference keeps
``_wavefield`` and its Devito functions.
"""

import logging

import numpy as np

from .._native import Solver
from ..problem import Grid, ScalarField, Traces
from .boundaries import SPONGE, PML
from .gradients import compat_finish, exact_gradients
from .plan import build_plan, adjoint_source_samples, RHO_VARIABLE

__all__ = ['ProblemTypeBase', 'IsoAcousticHIP', 'IsoAcousticDevito']

logger = logging.getLogger('stride_amd')


def _next_in_mro(obj, after, attr):
    """The first definition of ``attr`` in the classes that follow ``after`` in the method
    resolution order of ``obj``'s class (``object`` excluded), or None."""
    mro = type(obj).__mro__
    for klass in mro[mro.index(after) + 1:]:
        if klass is object:
            continue
        if attr in klass.__dict__:
            return klass.__dict__[attr]
    return None


class ProblemTypeBase:
    """before/run/after template and per-variable gradient dispatch
    (stride/physics/problem_type.py:58-321).

    The class also works as a mixin in front of Stride's own ``ProblemTypeBase``
    (``class X(IsoAcousticHIP, stride.ProblemTypeBase)``, what a Stride maintainer writes to
    make it a Mosaic tessera):

    - ``__init__`` is cooperative: when another class follows this one in the method order
      (Stride's ``ProblemTypeBase``, which initialises ``Gridded`` and ``Operator``,
      problem_type.py:51-53), it is called with the same keywords, so the operator gets
      its name, uid and grid the way Stride builds them;
    - ``__call__`` hands over to the next ``__call__`` in the method order when there is
      one: Stride's asynchronous ``Operator.__call__`` (stride/core.py:728-793), which
      records the adjoint graph nodes on the outputs and calls ``forward``. Standing alone,
      calling the operator runs ``forward`` directly;
    - ``grid``, ``space`` and ``time`` read the grid object as ``Gridded`` does
      (stride/problem/base.py:11-73).
    """

    def __init__(self, name=None, grid=None, space=None, time=None, **kwargs):
        if _next_in_mro(self, ProblemTypeBase, '__init__') is not None:
            super().__init__(**({'name': name} if name is not None else {}), grid=grid, space=space, time=time,
                             **kwargs)
        if getattr(self, '_grid', None) is None:
            if grid is not None:
                space = space or grid.space
                time = time or grid.time
            self._grid = Grid(space, time)
        if getattr(self, 'name', None) is None:
            self.name = name or self.__class__.__name__.lower()

    @property
    def grid(self):
        return self._grid

    @property
    def space(self):
        return self._grid.space

    @property
    def time(self):
        return self._grid.time

    def forward(self, *args, **kwargs):
        self.before_forward(*args, **kwargs)
        fw = self.run_forward(*args, **kwargs)
        out = self.after_forward(*args, **kwargs)
        return out if out is not None else fw

    def adjoint(self, *args, **kwargs):
        self.before_adjoint(*args, **kwargs)
        ad = self.run_adjoint(*args, **kwargs)
        out = self.after_adjoint(*args, **kwargs)
        return out if out is not None else ad

    def __call__(self, *args, **kwargs):
        parent = _next_in_mro(self, ProblemTypeBase, '__call__')
        if parent is not None:
            return parent(self, *args, **kwargs)
        return self.forward(*args, **kwargs)

    def _dispatch(self, prefix, variable):
        methods = [getattr(self, prefix + part, None) for part in variable.name.split('_')]
        methods = [m for m in methods if m is not None]
        if not methods:
            raise ValueError('Variable %s not implemented' % variable.name)
        return methods[0]

    def prepare_grad(self, *wrt, **kwargs):
        updates = []
        for variable in wrt:
            if variable is None or not variable.needs_grad:
                continue
            update = self._dispatch('prepare_grad_', variable)(variable, wrt=wrt, **kwargs)
            updates += list(update) if isinstance(update, tuple) else [update]
        return updates

    def init_grad(self, *wrt, **kwargs):
        for variable in wrt:
            if variable is None or not variable.needs_grad:
                continue
            self._dispatch('init_grad_', variable)(variable, wrt=wrt, **kwargs)

    def get_grad(self, *wrt, **kwargs):
        grads = []
        for variable in wrt:
            if variable is None or not variable.needs_grad:
                grads.append(None)
                continue
            grad = self._dispatch('get_grad_', variable)(variable, wrt=wrt, **kwargs)
            data = grad.data if hasattr(grad, 'data') else grad
            if np.any(np.isnan(data)) or np.any(np.isinf(data)):
                msg = 'Nan or inf detected in %s' % self.name
                problem = kwargs.get('problem')
                if problem is not None:
                    msg = '(ShotID %d) ' % problem.shot.id + msg
                logger.warning(msg)
                return None
            grads.append(grad)
        return tuple(grads)


# grids from this size up are autotuned by default (the tuning runs take about a second)
_AUTOTUNE_MIN_POINTS = 1 << 24
# raw snapshots above this many bytes per shot are recomputed from checkpoints instead of stored
SAVE_BUDGET_BYTES = 64 << 30


def raw_checkpoint(plan, checkpoint=None, budget=SAVE_BUDGET_BYTES):
    """Snapshots per recomputation segment for the raw saves of the rho / alpha / wavelet
    gradients (0: store them all, as devito.py:366-437 does). `checkpoint`: an explicit segment
    length, or None for automatic: segments of K = ceil(sqrt(2 N)) snapshots (N snapshots), which
    keeps 2 N / K + K + 2 snapshot-sized buffers, once the N stored snapshots would take more
    than `budget` bytes. The recomputation needs the two time levels alone to restart, so the
    CFS-PML and power-2 attenuation schemes always store."""
    if plan.boundary == PML or plan.att_power == 2:
        return 0
    nsnap = plan.tb1 // plan.undersampling + 1
    if checkpoint is not None:
        return max(0, int(checkpoint))
    if nsnap * int(np.prod(plan.shape)) * 4 <= budget:
        return 0
    return int(np.ceil(np.sqrt(2 * nsnap)))


def _dump_factor(option, plan):
    # an int samples every `option` steps; True the save undersampling (devito.py:421-424)
    return option if type(option) is int else plan.undersampling


def _needs_grad(*variables):
    return any(v is not None and getattr(v, 'needs_grad', False) for v in variables)


class IsoAcousticHIP(ProblemTypeBase):
    space_order = 10
    time_order = 2

    def __init__(self, name=None, grid=None, space=None, time=None, deviceid=0, **kwargs):
        super().__init__(name=name, grid=grid, space=space, time=time, **kwargs)
        self.kernel = 'OT4'
        self.undersampling_factor = 4
        self.boundary_type = 'sponge_boundary_2'
        self.interpolation_type = 'linear'
        self.attenuation_power = 0
        self.deviceid = deviceid
        self._solver = None
        self._plan = None
        self._traces = None
        self._grads = None
        self._save_mode = 0
        self._bandwidth = (0., 0., 0.)
        self._cache = None  # cache_forward: the snapshots (array) or the file holding them
        self._cache_folder = None

    # operator lifetime -------------------------------------------------------------------

    def clear_operators(self):
        if self._solver is not None:
            self._solver.close()
        self._solver = None

    def deallocate_wavefield(self, platform='gpu', deallocate=False, **kwargs):
        if self._solver is not None:
            self._solver.free_saved()
        self._cache = None
        if deallocate:
            self.clear_operators()

    @property
    def wavefield(self):
        """The stored forward snapshots, (nsave, stored box) (devito.py:198-212); None when the
        last forward run kept none or recomputes them from checkpoints."""
        if self._solver is None or self._save_mode not in (1, 2) or self._cache is not None:
            return None
        data = self._solver.saved_to_host(release=False)
        plan = self._plan
        box = plan.shape if self._save_mode == 2 else tuple(n - 2 * a for n, a in zip(plan.shape, plan.absorbing))
        return ScalarField('p', data=data.reshape((-1,) + tuple(box)), shape=(data.size // int(np.prod(box)),) + box,
                           extended_shape=(data.size // int(np.prod(box)),) + box)

    @property
    def solver(self):
        return self._solver

    def _get_solver(self, plan, device):
        key = (plan.dim, plan.shape, plan.spacing, plan.absorbing, float(plan.dt), int(plan.nt), int(device))
        if self._solver is None or self._solver.key != key:
            self.clear_operators()
            self._solver = Solver(plan.dim, plan.shape, plan.spacing, plan.absorbing, plan.dt, plan.nt,
                                  device=device)
        return self._solver

    # forward -----------------------------------------------------------------------------

    def before_forward(self, wavelets, vp, rho=None, alpha=None, **kwargs):
        problem = kwargs.get('problem')
        shot = problem.shot
        platform = kwargs.get('platform', 'gpu')
        if platform not in (None, 'gpu', 'amd-hip', 'cpu'):
            raise ValueError('Unknown platform %s' % platform)
        device = kwargs.get('deviceid', self.deviceid)
        mode = kwargs.get('adjoint_mode', 'compat')
        if mode not in ('compat', 'exact'):
            raise ValueError('adjoint_mode must be "compat" or "exact"')

        save = kwargs.get('save_wavefield', False) or _needs_grad(vp, rho, alpha)
        plan = build_plan(self.space, self.time, wavelets.data, vp.extended_data,
                          rho=None if rho is None else rho.extended_data,
                          alpha=None if alpha is None else alpha.extended_data,
                          source_coordinates=shot.source_coordinates,
                          receiver_coordinates=shot.receiver_coordinates,
                          kernel=kwargs.get('kernel'),
                          boundary_type=kwargs.get('boundary_type', 'sponge_boundary_2'),
                          interpolation_type=kwargs.get('interpolation_type', 'linear'),
                          attenuation_power=kwargs.get('attenuation_power', 0),
                          save_undersampling=kwargs.get('save_undersampling'),
                          time_bounds=kwargs.get('time_bounds'),
                          diff_source=kwargs.get('diff_source', False),
                          grad_vp=bool(vp.needs_grad), grad_rho=bool(rho is not None and rho.needs_grad),
                          grad_alpha=bool(alpha is not None and alpha.needs_grad),
                          grad_src=bool(getattr(wavelets, 'needs_grad', False)))
        plan.adjoint_mode = mode
        for msg in plan.conditions.messages():
            logger.debug('(ShotID %d) %s', shot.id, msg)
        if 'OT4' in plan.kernel and plan.boundary == PML:
            logger.warning('(ShotID %d) Higdon and PML boundary conditions are unstable beyond OT2 limits',
                           shot.id)
        self.kernel = plan.kernel
        self.undersampling_factor = plan.undersampling
        self.boundary_type = plan.boundary_type
        self.interpolation_type = plan.interpolation_type
        self.attenuation_power = plan.attenuation_power
        self._bandwidth = (plan.conditions.f_min, plan.conditions.f_centre, plan.conditions.f_max)

        if not save:
            save_mode = 0
        elif mode == 'exact':
            save_mode = 3
        elif _needs_grad(wavelets, rho, alpha):
            save_mode = 2
        else:
            save_mode = 1
        checkpoint = 0
        if save_mode == 2:
            checkpoint = raw_checkpoint(plan, kwargs.get('checkpoint_snapshots'),
                                        kwargs.get('save_budget', SAVE_BUDGET_BYTES))
            if checkpoint:
                save_mode = 5
                logger.debug('(ShotID %d) raw snapshots recomputed in segments of %d', shot.id, checkpoint)
        self._save_mode = save_mode
        self._cache = None

        solver = self._get_solver(plan, device)
        solver.set_scheme(plan.kernel_order, plan.rho_mode, plan.att_power, plan.boundary, plan.diff_source,
                          plan.undersampling, save_mode, plan.tb0, plan.tb1, plan.grad_vp, plan.grad_rho,
                          plan.grad_src, checkpoint=checkpoint)
        dump = kwargs.get('dump_forward_wavefield', False)
        dump_id = kwargs.get('dump_wavefield_id', shot.id)
        solver.set_dump(_dump_factor(dump, plan) if dump and dump_id == shot.id else 0, 0)
        solver.set_medium(plan.vp, plan.rho, plan.alpha)
        if plan.boundary == SPONGE:
            solver.set_sponge(plan.sponge_tables)
        elif plan.boundary == PML:
            solver.set_pml(plan.pml)
        solver.set_sources(plan.src_idx, plan.src_w)
        solver.set_receivers(plan.rec_idx, plan.rec_w)
        tune = kwargs.get('autotune')
        if tune is None:
            tune = plan.dim == 3 and int(np.prod(plan.shape)) >= _AUTOTUNE_MIN_POINTS
        if tune:
            from ..autotune import tune_solver
            tile = tune_solver(solver, plan, save_mode, device=device)
            if tile is not None:
                logger.debug('(ShotID %d) update tile %s', shot.id, tile)
        self._plan = plan

    def run_forward(self, wavelets, vp, rho=None, alpha=None, **kwargs):
        self._traces = None
        if np.linalg.norm(wavelets.data) < 1e-31:
            logger.warning('(ShotID %d) Empty wavelets, not running forward', kwargs['problem'].shot_id)
            return
        self._traces = self._solver.forward(self._plan.src_data)

    def after_forward(self, wavelets, vp, rho=None, alpha=None, **kwargs):
        problem = kwargs['problem']
        shot = problem.shot
        plan = self._plan
        dump = kwargs.get('dump_forward_wavefield', False)
        if dump and kwargs.get('dump_wavefield_id', shot.id) == shot.id and self._traces is not None:
            _dump_wavefield(self._solver.dump_to_host(0), 'forward_wavefield-Shot%05d' % shot.id, self.space,
                            problem, kwargs)
        if kwargs.get('cache_forward', False) and self._save_mode in (1, 2) and self._traces is not None:
            self._cache_snapshots(problem, kwargs.get('cache_location'))
        data = self._traces.T if self._traces is not None else np.zeros((plan.nrec, plan.nt), np.float32)
        traces = shot.observed.alike(name='modelled', data=np.ascontiguousarray(data, dtype=np.float32))
        if kwargs.get('deallocate', False) and self._save_mode == 0:
            self.clear_operators()
        return traces

    # adjoint -----------------------------------------------------------------------------

    def before_adjoint(self, adjoint_source, wavelets, vp, rho=None, alpha=None, **kwargs):
        if self._plan is None or self._solver is None:
            raise RuntimeError('adjoint called before a forward run of this shot')
        data = adjoint_source.data if hasattr(adjoint_source, 'data') else adjoint_source
        self._res = adjoint_source_samples(self._plan, data)
        self._res_norm = np.linalg.norm(data)
        if self._cache is not None:
            self._restore_snapshots()
        problem = kwargs.get('problem')
        dump = kwargs.get('dump_adjoint_wavefield', False)
        shot_id = problem.shot_id if problem is not None else None
        self._solver.set_dump(0, _dump_factor(dump, self._plan) if dump and
                              kwargs.get('dump_wavefield_id', shot_id) == shot_id else 0)
        self.init_grad(wavelets, vp, rho, alpha, **kwargs)

    def run_adjoint(self, adjoint_source, wavelets, vp, rho=None, alpha=None, **kwargs):
        plan = self._plan
        self._grads = None
        if self._res_norm < 1e-31:
            logger.warning('(ShotID %d) Empty adjoint source, not running adjoint', kwargs['problem'].shot_id)
            return
        if plan.adjoint_mode == 'exact':
            acc = self._solver.adjoint_exact(self._res, want_m=plan.kernel_order == 4,
                                             want_b=plan.rho_mode == RHO_VARIABLE)
            self._grads = ('exact', exact_gradients(plan, acc, acc['lam_src']))
        else:
            out = self._solver.adjoint(self._res, want_vp=bool(vp.needs_grad),
                                       want_rho=bool(rho is not None and rho.needs_grad),
                                       want_src=bool(getattr(wavelets, 'needs_grad', False)))
            self._grads = ('compat', out)

    # cache_forward (devito.py:632-668, 820-842): the interior of the snapshots in host memory, or
    # in a .npy file under cache_location, between the forward and the adjoint run
    def _cache_snapshots(self, problem, location):
        extra = tuple(self.space.extra)
        data = self._solver.saved_to_host(margin=extra, release=True)
        if location is None:
            self._cache = data
            return
        import glob
        import os
        import tempfile
        if self._cache_folder is None:
            prev = glob.glob(os.path.join(location, 'stride-*'))
            self._cache_folder = prev[0] if prev else tempfile.mkdtemp(prefix='stride-', dir=location)
        filename = os.path.join(self._cache_folder, '%s-%s-%05d.npy' % (problem.name, 'P', problem.shot_id))
        np.save(filename, data)
        self._cache = filename

    def _restore_snapshots(self):
        import os
        data = self._cache
        if isinstance(data, str):
            filename, data = data, np.load(data)
            os.remove(filename)
        self._solver.saved_from_host(data, margin=tuple(self.space.extra))
        self._cache = None

    def after_adjoint(self, adjoint_source, wavelets, vp, rho=None, alpha=None, **kwargs):
        problem = kwargs.get('problem')
        dump = kwargs.get('dump_adjoint_wavefield', False)
        if dump and problem is not None and kwargs.get('dump_wavefield_id', problem.shot_id) == problem.shot_id \
                and self._res_norm >= 1e-31:
            _dump_wavefield(self._solver.dump_to_host(1), 'adjoint_wavefield-Shot%05d' % problem.shot_id, self.space,
                            problem, kwargs)
        grads = self.get_grad(wavelets, vp, rho, alpha, **kwargs)
        self.deallocate_wavefield(deallocate=kwargs.get('deallocate', False))
        return grads

    # gradients ---------------------------------------------------------------------------

    def _inner(self, field):
        return np.ascontiguousarray(field[self.space.inner], dtype=np.float32)

    def prepare_grad_vp(self, vp, **kwargs):
        return ()

    def init_grad_vp(self, vp, **kwargs):
        pass

    def get_grad_vp(self, vp, **kwargs):
        kind, g = self._grads if self._grads is not None else (None, None)
        shape = self.space.shape
        if kind == 'exact':
            grad = self._inner(g['vp'])
            prec = None
        elif kind == 'compat':
            grad, prec = compat_finish(self._inner(g['grad_vp']), self._inner(g['prec_vp']),
                                       vp.data, transform=vp.transform)
        else:
            grad, prec = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        out = _gradient_field(vp, 'vp', grad, prec)
        _dump_local(out, 'vp', kwargs)
        return out

    def prepare_grad_rho(self, rho, **kwargs):
        return ()

    def init_grad_rho(self, rho, **kwargs):
        pass

    def get_grad_rho(self, rho, **kwargs):
        kind, g = self._grads if self._grads is not None else (None, None)
        shape = self.space.shape
        if kind == 'exact':
            if 'rho' not in g:
                raise ValueError('exact density gradient needs variable density')
            grad, prec = self._inner(g['rho']), None
        elif kind == 'compat':
            grad, prec = self._inner(g['grad_rho']), self._inner(g['prec_rho'])
        else:
            grad, prec = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        return _gradient_field(rho, 'rho', grad, prec)

    # the reference has no alpha gradient (problem_type.py:226-231 raises for it); the exact
    # adjoint provides one
    def _check_alpha(self):
        if self._plan is None or self._plan.adjoint_mode != 'exact':
            raise ValueError('Variable alpha not implemented')

    def prepare_grad_alpha(self, alpha, **kwargs):
        self._check_alpha()
        return ()

    def init_grad_alpha(self, alpha, **kwargs):
        self._check_alpha()

    def get_grad_alpha(self, alpha, **kwargs):
        kind, g = self._grads if self._grads is not None else (None, None)
        grad = self._inner(g['alpha']) if kind == 'exact' else np.zeros(self.space.shape, np.float32)
        return _gradient_field(alpha, 'alpha', grad, None)

    def prepare_grad_wavelets(self, wavelets, **kwargs):
        return ()

    def init_grad_wavelets(self, wavelets, **kwargs):
        if self._plan is not None and self._plan.adjoint_mode == 'exact' and self._plan.diff_source:
            raise ValueError('exact wavelet gradient with diff_source is not supported')

    def get_grad_wavelets(self, wavelets, **kwargs):
        kind, g = self._grads if self._grads is not None else (None, None)
        plan = self._plan
        if kind == 'exact':
            window = np.pad(_tukey_for(plan), ((0, plan.nt - plan.tb1),), mode='constant')
            data = (g['src_samples'] * window[:, None]).T
        elif kind == 'compat':
            data = g['grad_src'].T
        else:
            data = np.zeros(wavelets.data.shape, np.float32)
        return Traces('wavelets_grad', data=np.ascontiguousarray(data, dtype=np.float32))

    # convenience -------------------------------------------------------------------------

    def wavefield(self, which=0):
        """Last (0) or previous (1) wavefield level left in the solver, extended grid."""
        return self._solver.wavefield(which)


def _gradient_field(variable, name, grad, prec):
    """The gradient as the variable's own kind of field, inner-shaped, with ``.prec``:
    ``vp.alike(name='vp_grad', data=..., shape=..., extended_shape=..., inner=None)``
    (stride/physics/iso_acoustic/devito.py:1088-1095, 1195-1202). With Stride's ScalarField
    that keeps its grid, transform and graph machinery."""
    out = variable.alike(name='%s_grad' % name, data=grad, shape=grad.shape, extended_shape=grad.shape,
                         inner=None)
    out.prec = None if prec is None else variable.alike(name='%s_prec' % name, data=prec, shape=prec.shape,
                                                        extended_shape=prec.shape, inner=None)
    return out


def _dump_local(grad, name, kwargs):
    """dump_local_grad / dump_local_prec (devito.py:1097-1111): the shot's gradient and
    preconditioner written to the problem's output folder, version = iteration.abs_id + 1."""
    problem = kwargs.get('problem')
    if problem is None:
        return
    iteration = kwargs.get('iteration')
    version = iteration.abs_id + 1 if iteration is not None else 0
    if kwargs.get('dump_local_grad', False):
        grad.dump(path=problem.output_folder, project_name=problem.name,
                  parameter='%s_local_grad-Shot%05d' % (name, problem.shot_id), version=version)
    if kwargs.get('dump_local_prec', False) and grad.prec is not None:
        grad.prec.dump(path=problem.output_folder, project_name=problem.name,
                       parameter='%s_local_src_prec-Shot%05d' % (name, problem.shot_id), version=version)


def _dump_wavefield(data, name, space, problem, kwargs):
    """The dumped levels (slots, extended shape) written as the reference's StructuredData dump
    (devito.py:617-623, 955-963), version = iteration.abs_id + 1."""
    iteration = kwargs.get('iteration')
    version = iteration.abs_id + 1 if iteration is not None else 0
    field = ScalarField(name, space=space, data=data, shape=data.shape, extended_shape=data.shape)
    return field.dump(path=problem.output_folder, project_name=problem.name, version=version)


def _tukey_for(plan):
    from ..utils import tukey
    return tukey(plan.tb1, 0.001)


# drop-in name: user code written for Stride's Devito operator imports this name
IsoAcousticDevito = IsoAcousticHIP
