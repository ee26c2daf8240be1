"""Device propagation driver: the hot-path part of the reference's run loop.

Mirrors python/run3d.py:105-225 for what lies on the H.psi -> step -> surface
path:
  * the step-count rule (``schedule``): T = last pulse end + post_time, or
    ``end_time`` (refused if it precedes the end of the pulses); A(T) must
    vanish; nsteps = ceil(T / dt) rounded up to a multiple of the surface
    stride, dt <- T / nsteps (run3d.py:105-112);
  * the stepping loop (``propagate``): the ``max_steps`` prefix cut back to a
    surface record (run3d.py:160-162), one Arnoldi or CF4-Pade step per
    iteration on the device engine (run3d.py:201-208), a surface frame every
    ``stride`` steps (run3d.py:209-212) and the observables every
    ``checkpoint_every`` steps and at the end (run3d.py:214-223);
  * the observables (``observables``): the inner-region norm with the real
    element weights (interior weight / |weight|, run3d.py:216-219), the
    ground-state population |<psi_0|psi>|^2 (run3d.py:220) and the exchange
    error of the state (python/helium3d.py:101-103);
  * the online tSURFF chain (``OnlineSurface``): device flux, channel
    projection onto the ionic duals chi(t) and coherent Volkov accumulation
    (python/streaming_surface.py:252-263 -> surface_contract.py:41-56 ->
    streaming_surface.py:54-64), the duals from the backward adjoint CF4 of
    ``laqphysics_amd.ionic``;
  * a run from a configuration dictionary (``setup``, ``run_config``): the
    real/complex operator pair (python/run3d.py:23-39), the device ground
    state on the real grid carried onto the ECS grid by sqrt(w_ecs / w_real)
    per electron (run3d.py:134-136) and the stepping loop above with the
    configuration's Krylov / linear-solver settings (run3d.py:198-208).
  * the run's files (``run``): config.json, history.json, checkpoint.npz and
    run.json with the reference's keys, atomic replacement and size cap, the
    configuration signature sha256(json.dumps(config, sort_keys=True)), the
    refusals of an existing output without ``resume`` and of a resume under a
    changed configuration (python/run3d.py:74-136, 163-171, 227-252,
    python/surface_storage.py:9-26), and with ``spectrum`` the extracted
    channel spectrum in spectrum.npz (python/run3d.py:281-296).
Signals, raw surface-history storage and the mixed gauge stay with the
reference.
"""
import hashlib
import json
import time
from pathlib import Path

import numpy as np
import torch

from . import hip_backend, implicit
from .ionic import Ionic, backward_duals
from .pulses import Pulse, vector_function
from .storage import DEFAULT_FILE_BYTES, HARD_FILE_BYTES, ShardedArray, atomic_json, capped_npz, file_limit  # noqa: F401
from .surface import SurfaceContractor, VolkovAccumulator, directions


def schedule(pulses, dt=.05, stride=1, end_time=None, post_time=60.):
    """(T, nsteps, dt) of a run, by the reference's rule (python/run3d.py:105-112)."""
    stride = int(stride)
    if stride < 1 or not dt > 0:
        raise ValueError('stride must be >= 1 and dt > 0')
    ends = max(p.start + p.duration for p, pol in pulses)
    T = ends + post_time
    if end_time is not None:
        if end_time < ends:
            raise ValueError('end_time precedes end of pulse')
        T = float(end_time)
    avec = vector_function(pulses)
    if np.max(abs(avec(T))) > 1e-10:
        raise ValueError('This production driver requires zero final vector potential for its bound-state spectra')
    nsteps = int(np.ceil(T / dt))
    nsteps += (-nsteps) % stride
    return T, nsteps, T / nsteps


def observables(h, psi, ground=None):
    """Inner-region norm, ground population and exchange error of a host state
    (python/run3d.py:214-220)."""
    u = np.asarray(psi).reshape(h.shape, order='F')
    f = h.grid.interior_weights / abs(h.grid.weights)
    row = {'real_region_norm': float(np.sum(abs(u) ** 2 * f[:, None, None] * f[None, :, None])),
           'exchange_error': h.exchange_error(psi)}
    if ground is not None:
        row['ground_population'] = float(abs(np.vdot(ground, psi)) ** 2)
    return row


class OnlineSurface:
    """tSURFF channel amplitudes accumulated on the device, one frame per record.

    ``labels`` are the ionic channels (n, l, m); ``times`` the frame times
    (index k at times[k]); ``radius`` the surface (python/run3d.py:97 prepares
    it on the propagated operator, the ionic side uses its product basis);
    ``substeps`` / ``propagator`` choose the dual propagation between frames
    (``laqphysics_amd.ionic.backward_duals``); ``ion_factory(base, magnetic)``
    and ``ionic_field(t)`` replace the ionic operator and its field, as the
    mixed gauge's ``MixedIonic`` with its 12 coefficients does
    (python/surface_projection.py:26-35); ``ionic_device`` propagates the duals
    on that device, keeping endpoints every ``ionic_block`` frames
    (python/streaming_surface.py:125-243; None: every frame while all frames
    take at most 8 GiB of device memory, else 128), with the reference's
    preparation files in ``ionic_folder`` when one is given."""

    def __init__(self, engine, radius, labels, times, avec, energy, angles=None, folded=True, device=None,
                 substeps=1, propagator='cf4', progress=None, ion_factory=None, ionic_field=None,
                 ionic_device=None, ionic_block=None, ionic_folder=None, configuration='', mode='spectrum',
                 transfer_times=(), max_file_bytes=DEFAULT_FILE_BYTES):
        h = engine.h
        base = getattr(h, 'uncoupled', h)
        h.prepare_surface(radius)
        if base is not h:
            base.prepare_surface(radius)
        self.engine = engine
        self.avec = avec
        self.times = np.asarray(times, float)
        key = (lambda n, l, m: (n, l, abs(m))) if folded else (lambda n, l, m: (n, l, m))
        unique = list(dict.fromkeys(key(*x) for x in labels))
        columns = [unique.index(key(*x)) for x in labels]
        magnetic = {m for n, l, m in unique} if folded else None
        self.ion = (ion_factory or Ionic)(base, magnetic)
        if ionic_block is None:
            frames = 16 * self.ion.na * self.ion.n * len(unique) * len(self.times)
            ionic_block = 1 if frames <= 8 << 30 else 128
        self.chis = backward_duals(self.ion, unique, self.times, ionic_field or avec, substeps, propagator, progress,
                                   ionic_device, int(ionic_block), ionic_folder, configuration)
        self.contractor = SurfaceContractor(base, self.ion, labels, unique, columns, folded,
                                            getattr(h, 'transform', None), device or engine.device)
        self.accumulator = VolkovAccumulator(self.contractor, energy,
                                             angles if angles is not None else directions(folded), avec)
        # storage modes of python/streaming_surface.py:153-154, 188-190: 'projected'
        # also keeps every frame's channel projection q in surface_online/projected
        if mode not in ('spectrum', 'projected'):
            raise ValueError('unknown online storage mode')
        self.mode = mode
        self.limit = file_limit(max_file_bytes)
        self.configuration = configuration
        self.projected = None
        if mode == 'projected':
            if ionic_folder is None:
                raise ValueError('projected storage needs an output folder')
            folder = Path(ionic_folder) / 'projected'
            self.projected = ShardedArray(folder, (len(self.times), *self.contractor.shape),
                                          mode='r+' if (folder / 'array.json').exists() else 'w+',
                                          max_file_bytes=self.limit)
        self.transfer_times = [float(t) for t in transfer_times]
        self.transfers = []
        self.generation = -1

    def record(self, index, state):
        t = self.times[index]
        frame = self.engine.flux_device(state, self.avec(t))
        chi = self.chis[index]
        q = self.contractor(chi, frame)
        self.accumulator.add(index, t, q)
        if self.projected is not None:
            self.projected[index] = q.cpu().numpy() if isinstance(q, torch.Tensor) else np.asarray(q)
        for requested in self.transfer_times:
            if int(np.argmin(abs(self.times - requested))) == index:
                self.transfers.append(self._transfer(requested, index, chi))

    def _transfer(self, requested, index, chi):
        """Amplitudes of the prepared He+ 1s in the recorded bound channels at a
        frame time (python/streaming_surface.py:256-263)."""
        chi = chi.cpu().numpy() if isinstance(chi, torch.Tensor) else np.asarray(chi)
        initial = self.ion.final_states([(1, 0, 0)])[:, 0]
        amplitude = (chi.conj().T @ initial)[self.contractor.columns]
        if self.contractor.folded:
            amplitude = np.where(np.array(self.contractor.labels)[:, 2] == 0, amplitude, 0.)
        return {'requested_time': requested, 'actual_time': float(self.times[index]),
                'labels': [list(x) for x in self.contractor.labels],
                'amplitude_real': amplitude.real.tolist(), 'amplitude_imag': amplitude.imag.tolist(),
                'probabilities': (abs(amplitude) ** 2).tolist(),
                'scope': 'transition from prepared ionic 1s into recorded final bound channels'}

    def stage_checkpoint(self, folder, kernel):
        """Write the accumulator to the bank the last checkpoint does not name
        (surface_online/accumulator_{0,1}.npz with the reference's keys,
        python/streaming_surface.py:265-270); returns the bank, which the
        wavefunction checkpoint records as surface_generation."""
        bank = 1 if self.generation == 0 else 0
        if self.projected is not None:
            self.projected.flush()
        Path(folder).mkdir(parents=True, exist_ok=True)
        capped_npz(Path(folder) / f'accumulator_{bank}.npz', self.limit, **self.accumulator.payload(),
                   signature=self.configuration, kernel=kernel, transfers=json.dumps(self.transfers),
                   maximum_ionic_residual=0.)
        return bank

    def restore_bank(self, folder, bank, frame):
        """Continue from accumulator bank ``bank`` written after surface frame
        ``frame`` (python/streaming_surface.py:272-278).  A bank of another
        configuration, or of another frame than the wavefunction checkpoint's,
        is refused; the writer's ``kernel`` string (the reference records a
        hash of its own sources there) is returned, not compared, so a bank the
        reference wrote resumes here."""
        with np.load(Path(folder) / f'accumulator_{int(bank)}.npz') as d:
            if str(d['signature']) != self.configuration:
                raise ValueError('spectrum checkpoint signature mismatch')
            index = int(d['sample_index'])
            if index != frame and not (frame == 0 and index == -1):
                raise ValueError('wavefunction/spectrum checkpoint mismatch')
            self.accumulator.restore(d)
            self.transfers = json.loads(str(d['transfers']))
            kernel = str(d['kernel'])
        self.generation = int(bank)
        return kernel

    def close(self):
        if self.projected is not None:
            self.projected.close()

    def amplitude(self):
        return self.accumulator.host_amplitude()


def propagate(engine, psi0, avec, dt, nsteps, integrator='arnoldi', tol=1e-10, maxdim=48, linear_tolerance=1e-10,
              reference_energy=0., restart=None, stride=1, max_steps=None, checkpoint_every=100, surface=None,
              ground=None, log=None, on_checkpoint=None, begin=0, totals=None, field=None):
    """Run steps begin .. stop-1 of an ``nsteps`` schedule on the device engine.

    Returns (device state, history rows).  A row holds step, time, the
    observables and the integrator's info dictionary, as the reference's
    history.json does.  ``on_checkpoint(row, host_state)`` is called at every
    history row.  ``begin`` > 0 continues a run whose state after step
    ``begin`` is ``psi0`` (the surface frame of that step is already recorded);
    ``totals`` (a dict) accumulates the CF4 steps' largest linear residual and
    their linear iterations (python/run3d.py:204).  ``field(t)`` is the
    propagation field (default ``avec``; the 12 mixed-gauge coefficients for a
    ``MixedHamiltonian``, python/run3d.py:182-188); surface frames use avec."""
    if integrator not in ('arnoldi', 'cf4-pade'):
        raise ValueError('time_integrator must be arnoldi or cf4-pade')
    if integrator == 'cf4-pade' and getattr(engine, 'separable', None) is None:
        engine.separable = implicit.SeparablePreconditioner(engine)
    h = engine.h
    state = engine.state(psi0)
    field = avec if field is None else field
    stop = nsteps if max_steps is None else min(nsteps, begin + int(max_steps))
    stop -= stop % stride
    if surface is not None and begin == 0:
        surface.record(0, state)
    history = []
    for i in range(begin, stop):
        t = i * dt
        if integrator == 'cf4-pade':
            state, info = implicit.cf4_step(engine, state, field, t, dt, reference_energy, tol=linear_tolerance,
                                            restart=restart)
            if totals is not None:
                totals['maximum_linear_residual'] = max(totals.get('maximum_linear_residual', 0.),
                                                        info['linear_residual'])
                totals['linear_iterations'] = totals.get('linear_iterations', 0) + info['linear_iterations']
        else:
            state, info = hip_backend.step(lambda s, y: engine.apply(y, field(s), True), state, t, dt, tol=tol,
                                           maxdim=maxdim)
        if surface is not None and (i + 1) % stride == 0:
            surface.record((i + 1) // stride, state)
        if ((i + 1) % checkpoint_every == 0 and (i + 1) % stride == 0) or i + 1 == stop:
            host = engine.host(state)
            row = {'step': i + 1, 'time': (i + 1) * dt, **observables(h, host, ground), **info}
            history.append(row)
            if log is not None:
                log(row)
            if on_checkpoint is not None:
                on_checkpoint(row, host)
    return state, history


def setup(config):
    """(operator on the real grid, operator on the ECS grid) of a run
    configuration (python/run3d.py:23-39): product, coupled (``total_Lmax``)
    or bipolar (``angular_representation``) channels, same refusals."""
    from .fedvr import make_grid
    from .helium3d import Helium
    gconfig = dict(config['radial'])
    angle = gconfig.pop('ecs_angle', 0.)
    real = make_grid(**gconfig)
    complex_grid = make_grid(**gconfig, ecs_angle=angle)
    args = {'lmax': config['lmax'], 'M': config.get('M', 0), 'cutoff_radii': config['cutoff_radii']}
    if config.get('angular_representation') == 'bipolar':
        from .bipolar import BipolarHelium
        natural = config.get('natural_parity', False)
        if natural and (config.get('M', 0) != 0 or any(p.get('polarization', 'z') != 'z'
                                                        for p in config.get('pulses', []))):
            raise ValueError('natural-parity restriction requires M=0 and z polarization')
        return tuple(BipolarHelium(g, total_Lmax=config['total_Lmax'], natural_parity=natural, **args)
                     for g in (real, complex_grid))
    pair = (Helium(real, **args), Helium(complex_grid, **args))
    if 'total_Lmax' in config:
        if config.get('M', 0) != 0 or any(p.get('polarization', 'z') != 'z' for p in config.get('pulses', [])):
            raise ValueError('total_Lmax compression currently supports M=0 and z polarization only')
        from .coupled_angular import CoupledHelium
        pair = tuple(CoupledHelium(h, config['total_Lmax']) for h in pair)
    return pair


def config_pulses(config):
    """[(Pulse, polarisation)] of a configuration (python/run3d.py:41-42)."""
    return [(Pulse(**p['pulse']), p.get('polarization', 'z')) for p in config['pulses']]


def transfer(hreal, h, ground):
    """Real-grid state carried onto the ECS grid: psi(i, j) * sqrt(w_i w_j)
    ratio of the two grids' weights (python/run3d.py:135-136)."""
    ratio = np.sqrt(h.grid.weights / hreal.grid.weights)
    u = np.asarray(ground).reshape(h.shape, order='F')
    return (u * ratio[:, None, None] * ratio[None, :, None]).ravel(order='F')


def spectrum_surface(engine, config, nsteps, dt, avec, progress=None, field=None, ionic_device=None,
                     ionic_folder=None):
    """The tSURFF channel spectrum of a run as the reference extracts it from
    its surface history (python/run3d.py:281-296 -> python/surface3d.py:50-111):
    frames every ``surface_stride`` steps, ionic channels ``ionic_channels``
    (default He+ 1s and 2p0 for M = 0, 1s and the three 2p for all M),
    energies ``spectrum_energy`` (default linspace(.3, .9, 121)), 61 polar
    angles on [0, pi] at phi = 0 for M = 0 or the full sphere (20
    Gauss-Legendre polar x 32 azimuthal angles) for all M, duals by the
    configuration's ``ionic_propagator`` (default expm) with one sub-step per
    propagation step, accumulated online on the device.  ``ionic_device`` (or a
    cuda ``storage.ionic_device`` of the configuration) propagates the CF4
    duals on the device, with endpoints every ``storage.ionic_block_frames``
    frames (by default every frame while all frames take at most 8 GiB, else
    128), and keeps the reference's preparation files in ``ionic_folder`` when
    one is given."""
    M = config.get('M', 0)
    if M not in (0, None):
        raise ValueError('the device spectrum covers M = 0 and all-M runs')
    stride = config.get('surface_stride', 1)
    default = [(1, 0, 0), (2, 1, 0)] if M == 0 else [(1, 0, 0), (2, 1, -1), (2, 1, 0), (2, 1, 1)]
    labels = [tuple(a) for a in config.get('ionic_channels', default)]
    if any(n <= l or l < 0 or l > config['lmax'] or abs(m) > l for n, l, m in labels):
        raise ValueError('ionic_channels must contain valid (n,l,m) within lmax')
    energy = np.linspace(*config.get('spectrum_energy', [.3, .9, 121]))
    if M == 0:
        theta = np.linspace(0, np.pi, 61)
        angles = (theta, np.zeros(61), np.ones(61))
    else:
        z, w = np.polynomial.legendre.leggauss(20)
        theta, phi = np.meshgrid(np.arccos(z), np.arange(32) * 2 * np.pi / 32, indexing='ij')
        angles = (theta.ravel(), phi.ravel(), np.repeat(w, 32) * 2 * np.pi / 32)
    linear = all(pol == 'z' for p, pol in config_pulses(config))
    times = np.arange(nsteps // stride + 1) * dt * stride
    factory = None
    gauge = config.get('gauge', {'type': 'velocity'})
    if gauge['type'] == 'mixed':
        # python/run3d.py:286-291: MixedIonic duals driven by the 12 coefficients
        from .mixed_gauge import MixedIonic

        def factory(base, magnetic):
            return MixedIonic(base, magnetic, gauge['inner'], gauge['outer'])
    storage = config.get('storage', {})
    if ionic_device is None and str(storage.get('ionic_device', '')).startswith('cuda'):
        ionic_device = storage['ionic_device']
    return OnlineSurface(engine, config['surface'], labels, times, avec, energy, angles, folded=linear,
                         substeps=stride, propagator=config.get('ionic_propagator', 'expm'), progress=progress,
                         ion_factory=factory, ionic_field=field if factory is not None else None,
                         ionic_device=ionic_device, ionic_block=storage.get('ionic_block_frames'),
                         ionic_folder=ionic_folder, configuration=signature(config),
                         mode=storage.get('mode', 'spectrum'), transfer_times=config.get('ionic_transfer_times', []),
                         max_file_bytes=storage.get('max_file_bytes', DEFAULT_FILE_BYTES))


def numerical_signature():
    """Identity of this engine's numerics: a hash of the package's Python and
    HIP sources (the reference records a hash of its own sources in the same
    place, python/streaming_surface.py:17-24)."""
    root = Path(__file__).resolve().parent
    h = hashlib.sha256(b'laqphysics_amd')
    for path in sorted(list(root.glob('*.py')) + list((root / 'csrc').glob('*'))):
        h.update(path.name.encode() + b'\0')
        h.update(path.read_bytes())
    return h.hexdigest()


def _engine_and_field(config, h, idx, pulses, device):
    """Device operator and propagation field of a configuration's gauge
    (python/run3d.py:98-102, 181-188): the velocity gauge with A(t), or the mixed
    gauge with its 12 coefficients, refused unless the switch ends inside the
    surface stencil."""
    gauge = config.get('gauge', {'type': 'velocity'})
    if gauge['type'] not in ('velocity', 'mixed'):
        raise ValueError('unknown gauge')
    avec = vector_function(pulses)
    if gauge['type'] == 'velocity':
        return hip_backend.HipHamiltonian(h, device), avec
    if idx is None or gauge['outer'] >= min(h.r[idx].real):
        raise ValueError('mixed gauge requires Torch and a switch completed before the surface stencil')
    from .mixed_gauge import MixedHamiltonian, parameters
    from .pulses import electric_function
    electric = electric_function(pulses)
    engine = MixedHamiltonian(h, device, gauge['inner'], gauge['outer'])
    return engine, (lambda t: parameters(avec, electric, t))


def run_config(config, device='cuda:0', max_steps=None, ground_cache=None, log=None, on_checkpoint=None,
               spectrum=False, dual_progress=None, ionic_device=None, preconditioner_precision='complex128'):
    """Ground state, transfer and propagation of one configuration on the device.

    Returns (device state, history rows, meta) with meta holding
    ``ground_energy``, ``ground_residual``, ``dt``, ``nsteps`` and ``nrad`` /
    ``channels`` as the reference's run.json does (python/run3d.py:227-229).
    With ``spectrum`` the tSURFF channel amplitudes (labels, energies, polar
    angles) of the whole run are accumulated on the way and returned in
    ``meta['amplitudes']`` with ``meta['energy']``, ``meta['theta']`` and ``meta['phi']``;
    ``dual_progress(done, total)`` reports the ionic dual preparation, which
    ``ionic_device`` moves to the device.  ``preconditioner_precision`` is the
    CF4 preconditioner's (python/run3d.py:193; 'complex64' runs the flexible
    GMRES at restart 16, as the reference's run3d does)."""
    hreal, h = setup(config)
    idx = h.prepare_surface(config['surface']) if 'surface' in config else None
    pulses = config_pulses(config)
    engine, field = _engine_and_field(config, h, idx, pulses, device)
    if config.get('M', 0) is not None and any(pol != 'z' for p, pol in pulses):
        raise ValueError('circular polarization requires M=null (all magnetic channels)')
    stride = config.get('surface_stride', 1)
    T, nsteps, dt = schedule(pulses, config.get('dt', .05), stride, config.get('end_time'),
                             config.get('post_time', 60.))
    E, gs, gres = hreal.ground(tol=config.get('ground_tolerance', 1e-9), cache_dir=ground_cache, device=device)
    psi = transfer(hreal, h, gs)
    avec = vector_function(pulses)
    surface = (spectrum_surface(engine, config, nsteps, dt, avec, dual_progress, field, ionic_device)
               if spectrum else None)
    integrator = config.get('time_integrator', 'arnoldi')
    if integrator == 'cf4-pade':
        engine.separable = implicit.SeparablePreconditioner(engine, precision=preconditioner_precision)
    state, history = propagate(engine, psi, avec, dt, nsteps, integrator,
                               tol=config.get('krylov_tolerance', 1e-10), maxdim=config.get('krylov_dimension', 48),
                               linear_tolerance=config.get('linear_tolerance', 1e-10), reference_energy=E,
                               stride=stride, max_steps=max_steps, checkpoint_every=config.get('checkpoint_every', 100),
                               ground=gs, log=log, on_checkpoint=on_checkpoint, surface=surface, field=field)
    meta = {'ground_energy': E, 'ground_residual': gres, 'nrad': h.n, 'channels': h.nc, 'dt': dt,
            'nsteps': nsteps, 'engine': engine, 'ground': gs}
    if surface is not None:
        meta.update(amplitudes=surface.amplitude(), energy=surface.accumulator.energy,
                    theta=surface.accumulator.theta, phi=surface.accumulator.phi, labels=surface.contractor.labels)
    return state, history, meta


def signature(config):
    """sha256 of the canonical JSON of a configuration (python/run3d.py:93)."""
    return hashlib.sha256(json.dumps(config, sort_keys=True).encode()).hexdigest()


def run(config, out, resume=False, max_steps=None, device='cuda:0', ground_cache=None, spectrum=False, log=None,
        ionic_device=None, preconditioner_precision='complex128'):
    """A configuration's whole run with the reference's output directory.

    Files: config.json (at the start), history.json and checkpoint.npz (psi,
    step, signature, ground_energy, ground, ground_residual,
    surface_generation, maximum_linear_residual_all_segments,
    linear_iterations_all_segments; with ``spectrum`` also the online
    accumulator) at every history row, run.json at the end of every
    invocation, spectrum.npz when a ``spectrum`` run completes; the spectrum's
    ionic duals run on ``ionic_device`` when one is given.  With ``spectrum``
    the accumulator is staged at every history row into the reference's
    double-banked surface_online/accumulator_{0,1}.npz, the checkpoint's
    surface_generation naming the bank written with it
    (python/run3d.py:162-171), and a resume continues from that bank.
    ``preconditioner_precision`` 'complex64' (the reference's production
    setting, python/implicit.py:27-29) runs the complex64 preconditioner with
    flexible GMRES, as the reference does; run.json records it.  Returns the
    run.json dictionary."""
    t0 = time.perf_counter()
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    storage = config.get('storage', {})
    max_file_bytes = file_limit(storage.get('max_file_bytes', DEFAULT_FILE_BYTES))
    sig = signature(config)
    checkpoint = out / 'checkpoint.npz'
    saved = None
    if resume:
        with np.load(checkpoint) as data:
            saved = {k: data[k] for k in data.files}
        if str(saved['signature']) != sig:
            raise ValueError('config changed; refusing unsafe resume')
    elif checkpoint.exists() or (out / 'flux.npy').exists() or (out / 'flux_shards').exists():
        raise FileExistsError('output exists; use --resume or another directory')
    hreal, h = setup(config)
    idx = h.prepare_surface(config['surface'])
    pulses = config_pulses(config)
    if config.get('M', 0) is not None and any(pol != 'z' for p, pol in pulses):
        raise ValueError('circular polarization requires M=null (all magnetic channels)')
    stride = config.get('surface_stride', 1)
    T, nsteps, dt = schedule(pulses, config.get('dt', .05), stride, config.get('end_time'),
                             config.get('post_time', 60.))
    if 2 * h.size * 16 + 65536 > max_file_bytes:
        raise ValueError('wavefunction checkpoint exceeds single-file cap; reduce the basis or add state partitioning')
    if saved is None:
        E, gs, gres = hreal.ground(tol=config.get('ground_tolerance', 1e-9), cache_dir=ground_cache, device=device)
        psi = transfer(hreal, h, gs)
        begin, history = 0, []
        prior = {'maximum_linear_residual': 0., 'linear_iterations': 0}
        atomic_json(out / 'config.json', config)
    else:
        psi = saved['psi']
        begin = int(saved['step'])
        E, gs, gres = float(saved['ground_energy']), saved['ground'], float(saved['ground_residual'])
        history = [r for r in json.loads((out / 'history.json').read_text()) if r['step'] <= begin]
        prior = {'maximum_linear_residual': float(saved.get('maximum_linear_residual_all_segments', 0.)),
                 'linear_iterations': int(saved.get('linear_iterations_all_segments', 0))}
    engine, field = _engine_and_field(config, h, idx, pulses, device)
    avec = vector_function(pulses)
    surface = (spectrum_surface(engine, config, nsteps, dt, avec, field=field, ionic_device=ionic_device,
                                ionic_folder=out / 'surface_online') if spectrum else None)
    kernel = numerical_signature()
    resumed_kernel = None
    if surface is not None and saved is not None:
        bank = int(saved.get('surface_generation', -1))
        if bank not in (0, 1):
            raise ValueError('checkpoint holds no spectrum accumulator; resume without spectrum')
        resumed_kernel = surface.restore_bank(out / 'surface_online', bank, begin // stride)
    totals = {'maximum_linear_residual': 0., 'linear_iterations': 0}
    if config.get('time_integrator', 'arnoldi') == 'cf4-pade':
        engine.separable = implicit.SeparablePreconditioner(engine, precision=preconditioner_precision)

    def save(step, host):
        bank = surface.stage_checkpoint(out / 'surface_online', kernel) if surface is not None else -1
        atomic_json(out / 'history.json', history)
        capped_npz(checkpoint, max_file_bytes, psi=host, step=step, signature=sig, ground_energy=E, ground=gs,
                   ground_residual=gres, surface_generation=bank,
                   maximum_linear_residual_all_segments=max(prior['maximum_linear_residual'],
                                                            totals['maximum_linear_residual']),
                   linear_iterations_all_segments=prior['linear_iterations'] + totals['linear_iterations'])
        if surface is not None:
            surface.generation = bank

    def row_saved(row, host):
        history.append(row)
        save(row['step'], host)
        if log is not None:
            log(row)
    if surface is not None and saved is None:
        save(0, psi)
    integrator = config.get('time_integrator', 'arnoldi')
    state, _ = propagate(engine, psi, avec, dt, nsteps, integrator, tol=config.get('krylov_tolerance', 1e-10),
                         maxdim=config.get('krylov_dimension', 48),
                         linear_tolerance=config.get('linear_tolerance', 1e-10), reference_energy=E, stride=stride,
                         max_steps=max_steps, checkpoint_every=config.get('checkpoint_every', 100), surface=surface,
                         ground=gs, on_checkpoint=row_saved, begin=begin, totals=totals, field=field)
    stop = history[-1]['step'] if history else begin
    cf4 = integrator == 'cf4-pade'
    meta = {'signature': sig, 'config': config, 'ground_energy': E, 'ground_residual': gres, 'nrad': h.n,
            'channels': h.nc, 'dt': dt, 'nsteps': nsteps, 'completed_steps': stop, 'complete': stop == nsteps,
            'surface_indices': [int(i) for i in idx], 'seconds_this_invocation': time.perf_counter() - t0,
            'backend': 'hip', 'device': str(device), 'precision': 'complex128',
            'ground_angular_basis': 'coupled total L=0, expanded without angular approximation',
            'time_integrator': integrator, 'gauge': config.get('gauge', {'type': 'velocity'}),
            'coulomb_contraction': engine.coulomb_backend, 'precomputed_coulomb_bytes': engine.coulomb_block_bytes,
            'preconditioner_precision': engine.separable.precision if cf4 else None,
            'requested_preconditioner_precision': preconditioner_precision if cf4 else None,
            'linear_solver': ('FGMRES' if engine.separable.flexible else 'GMRES') if cf4 else None,
            'gmres_restart': (16 if engine.separable.flexible else 32) if cf4 else None,
            'maximum_linear_residual_this_invocation': totals['maximum_linear_residual'],
            'linear_iterations_this_invocation': totals['linear_iterations'],
            'surface_storage': surface.mode if surface is not None else None, 'raw_flux_written': False,
            'ionic_transfers': surface.transfers if surface is not None else [],
            'numerical_signature': kernel, 'resumed_accumulator_kernel': resumed_kernel,
            'max_file_bytes': max_file_bytes, 'phase': 'complete' if stop == nsteps else 'propagation',
            'maximum_linear_residual_all_segments': max(prior['maximum_linear_residual'],
                                                        totals['maximum_linear_residual']),
            'linear_iterations_all_segments': prior['linear_iterations'] + totals['linear_iterations']}
    atomic_json(out / 'run.json', meta)
    if surface is not None:
        surface.close()
    if surface is not None and stop == nsteps:
        from scipy.integrate import simpson
        acc = surface.accumulator
        b = surface.amplitude()
        pes = abs(b) ** 2 * acc.k[None, :, None]
        if config.get('M', 0) == 0:
            total = 2 * np.pi * simpson(pes * np.sin(acc.theta), x=acc.theta, axis=2)
        else:
            total = np.sum(pes * acc.solid_angle_weights[None, None, :], axis=2)
        capped_npz(out / 'spectrum.npz', max_file_bytes, energy=acc.energy, theta=acc.theta, phi=acc.phi,
                   amplitudes=b, pes=pes, angle_integrated=total, labels=np.array(surface.contractor.labels),
                   ionic_projection='analytic radial c-dual', extraction_end_time=float(nsteps * dt),
                   source_signature=sig)
    return meta
