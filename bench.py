"""Propagation benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode metric|plain|batched|...]
                    [--integrator cf4|arnoldi] [--no-cpu] [--cpu-threads C]

Default line (--mode metric): the metric's own measurement.  value is the
wavefunction-step rate of the B = 16 pump-probe scan (batched_l4_L3_n441)
propagated by the batched engine over a window of the scan's uneven step
counts (scaled by K/20000), with the plain engine timed on the same
member-steps (batched_over_plain_per_wavefunction).  Fields: the adjoint
benchmark's gradient/forward time over 64 K steps, and the production CF4
step (production_l6_L4_n441) with its per-step tSURFF frame.  All state and
tables are resident in HBM before timing starts.

roofline: average launch time of the dominant hand-written kernel, measured
with HIP events on the stream it launches on (torch's current stream) -- the
16-member H.psi (32 N B algorithmic bytes, HBM peak); the production line
adds the plain H.psi and the rocBLAS preconditioner (32 n^3 nc flops, FP64
MFMA peak).  cpu_baseline: the reference CPU path (its own Fortran H.psi
compiled from source into oracle/_ref, the NumPy Arnoldi restatement) on the
host cores, OpenMP threads fixed before the runtimes load, on one Krylov
attempt, extrapolated by the device run's attempts per step.

Other modes: plain (one wavefunction: CF4 or Arnoldi on --workload), batched
(B members from a common start), adjoint, reference-run,
reference-production, ionic-duals.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def host_cores():
    """Host threads of one GPU's share of the box: the process's CPU affinity,
    at most 16 (the share a one-GPU box gives a job)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


# The CPU baseline's OpenMP runtimes (libgomp for the C restatement, LLVM
# libomp for the reference's Fortran kernel) read OMP_NUM_THREADS when they
# start: set it before anything loads them.
if '--cpu-threads' in sys.argv:
    CPU_THREADS = int(sys.argv[sys.argv.index('--cpu-threads') + 1]) or host_cores()
else:
    CPU_THREADS = host_cores()
os.environ['OMP_NUM_THREADS'] = str(CPU_THREADS)

METRIC = 'prop steps/s (×batch) + adjoint grad time vs plain GPU at ‖Δψ‖≤1e-10'
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6
FP32_MFMA_PEAK_TFS = 157.3
# The reference's own record of its 3D run on its CPU host (Fortran/OpenMP
# H.psi + NumPy Arnoldi, Xeon Silver 4416+): 2952 steps in 307.54 s
# (results/3d_dt_refined/run.json).
REFERENCE_RUN_STEPS, REFERENCE_RUN_SECONDS = 2952, 307.54033255483955


def copy_probe(device):
    """Copy a 256 MiB device buffer three times: the copy kernel of known size
    that calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(1 << 24, dtype=torch.complex128, device=device)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def apply_kernels(shape):
    """The kernels of one H.psi at radial launch shape ``shape``."""
    if shape == 16:
        return 'k_dipole_coefs + k_radial_mfma (FP64 MFMA walks of the mixed field-coupling sources) + k_angular'
    if shape in (1, 2, 4, 8):
        return f'k_radial ({shape} rows per workgroup) + k_dipole_coefs + k_angular'
    return 'k_radial_staged + k_dipole_coefs + k_angular'


def progress(*what):
    """One line per finished phase on stderr (the JSON line stays alone on stdout)."""
    print('bench:', *what, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=1)
    p.add_argument('--workload', default='production_l6_L4_n441')
    p.add_argument('--integrator', choices=['cf4', 'arnoldi'], default=None)
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--cpu-threads', type=int, default=0)
    p.add_argument('--mode', choices=['metric', 'plain', 'batched', 'adjoint', 'reference-run',
                                      'reference-production', 'ionic-duals'],
                   default='metric',
                   help='metric (default line): the B = 16 pump-probe scan batched against the plain engine, '
                        'with the adjoint gradient/forward ratio and the production CF4 step as fields; '
                        'plain: one wavefunction; batched: B-member scan from a common start; '
                        'adjoint: gradient/forward time ratio; reference-run: the reference\'s recorded '
                        '2952-step run, whole, against its published wall time; reference-production: the '
                        'reference\'s production-step benchmark against its published seconds per step')
    p.add_argument('--production-lmax', type=int, choices=[4, 6], default=4)
    p.add_argument('--preconditioner-precision', choices=['complex128', 'complex64'], default='complex64',
                   help='reference-production only: the separable preconditioner\'s precision; complex64 (the '
                        'default) is the reference\'s own setting for that benchmark and runs with flexible GMRES '
                        'restart 16, as there')
    p.add_argument('--traffic-probe', action='store_true',
                   help='before timing, copy a 256 MiB device buffer three times: the copy kernel of known '
                        'size that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE counters')
    p.add_argument('--members', type=int, default=16)
    p.add_argument('--attempts', type=int, default=64)
    return p.parse_args()


def measured_traffic(workload, kernel):
    """HBM bytes per launch of ``kernel`` on ``workload`` as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised under profiles/
    (lines 'traffic_bytes: <workload> <kernel> <bytes>'), or None."""
    for path in sorted((ROOT / 'profiles').glob('*.md')):
        for line in path.read_text().splitlines():
            parts = line.split()
            if len(parts) == 4 and parts[0] == 'traffic_bytes:' and parts[1] == workload and parts[2] == kernel:
                return float(parts[3])
    return None


def measured_mfma(workload, kernel):
    """(FP64 matrix flops, duration in s) per launch of ``kernel`` on
    ``workload`` from the rocprofv3 SQ_INSTS_VALU_MFMA_MOPS_F64 pass summarised
    under profiles/ (lines 'mfma_flops: <workload> <kernel> <flops> <ns>'), or None."""
    for path in sorted((ROOT / 'profiles').glob('*.md')):
        for line in path.read_text().splitlines():
            parts = line.split()
            if len(parts) == 5 and parts[0] == 'mfma_flops:' and parts[1] == workload and parts[2] == kernel:
                return float(parts[3]), float(parts[4]) * 1e-9
    return None


def _cpu_apply_sample(h, field, threads):
    import numpy as np
    import oracle
    use_ref = oracle.have_reference_kernel()
    apply = oracle.ref_tensor_apply if use_ref else oracle.tensor_apply
    rng = np.random.default_rng(1)
    x = rng.normal(size=h.size) + 1j * rng.normal(size=h.size)
    x /= np.linalg.norm(x)
    h.diag, h.rad  # full tables, built outside the timed sample
    apply(h, x, field, True)  # thread start, first touch
    t = time.perf_counter()
    y = apply(h, x, field, True)
    t_apply = time.perf_counter() - t
    kind = 'reference' if use_ref else 'port'
    label = ('H.psi by the reference Fortran kernel built from src/kernels.f90' if use_ref
             else 'H.psi by the C restatement')
    return x, y, t_apply, kind, label


def cpu_baseline_arnoldi(h, field, dimension, threads):
    import numpy as np
    from scipy.linalg.blas import zgemv
    x, y, t_apply, kind, label = _cpu_apply_sample(h, field, threads)
    k = max(1, dimension // 2)
    Q = np.empty((h.size, k), complex, order='F')
    for m in range(k):
        Q[:, m] = x * (m + 1)
    t = time.perf_counter()
    v = y.copy()
    for _ in range(2):
        z = zgemv(1., Q, v, trans=2)
        v = zgemv(-1., Q, z, beta=1., y=v, overwrite_y=1)
    np.linalg.norm(v)
    t_cgs = time.perf_counter() - t
    step = dimension * (t_apply + t_cgs)
    return {'value': 1.0 / step, 'unit': 'steps/s', 'cores': threads, 'kind': kind,
            'sample': f'one full-size {label} ({t_apply:.2f} s, OpenMP {threads} threads) + one CGS2 column '
                      f'over {k} vectors (NumPy/BLAS zgemv, {t_cgs:.2f} s), extrapolated to a '
                      f'{dimension}-dimensional Arnoldi step'}


def cpu_baseline_cf4(h, field, linear_iterations, threads):
    import numpy as np
    from oracle import implicit as oi
    x, y, t_apply, kind, label = _cpu_apply_sample(h, field, threads)
    pre = oi.SeparablePreconditioner(oi.OracleEngine(h))
    alpha = .01 / (3 + 1j * np.sqrt(3))
    t = time.perf_counter()
    pre.apply(x, alpha, 2.9)
    t_pre = time.perf_counter() - t
    # per solve: one residual apply, `iterations` (preconditioner + apply), one
    # preconditioner + apply for the true residual; four solves per CF4 step
    per_iteration = t_apply + t_pre
    step = linear_iterations * per_iteration + 4 * (2 * t_apply + t_pre)
    return {'value': 1.0 / step, 'unit': 'steps/s', 'cores': threads, 'kind': kind,
            'sample': f'one full-size {label} ({t_apply:.2f} s, OpenMP {threads} threads) + one separable '
                      f'preconditioner application (NumPy/BLAS per-channel complex GEMMs, {t_pre:.2f} s), '
                      f'extrapolated to a CF4 step of {linear_iterations} GMRES iterations (as measured on '
                      'the GPU) plus 4 residual applies and 4 candidate preconditioner+apply pairs; CGS2 '
                      'omitted (lower bound on the CPU step time)'}


def scan_members(B, seed=2024):
    """B pump-probe members: delays U[0, 400] a.u. after the pump ends and probe
    CEPs U[0, 2 pi) from one seed; step counts 2500 + 2500 * delay / 400."""
    import numpy as np
    from laqphysics_amd.pulses import Pulse
    rng = np.random.default_rng(seed)
    delays = rng.uniform(0., 400., B)
    ceps = rng.uniform(0., 2 * np.pi, B)
    pump = Pulse(omega=1.5, cycles=8, field=.03)
    members = [[(pump, 'z'), (Pulse(omega=.114, cycles=4, field=.01, start=pump.duration + d, cep=c), 'z')]
               for d, c in zip(delays, ceps)]
    counts = [int(2500 + round(2500 * d / 400.)) for d in delays]
    return members, counts, delays, ceps


def bench_batched(a):
    """B members, each running the same number of whole steps from a common
    start time inside the pump (identical work per member), timed to
    completion; the plain engine runs the same steps for one member."""
    import numpy as np
    import torch
    from laqphysics_amd import hip_backend, workloads
    from laqphysics_amd.batched import BatchedPropagator
    from laqphysics_amd.hip_backend import HipHamiltonian
    from laqphysics_amd.pulses import vector_function
    name = a.workload if a.workload != 'production_l6_L4_n441' else 'batched_l4_L3_n441'
    w = workloads.ALL[name]()
    h = w.h
    op = HipHamiltonian(h, 'cuda:0')
    if a.traffic_probe:
        copy_probe('cuda:0')
    B = a.members
    members, counts, delays, ceps = scan_members(B)
    if name != 'batched_l4_L3_n441':
        members = [w.pulses] * B
    t0 = 10.0   # inside the pump for every member
    k = max(1, a.steps)
    x0 = workloads.initial_state(h)
    # warm-up batch (graph capture, autotune) on one step, then the timed one
    BatchedPropagator(op, [x0] * B, members, [t0] * B, [1] * B, w.dt, tol=w.tolerance,
                      maxdim=w.krylov_dimension, tune=True).run(chunk=4)
    batch = BatchedPropagator(op, [x0] * B, members, [t0] * B, [k] * B, w.dt, tol=w.tolerance,
                              maxdim=w.krylov_dimension, tune=True)
    torch.cuda.synchronize()
    start = time.perf_counter()
    attempts = batch.run(chunk=4)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    s1 = batch.status()
    if np.any(s1['failures']) or list(s1['steps']) != [k] * B:
        raise RuntimeError('batched run: a member failed to converge')
    steps = B * k
    rate = steps / elapsed
    # plain engine, one member, the same steps
    avec = vector_function(members[0])
    state = op.state(x0)
    state, _ = hip_backend.step(lambda tt, v: op.apply(v, avec(tt), True), state, t0, w.dt, tol=w.tolerance,
                                maxdim=w.krylov_dimension)
    state = op.state(x0)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for i in range(k):
        state, _ = hip_backend.step(lambda tt, v: op.apply(v, avec(tt), True), state, t0 + i * w.dt, w.dt,
                                    tol=w.tolerance, maxdim=w.krylov_dimension)
    torch.cuda.synchronize()
    plain = k / (time.perf_counter() - start)
    diff = float(torch.linalg.norm(batch.psi[0] - state))
    # roofline of the batched H.psi (the B-member apply the graph launches
    # every column), timed with HIP events on torch's current stream
    xs = batch.psi.clone()
    fields = [avec(t0)] * B
    outb = op.apply_batched(xs, fields, True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(5):
        op.apply_batched(xs, fields, True, out=outb)
    ev1.record()
    torch.cuda.synchronize()
    t_apply = ev0.elapsed_time(ev1) / 1e3 / 5
    bytes_apply = 32.0 * h.size * B
    roofline = {'bound': 'hbm', 'kernel': f'{B}-member H.psi ({apply_kernels(batch.launch)})',
                'launch_shape': batch.launch,
                'achieved': bytes_apply / t_apply / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': bytes_apply / t_apply / 1e9 / HBM_PEAK_GBS,
                # the traffic profile was measured on the autotuned 16-member shape of this workload
                'traffic': measured_traffic(name, f'hpsi_B{B}') if batch.launch == 16 else None,
                'launch_ms': 1e3 * t_apply, 'algorithmic_bytes': bytes_apply}
    line = {'metric': METRIC, 'value': rate, 'unit': 'wavefunction steps/s', 'n_gpus': 1, 'steps': steps,
            'warmup': 1, 'ms_per_step': 1e3 * elapsed / steps, 'higher_is_better': True,
            'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex128', 'data': 'synthetic',
            'config': {'workload': f'{w.name} batched propagation, B={B} members x {k} steps from t = {t0} '
                                   f'(pump-probe scan members), midpoint Arnoldi order {w.krylov_dimension}, dt {w.dt}, '
                                   'recursive subdivision on device', 'n': h.n, 'channels': h.nc, 'N': int(h.size),
                       'batch': B, 'attempts': int(attempts), 'substeps': int(s1['substeps'].sum()),
                       'parallelism': 'single GPU'},
            'roofline': roofline,
            'plain_engine_steps_per_s': plain, 'batched_over_plain_per_wavefunction': rate / plain,
            'member0_vs_plain_norm_diff': diff}
    print(json.dumps(line), flush=True)


def omp_threads_in_use():
    """Thread counts the two OpenMP runtimes of the CPU baseline will run with
    (omp_get_max_threads of the loaded libgomp and LLVM libomp)."""
    import ctypes
    out = {}
    for name, path in (('libgomp', 'libgomp.so.1'), ('libomp', '/opt/rocm/lib/llvm/lib/libomp.so')):
        try:
            out[name] = int(ctypes.CDLL(path).omp_get_max_threads())
        except OSError:
            out[name] = None
    return out


def cpu_baseline_attempts(h, avec, t, dt, maxdim, attempts_per_step, threads):
    """The reference CPU path on one Krylov attempt of the scan: the reference's
    Fortran H.psi (oracle/_ref; the C restatement when absent) with OpenMP on
    ``threads`` cores and the NumPy/BLAS CGS2 Arnoldi of python/propagate.py
    at fixed dimension ``maxdim`` (12 applies, two Gram-Schmidt passes per
    column, the small exponential), timed with single-threaded BLAS (the
    reference's convention) and with BLAS on all ``threads`` cores; one step
    costs ``attempts_per_step`` such attempts, as the parity-identical device
    run of the same members measured."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    import oracle
    from oracle import arnoldi as oa
    use_ref = oracle.have_reference_kernel()
    apply = oracle.ref_tensor_apply if use_ref else oracle.tensor_apply
    rng = np.random.default_rng(1)
    x = rng.normal(size=h.size) + 1j * rng.normal(size=h.size)
    x /= np.linalg.norm(x)
    A = avec(t + dt / 2)
    h.diag, h.rad   # full tables, built outside the timed sample
    apply(h, x, A, True)   # thread start, first touch
    seconds = {}
    for blas in (1, threads):
        with threadpool_limits(blas, user_api='blas'):
            start = time.perf_counter()
            oa.exponential_action_fixed(lambda y: apply(h, y, A, True), x, dt, maxdim)
            seconds[blas] = time.perf_counter() - start
    t_attempt = seconds[1]
    return {'value': 1.0 / (attempts_per_step * t_attempt), 'unit': 'wavefunction steps/s', 'cores': threads,
            'kind': 'reference' if use_ref else 'port', 'openmp_threads': omp_threads_in_use(),
            'sample': f'one Krylov attempt of dimension {maxdim} of a scan member '
                      f"({'the reference Fortran H.psi built from src/kernels.f90' if use_ref else 'the C restatement'}"
                      f' on {threads} OpenMP threads + NumPy/BLAS CGS2 and expm): {seconds[1]:.2f} s with 1 BLAS '
                      f'thread (the reference convention, used for value), {seconds[threads]:.2f} s with '
                      f'{threads}; extrapolated to {attempts_per_step:.1f} attempts per wavefunction step (the '
                      'count the device run of the same members took)',
            'attempt_s_blas1': seconds[1], f'attempt_s_blas{threads}': seconds[threads],
            'value_all_core_blas': 1.0 / (attempts_per_step * seconds[threads])}


def _scan_counts(full, steps):
    """Per-member step counts of the timed window: the scan's counts
    (2500-5000) scaled by steps/20000, at least one -- as uneven as the scan."""
    return [max(1, int(round(c * steps / 20000.))) for c in full]


def metric_production_cf4(steps, warmup):
    """production_l6_L4_n441: CF4-Pade steps from the ground state at mid-pump,
    each followed by its tSURFF frame at r = 80 (flux, projection on the He+
    n <= 2 duals, Volkov accumulation); the duals over the window's frames are
    prepared before timing, as a run prepares them before propagating."""
    import numpy as np
    import torch
    from laqphysics_amd import implicit, workloads
    from laqphysics_amd.hip_backend import HipHamiltonian
    from laqphysics_amd.run import OnlineSurface
    w = workloads.production_l6_L4_n441()
    h = w.h
    E, x0 = workloads.ground_state(h)
    op = HipHamiltonian(h, 'cuda:0')
    op.separable = implicit.SeparablePreconditioner(op)
    restart = w.extra['gmres_restart']
    t0 = 8 * np.pi / 1.5 / 2
    times = t0 + w.dt * np.arange(warmup + steps + 1)
    surface = OnlineSurface(op, w.surface, [(1, 0, 0), (2, 0, 0), (2, 1, 0)], times, w.avec,
                            np.linspace(.05, 2., 40))
    state = op.state(x0)
    surface.record(0, state)
    iters = []
    torch.cuda.synchronize()
    for k in range(warmup + steps):
        if k == warmup:
            torch.cuda.synchronize()
            start = time.perf_counter()
        state, info = implicit.cf4_step(op, state, w.avec, times[k], w.dt, E, tol=w.tolerance, restart=restart)
        surface.record(k + 1, state)
        if k >= warmup:
            iters.append(info['linear_iterations'])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    # roofline of the dominant kernels of this step: the rocBLAS preconditioner
    # GEMMs and the hand-written H.psi, each timed with HIP events
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = torch.empty_like(state)
    alpha = w.dt / 2 / (3 + 1j * np.sqrt(3))
    field = w.avec(t0)
    for _ in range(2):
        op.separable.apply(state, alpha, -E, out=out)
        op.apply(state, field, True, out=out)
    ev0.record()
    for _ in range(10):
        op.separable.apply(state, alpha, -E, out=out)
    ev1.record()
    torch.cuda.synchronize()
    t_pre = ev0.elapsed_time(ev1) / 1e3 / 10
    ev0.record()
    for _ in range(10):
        op.apply(state, field, True, out=out)
    ev1.record()
    torch.cuda.synchronize()
    t_apply = ev0.elapsed_time(ev1) / 1e3 / 10
    flops = 32.0 * h.n ** 3 * h.nc
    return {'workload': f'{w.name} CF4-Pade step (GMRES restart {restart}, tol {w.tolerance}, separable '
                        f'preconditioner) + tSURFF frame at r = {w.surface}, dt {w.dt}, from mid-pump',
            'steps_per_s': steps / elapsed, 'ms_per_step': 1e3 * elapsed / steps, 'steps': steps,
            'linear_iterations_per_step': iters, 'ground_energy': E,
            'roofline_hpsi': {'bound': 'hbm', 'kernel': 'H.psi (k_dipole_coefs + k_radial_staged + k_angular)',
                              'achieved': 32.0 * h.size / t_apply / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                              'frac': 32.0 * h.size / t_apply / 1e9 / HBM_PEAK_GBS, 'launch_ms': 1e3 * t_apply,
                              'algorithmic_bytes': 32.0 * h.size,
                              'traffic': measured_traffic('production_l6_L4_n441', 'hpsi')},
            'roofline_preconditioner': {'bound': 'mfma', 'kernel': 'separable preconditioner (rocBLAS complex128 '
                                        'GEMMs + resolvent)', 'achieved': flops / t_pre / 1e12,
                                        'peak': FP64_MFMA_PEAK_TFS, 'unit': 'TFLOP/s',
                                        'frac': flops / t_pre / 1e12 / FP64_MFMA_PEAK_TFS,
                                        'launch_ms': 1e3 * t_pre, 'algorithmic_flops': flops,
                                        'traffic': measured_traffic('production_l6_L4_n441',
                                                                    'separable_preconditioner')}}


def metric_adjoint(steps):
    """The adjoint benchmark's setup (512 samples over 1000 a.u., J = He+(n=2)
    yield + ionisation, checkpoints every 64 steps) over a window of 64 * steps
    steps from t = 468: gradient time / forward time (the forward includes
    evaluating J)."""
    import numpy as np
    import torch
    from laqphysics_amd import workloads
    from laqphysics_amd.adjoint import PiecewiseLinearControl, Problem
    from laqphysics_amd.hip_backend import HipHamiltonian
    w = workloads.plain_l2_n85()
    h = w.h
    op = HipHamiltonian(h, 'cuda:0')
    E, psi0 = workloads.ground_state(h)
    control = PiecewiseLinearControl(np.random.default_rng(7).normal(0., .01, 512), 1000.)
    nsteps = 64 * max(2, steps)
    prob = Problem(op, psi0, control, w.dt, nsteps, t0=468., tol=w.tolerance, maxdim=w.krylov_dimension,
                   checkpoint_every=64, radius=w.surface)
    prob.add_yield(w.surface, labels=[(2, 0, 0), (2, 1, 0)])
    prob.objective(prob.forward(keep=False)[0])   # warm-up
    torch.cuda.synchronize()
    start = time.perf_counter()
    psiN, checkpoints = prob.forward()
    J = prob.objective(psiN)
    torch.cuda.synchronize()
    t_forward = time.perf_counter() - start
    start = time.perf_counter()
    g, _ = prob.gradient(psiN, checkpoints)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - start
    return {'workload': f'{w.name} adjoint gradient over 512 piecewise-linear field samples (1000 a.u., sin^2 '
                        f'window, N(0, 0.01^2) from a seed), window of {nsteps} steps from t = 468, checkpoint '
                        f'every 64 steps, J = He+(n=2) tSURFF yield at r = {w.surface} + ionisation probability',
            'gradient_over_forward': t_grad / t_forward, 'forward_s': t_forward, 'gradient_s': t_grad,
            'steps': nsteps, 'objective': J, 'gradient_norm': float(np.linalg.norm(g))}


def bench_metric(a):
    """Default line.  value: wavefunction steps/s of the B = 16 pump-probe scan
    (batched_l4_L3_n441: coupled l <= 4, L <= 3, n = 441, midpoint Arnoldi
    order 12, dt 0.02) propagated by the batched engine over a window of the
    scan's uneven step counts, every member from the ground state at t = 0;
    every member starts 10 a.u. into its own probe pulse; the plain engine
    runs the same member-steps one member after another.
    Fields: the adjoint benchmark's gradient/forward ratio and the production
    CF4 step with its tSURFF frame."""
    import numpy as np
    import torch
    from laqphysics_amd import hip_backend, workloads
    from laqphysics_amd.batched import BatchedPropagator
    from laqphysics_amd.hip_backend import HipHamiltonian
    from laqphysics_amd.pulses import vector_function
    w = workloads.batched_l4_L3_n441()
    h = w.h
    op = HipHamiltonian(h, 'cuda:0')
    B = 16
    members, full, delays, ceps = scan_members(B)
    counts = _scan_counts(full, a.steps)
    E, x0 = workloads.ground_state(h)
    # every member's window starts 10 a.u. into its own probe pulse (its own
    # field and phase), where the steps subdivide as they do through the scan
    pump = members[0][0][0]
    t0 = [pump.duration + d + 10. for d in delays]
    # warm-up batch: graph capture and the autotuned launch shape
    BatchedPropagator(op, [x0] * B, members, t0, [1] * B, w.dt, tol=w.tolerance, maxdim=w.krylov_dimension,
                      tune=True).run(chunk=4)
    batch = BatchedPropagator(op, [x0] * B, members, t0, counts, w.dt, tol=w.tolerance, maxdim=w.krylov_dimension,
                              tune=True)
    torch.cuda.synchronize()
    start = time.perf_counter()
    batch.run(chunk=4)
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - start
    progress(f'batched scan {t_batched:.2f} s')
    status = batch.status()
    if np.any(status['failures']) or list(status['steps']) != counts:
        raise RuntimeError('batched scan: a member failed to converge')
    wf_steps = int(sum(counts))
    # the plain engine on the same member-steps
    avecs = [vector_function(m) for m in members]
    state = op.state(x0)
    hip_backend.step(lambda tt, v: op.apply(v, avecs[0](tt), True), state, t0[0], w.dt, tol=w.tolerance,
                     maxdim=w.krylov_dimension)
    finals = []
    torch.cuda.synchronize()
    start = time.perf_counter()
    for b in range(B):
        state = op.state(x0)
        for k in range(counts[b]):
            state, _ = hip_backend.step(lambda tt, v: op.apply(v, avecs[b](tt), True), state, t0[b] + k * w.dt,
                                        w.dt, tol=w.tolerance, maxdim=w.krylov_dimension)
        finals.append(state)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - start
    progress(f'plain engine on the same member-steps {t_plain:.2f} s')
    worst = max(float(torch.linalg.norm(batch.psi[b] - finals[b])) for b in range(B))
    del finals
    # roofline: the 16-member H.psi the attempt graph launches every Krylov
    # column, timed with HIP events on the stream it launches on (torch's current one)
    xs = batch.psi.clone()
    fields = [avecs[b](t0[b]) for b in range(B)]
    outb = op.apply_batched(xs, fields, True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(5):
        op.apply_batched(xs, fields, True, out=outb)
    ev1.record()
    torch.cuda.synchronize()
    t_apply = ev0.elapsed_time(ev1) / 1e3 / 5
    bytes_apply = 32.0 * h.size * B
    attempts = float(status['attempts'].sum())
    line = {'metric': METRIC, 'value': wf_steps / t_batched, 'unit': 'wavefunction steps/s', 'n_gpus': 1,
            'steps': wf_steps, 'warmup': 1, 'ms_per_step': 1e3 * t_batched / wf_steps, 'higher_is_better': True,
            'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex128',
            'data': 'synthetic: random-free problem tables; every member from the device ground state',
            'config': {'workload': f'{w.name} pump-probe scan, B={B} batched members (probe delays U[0,400] '
                                   'a.u. after the pump, CEPs U[0,2pi) from seed 2024; full counts 2500-5000 '
                                   f'steps scaled by {a.steps}/20000), midpoint Arnoldi order '
                                   f'{w.krylov_dimension}, tol {w.tolerance}, dt {w.dt}; each member from 10 a.u. into its probe',
                       'n': h.n, 'channels': h.nc, 'N': int(h.size), 'batch': B, 'member_steps': counts,
                       'attempts': int(attempts), 'substeps': int(status['substeps'].sum()),
                       'parallelism': 'single GPU'},
            'batched_over_plain_per_wavefunction': t_plain / t_batched,
            'plain_engine_wavefunction_steps_per_s': wf_steps / t_plain,
            'max_member_vs_plain_norm_diff': worst,
            'roofline': {'bound': 'hbm', 'kernel': f'{B}-member H.psi ({apply_kernels(batch.launch)})',
                         'launch_shape': batch.launch,
                         'achieved': bytes_apply / t_apply / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                         'frac': bytes_apply / t_apply / 1e9 / HBM_PEAK_GBS,
                         'traffic': measured_traffic(w.name, f'hpsi_B{B}'), 'launch_ms': 1e3 * t_apply,
                         'algorithmic_bytes': bytes_apply}}
    # the Krylov side: one CGS2 column at the attempt's last column (12 basis
    # vectors) on one member's state, hex_cgs2_commit on torch's current stream
    # (its four kernels read the basis three times and v four times, and
    # write v twice and the next basis vector once)
    line['roofline_krylov'] = krylov_column_roofline(op, h.size, w.krylov_dimension)
    mf = measured_mfma(w.name, f'radial_B{B}') if batch.launch == 16 else None
    if mf:
        line['mfma_batched_radial'] = {'kernel': 'k_radial_mfma (16-member launch)', 'source': 'rocprofv3 counters '
                                       'SQ_INSTS_VALU_MFMA_MOPS_F64 x 512 over the kernel duration, '
                                       'profiles/mfma_batched_l4_L3_n441_B16.md',
                                       'achieved': mf[0] / mf[1] / 1e12, 'peak': FP64_MFMA_PEAK_TFS,
                                       'unit': 'TFLOP/s', 'frac': mf[0] / mf[1] / 1e12 / FP64_MFMA_PEAK_TFS}
    del batch, xs, outb
    torch.cuda.empty_cache()
    line['adjoint'] = metric_adjoint(a.steps)
    progress('adjoint gradient/forward done')
    torch.cuda.empty_cache()
    line['production_cf4'] = metric_production_cf4(max(2, a.steps // 4), 1)
    progress('production CF4 step done')
    if not a.no_cpu:
        line['cpu_baseline'] = cpu_baseline_attempts(h, avecs[0], t0[0], w.dt, w.krylov_dimension,
                                                     attempts / wf_steps, CPU_THREADS)
    print(json.dumps(line), flush=True)


def krylov_column_roofline(op, N, k):
    """HBM roofline of one CGS2 column with k basis vectors (hex_cgs2_commit:
    dots, update + dots, update + norm, scale into the next basis vector),
    timed with HIP events over 10 columns on torch's current stream."""
    import torch
    from laqphysics_amd import native
    from laqphysics_amd.hip_backend import _stream, _vp
    dev = op.device
    g = torch.Generator(device=dev).manual_seed(5)
    Q = torch.randn((k + 1, N), dtype=torch.complex128, device=dev, generator=g) / N ** .5
    v = torch.randn(N, dtype=torch.complex128, device=dev, generator=g)
    hdev = torch.empty(64, dtype=torch.complex128, device=dev)
    lib = native.lib()

    def column():
        native.check(lib.hex_cgs2_commit(op._engine, _vp(Q), N, k, _vp(v), N, _vp(hdev), _vp(Q[k]), _stream(dev)),
                     'cgs2')
    column()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        column()
    ev1.record()
    torch.cuda.synchronize()
    t = ev0.elapsed_time(ev1) / 1e3 / 10
    nbytes = 16.0 * N * (3 * k + 7)
    del Q, v
    torch.cuda.empty_cache()
    return {'bound': 'hbm', 'kernel': f'one CGS2 column, {k} basis vectors (k_dots + k_update_dots + k_update + '
                                      'k_scale_by)',
            'achieved': nbytes / t / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': nbytes / t / 1e9 / HBM_PEAK_GBS,
            'launch_ms': 1e3 * t, 'algorithmic_bytes': nbytes,
            'bytes_note': 'basis read 3 times, v read 4 times and written 3 times (the scale writes the next basis '
                          'vector): (3k + 7) 16 N'}


def bench_adjoint(a):
    import numpy as np
    import torch
    from laqphysics_amd import workloads
    from laqphysics_amd.adjoint import PiecewiseLinearControl, Problem
    from laqphysics_amd.hip_backend import HipHamiltonian
    w = workloads.plain_l2_n85()
    h = w.h
    op = HipHamiltonian(h, 'cuda:0')
    rng = np.random.default_rng(7)
    control = PiecewiseLinearControl(rng.normal(0., .01, 512), 1000.)
    nsteps = 64 * max(2, a.steps)
    prob = Problem(op, workloads.initial_state(h), control, w.dt, nsteps, t0=500., tol=w.tolerance,
                   maxdim=w.krylov_dimension, checkpoint_every=64, radius=w.surface)
    # J = He+(n=2) tSURFF channel yield + ionisation probability
    prob.add_yield(w.surface, labels=[(2, 0, 0), (2, 1, 0)])
    prob.objective(prob.forward(keep=False)[0])
    torch.cuda.synchronize()
    start = time.perf_counter()
    psiN, checkpoints = prob.forward()
    J0 = prob.objective(psiN)
    torch.cuda.synchronize()
    t_forward = time.perf_counter() - start
    start = time.perf_counter()
    g, J = prob.gradient(psiN, checkpoints)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - start
    line = {'metric': METRIC, 'value': t_grad / t_forward, 'unit': 'gradient time / forward time',
            'n_gpus': 1, 'steps': nsteps, 'warmup': 1, 'ms_per_step': 1e3 * t_forward / nsteps,
            'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex128',
            'data': 'synthetic',
            'config': {'workload': f'{w.name} adjoint gradient over 512 piecewise-linear field samples '
                                   f'(T = 1000 a.u., sin^2 window, N(0, 0.01^2) from a seed), window of {nsteps} '
                                   f'steps from t = 500, checkpoint every 64 steps; J = He+(n=2) tSURFF yield at '
                                   f'r = {w.surface} + ionisation probability; forward time includes evaluating J',
                       'n': h.n, 'channels': h.nc,
                       'N': int(h.size), 'parallelism': 'single GPU'},
            'forward_s': t_forward, 'gradient_s': t_grad, 'objective': J,
            'gradient_norm': float(np.linalg.norm(g))}
    print(json.dumps(line), flush=True)


def bench_reference_run(a):
    """The reference's recorded run (results/3d_dt_refined), every step: device
    ground state, transfer, 2952 midpoint Arnoldi steps (dimension 48, tol
    1e-10), observables every 100 steps.  vs_baseline = this rate over the
    reference's published rate for the same run on its CPU host."""
    import numpy as np
    import torch
    from laqphysics_amd import workloads
    from laqphysics_amd.run import run_config
    config = workloads.reference_dt_refined_l1_n88()
    run_config(config, max_steps=8)               # warm-up: library, graphs of the solver, ground
    torch.cuda.synchronize()
    start = time.perf_counter()
    state, history, meta = run_config(config)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    steps = meta['nsteps']
    rate = steps / elapsed
    published = REFERENCE_RUN_STEPS / REFERENCE_RUN_SECONDS
    engine = meta['engine']
    h = engine.h
    x = state
    out = torch.empty_like(x)
    field = np.array([0., 0., .005])
    for _ in range(3):
        engine.apply(x, field, True, out=out)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(20):
        engine.apply(x, field, True, out=out)
    ev1.record()
    torch.cuda.synchronize()
    t_apply = ev0.elapsed_time(ev1) / 1e3 / 20
    bytes_apply = 32.0 * h.size
    line = {'metric': METRIC, 'value': rate, 'unit': 'steps/s', 'n_gpus': 1, 'steps': steps, 'warmup': 8,
            'ms_per_step': 1e3 * elapsed / steps, 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': rate / published, 'dtype': 'complex128',
            'data': 'the reference\'s recorded run configuration (results/3d_dt_refined/config.json), '
                    'ground state solved on the device; wall time includes the ground solve',
            'config': {'workload': 'reference_dt_refined_l1_n88: product lmax 1 (6 channels), n 88, z pulse '
                                   '(1.5, 8 cycles, 0.005), midpoint Arnoldi dimension 48 tol 1e-10, 2952 steps',
                       'n': h.n, 'channels': h.nc, 'N': int(h.size), 'parallelism': 'single GPU',
                       'ground_energy': meta['ground_energy'],
                       'krylov_dimensions': sorted({r['dimension'] for r in history}),
                       'final_real_region_norm': history[-1]['real_region_norm'],
                       'final_ground_population': history[-1]['ground_population']},
            'baseline': {'value': published, 'unit': 'steps/s', 'source': 'results/3d_dt_refined/run.json '
                         '(2952 steps in 307.5 s, Fortran/OpenMP + NumPy on a Xeon Silver 4416+)'},
            'roofline': {'bound': 'hbm', 'kernel': 'H.psi (k_radial_staged + k_angular), cache-resident '
                         '(1.5 MB state)', 'achieved': bytes_apply / t_apply / 1e9, 'peak': HBM_PEAK_GBS,
                         'unit': 'GB/s', 'frac': bytes_apply / t_apply / 1e9 / HBM_PEAK_GBS, 'traffic': None,
                         'launch_ms': 1e3 * t_apply, 'algorithmic_bytes': bytes_apply}}
    if not a.no_cpu:
        line['cpu_baseline'] = cpu_baseline_reference_run(config, CPU_THREADS)
    print(json.dumps(line), flush=True)


def cpu_baseline_reference_run(config, threads, steps=100):
    """The same run on the host cores for its first ``steps`` steps: the
    reference's Fortran H.psi (oracle/_ref; the C restatement when absent)
    and the NumPy Arnoldi restatement, extrapolated to the whole run."""
    import oracle
    from oracle import arnoldi as oracle_arnoldi
    from oracle import ground as oracle_ground
    from laqphysics_amd.ground import GroundS
    from laqphysics_amd.pulses import vector_function
    from laqphysics_amd.run import config_pulses, schedule, setup, transfer
    from threadpoolctl import threadpool_limits
    oracle.set_threads(threads)
    # single-threaded BLAS beside the OpenMP kernel, the reference's convention
    # (OPENBLAS_NUM_THREADS=1); a BLAS pool spinning on the same cores slows both
    blas = threadpool_limits(1, user_api='blas')
    use_ref = oracle.have_reference_kernel()
    apply = oracle.ref_tensor_apply if use_ref else oracle.tensor_apply
    hreal, h = setup(config)
    pulses = config_pulses(config)
    T, nsteps, dt = schedule(pulses, config['dt'], config['surface_stride'], None, config['post_time'])
    start = time.perf_counter()
    g = GroundS(hreal)
    E, small, res, its = oracle_ground.solve(g, tol=config['ground_tolerance'])
    x = transfer(hreal, h, g.expand(small))
    avec = vector_function(pulses)
    for i in range(steps):
        x, info = oracle_arnoldi.step(lambda s, y: apply(h, y, avec(s), True), x, i * dt, dt,
                                      tol=config['krylov_tolerance'], maxdim=config['krylov_dimension'])
    elapsed = time.perf_counter() - start
    blas.restore_original_limits()
    return {'value': steps / elapsed, 'unit': 'steps/s', 'cores': threads,
            'kind': 'reference' if use_ref else 'port',
            'sample': f'ground solve + the first {steps} of {nsteps} steps ({elapsed:.1f} s) with the '
                      + ('reference Fortran H.psi built from src/kernels.f90' if use_ref else 'C restatement of H.psi')
                      + f' (OpenMP {threads} threads) and the NumPy Arnoldi restatement (single-threaded BLAS)'}


def bench_reference_production(a):
    """The reference's production-step benchmark (scripts/benchmark_step.py):
    the field-free ground state, one untimed CF4 step at the middle of the
    last pulse, then ``steps`` timed CF4 steps, each followed by the surface
    flux brought to the host, as that script times them.  vs_baseline = this
    rate over the reference's published rate for the same configuration."""
    import numpy as np
    import torch
    from laqphysics_amd import implicit, workloads
    from laqphysics_amd.hip_backend import HipHamiltonian
    from laqphysics_amd.pulses import vector_function
    from laqphysics_amd.run import config_pulses, schedule, setup, transfer
    lmax = a.production_lmax
    config = workloads.reference_production(lmax)
    pulses = config_pulses(config)
    T, nsteps, dt = schedule(pulses, config['dt'], 1, config['end_time'])
    avec = vector_function(pulses)
    last = max(pulses, key=lambda x: x[0].start)[0]
    t0 = last.start + last.duration / 2
    started = time.perf_counter()
    hreal, h = setup(config)
    h.prepare_surface(config['surface'])
    E, gs, gres = hreal.ground(tol=config['ground_tolerance'])
    engine = HipHamiltonian(h, 'cuda:0')
    engine.separable = implicit.SeparablePreconditioner(engine, precision=a.preconditioner_precision)
    single = engine.separable.flexible
    state = engine.state(transfer(hreal, h, gs))
    setup_s = time.perf_counter() - started
    tol = config['linear_tolerance']
    state, info = implicit.cf4_step(engine, state, avec, t0, dt, E, tol=tol)
    torch.cuda.synchronize()
    steps = max(1, a.steps)
    rows = []
    start = time.perf_counter()
    for i in range(steps):
        state, info = implicit.cf4_step(engine, state, avec, t0 + (i + 1) * dt, dt, E, tol=tol)
        engine.flux(state, avec(t0 + (i + 2) * dt))
        torch.cuda.synchronize()
        rows.append({'linear_iterations': info['linear_iterations'], 'linear_residual': info['linear_residual']})
    per_step = (time.perf_counter() - start) / steps
    published = workloads.REFERENCE_PRODUCTION_SECONDS[lmax]
    x = state
    out = torch.empty_like(x)
    alpha = dt / 2 / (3 + 1j * np.sqrt(3))
    for _ in range(2):
        engine.separable.apply(x, alpha, -E, out=out)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        engine.separable.apply(x, alpha, -E, out=out)
    ev1.record()
    torch.cuda.synchronize()
    t_kernel = ev0.elapsed_time(ev1) / 1e3 / 10
    flops = 32.0 * h.n ** 3 * h.nc
    peak = FP32_MFMA_PEAK_TFS if single else FP64_MFMA_PEAK_TFS
    line = {'metric': METRIC, 'value': 1.0 / per_step, 'unit': 'steps/s', 'n_gpus': 1, 'steps': steps, 'warmup': 1,
            'ms_per_step': 1e3 * per_step, 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': published / per_step, 'dtype': 'complex128',
            'data': 'the reference\'s production configuration, ground state solved on the device, steps from the '
                    'middle of the probe pulse',
            'config': {'workload': f'reference production step benchmark, bipolar lmax {lmax}, total L <= 4, all M '
                                   f'(n {h.n}, {h.nc} channels), CF4-Pade dt {dt:.6g}, linear tol {tol}, '
                                   f'{engine.separable.precision} separable preconditioner, '
                                   + ('flexible GMRES restart 16' if single else 'GMRES restart 32')
                                   + '; each step + host surface flux',
                       'n': h.n, 'channels': h.nc, 'N': int(h.size), 'parallelism': 'single GPU',
                       'timed_at': t0, 'ground_energy': E, 'setup_seconds_including_ground_state': setup_s,
                       'linear_steps': rows},
            'baseline': {'value': 1.0 / published, 'unit': 'steps/s',
                         'source': f'results/probe_20260924/validation/production_step_benchmark_l{lmax}.json: '
                                   f'{published:.3f} s per step on a shared RTX 6000 Ada (complex64 preconditioner, '
                                   'FGMRES restart 16, 20-22 linear iterations per step)'},
            'roofline': {'bound': 'mfma', 'kernel': f'separable preconditioner (rocBLAS {engine.separable.precision} '
                                                    'GEMMs + resolvent)',
                         'achieved': flops / t_kernel / 1e12, 'peak': peak, 'unit': 'TFLOP/s',
                         'frac': flops / t_kernel / 1e12 / peak, 'traffic': None,
                         'launch_ms': 1e3 * t_kernel, 'algorithmic_flops': flops}}
    if not a.no_cpu:
        its = int(round(np.mean([r['linear_iterations'] for r in rows])))
        line['cpu_baseline'] = cpu_baseline_cf4(h, avec(t0), its, CPU_THREADS)
    print(json.dumps(line), flush=True)


def cpu_baseline_ionic(ion, unique, avec, t0, dt, steps=4):
    """The oracle's sparse-LU backward CF4 step (the reference's SparseCF4,
    python/implicit.py:111-123) on the same system, single-threaded SuperLU."""
    from oracle import ionic as oracle_ionic
    chi = ion.dual_final_states(unique)
    start = time.perf_counter()
    for i in range(steps):
        t = t0 - i * dt
        chi = oracle_ionic.cf4_step(ion.matrix(avec(t - oracle_ionic.C1 * dt)).conj().T,
                                    ion.matrix(avec(t - oracle_ionic.C2 * dt)).conj().T, chi, -dt)
    seconds = (time.perf_counter() - start) / steps
    return {'value': 1. / seconds, 'unit': 'steps/s', 'cores': 1, 'kind': 'port',
            'sample': f'{steps} backward CF4 steps of the oracle (SuperLU, MMD ordering) from t = {t0:.6g}'}


def bench_ionic_duals(a):
    """Backward ionic CF4 steps of the reference production configuration's
    tSURFF duals (scripts/benchmark_step.py: AdjointStepper, 20 channels):
    He+ on n = 264 with all 25 (l, m) of l <= 4, 20 dual columns, one untimed
    step at the middle of the last pulse, then ``steps`` timed steps at
    linear tol 1e-12.  vs_baseline = this rate over the reference's published
    CPU rate for the same system."""
    import numpy as np
    import torch
    from laqphysics_amd import workloads
    from laqphysics_amd.ionic_adjoint import IonicAdjoint
    from laqphysics_amd.pulses import vector_function
    from laqphysics_amd.run import config_pulses, schedule
    config = workloads.reference_production(4)
    pulses = config_pulses(config)
    T, nsteps, dt = schedule(pulses, config['dt'], 1, config['end_time'])
    avec = vector_function(pulses)
    last = max(pulses, key=lambda x: x[0].start)[0]
    t0 = last.start + last.duration / 2
    ion = workloads.reference_production_ion(4)
    unique = list(dict.fromkeys(workloads.PRODUCTION_IONIC_CHANNELS))
    if a.traffic_probe:
        copy_probe('cuda:0')
    started = time.perf_counter()
    engine = IonicAdjoint(ion, len(unique), 'cuda:0')
    state = engine.state(ion.dual_final_states(unique))
    setup_s = time.perf_counter() - started
    state, info = engine.step(state, avec, t0, -dt)
    torch.cuda.synchronize()
    steps = max(1, a.steps)
    iterations = []
    start = time.perf_counter()
    for i in range(steps):
        state, info = engine.step(state, avec, t0 - (i + 1) * dt, -dt)
        iterations.append(info['linear_iterations'])
    torch.cuda.synchronize()
    per_step = (time.perf_counter() - start) / steps
    # the sparse multi-column product, timed alone on the engine's stream
    x = state
    out = torch.empty_like(x)
    coef = engine.coefficients(avec(t0))
    for _ in range(3):
        engine.apply_affine(x, coef, True, 1j * dt, 1., out=out)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(50):
        engine.apply_affine(x, coef, True, 1j * dt, 1., out=out)
    ev1.record()
    torch.cuda.synchronize()
    t_apply = ev0.elapsed_time(ev1) / 1e3 / 50
    nnz = len(engine._col)
    traffic = 20. * nnz + 32. * engine.size
    published = workloads.REFERENCE_IONIC_STEP_SECONDS
    line = {'metric': 'backward ionic CF4 steps per second', 'value': 1.0 / per_step, 'unit': 'steps/s',
            'n_gpus': 1, 'steps': steps, 'warmup': 1, 'ms_per_step': 1e3 * per_step, 'higher_is_better': True,
            'scaling': 'weak', 'vs_baseline': published / per_step, 'dtype': 'complex128',
            'data': 'the reference production configuration\'s He+ operator and dual final states, steps from the '
                    'middle of the probe pulse',
            'config': {'workload': f'tSURFF ionic duals of the reference production run: He+ n {ion.n}, '
                                   f'{ion.na} (l, m) states, {len(unique)} dual columns, CF4-Pade dt {dt:.6g}, '
                                   'GMRES restart 32 at tol 1e-12, field-free eigenbasis preconditioner',
                       'ndof': engine.ndof, 'columns': len(unique), 'nnz': nnz, 'parallelism': 'single GPU',
                       'timed_at': t0, 'setup_seconds': setup_s,
                       'linear_iterations_per_step': float(np.mean(iterations))},
            'baseline': {'value': 1.0 / published, 'unit': 'steps/s',
                         'source': 'results/probe_20260924/validation/production_step_benchmark_l4.json: '
                                   f'{published:.3f} s per backward ionic step on 8 CPU threads'},
            'roofline': {'bound': 'hbm', 'kernel': 'k_ionic_apply (multi-column sparse product)',
                         'achieved': traffic / t_apply / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                         'frac': traffic / t_apply / 1e9 / HBM_PEAK_GBS,
                         'traffic': measured_traffic('ionic_duals_production_l4', 'k_ionic_apply'),
                         'launch_ms': 1e3 * t_apply, 'algorithmic_bytes': traffic}}
    if not a.no_cpu:
        line['cpu_baseline'] = cpu_baseline_ionic(ion, unique, avec, t0, dt)
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({'error': 'this build measures one GPU; multi-GPU runs are not part of it',
                          'n_gpus': a.gpus}))
        sys.exit(2)
    if a.mode == 'metric':
        return bench_metric(a)
    if a.mode == 'batched':
        return bench_batched(a)
    if a.mode == 'adjoint':
        return bench_adjoint(a)
    if a.mode == 'reference-run':
        return bench_reference_run(a)
    if a.mode == 'ionic-duals':
        return bench_ionic_duals(a)
    if a.mode == 'reference-production':
        return bench_reference_production(a)
    import numpy as np
    import torch
    from laqphysics_amd import hip_backend, implicit, workloads
    from laqphysics_amd.hip_backend import HipHamiltonian

    w = workloads.ALL[a.workload]()
    integrator = a.integrator or ('cf4' if w.integrator == 'cf4-pade' else 'arnoldi')
    h = w.h
    dev = torch.device('cuda:0')
    op = HipHamiltonian(h, dev)
    launch = op.autotune()
    avec = w.avec
    pulse = w.pulses[0][0]
    t0 = pulse.start + 0.5 * pulse.duration  # mid-pump: dipole terms active
    # the field-free ground state (device LOBPCG on the real grid, carried onto
    # the ECS grid); its energy is the CF4 reference energy, as in a full run
    ground_energy, x0 = workloads.ground_state(h)
    state = op.state(x0)
    infos = []
    restart = w.extra.get('gmres_restart', 30)
    if integrator == 'cf4':
        op.separable = implicit.SeparablePreconditioner(op)

    def one(k, s):
        t = t0 + k * w.dt
        if integrator == 'cf4':
            s, info = implicit.cf4_step(op, s, avec, t, w.dt, ground_energy, tol=w.tolerance, restart=restart)
        else:
            s, info = hip_backend.step(lambda tt, v: op.apply(v, avec(tt), True), s, t, w.dt,
                                       tol=w.tolerance, maxdim=w.krylov_dimension)
        infos.append(info)
        return s

    if a.traffic_probe:
        copy_probe(dev)
    for k in range(a.warmup):
        state = one(k, state)
    infos.clear()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for k in range(a.steps):
        state = one(a.warmup + k, state)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    rate = a.steps / elapsed

    field = avec(t0)
    x = state
    out = torch.empty_like(x)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    if integrator == 'cf4':
        alpha = w.dt / 2 / (3 + 1j * np.sqrt(3))
        for _ in range(2):
            op.separable.apply(x, alpha, -ground_energy, out=out)
        ev0.record()
        for _ in range(reps):
            op.separable.apply(x, alpha, -ground_energy, out=out)
        ev1.record()
        torch.cuda.synchronize()
        t_kernel = ev0.elapsed_time(ev1) / 1e3 / reps
        flops = 32.0 * h.n ** 3 * h.nc
        achieved = flops / t_kernel / 1e12
        roofline = {'bound': 'mfma', 'kernel': 'separable preconditioner (rocBLAS complex128 GEMMs + resolvent)',
                    'achieved': achieved, 'peak': FP64_MFMA_PEAK_TFS, 'unit': 'TFLOP/s',
                    'frac': achieved / FP64_MFMA_PEAK_TFS,
                    'traffic': measured_traffic(a.workload, 'separable_preconditioner'), 'launch_ms': 1e3 * t_kernel,
                    'algorithmic_flops': flops, 'gemm_streams': op.separable.lanes}
        iters = [i['linear_iterations'] for i in infos]
        cfg_extra = {'linear_iterations_per_step': iters, 'gmres_restart': restart, 'tol': w.tolerance}
    else:
        for _ in range(3):
            op.apply(x, field, True, out=out)
        ev0.record()
        for _ in range(reps):
            op.apply(x, field, True, out=out)
        ev1.record()
        torch.cuda.synchronize()
        t_kernel = ev0.elapsed_time(ev1) / 1e3 / reps
        bytes_apply = 32.0 * h.size
        achieved = bytes_apply / t_kernel / 1e9
        roofline = {'bound': 'hbm', 'kernel': f'H.psi ({apply_kernels(launch)})', 'launch_shape': launch,
                    'achieved': achieved,
                    'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                    # the traffic profiles were measured on the staged walks
                    'traffic': measured_traffic(a.workload, 'hpsi') if launch == 0 else None,
                    'launch_ms': 1e3 * t_kernel, 'algorithmic_bytes': bytes_apply}
        cfg_extra = {'krylov_dimensions': sorted({i['dimension'] for i in infos}), 'tol': w.tolerance}

    name = ('CF4-Pade step (GMRES restart %d, separable preconditioner)' % restart if integrator == 'cf4'
            else 'midpoint Arnoldi step (order %d)' % w.krylov_dimension)
    line = {
        'metric': METRIC, 'value': rate, 'unit': 'steps/s', 'n_gpus': 1, 'steps': a.steps,
        'warmup': a.warmup, 'ms_per_step': 1e3 * elapsed / a.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex128',
        'data': 'synthetic: random-free problem tables; initial state = device ground state, steps from mid-pump',
        'config': {'workload': f'{w.name} {name}, dt {w.dt}', 'n': h.n, 'channels': h.nc, 'N': int(h.size),
                   'batch': 1, 'parallelism': 'single GPU', 'ground_energy': ground_energy, **cfg_extra},
        'roofline': roofline,
    }
    if not a.no_cpu:
        threads = CPU_THREADS
        if integrator == 'cf4':
            its = int(round(np.mean([i['linear_iterations'] for i in infos])))
            line['cpu_baseline'] = cpu_baseline_cf4(h, field, its, threads)
        else:
            line['cpu_baseline'] = cpu_baseline_arnoldi(h, field, max(i['dimension'] for i in infos), threads)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
