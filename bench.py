"""Propagation benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one midpoint Arnoldi time step (python/torch_backend.py:160-167
semantics: adaptive Krylov dimension up to the workload's order, host-side
check cadence) of the whole wavefunction, with the operator tables and the
state resident in HBM before timing starts.  Prints one JSON line.

roofline: the dominant kernel's average launch time is measured with HIP
events on the stream the engine launches on (torch's current stream), and its
algorithmic bytes (read psi once, write H psi once: 32 N bytes for N complex128
entries) are divided by it.  cpu_baseline: the reference's own Fortran kernel
(compiled from its source into oracle/_ref; the restated C kernel when that is
absent) timed on the host cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = 'prop steps/s (×batch) + adjoint grad time vs plain GPU at ‖Δψ‖≤1e-10'
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--workload', default='production_l6_L4_n441')
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--cpu-threads', type=int, default=0)
    return p.parse_args()


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(h, field, dimension, threads):
    """Bounded CPU sample: 2 full-size applies by the reference kernel and one
    CGS2 column at the mid Krylov index, extrapolated to a step of the GPU's
    Krylov dimension.  Runs in a child-free way: OMP threads fixed first."""
    os.environ['OMP_NUM_THREADS'] = str(threads)
    import numpy as np
    import oracle
    from scipy.linalg.blas import zgemv
    use_ref = oracle.have_reference_kernel()
    apply = oracle.ref_tensor_apply if use_ref else oracle.tensor_apply
    rng = np.random.default_rng(1)
    x = rng.normal(size=h.size) + 1j * rng.normal(size=h.size)
    x /= np.linalg.norm(x)
    h.diag, h.rad  # full tables, built outside the timed sample
    apply(h, x, field, True)  # first-touch / thread start
    t = time.perf_counter()
    reps = 2
    for _ in range(reps):
        y = apply(h, x, field, True)
    t_apply = (time.perf_counter() - t) / reps
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
    return {
        'value': 1.0 / step, 'unit': 'steps/s', 'cores': threads,
        'kind': 'reference' if use_ref else 'port',
        'sample': (f'{reps} full-size H.psi by the reference Fortran kernel built from src/kernels.f90 '
                   if use_ref else f'{reps} full-size H.psi by the C restatement ')
        + f'({t_apply:.2f} s each, OpenMP {threads} threads) + one CGS2 column over {k} vectors '
          f'(NumPy/BLAS zgemv, {t_cgs:.2f} s), extrapolated to a {dimension}-dimensional Arnoldi step',
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({'error': 'this build measures one GPU; multi-GPU runs are not part of it',
                          'n_gpus': a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from laqphysics_amd import hip_backend, workloads
    from laqphysics_amd.hip_backend import HipHamiltonian

    w = workloads.ALL[a.workload]()
    h = w.h
    dev = torch.device('cuda:0')
    op = HipHamiltonian(h, dev)
    avec = w.avec
    pulse = w.pulses[0][0]
    t0 = pulse.start + 0.5 * pulse.duration  # mid-pump: dipole terms active
    state = op.state(workloads.initial_state(h))
    dims = []

    def one(k, s):
        t = t0 + k * w.dt
        s, info = hip_backend.step(lambda tt, v: op.apply(v, avec(tt), True), s, t, w.dt,
                                   tol=w.tolerance, maxdim=w.krylov_dimension)
        dims.append(info['dimension'])
        return s

    for k in range(a.warmup):
        state = one(k, state)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for k in range(a.steps):
        state = one(a.warmup + k, state)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    rate = a.steps / elapsed

    # dominant kernel: the fused H.psi apply, timed alone with HIP events
    field = avec(t0)
    x = state
    out = torch.empty_like(x)
    for _ in range(3):
        op.apply(x, field, True, out=out)
    reps = 20
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        op.apply(x, field, True, out=out)
    ev1.record()
    torch.cuda.synchronize()
    t_apply = ev0.elapsed_time(ev1) / 1e3 / reps
    bytes_apply = 32.0 * h.size
    achieved = bytes_apply / t_apply / 1e9

    line = {
        'metric': METRIC, 'value': rate, 'unit': 'steps/s', 'n_gpus': 1, 'steps': a.steps,
        'warmup': a.warmup, 'ms_per_step': 1e3 * elapsed / a.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex128', 'data': 'synthetic',
        'config': {'workload': f'{w.name} midpoint-Arnoldi step (order {w.krylov_dimension}, dt {w.dt})',
                   'n': h.n, 'channels': h.nc, 'N': int(h.size), 'batch': 1,
                   'krylov_dimensions': sorted(set(dims)), 'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'kernel': 'apply_direct (fused H.psi)', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                     'launch_ms': 1e3 * t_apply, 'algorithmic_bytes': bytes_apply},
    }
    if not a.no_cpu:
        threads = a.cpu_threads or host_cores()
        line['cpu_baseline'] = cpu_baseline(h, field, max(dims), threads)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
