"""Propagation benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--integrator cf4|arnoldi] [--no-cpu]

A step is one time step of the whole wavefunction with the operator tables,
preconditioner and state resident in HBM before timing starts:
  cf4      one CF4-Pade step (python/implicit.py:91-109): two Pade factor
           pairs, four preconditioned GMRES solves (restart 30, tol 1e-12);
  arnoldi  one midpoint Arnoldi step (python/torch_backend.py:160-167).
Prints one JSON line.

roofline: the dominant kernel's average launch time is measured with HIP
events on the stream the engine launches on (torch's current stream) -- the
separable preconditioner's complex128 GEMMs (32 n^3 nc FP64 flops per
application, FP64 MFMA peak) for cf4, the fused H.psi (32 N algorithmic bytes,
HBM peak) for arnoldi.  cpu_baseline: the same algorithm on the host cores --
the reference's own Fortran H.psi kernel (compiled from its source into
oracle/_ref; the C restatement when absent) and the NumPy restatement of the
integrator -- on a bounded sample, extrapolated to one step.
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
FP64_MFMA_PEAK_TFS = 78.6
REFERENCE_ENERGY = -2.9


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=5)
    p.add_argument('--warmup', type=int, default=1)
    p.add_argument('--workload', default='production_l6_L4_n441')
    p.add_argument('--integrator', choices=['cf4', 'arnoldi'], default=None)
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--cpu-threads', type=int, default=0)
    return p.parse_args()


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


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


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({'error': 'this build measures one GPU; multi-GPU runs are not part of it',
                          'n_gpus': a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from laqphysics_amd import hip_backend, implicit, workloads
    from laqphysics_amd.hip_backend import HipHamiltonian

    w = workloads.ALL[a.workload]()
    integrator = a.integrator or ('cf4' if w.integrator == 'cf4-pade' else 'arnoldi')
    h = w.h
    dev = torch.device('cuda:0')
    op = HipHamiltonian(h, dev)
    avec = w.avec
    pulse = w.pulses[0][0]
    t0 = pulse.start + 0.5 * pulse.duration  # mid-pump: dipole terms active
    state = op.state(workloads.initial_state(h))
    infos = []
    restart = w.extra.get('gmres_restart', 30)
    if integrator == 'cf4':
        op.separable = implicit.SeparablePreconditioner(op)

    def one(k, s):
        t = t0 + k * w.dt
        if integrator == 'cf4':
            s, info = implicit.cf4_step(op, s, avec, t, w.dt, REFERENCE_ENERGY, tol=w.tolerance, restart=restart)
        else:
            s, info = hip_backend.step(lambda tt, v: op.apply(v, avec(tt), True), s, t, w.dt,
                                       tol=w.tolerance, maxdim=w.krylov_dimension)
        infos.append(info)
        return s

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
            op.separable.apply(x, alpha, 2.9, out=out)
        ev0.record()
        for _ in range(reps):
            op.separable.apply(x, alpha, 2.9, out=out)
        ev1.record()
        torch.cuda.synchronize()
        t_kernel = ev0.elapsed_time(ev1) / 1e3 / reps
        flops = 32.0 * h.n ** 3 * h.nc
        achieved = flops / t_kernel / 1e12
        roofline = {'bound': 'mfma', 'kernel': 'separable preconditioner (rocBLAS complex128 GEMMs + resolvent)',
                    'achieved': achieved, 'peak': FP64_MFMA_PEAK_TFS, 'unit': 'TFLOP/s',
                    'frac': achieved / FP64_MFMA_PEAK_TFS, 'traffic': None, 'launch_ms': 1e3 * t_kernel,
                    'algorithmic_flops': flops}
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
        roofline = {'bound': 'hbm', 'kernel': 'apply_direct (fused H.psi)', 'achieved': achieved,
                    'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                    'launch_ms': 1e3 * t_kernel, 'algorithmic_bytes': bytes_apply}
        cfg_extra = {'krylov_dimensions': sorted({i['dimension'] for i in infos}), 'tol': w.tolerance}

    name = ('CF4-Pade step (GMRES restart %d, separable preconditioner)' % restart if integrator == 'cf4'
            else 'midpoint Arnoldi step (order %d)' % w.krylov_dimension)
    line = {
        'metric': METRIC, 'value': rate, 'unit': 'steps/s', 'n_gpus': 1, 'steps': a.steps,
        'warmup': a.warmup, 'ms_per_step': 1e3 * elapsed / a.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex128', 'data': 'synthetic',
        'config': {'workload': f'{w.name} {name}, dt {w.dt}', 'n': h.n, 'channels': h.nc, 'N': int(h.size),
                   'batch': 1, 'parallelism': 'single GPU', **cfg_extra},
        'roofline': roofline,
    }
    if not a.no_cpu:
        threads = a.cpu_threads or host_cores()
        if integrator == 'cf4':
            its = int(round(np.mean([i['linear_iterations'] for i in infos])))
            line['cpu_baseline'] = cpu_baseline_cf4(h, field, its, threads)
        else:
            line['cpu_baseline'] = cpu_baseline_arnoldi(h, field, max(i['dimension'] for i in infos), threads)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
