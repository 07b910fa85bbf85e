"""Benchmark: 2D Euler Kelvin-Helmholtz, 1024x1024 periodic, Roe + plm cons
(minmod), RK2-TVD, CFL 0.4, FP64 — the plain single-grid form of the
finite-volume hot path on one MI355X.

A step is one full RK2-TVD update of the 1024^2 grid (two fused stage
launches; dt stays on the device). The state is resident in HBM before the
timed region. Prints one JSON line with Gcell-updates/s, the roofline of the
stage kernel (HIP events around each launch on the solver stream) and the CPU
oracle timed on the host cores on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, 'hydro-cl-lua_amd'))

HBM_PEAK_GBS = 8000.0
N = 1024
NQ = 5                      # integrated reals per cell (Euler)


def algorithmic_bytes_per_cell_update(w, stages_rw):
    """bytes of state that must cross HBM per cell-update (one RK2-TVD step):
    stage 1 reads U^n and writes U^1, stage 2 reads U^1 and U^n and writes
    U^{n+1}: 5 q w"""
    return stages_rw * NQ * w


def cpu_baseline(seconds=10.0):
    """The CPU oracle (FP64, reference kernel sequence, OpenMP) on the same
    KH 1024^2 input, timed for about `seconds` of steps."""
    sys.path.insert(0, os.path.join(REPO, 'oracle'))
    import pyoracle
    from hydro.init import euler as init_euler
    threads = min(16, os.cpu_count() or 1)
    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(dim=2, size=(N, N), mins=(-.5, -.5), maxs=(.5, .5), plm='plm cons',
                        slope_limiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4,
                        boundary='periodic')
    U0 = init_euler.to_ghosted(o.cfg, init_euler.kelvin_helmholtz(o.cfg, seed=0))
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    o.update(1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.update(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds and steps >= 2:
            break
    rate = N * N * steps / el / 1e9
    cpu = ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                cpu = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return {'value': rate, 'unit': 'Gcell-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'{steps} RK2-TVD steps of the same KH 1024^2 FP64 input ({el:.1f} s), '
                      f'oracle = reference kernel sequence, AoS, unfused, OpenMP; cpu: {cpu}'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--real', default='double', choices=['double', 'float'])
    ap.add_argument('--size', type=int, default=N)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--exact-math', action='store_true',
                    help='IEEE division/sqrt (bitwise reference arithmetic) instead of the faithful fast sequences')
    ap.add_argument('--rows', type=int, default=0, help='rows per wave of the stage kernel (0: autotuned default)')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (got --gpus {args.gpus}); the grid path is not '
              'sharded across GPUs in this build', file=sys.stderr)
        sys.exit(2)

    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler

    n = args.size
    s = HipFVSolver(dim=2, gridSize=(n, n), mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons',
                    slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4,
                    boundary='periodic', real=args.real, fastMath=not args.exact_math)
    if args.rows:
        s.setLaunch(args.rows, 0)
    U0 = init_euler.to_ghosted(s.cfg, init_euler.kelvin_helmholtz(s.cfg, seed=0))
    s.resetState(U0.astype(s.dtype))
    s.update(max(args.warmup, 1))
    s.sync()

    # timed region: exactly K steps, synchronized on both sides
    s.sync()
    t0 = time.perf_counter()
    s.update(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    cells = n * n
    value = cells * args.steps / wall / 1e9

    # stage kernel duration: HIP events around each launch on the solver stream
    kt = s.timeUpdates(min(args.steps, 100), perKernel=True)
    avg_ms = kt.stage_ms / max(kt.n_stage_launches, 1)
    w = 8 if args.real == 'double' else 4
    bytes_per_update = algorithmic_bytes_per_cell_update(w, 5)
    bytes_per_launch = bytes_per_update / 2 * cells      # two stage launches per step
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(REPO, 'profiles', 'stage_kernel_traffic.json')
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(f'kh{n}_{args.real}')
        except (OSError, ValueError):
            traffic = None

    out = {
        'metric': 'Gcell-updates/s (2D Euler Roe+PLM) + batched solves/s & p99 latency vs plain',
        'value': value,
        'unit': 'Gcell-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64' if args.real == 'double' else 'f32',
        'data': 'synthetic (seeded Kelvin-Helmholtz shear layer)',
        'config': {
            'workload': f'2D Euler Kelvin-Helmholtz {n}x{n} periodic, Roe + plm cons/minmod, '
                        f'RK2-TVD, CFL 0.4, {"FP64" if args.real == "double" else "FP32 bulk"}',
            'cells': cells, 'stage_launches_per_step': 2,
            'rows_per_wave': s.getLaunch()[0],
            'math': 'ieee' if args.exact_math else 'fast (faithfully rounded div/sqrt)',
        },
        'roofline': {
            'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
            'kernel': 'hydro::stage_kernel<double, Euler, 2, PLM, minmod>' if args.real == 'double'
                      else 'hydro::stage_kernel<float, Euler, 2, PLM, minmod>',
            'avg_launch_ms': avg_ms,
            'algorithmic_bytes_per_launch': bytes_per_launch,
        },
        't_sim': s.t,
    }
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(args.cpu_seconds)
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out))


if __name__ == '__main__':
    main()
