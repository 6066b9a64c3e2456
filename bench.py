"""Benchmark: parallel-beam projector fwd+back throughput on one MI355X.

Workload (the GPU configuration of the headline metric): 720 views over
[0, pi), detector 64 rows x 512 channels, volume 512 x 512 x 64 FP32, seeded
40-ellipsoid phantom, sinogram = its forward projection + 1% Gaussian noise.
A step is one full forward_project plus one full back_project through the
public model API with device-resident inputs (output_sharded=True).  The
value is the whole-grid work, 2 * rows*cols*slices*views voxel-views per
step, divided by the step time: Gvox.views/s.  Ten seeded VCD iterations on
the same data are timed once after the projector steps and a one-iteration
warm-up reconstruction (vcd_ms).

The roofline entry is for the dominant kernel of the step, timed live with
HIP events on the stream it runs on.  The CPU baseline times the oracle
(test infrastructure, FP64 sparse-matrix restatement of the reference) on a
bounded sample of the same workload on this host.

Prints ONE JSON line.  Only --gpus 1 is supported (one GPU per box).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

V, R, C = 720, 64, 512
NX = 512
FP32_PEAK_TFLOPS = 157.3
METRIC = ('Gvox\u00b7views/s fwd+back; recons/s + p50/p99 latency batched vs plain '
          '@NRMSE\u22641e-3')
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--vcd-iters', type=int, default=10)
    ap.add_argument('--no-vcd', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--serving-requests', type=int, default=32,
                    help='requests of the mixed serving stream (0 skips it)')
    ap.add_argument('--serving-rate', type=float, default=40.0,
                    help='Poisson arrival rate of the serving stream (requests/s)')
    ap.add_argument('--serving-slots', type=int, default=8)
    ap.add_argument('--serving-iters', type=int, default=20)
    ap.add_argument('--helical', action='store_true',
                    help='also run the helical mixed-precision configuration (2048 views over '
                         '4 turns, 64x736 detector, 512x512x128 volume): float32 vs bfloat16 '
                         'storage VCD')
    ap.add_argument('--helical-iters', type=int, default=3)
    ap.add_argument('--copy-calibration', action='store_true',
                    help='also run a plain 256 MiB device copy (the reference kernel for '
                         'scaling HBM counters) three times')
    ap.add_argument('--cpu-views', type=int, default=24,
                    help='views of the workload the CPU baseline projects')
    return ap.parse_args()


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a plain
    copy of known size), or None when it has not been measured."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        'hbm_traffic.json')
    try:
        with open(path) as f:
            entry = json.load(f)['kernels'].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if entry is None else entry['bytes_per_launch']


def serving_stream(n, rate, iters, seed=0):
    """The mixed request stream: volume drawn from {128^2 x 8, 256^2 x 16,
    512^2 x 32}, views uniform in 90..720 over [0, pi), detector channels =
    width and rows = slices, seeded ellipsoid phantom + 1% noise, Poisson
    arrivals at `rate`."""
    import mbirtorch_amd as mb
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    from mbirtorch_amd.serving import ReconRequest
    rng = np.random.default_rng(seed)
    sizes = [(128, 8), (256, 16), (512, 32)]
    phantoms = {}
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        w, S = sizes[int(rng.integers(3))]
        V = int(rng.integers(90, 721))
        angles = np.linspace(0, np.pi, V, endpoint=False)
        model = mb.ParallelBeamModel((V, S, w), angles)
        rs = tuple(model.get_params('recon_shape'))
        if rs not in phantoms:
            phantoms[rs] = random_ellipsoids(rs, 10, seed=len(phantoms))
        sino = add_noise(model.forward_project(phantoms[rs]), 0.01, seed=1000 + i)
        reqs.append(ReconRequest(sino, angles, max_iterations=iters, seed=i,
                                 params=dict(p=1.2, q=2.0, T=1.0), tag=i,
                                 arrival=float(arrivals[i])))
    return reqs


def serving_bench(args):
    """recons/s and p50/p99 latency of the request engine (continuous
    batching) against the one-at-a-time plain path on the same stream."""
    import dataclasses
    import torch
    from mbirtorch_amd.serving import ReconServer, run_plain, latency_summary
    reqs = serving_stream(args.serving_requests, args.serving_rate, args.serving_iters)
    torch.cuda.synchronize()
    log(f'serving: plain path, {len(reqs)} requests')
    plain = run_plain([dataclasses.replace(r) for r in reqs])
    log('serving: request engine')
    with ReconServer(max_active=args.serving_slots) as srv:
        batched = srv.run_stream([dataclasses.replace(r) for r in reqs])
    by_tag = {r.tag: r for r in plain}
    worst = 0.0
    for r in batched:
        ref = by_tag[r.tag].recon
        worst = max(worst, float(np.sqrt(np.mean((r.recon - ref) ** 2))
                                 / max(np.sqrt(np.mean(ref ** 2)), 1e-30)))
    b, p = latency_summary(batched), latency_summary(plain)
    return {'requests': len(reqs), 'rate_per_s': args.serving_rate,
            'slots': args.serving_slots, 'iterations': args.serving_iters,
            'mix': '128^2x8 / 256^2x16 / 512^2x32, 90-720 views',
            'batched': b, 'plain': p,
            'speedup_recons_per_s': b['recons_per_s'] / p['recons_per_s'],
            'max_nrmse_batched_vs_plain': worst}


def helical_bench(iters):
    """Helical cone beam, 2048 views over 4 turns at pitch 1.0, detector 64 x
    736, magnification 2 (SID 1472, SDD 2944 ALU), volume 512 x 512 x 128
    (voxel 0.5 ALU, slice aspect 2.5): `iters` VCD iterations with float32 and
    with bfloat16 storage of the error sinogram from the same seed; time per
    run, NRMSE between them and the error-sinogram bytes."""
    import torch
    import mbirtorch_amd as mb
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    V, R, C, turns = 2048, 64, 736, 4
    angles = -np.pi + 2 * np.pi * np.arange(V) / (V // turns)
    h_iso = R / 2.0
    z = np.linspace(-turns * h_iso / 2, turns * h_iso / 2, V)
    out = {'views': V, 'detector': [R, C], 'volume': [512, 512, 128], 'iterations': iters}
    recons = {}
    sino = None
    for storage in ('float32', 'bfloat16'):
        m = mb.ConeBeamModel((V, R, C), angles, 2944.0, 1472.0, helical_z_shifts=z)
        m.set_params(recon_shape=(512, 512, 128), delta_voxel=0.5, voxel_slice_aspect=2.5,
                     sinogram_storage=storage)
        if sino is None:
            vol = random_ellipsoids((512, 512, 128), 40, seed=0)
            sino = add_noise(m.forward_project(vol), 0.01, seed=1)
        torch.cuda.synchronize()
        log(f'helical {storage}')
        t0 = time.perf_counter()
        np.random.seed(0)
        recons[storage], d = m.recon(sino, max_iterations=iters, stop_threshold_change_pct=0.0,
                                     print_logs=False, logfile_path=None, output_sharded=True)
        torch.cuda.synchronize()
        out[f'{storage}_s'] = round(time.perf_counter() - t0, 3)
        out[f'{storage}_fm_rmse'] = d['recon_params']['fm_rmse'][-1]
    a, b = recons['bfloat16'], recons['float32']
    out['nrmse_bf16_vs_f32'] = float(torch.sqrt(torch.mean((a - b) ** 2))
                                     / torch.sqrt(torch.mean(b ** 2)))
    out['error_sinogram_bytes'] = {'float32': 4 * V * R * C, 'bfloat16': 2 * V * R * C}
    return out


def cpu_baseline(angles, vol, sino, views):
    """Oracle fwd+back on `views` views of the workload (FP64, 1 core)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import parallel as op
    rs = vol.shape
    idx = op.full_indices(rs)
    geom = dict(num_rows=rs[0], num_cols=rs[1], num_channels=C, delta_det_channel=1.0,
                det_channel_offset=0.0, delta_voxel=1.0, delta_voxel_row=1.0)
    sel = np.linspace(0, V - 1, views).astype(int)
    vals = vol.reshape(-1, rs[2])[idx]
    t0 = time.perf_counter()
    op.forward(vals, idx, angles[sel], geom)
    op.back(sino[sel], idx, angles[sel], geom)
    dt = time.perf_counter() - t0
    work = 2.0 * rs[0] * rs[1] * rs[2] * views
    return dict(value=work / dt / 1e9, unit='Gvox\u00b7views/s', cores=1, kind='port',
                sample=f'oracle forward+back (FP64 sparse matrix, scipy, 1 thread) of {views} '
                       f'of the {V} views, full 512x512x64 grid: {dt:.1f} s')


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({'error': 'mbirtorch_amd runs on one GPU; --gpus must be 1 '
                                   '(multi-GPU sharding is designed, not built)'}))
        sys.exit(2)
    import torch
    import mbirtorch_amd as mb
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    from mbirtorch_amd import _lib

    torch.cuda.set_device(0)
    angles = np.linspace(0, np.pi, V, endpoint=False)
    model = mb.ParallelBeamModel((V, R, C), angles)
    rs = tuple(model.get_params('recon_shape'))
    log(f'phantom {rs}')
    vol_np = random_ellipsoids(rs, 40, seed=0)
    vol = torch.as_tensor(vol_np, device='cuda')
    clean = model.forward_project(vol, output_sharded=True)
    sino_np = add_noise(clean.cpu().numpy(), 0.01, seed=1)
    sino = torch.as_tensor(sino_np, device='cuda')
    torch.cuda.synchronize()

    def step():
        model.forward_project(vol, output_sharded=True)
        model.back_project(sino, output_sharded=True)

    log(f'warmup {args.warmup}')
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f'timing {args.steps} steps')
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    work = 2.0 * rs[0] * rs[1] * rs[2] * V
    value = work / (ms_step * 1e-3) / 1e9

    # Dominant kernels, timed alone on the launch stream.
    ops = model.projector()
    idx = model.full_indices_device()
    vals = vol.reshape(-1, rs[2])[idx].contiguous()
    sino_vcr = model._vrc_to_vcr(sino)

    def time_kernel(fn, n=5):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    runs = model.full_index_runs_device()
    fwd_ms = time_kernel(lambda: ops.forward(vals, idx, layout=_lib.MBIR_LAYOUT_VRC, runs=runs))
    back_ms = time_kernel(lambda: ops.back(sino_vcr, idx))
    P = int(idx.shape[0])
    ror_vv = P * rs[2] * V
    # Algorithmic FLOPs per ROR voxel-view: two taps x (mul + add).
    flops = 4.0 * ror_vv
    dom_name, dom_ms = ('back', back_ms) if back_ms >= fwd_ms else ('forward', fwd_ms)
    achieved = flops / (dom_ms * 1e-3) / 1e12
    bytes_alg = 4.0 * (P * rs[2] + V * R * C)
    roofline = dict(bound='mfma', achieved=round(achieved, 3), peak=FP32_PEAK_TFLOPS,
                    unit='TFLOP/s', frac=round(achieved / FP32_PEAK_TFLOPS, 4),
                    traffic=measured_traffic(f'par_{dom_name}'),
                    kernel=f'par_{dom_name}', hbm_gbs=round(bytes_alg / (dom_ms * 1e-3) / 1e9, 1))

    if args.copy_calibration:
        log('copy calibration')
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device='cuda')  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    vcd_ms = None
    if not args.no_vcd:
        log('vcd')
        model.set_params(p=1.2, q=2.0, T=1.0)
        # Warm-up: the first reconstruction in a process loads the VCD kernels
        # and grows the allocator's pools.
        model.recon(sino, max_iterations=1, stop_threshold_change_pct=0.0,
                    print_logs=False, logfile_path=None, output_sharded=True)
        np.random.seed(0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        model.recon(sino, max_iterations=args.vcd_iters, stop_threshold_change_pct=0.0,
                    print_logs=False, logfile_path=None, output_sharded=True)
        torch.cuda.synchronize()
        vcd_ms = (time.perf_counter() - t1) * 1e3

    serving = None
    if args.serving_requests > 0:
        serving = serving_bench(args)
    helical = helical_bench(args.helical_iters) if args.helical else None

    cpu = None
    if not args.no_cpu:
        log('cpu baseline')
        cpu = cpu_baseline(angles, vol_np, sino_np, args.cpu_views)

    out = {
        'metric': METRIC,
        'value': round(value, 2), 'unit': 'Gvox\u00b7views/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms_step, 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded 40-ellipsoid phantom, forward projection + 1% noise)',
        'config': {'workload': 'parallel-beam fwd+back: 720 views [0,pi), 64x512 detector, '
                               '512x512x64 volume', 'views': V, 'det_rows': R,
                   'det_channels': C, 'volume': list(rs), 'parallelism': 'single GPU'},
        'roofline': roofline, 'cpu_baseline': cpu, 'serving': serving, 'helical_bf16': helical,
        'fwd_ms': round(fwd_ms, 4), 'back_ms': round(back_ms, 4),
        'vcd_iters': None if vcd_ms is None else args.vcd_iters,
        'vcd_ms': None if vcd_ms is None else round(vcd_ms, 1),
        'wall_s': round(wall, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
