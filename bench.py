#!/usr/bin/env python3
"""Throughput benchmark of the photon-packet transport engine on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One step is one simulation run (the C-ABI call behind Mc.run: kernel launch,
accumulator download) over the workload's packet budget, with the voxel grid
already resident on the device. Prints one JSON line.

Roofline: the transport kernels do scalar binary32 work with no matrix
contraction, so the bound is the VALU issue rate: 256 CUs x 64 lanes/clock
x 2.4 GHz = 39.3 T lane-op slots/s (transcendental-class operations — log,
sin/cos, sqrt, division — take 4 slots). The algorithmic work per transport
iteration is a fixed hand count of the reference's operations (an interior
albedo-weight + HG iteration 85 lane-ops with 6 transcendental-class, a
boundary iteration 43 with 6, the voxel intersection adds 12 with 3), so
achieved = slots per launch / kernel time. The kernel time is measured with
HIP events on the kernel's own stream inside the library.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_SLOTS = 256 * 64 * 2.4e9          # lane-op issue slots per second
SLOTS_INTERIOR = 85 - 6 + 6 * 4         # 103
SLOTS_BOUNDARY = 43 - 6 + 6 * 4         # 61
SLOTS_VOXEL_EXTRA = 12 - 3 + 3 * 4      # 21

SERVING_QPS = 250.0                     # requests/s of the serving run

# HBM bytes per launch of the dominant kernel: rocprofv3 FETCH_SIZE and
# WRITE_SIZE of one launch, each scaled by what the counter reads for a
# 1 GiB device-to-device copy (profiles/batch_kernel_layered_batch_4096.md,
# profiles/mcvox_kernel_skin_vessel_200.md). Offline figures: measured for
# the kernels as they are in this tree.
TRAFFIC_BYTES_PER_LAUNCH = {'layered_batch_4096': 1.784e11,
                            'skin_vessel_200': 7.844e11}

METRIC = ('photon packets/s (1 GPU) batched vs plain; p50/p99 solve latency '
          'at fixed QPS')


def algorithmic_slots(rep, voxel: bool) -> float:
    interior = rep.steps - rep.boundary_steps
    slots = interior * SLOTS_INTERIOR + rep.boundary_steps * SLOTS_BOUNDARY
    if voxel:
        slots += rep.steps * SLOTS_VOXEL_EXTRA
    return float(slots)


def make_workload(name: str, packets, launch=None, lut_processes=None):
    from pyxopto_amd import workloads
    if name == 'layered_batch_4096':
        from pyxopto_amd.mcml.batch import Batch
        scale = 1.0 if not packets else packets / 8.77e8
        insts = workloads.layered_batch(
            budget_scale=scale, processes=lut_processes,
            cache_dir=os.path.join(REPO, 'pyxopto_amd', 'data'))
        batch = Batch(insts, rnginit=0x5EED000000000001)
        packed = batch.pack()
        if launch:
            fields = [int(v) for v in launch.split(':')]
            opts = batch.opts(*fields)
        else:
            from pyxopto_amd.mcml.autotune import tuned_opts
            opts = tuned_opts(batch, packed)
        arr, lut, naccum = packed
        npk = sum(int(i.nphotons) for i in insts)
        ctx = batch.context

        def step(first):
            return ctx.run_batch(arr, lut, opts, naccum)[1]

        config = {'workload': 'layered_batch_4096', 'instances': len(insts),
                  'layers': '1-5 turbid + 2 ambient', 'packets': npk,
                  'phase_functions': 'HG or 2000-point LUT of MHG',
                  'detectors': '100-bin radial + total R/T per instance',
                  'launch': {'chunk': opts.chunk, 'block': opts.block,
                             'blocks_per_cu': opts.blocks_per_cu,
                             'refill_min': opts.refill_min or 4,
                             'defer_min': opts.defer_min or 16}}
        batch.bench_plain = lambda: plain_rate(batch, insts)
        batch.bench_insts = insts
        return batch, npk, step, False, config
    if name == 'skin_vessel_200':
        sim = workloads.skin_vessel()
        npk = packets or workloads.SKIN_VESSEL_PACKETS
        cfg, lut, naccum = sim.pack()
        ctx = sim.context
        vox = sim.voxels.material

        def step(first):
            return ctx.run_mcvox(cfg, vox if first else None, lut, npk,
                                 naccum)[1]

        config = {'workload': 'skin_vessel_200',
                  'voxels': '200x200x200 x 10 um', 'packets': npk,
                  'source': 'gaussian sigma 50 um',
                  'accumulator': 'uint64 deposition per voxel'}
        return sim, npk, step, True, config
    if name == 'slab_1mm':
        sim = workloads.slab_1mm()
        npk = packets or workloads.SLAB_PACKETS
        cfg, lut, naccum = sim.pack()
        ctx = sim.context

        def step(first):
            return ctx.run_mcml(cfg, lut, npk, naccum)[1]

        config = {'workload': 'slab_1mm', 'layers': 3, 'packets': npk,
                  'detectors': 'radial 100 bins + total R/T'}
        return sim, npk, step, False, config
    raise SystemExit('unknown workload ' + name)


def ceilings(ctx):
    """Measured ceilings on this device: streaming copy bandwidth (1 GiB,
    read + write) and random-address u64 atomic adds over a 64 MiB array."""
    nbytes = 1 << 30
    ms = ctx.probe_copy(nbytes, reps=10)
    natom = 1 << 28
    ams, total = ctx.probe_atomic_u64(64 << 20, natom, seed=7)
    if total != natom:
        raise RuntimeError('atomic probe lost updates: {} of {}'.format(
            total, natom))
    vms, vops = ctx.probe_valu(20000)
    return {'valu_fma_lane_ops_per_s': vops / (vms * 1e-3),
            'copy_GBps': 2 * nbytes / (ms * 1e-3) / 1e9,
            'atomic_u64_per_s': natom / (ams * 1e-3),
            'atomic_u64_GBps': 8 * natom / (ams * 1e-3) / 1e9,
            'probe': 'fp32 FMA chains 8 x 20000 per lane; copy 1 GiB x10; '
                     '2^28 u64 atomics, 64 MiB array'}


def serving(insts, qps: float, duration: float, seed: int = 20261203,
            voxel_fraction: float = 0.1, saturation_requests: int = 400,
            tick_packets: int = None):
    """Poisson arrivals at a fixed rate for `duration` seconds: 90 % of the
    requests are instances of the layered batch (in a seeded random order),
    10 % are 64^3 two-layer voxel simulations of 1e6 packets. Requests are
    packed before the run; the clock runs from the first submit to the last
    result. Latency is submit to result. Before it, a burst of requests
    submitted at once measures the capacity of the same mix."""
    import numpy as np
    from pyxopto_amd import serve, workloads
    from pyxopto_amd.mcml.batch import Batch
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(insts))
    packed = []
    for i in order:
        arr, lut, _ = Batch([insts[i]], rnginit=0x5EED000000000001).pack()
        packed.append((arr[0], lut))
    vox = workloads.skin_two_layer()
    vcfg, vlut, vnaccum = vox.pack()
    vmat = vox.voxels.material
    nvox = workloads.VOXEL_REQUEST_PACKETS
    is_vox = rng.random(1 << 20) < voxel_fraction

    def make(srv, k):
        if is_vox[k % is_vox.size]:
            return srv.submit_packed_voxel(vcfg, vmat, vlut, vnaccum, nvox,
                                           seed=0xA5A5 + k)
        inst, lut = packed[k % len(packed)]
        return srv.submit_packed_layered(inst, lut)

    opts = {} if tick_packets is None else {'tick_packets': tick_packets}
    with serve.Server(**opts) as srv:
        # first use of the server (graph, tables) outside the clock
        warm = [make(srv, k) for k in range(4)]
        srv.drain(120)
        # capacity: a burst of requests submitted at once
        t0 = time.perf_counter()
        burst = [make(srv, (1 << 19) + k) for k in range(saturation_requests)]
        srv.drain(600)
        t_burst = time.perf_counter() - t0
        s0 = srv.stats()
        reqs, span = serve.poisson_load(srv, lambda sv, k: make(sv, k + 4),
                                        qps, duration, seed=seed)
        s1 = srv.stats()
    del warm
    burst_packets = sum(r.npackets for r in burst)
    packets = sum(r.npackets for r in reqs)
    out = {'capacity': {'requests': len(burst),
                        'requests_per_s': len(burst) / t_burst,
                        'packets_per_s': burst_packets / t_burst},
           'qps': qps, 'arrivals_s': duration, 'requests': len(reqs),
           'voxel_requests': sum(r.kind == 1 for r in reqs),
           'packets': packets, 'packets_per_s': packets / span,
           'steps_per_s': (s1['steps'] - s0['steps']) / span,
           'ticks': s1['ticks'] - s0['ticks'],
           'mix': '90% layered-batch instances, 10% 64^3 two-layer voxel '
                  'requests of 1e6 packets'}
    out.update(serve.latency_summary(reqs))
    for kind, name in ((0, 'layered'), (1, 'voxel')):
        lat = serve.latency_summary([r for r in reqs if r.kind == kind])
        out[name + '_p50_ms'] = lat['p50_ms']
        out[name + '_p99_ms'] = lat['p99_ms']
    return out


def surrogate_rates(ctx, rows=(65536, 1048576), batch: int = 4096,
                    steps: int = 50):
    """MFMA throughput of the surrogate network (5-256-256-256-100, bf16
    products, fp32 accumulation): fused inference over each count of
    synthetic input rows in `rows` (the figures without a suffix are the
    largest), and `steps` captured Adam steps on minibatches of `batch`."""
    import numpy as np
    from pyxopto_amd import surrogate as S
    dims = S.DIMS
    flops_row = 2 * sum(i * o for i, o in zip(dims[:-1], dims[1:]))
    rng = np.random.default_rng(2)
    out = {'network': '5-256-256-256-100 relu, bf16 MFMA, fp32 accumulate'}
    infer = None
    for n in sorted(rows):
        net = S.Network(dims, ctx, max_batch=n)
        net.set_params(S.init_params(dims, seed=1))
        x = rng.uniform(-1, 1, (n, dims[0])).astype(np.float32)
        net.forward(x)
        _, ms = net.forward(x)
        net.close()
        infer = flops_row * n / (ms * 1e-3) / 1e12
        out['inference_rows_%d_per_s' % n] = n / (ms * 1e-3)
        out['inference_TFLOPs_%d_rows' % n] = infer
    out['inference_rows'] = max(rows)
    out['inference_TFLOPs'] = infer
    n = 65536
    net = S.Network(dims, ctx, max_batch=batch)
    net.set_params(S.init_params(dims, seed=1))
    x = rng.uniform(-1, 1, (n, dims[0])).astype(np.float32)
    t = rng.normal(0, 1, (n, dims[-1])).astype(np.float32)
    net.set_data(x, t)
    net.train(2, batch, 1e-3, seed=3)
    _, tms = net.train(steps, batch, 1e-3, seed=3)
    net.close()
    train = 3 * flops_row * batch * steps / (tms * 1e-3) / 1e12
    out.update({'train_step_ms': tms / steps, 'train_batch': batch,
                'train_TFLOPs': train,
                'roofline': {'bound': 'mfma', 'peak': 2500.0, 'unit': 'TFLOP/s',
                             'frac_inference': infer / 2500.0,
                             'frac_train': train / 2500.0}})
    return out


PUBLISHED_FIBER_S = 5.7   # RTX A6000, test/performance.py:134-135 (4e9 packets)


def published_benchmark(packets: int = None):
    """The reference's published benchmark (test/performance.py with
    validate.SingleLayerUniformFiberRadial): 400 runs of 1e7 packets of a
    single layer under a multimode fiber, timed from simulator construction
    to the last result as the reference does (t2 - t1). Plain: one Mc.run per
    grid point, as the reference. Batched: the 400 runs as one batch."""
    from pyxopto_amd import workloads as W
    from pyxopto_amd.mcml.batch import Batch
    npk = packets or W.FIBER_PACKETS
    mua, musr = W.fiber_grid()
    # plain, as the reference test loops
    t0 = time.perf_counter()
    sim = W.fiber_sim()
    for m, s in zip(mua, musr):
        sim.layers[1].mua = m
        sim.layers[1].mus = s / (1.0 - 0.85)
        sim.run(npk)
    t_plain = time.perf_counter() - t0
    # batched
    t0 = time.perf_counter()
    b = Batch(W.fiber_batch(packets=npk), rnginit=0x5EED000000000001)
    b._ctx = sim.context
    res, rep = b.run(b.opts(chunk=8192, block=256))
    t_batch = time.perf_counter() - t0
    total = 400 * npk
    return {'workload': '400 runs x {} packets, UniformFiberNI NA 0.22 '
                        '200 um, 8 mm layer n 1.33 HG 0.85, rmax 5 mm'
                        .format(npk),
            'published_s': PUBLISHED_FIBER_S * npk / 1e7,
            'published_device': 'NVIDIA RTX A6000',
            'plain_s': t_plain, 'batched_s': t_batch,
            'plain_packets_per_s': total / t_plain,
            'batched_packets_per_s': total / t_batch,
            'batched_kernel_s': rep.kernel_ms / 1e3,
            'speedup_vs_published_plain': PUBLISHED_FIBER_S * npk / 1e7 /
            t_plain,
            'speedup_vs_published_batched': PUBLISHED_FIBER_S * npk / 1e7 /
            t_batch}


def plain_rate(batch, insts, every: int = 16):
    """The same instances one launch each (Mc.run: xo_mcml_run with
    lane-owned generators), on every 16th instance: packets/s."""
    from pyxopto_amd.mcml import mc
    sub = insts[::every]
    sims = [mc.Mc(i.layers, i.source, i.detectors, rnginit=0x5EED000000000001)
            for i in sub]
    for s in sims:             # one context, created outside the timing
        s._ctx = batch.context
    packed = [s.pack() for s in sims]
    ctx = batch.context
    npk = sum(int(i.nphotons) for i in sub)
    t0 = time.perf_counter()
    for s, (cfg, lut, naccum), i in zip(sims, packed, sub):
        ctx.run_mcml(cfg, lut, i.nphotons, naccum)
    dt = time.perf_counter() - t0
    return {'packets_per_s': npk / dt, 'instances': len(sub),
            'packets': npk, 'seconds': dt}


def cpu_baseline(sim, voxel: bool, budget_s: float = 12.0):
    """The oracle (C, OpenMP, C-library math) on a bounded sample of the
    same workload, timed on this host's cores."""
    import numpy as np
    from pyxopto_amd import _native as N
    from oracle import pyoracle as O
    cores = len(os.sched_getaffinity(0))
    if hasattr(sim, 'instances'):
        return cpu_baseline_batch(sim, cores, budget_s)
    cfg, lut, naccum = sim.pack()
    lanes = 4 * cores * 64
    x, a = N.init_rng(lanes, 0x0123456789ABCDEF)
    chunk = 20000 if voxel else 200000
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        if voxel:
            _, _, x = O.mcvox_run(cfg, sim.voxels.material, lut, x, a, chunk,
                                  naccum, O.MATH_LIBM, cores)
        else:
            _, _, x = O.mcml_run(cfg, lut, x, a, chunk, naccum, O.MATH_LIBM,
                                 cores)
        done += chunk
    dt = time.perf_counter() - t0
    return {'value': done / dt, 'unit': 'photon packets/s', 'cores': cores,
            'kind': 'port',
            'sample': '{} packets of the same workload in {:.1f} s'.format(
                done, dt)}


def cpu_baseline_batch(batch, cores: int, budget_s: float):
    """The oracle on a seeded 256-instance subset of the batch (budgets cut
    to 2000 packets each), repeated until the time budget is spent."""
    from pyxopto_amd.mcml.batch import Batch
    from oracle import pyoracle as O
    import ctypes
    sub = Batch(batch.instances[::16], rnginit=0x5EED000000000001)
    arr, lut, naccum = sub.pack()
    for i in range(len(arr)):
        arr[i].npackets = min(int(arr[i].npackets), 2000)
    per = sum(int(arr[i].npackets) for i in range(len(arr)))
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        O.batch_run(arr, lut, naccum, math_mode=O.MATH_LIBM, nthreads=cores)
        done += per
        for i in range(len(arr)):
            arr[i].seed += 1
    dt = time.perf_counter() - t0
    return {'value': done / dt, 'unit': 'photon packets/s', 'cores': cores,
            'kind': 'port',
            'sample': '{} packets of a 256-instance subset (2000 packets '
                      'each) in {:.1f} s'.format(done, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='layered_batch_4096',
                    choices=['layered_batch_4096', 'skin_vessel_200',
                             'slab_1mm'])
    ap.add_argument('--packets', type=int, default=None)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-plain', action='store_true',
                    help='skip the one-launch-per-instance comparison')
    ap.add_argument('--no-serving', action='store_true',
                    help='skip the fixed-QPS serving run')
    ap.add_argument('--no-published', action='store_true',
                    help='skip the reference\'s published fiber benchmark')
    ap.add_argument('--no-surrogate', action='store_true',
                    help='skip the surrogate network throughput')
    ap.add_argument('--qps', type=float, default=SERVING_QPS)
    ap.add_argument('--tick-packets', type=int, default=None,
                    help='layered packets per serving tick')
    ap.add_argument('--serving-seconds', type=float, default=8.0)
    ap.add_argument('--lut-processes', type=int, default=None,
                    help='LUT builder processes (1 under a profiler: no '
                         'fork after the GPU is initialised)')
    ap.add_argument('--launch', default=None,
                    help='CHUNK:BLOCK:BLOCKS_PER_CU[:REFILL_MIN[:DEFER_MIN]] '
                         'instead of the '
                         'autotuner')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this engine has no multi-GPU '
              'path (the packet loop would shard by seed rows with one final '
              'integer all-reduce); got --gpus {}'.format(args.gpus),
              file=sys.stderr)
        sys.exit(2)

    sim, npk, step, voxel, config = make_workload(
        args.workload, args.packets, args.launch, args.lut_processes)
    print('bench: workload {} packets/step {} warmup {} steps {}'.format(
        config['workload'], npk, args.warmup, args.steps), file=sys.stderr,
        flush=True)
    first = True
    for _ in range(args.warmup):
        step(first)
        first = False
    reps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reps.append(step(first))
        first = False
    wall = time.perf_counter() - t0
    print('bench: timed {} steps in {:.3f} s'.format(args.steps, wall),
          file=sys.stderr, flush=True)

    kernel_s = sum(r.kernel_ms for r in reps) * 1e-3
    steps_total = sum(r.steps for r in reps)
    slots = sum(algorithmic_slots(r, voxel) for r in reps)
    achieved = slots / kernel_s / 1e12
    value = npk * args.steps / wall
    out = {
        'metric': METRIC,
        'value': value,
        'unit': 'photon packets/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': config,
        'transport_steps_per_s': steps_total / wall,
        'kernel_ms_per_step': kernel_s / args.steps * 1e3,
        'roofline': {
            'bound': 'valu',
            'achieved': achieved,
            'peak': PEAK_SLOTS / 1e12,
            'unit': 'T lane-op slots/s',
            'frac': achieved / (PEAK_SLOTS / 1e12),
            'traffic': TRAFFIC_BYTES_PER_LAUNCH.get(config['workload']),
        },
    }
    ceil = ceilings(sim.context)
    out['measured_ceilings'] = ceil
    out['roofline']['hbm_copy_GBps'] = ceil['copy_GBps']
    if voxel:
        dep = sum(r.deposits for r in reps)
        out['atomic_deposits_per_s'] = dep / kernel_s
        out['atomic_GBps'] = dep * 8 / kernel_s / 1e9
        out['atomic_frac_of_probe'] = dep / kernel_s / ceil['atomic_u64_per_s']
        out['voxel_crossings_per_s'] = sum(r.voxel_crossings
                                           for r in reps) / kernel_s
    if hasattr(sim, 'bench_plain') and not args.no_plain:
        plain = sim.bench_plain()
        out['plain_one_launch_per_instance'] = plain
        out['batched_vs_plain'] = value / plain['packets_per_s']
    if hasattr(sim, 'bench_insts') and not args.no_published:
        print('bench: published fiber benchmark', file=sys.stderr, flush=True)
        out['published_benchmark'] = published_benchmark()
    if not args.no_surrogate:
        out['surrogate'] = surrogate_rates(sim.context)
    if hasattr(sim, 'bench_insts') and not args.no_serving:
        print('bench: serving at {} requests/s for {} s'.format(
            args.qps, args.serving_seconds), file=sys.stderr, flush=True)
        out['serving'] = serving(sim.bench_insts, args.qps,
                                 args.serving_seconds,
                                 tick_packets=args.tick_packets)
        print('bench: serving done', file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(sim, voxel)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
