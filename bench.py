#!/usr/bin/env python3
"""Throughput benchmark of the photon-packet transport engine on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One step is one simulation run over the workload's packet budget with its
inputs already resident on the device: for the layered batch a replay of
the captured batch launch (xo_batch_plan_run) and the accumulator download,
for the single simulations the C-ABI call behind Mc.run (kernel launch,
accumulator download) with the voxel grid kept on the device. Prints one
JSON line.

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
# profiles/mcvox_kernel_skin_vessel_200.md,
# profiles/mcml_kernel_slab_1mm_and_surrogate.md). Offline figures: measured
# for the kernels as they are in this tree, hardware arithmetic.
TRAFFIC_BYTES_PER_LAUNCH = {'layered_batch_4096': 2.585e11,
                            'skin_vessel_200': 7.52e11,
                            'slab_1mm': 5.3e6}

METRIC = ('photon packets/s (1 GPU) batched vs plain; p50/p99 solve latency '
          'at fixed QPS')

# The production arithmetic: the hardware approximations the reference's own
# benchmark asks for with -cl-fast-relaxed-math (test/performance.py:104);
# --exact-math times the build that matches the oracle bit for bit.
FAST_MATH = ['-cl-fast-relaxed-math', '-cl-mad-enable']
BUILD_OPTS = FAST_MATH


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
        batch = Batch(insts, rnginit=0x5EED000000000001,
                      cl_build_options=BUILD_OPTS)
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
        # inputs resident on the device, the launch a captured graph: each
        # step replays it and downloads the accumulators
        plan = ctx.batch_plan(arr, lut, opts, naccum)

        def step(first):
            return plan.run()[1]

        def upload_inclusive():
            """One xo_batch_run: upload of the packed instances and LUTs,
            launch, download."""
            t0 = time.perf_counter()
            ctx.run_batch(arr, lut, opts, naccum)
            return npk / (time.perf_counter() - t0)

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
        batch.bench_upload_inclusive = upload_inclusive
        return batch, npk, step, False, config
    if name == 'skin_vessel_200':
        sim = workloads.skin_vessel(cl_build_options=BUILD_OPTS)
        npk = packets or workloads.SKIN_VESSEL_PACKETS
        cfg, lut, naccum = sim.pack()
        ctx = sim.context
        vox = sim.voxels.material
        flags = sim._run_flags(False)     # the simulator's arithmetic

        def step(first):
            return ctx.run_mcvox(cfg, vox if first else None, lut, npk,
                                 naccum, 0, flags)[1]

        config = {'workload': 'skin_vessel_200',
                  'voxels': '200x200x200 x 10 um', 'packets': npk,
                  'source': 'gaussian sigma 50 um',
                  'accumulator': 'uint64 deposition per voxel'}
        return sim, npk, step, True, config
    if name == 'slab_1mm':
        sim = workloads.slab_1mm(cl_build_options=BUILD_OPTS)
        npk = packets or workloads.SLAB_PACKETS
        cfg, lut, naccum = sim.pack()
        ctx = sim.context
        flags = sim._run_flags(False)     # the simulator's arithmetic

        def step(first):
            return ctx.run_mcml(cfg, lut, npk, naccum, 0, flags)[1]

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
    # the same atomics into 256 KiB, and into one 64 MiB copy per XCD with
    # workgroup-scope atomics: the rate does not depend on either
    hms, htot = ctx.probe_atomic_u64_xcd(64 << 20, 256 << 10, natom, False,
                                         seed=7)
    xms, xtot = ctx.probe_atomic_u64_xcd(64 << 20, 64 << 20, natom, True,
                                         seed=7)
    if htot != natom or xtot != natom:
        raise RuntimeError('atomic probes lost updates')
    return {'valu_fma_lane_ops_per_s': vops / (vms * 1e-3),
            'copy_GBps': 2 * nbytes / (ms * 1e-3) / 1e9,
            'atomic_u64_per_s': natom / (ams * 1e-3),
            'atomic_u64_GBps': 8 * natom / (ams * 1e-3) / 1e9,
            'atomic_u64_per_s_256KiB': natom / (hms * 1e-3),
            'atomic_u64_per_s_per_xcd_copies': natom / (xms * 1e-3),
            'probe': 'fp32 FMA chains 8 x 20000 per lane; copy 1 GiB x10; '
                     '2^28 u64 atomics, 64 MiB array (also: targets in '
                     '256 KiB; one 64 MiB array per XCD, workgroup scope)'}


def serving(insts, rates, duration: float, seed: int = 20261203,
            voxel_fraction: float = 0.1, saturation_requests: int = 400,
            tick_packets: int = None):
    """Poisson arrivals at each fixed rate of `rates` for `duration` seconds:
    90 % of the requests are instances of the layered batch (in a seeded
    random order), 10 % are 64^3 two-layer voxel simulations of 1e6
    packets. Requests are packed before the run; the clock runs from the
    first submit to the last result. Latency is submit to result. Before
    them, a burst of requests submitted at once measures the capacity of
    the same mix. The top-level figures are those of the first rate; every
    rate is listed under 'sweep'."""
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

    opts = {'fast_math': 0 if BUILD_OPTS is None else 1}
    if tick_packets is not None:
        opts['tick_packets'] = tick_packets
    sweep = []
    with serve.Server(**opts) as srv:
        # first use of the server (graph, tables) outside the clock
        warm = [make(srv, k) for k in range(4)]
        srv.drain(120)
        # capacity: a burst of requests submitted at once
        t0 = time.perf_counter()
        burst = [make(srv, (1 << 19) + k) for k in range(saturation_requests)]
        srv.drain(600)
        t_burst = time.perf_counter() - t0
        for j, qps in enumerate(rates):
            s0 = srv.stats()
            reqs, span = serve.poisson_load(
                srv, lambda sv, k, j=j: make(sv, (j << 16) + k + 4), qps,
                duration, seed=seed + j)
            s1 = srv.stats()
            packets = sum(r.npackets for r in reqs)
            point = {'qps': qps, 'arrivals_s': duration,
                     'requests': len(reqs),
                     'voxel_requests': sum(r.kind == 1 for r in reqs),
                     'packets': packets, 'packets_per_s': packets / span,
                     'steps_per_s': (s1['steps'] - s0['steps']) / span,
                     'ticks': s1['ticks'] - s0['ticks']}
            point.update(serve.latency_summary(reqs))
            for kind, name in ((0, 'layered'), (1, 'voxel')):
                lat = serve.latency_summary([r for r in reqs
                                             if r.kind == kind])
                point[name + '_p50_ms'] = lat['p50_ms']
                point[name + '_p99_ms'] = lat['p99_ms']
            print('bench: serving {} requests/s: p50 {:.1f} ms p99 {:.1f} '
                  'ms'.format(qps, point['p50_ms'], point['p99_ms']),
                  file=sys.stderr, flush=True)
            sweep.append(point)
    del warm
    burst_packets = sum(r.npackets for r in burst)
    out = {'capacity': {'requests': len(burst),
                        'requests_per_s': len(burst) / t_burst,
                        'packets_per_s': burst_packets / t_burst},
           'mix': '90% layered-batch instances, 10% 64^3 two-layer voxel '
                  'requests of 1e6 packets'}
    out.update(sweep[0])
    out['sweep'] = sweep
    return out


def surrogate_rates(ctx, rows=(65536, 1048576), batch: int = 4096,
                    steps: int = 1024):
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
    sim = W.fiber_sim(cl_build_options=BUILD_OPTS)
    for m, s in zip(mua, musr):
        sim.layers[1].mua = m
        sim.layers[1].mus = s / (1.0 - 0.85)
        sim.run(npk)
    t_plain = time.perf_counter() - t0
    # batched
    t0 = time.perf_counter()
    b = Batch(W.fiber_batch(packets=npk), rnginit=0x5EED000000000001,
              cl_build_options=BUILD_OPTS)
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
    sims = [mc.Mc(i.layers, i.source, i.detectors, rnginit=0x5EED000000000001,
                  cl_build_options=BUILD_OPTS)
            for i in sub]
    for s in sims:             # one context, created outside the timing
        s._ctx = batch.context
    packed = [s.pack() for s in sims]
    ctx = batch.context
    npk = sum(int(i.nphotons) for i in sub)
    t0 = time.perf_counter()
    flags = sims[0]._run_flags(False)   # the batch's arithmetic
    for s, (cfg, lut, naccum), i in zip(sims, packed, sub):
        ctx.run_mcml(cfg, lut, i.nphotons, naccum, 0, flags)
    dt = time.perf_counter() - t0
    return {'packets_per_s': npk / dt, 'instances': len(sub),
            'packets': npk, 'seconds': dt}


_NATIVE_ORACLE = None


def native_oracle():
    """Load the oracle built for this host's cores (gcc -O3 -march=native
    -fopenmp, oracle/Makefile `native`) for the CPU baselines; 'port' is the
    prebuilt portable build if that compile is not possible here."""
    global _NATIVE_ORACLE
    from oracle import pyoracle as O
    if _NATIVE_ORACLE is None:
        import tempfile
        out = os.path.join(tempfile.mkdtemp(prefix='xo_oracle_'),
                           'liboracle_native.so')
        try:
            O.build_native(out)
            O.use_library(out)
            _NATIVE_ORACLE = 'gcc -O3 -march=native -fopenmp'
        except Exception as e:        # no compiler: the prebuilt oracle
            print('bench: native oracle build failed ({}); using the '
                  'prebuilt one'.format(e), file=sys.stderr, flush=True)
            _NATIVE_ORACLE = 'prebuilt (gcc -O2 -march=x86-64-v3 -fopenmp)'
    return O, _NATIVE_ORACLE


def cpu_baseline(sim, voxel: bool, npackets: int, budget_s: float = 30.0):
    """The oracle (C, OpenMP on every host core, C-library math) on the
    same workload: npackets packets, in chunks, stopping early once
    budget_s has passed. Returns the measured rate."""
    from pyxopto_amd import _native as N
    O, build = native_oracle()
    cores = len(os.sched_getaffinity(0))
    if hasattr(sim, 'instances'):
        return cpu_baseline_batch(sim, cores, budget_s, build)
    cfg, lut, naccum = sim.pack()
    lanes = 4 * cores * 64
    x, a = N.init_rng(lanes, 0x0123456789ABCDEF)
    chunk = min(npackets, 50000 if voxel else 250000)
    done, steps, t0 = 0, 0, time.perf_counter()
    while done < npackets and time.perf_counter() - t0 < budget_s:
        if voxel:
            _, counts, x = O.mcvox_run(cfg, sim.voxels.material, lut, x, a,
                                       chunk, naccum, O.MATH_LIBM, cores)
        else:
            _, counts, x = O.mcml_run(cfg, lut, x, a, chunk, naccum,
                                      O.MATH_LIBM, cores)
        done += chunk
        steps += int(counts[0])
    dt = time.perf_counter() - t0
    return {'value': done / dt, 'unit': 'photon packets/s', 'cores': cores,
            'kind': 'port', 'build': build,
            'transport_steps_per_s': steps / dt,
            'sample': '{} packets of the same workload in {:.1f} s'.format(
                done, dt)}


def cpu_baseline_batch(batch, cores: int, budget_s: float, build: str):
    """The oracle on the seeded 256-instance subset of the batch (every
    16th instance, full budgets, 5.3e7 packets), 16 instances at a time,
    stopping early once budget_s has passed."""
    from pyxopto_amd.mcml.batch import Batch
    O, _ = native_oracle()
    sub = Batch(batch.instances[::16], rnginit=0x5EED000000000001)
    arr, lut, naccum = sub.pack()
    n = len(arr)
    done, steps, ninst, t0 = 0, 0, 0, time.perf_counter()
    for first in range(0, n, 16):
        if time.perf_counter() - t0 >= budget_s:
            break
        part = (type(arr[0]) * min(16, n - first))()
        for j in range(len(part)):
            part[j] = arr[first + j]
        _, counts = O.batch_run(part, lut, naccum, math_mode=O.MATH_LIBM,
                                nthreads=cores)
        done += sum(int(part[j].npackets) for j in range(len(part)))
        steps += int(counts[0])
        ninst += len(part)
    dt = time.perf_counter() - t0
    return {'value': done / dt, 'unit': 'photon packets/s', 'cores': cores,
            'kind': 'port', 'build': build,
            'transport_steps_per_s': steps / dt,
            'sample': '{} instances of the 256-instance subset (every 16th, '
                      'full budgets): {} packets in {:.1f} s'.format(
                          ninst, done, dt)}


def timed(step, steps: int, warmup: int):
    """Warm-up steps, then `steps` timed steps: (reports, wall seconds)."""
    first = True
    for _ in range(warmup):
        step(first)
        first = False
    reps = []
    t0 = time.perf_counter()
    for _ in range(steps):
        reps.append(step(first))
        first = False
    return reps, time.perf_counter() - t0


def roofline(reps, voxel: bool, workload: str, ceil=None):
    kernel_s = sum(r.kernel_ms for r in reps) * 1e-3
    achieved = sum(algorithmic_slots(r, voxel) for r in reps) / kernel_s / 1e12
    out = {'bound': 'valu', 'achieved': achieved, 'peak': PEAK_SLOTS / 1e12,
           'unit': 'T lane-op slots/s', 'frac': achieved / (PEAK_SLOTS / 1e12),
           'traffic': TRAFFIC_BYTES_PER_LAUNCH.get(workload)}
    if voxel and ceil:
        dep = sum(r.deposits for r in reps)
        out['atomic_deposits_per_s'] = dep / kernel_s
        out['atomic_GBps'] = dep * 8 / kernel_s / 1e9
        out['atomic_frac_of_probe'] = dep / kernel_s / ceil['atomic_u64_per_s']
    return out


def config_leg(name: str, steps: int, warmup: int, ceil, cpu: bool,
               cpu_packets: int):
    """One single-simulation configuration timed on its own: packets/s,
    steps/s, its roofline and its CPU baseline."""
    sim, npk, step, voxel, config = make_workload(name, None)
    print('bench: leg {} packets/step {}'.format(name, npk), file=sys.stderr,
          flush=True)
    reps, wall = timed(step, steps, warmup)
    kernel_s = sum(r.kernel_ms for r in reps) * 1e-3
    out = {'config': config, 'steps': steps, 'warmup': warmup,
           'packets_per_s': npk * steps / wall,
           'kernel_packets_per_s': npk * steps / kernel_s,
           'ms_per_step': wall / steps * 1e3,
           'kernel_ms_per_step': kernel_s / steps * 1e3,
           'transport_steps_per_s': sum(r.steps for r in reps) / wall,
           'lanes': int(reps[-1].lanes),
           'roofline': roofline(reps, voxel, name, ceil)}
    if voxel:
        out['voxel_crossings_per_s'] = sum(r.voxel_crossings
                                           for r in reps) / kernel_s
    if cpu:
        out['cpu_baseline'] = cpu_baseline(sim, voxel, cpu_packets)
        if voxel:
            out['cpu_baseline']['extrapolated_s_per_step'] = \
                npk / out['cpu_baseline']['value']
    sim.context.close()
    return out


def main():
    global BUILD_OPTS
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='layered_batch_4096',
                    choices=['layered_batch_4096', 'skin_vessel_200',
                             'slab_1mm'])
    ap.add_argument('--packets', type=int, default=None)
    ap.add_argument('--exact-math', action='store_true',
                    help='the exact (bit-exact with the oracle) arithmetic '
                         'instead of the hardware approximations')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-legs', action='store_true',
                    help='skip the 1 mm slab and 200^3 voxel configurations')
    ap.add_argument('--no-plain', action='store_true',
                    help='skip the one-launch-per-instance comparison')
    ap.add_argument('--no-serving', action='store_true',
                    help='skip the fixed-QPS serving run')
    ap.add_argument('--no-published', action='store_true',
                    help='skip the reference\'s published fiber benchmark')
    ap.add_argument('--no-surrogate', action='store_true',
                    help='skip the surrogate network throughput')
    ap.add_argument('--qps', type=float, default=SERVING_QPS)
    ap.add_argument('--qps-sweep', default='400,550,800,1000',
                    help='further request rates of the serving run '
                         '(comma-separated, empty for none)')
    ap.add_argument('--tick-packets', type=int, default=None,
                    help='layered packets per serving tick')
    ap.add_argument('--serving-seconds', type=float, default=8.0)
    ap.add_argument('--lut-processes', type=int, default=None,
                    help='LUT builder processes')
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
    if args.exact_math:
        BUILD_OPTS = None

    sim, npk, step, voxel, config = make_workload(
        args.workload, args.packets, args.launch, args.lut_processes)
    config['arithmetic'] = 'exact' if args.exact_math else \
        'hardware approximations (-cl-fast-relaxed-math)'
    print('bench: workload {} packets/step {} warmup {} steps {}'.format(
        config['workload'], npk, args.warmup, args.steps), file=sys.stderr,
        flush=True)
    reps, wall = timed(step, args.steps, args.warmup)
    print('bench: timed {} steps in {:.3f} s'.format(args.steps, wall),
          file=sys.stderr, flush=True)

    kernel_s = sum(r.kernel_ms for r in reps) * 1e-3
    steps_total = sum(r.steps for r in reps)
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
    }
    ceil = ceilings(sim.context)
    out['roofline'] = roofline(reps, voxel, config['workload'], ceil)
    out['roofline']['hbm_copy_GBps'] = ceil['copy_GBps']
    out['measured_ceilings'] = ceil
    if voxel:
        for key in ('atomic_deposits_per_s', 'atomic_GBps',
                    'atomic_frac_of_probe'):
            out[key] = out['roofline'][key]
        out['voxel_crossings_per_s'] = sum(r.voxel_crossings
                                           for r in reps) / kernel_s
    if hasattr(sim, 'bench_upload_inclusive'):
        out['upload_inclusive_packets_per_s'] = sim.bench_upload_inclusive()
    if hasattr(sim, 'bench_plain') and not args.no_plain:
        plain = sim.bench_plain()
        out['plain_one_launch_per_instance'] = plain
        out['batched_vs_plain'] = value / plain['packets_per_s']
    if hasattr(sim, 'bench_insts') and not args.no_legs:
        out['configs'] = {
            'slab_1mm': config_leg('slab_1mm', 20, 3, ceil,
                                   not args.no_cpu_baseline, 1_000_000),
            'skin_vessel_200': config_leg('skin_vessel_200', 3, 1, ceil,
                                          not args.no_cpu_baseline,
                                          1_000_000)}
    if hasattr(sim, 'bench_insts') and not args.no_published:
        print('bench: published fiber benchmark', file=sys.stderr, flush=True)
        out['published_benchmark'] = published_benchmark()
    if not args.no_surrogate:
        out['surrogate'] = surrogate_rates(sim.context)
    if hasattr(sim, 'bench_insts') and not args.no_serving:
        rates = [args.qps] + [float(v) for v in args.qps_sweep.split(',')
                              if v.strip()]
        print('bench: serving at {} requests/s for {} s each'.format(
            rates, args.serving_seconds), file=sys.stderr, flush=True)
        out['serving'] = serving(sim.bench_insts, rates,
                                 args.serving_seconds,
                                 tick_packets=args.tick_packets)
        print('bench: serving done', file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print('bench: cpu baseline', file=sys.stderr, flush=True)
        out['cpu_baseline'] = cpu_baseline(sim, voxel, npk)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
