"""Launch-shape autotuner of the batched engine.

Tunables: packets per work-queue item (chunk), threads per block and blocks
per CU (persistent grid). Budgets are never scaled down: a wave keeps the
instance of its item for a whole chunk, so the coherence a chunk buys
depends on how it compares with an instance's budget. Instead, each
candidate runs two calibration batches of instances drawn evenly from the
batch with their full budgets (about 1e8 and 2e8 packets); the time of a
launch is modelled as packets / rate + tail, where the tail (waves finishing
their last items) depends on the chunk but not on the batch size. The two
runs give rate and tail, the model predicts the whole batch, and the
candidate with the shortest predicted time wins. A plain "fastest
calibration run" would pick chunks far too small: the tail weighs much
more in a short run. The model can still misjudge the tail of a large
chunk, so the two best predictions of different chunks are then timed on
the whole batch (when it holds at most VERIFY_PACKETS packets) and the
faster one is kept. Results are cached per device (CU count) and batch
signature in pyxopto_amd/data/autotune.json so later runs skip the sweep.
"""
import hashlib
import json
import os

from .. import _native as N

CANDIDATES = [(chunk, block, bpc)
              for chunk in (2048, 8192, 32768)
              for block in (64, 256)
              for bpc in (0,)]
CACHE = os.path.join(N.DATA_DIR, 'autotune.json')
CALIBRATION_PACKETS = 200_000_000
VERIFY_PACKETS = 2_000_000_000


def _signature(arr) -> str:
    n = len(arr)
    kinds = sorted({int(arr[i].pf_kind) for i in range(n)})
    layers = sorted({int(arr[i].num_layers) for i in range(n)})
    total = sum(int(arr[i].npackets) for i in range(n))
    key = 'n{}-k{}-l{}-p{}'.format(n, kinds, layers, total // 1_000_000)
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache):
    os.makedirs(N.DATA_DIR, exist_ok=True)
    tmp = CACHE + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def calibration_batch(arr, packets: int = CALIBRATION_PACKETS):
    """Copies of evenly spaced instances, full budgets, until about
    `packets` packets (the whole batch if it is smaller)."""
    n = len(arr)
    total = sum(int(arr[i].npackets) for i in range(n))
    frac = min(1.0, packets / max(total, 1))
    step = max(1, int(round(1.0 / frac)))
    picks = list(range(0, n, step))
    cal = (N.XoBatchInstance * len(picks))()
    for j, i in enumerate(picks):
        cal[j] = arr[i]
    return cal


def _packets(arr) -> int:
    return sum(int(arr[i].npackets) for i in range(len(arr)))


def sweep(ctx, arr, lut, naccum, base_opts, candidates=CANDIDATES):
    """[(predicted_ms, chunk, block, bpc, rate_packets_per_ms, tail_ms)] for
    every candidate, best first."""
    big = calibration_batch(arr)
    small = (N.XoBatchInstance * ((len(big) + 1) // 2))()
    for j in range(len(small)):
        small[j] = big[2 * j]
    w1, w2, wf = _packets(small), _packets(big), _packets(arr)
    res = []
    for chunk, block, bpc in candidates:
        o = N.XoBatchOpts()
        o.chunk, o.block, o.blocks_per_cu = chunk, block, bpc
        o.w_min, o.lottery_chance = base_opts.w_min, base_opts.lottery_chance
        o.use_lottery = base_opts.use_lottery
        o.method = base_opts.method
        o.flags = base_opts.flags
        _, r1 = ctx.run_batch(small, lut, o, naccum)
        _, r2 = ctx.run_batch(big, lut, o, naccum)
        t1, t2 = r1.kernel_ms, r2.kernel_ms
        if w2 > w1 and t2 > t1:
            rate = (w2 - w1) / (t2 - t1)
            tail = max(0.0, t1 - w1 / rate)
            predicted = wf / rate + tail
        else:                       # batch too small to separate the two
            rate, tail, predicted = w2 / max(t2, 1e-9), 0.0, t2
        res.append((predicted, chunk, block, bpc, rate, tail))
    res.sort()
    return res


def tuned_opts(batch, packed, force: bool = False):
    arr, lut, naccum = packed
    ctx = batch.context
    base = batch.opts()
    key = '{}cu-{}{}'.format(ctx.compute_units, _signature(arr),
                             '-fast' if base.flags & N.XO_FLAG_FAST_MATH else '')
    cache = _load_cache()
    if not force and key in cache:
        chunk, block, bpc = cache[key]['best']
    else:
        res = sweep(ctx, arr, lut, naccum, base)
        _, chunk, block, bpc = res[0][:4]
        entry = {'sweep': [list(r) for r in res]}
        if _packets(arr) <= VERIFY_PACKETS:
            finalists = [res[0]] + [r for r in res[1:] if r[1] != res[0][1]][:1]
            timed = []
            for r in finalists:
                o = batch.opts(*r[1:4])
                _, rep = ctx.run_batch(arr, lut, o, naccum)
                timed.append((rep.kernel_ms, r[1], r[2], r[3]))
            timed.sort()
            _, chunk, block, bpc = timed[0]
            entry['verified'] = [list(t) for t in timed]
        entry['best'] = [chunk, block, bpc]
        cache[key] = entry
        _store_cache(cache)
    return batch.opts(chunk, block, bpc)
