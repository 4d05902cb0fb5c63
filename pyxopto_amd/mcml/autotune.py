"""Launch-shape autotuner of the batched engine.

Tunables: packets per work-queue item (chunk), threads per block and blocks
per CU (persistent grid). A calibration batch — up to 256 instances drawn
evenly from the batch, each with its budget scaled so the whole calibration
is about 2e7 packets — is run once per candidate; the fastest by device time
wins. Results are cached per device (CU count) and batch signature in
pyxopto_amd/data/autotune.json so later runs skip the sweep."""
import hashlib
import json
import os

from .. import _native as N

CANDIDATES = [(chunk, block, bpc)
              for chunk in (512, 2048, 8192)
              for block in (64, 256)
              for bpc in (0, 4)]
CACHE = os.path.join(N.DATA_DIR, 'autotune.json')
CALIBRATION_PACKETS = 20_000_000


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


def calibration_batch(arr):
    """Copies of up to 256 evenly spaced instances with scaled budgets."""
    n = len(arr)
    step = max(1, n // 256)
    picks = list(range(0, n, step))[:256]
    total = sum(int(arr[i].npackets) for i in picks)
    scale = min(1.0, CALIBRATION_PACKETS / max(total, 1))
    cal = (N.XoBatchInstance * len(picks))()
    for j, i in enumerate(picks):
        cal[j] = arr[i]
        cal[j].npackets = max(64, int(arr[i].npackets * scale))
    return cal


def sweep(ctx, arr, lut, naccum, base_opts, candidates=CANDIDATES):
    """[(kernel_ms, chunk, block, bpc)] for every candidate, fastest first."""
    cal = calibration_batch(arr)
    res = []
    for chunk, block, bpc in candidates:
        o = N.XoBatchOpts()
        o.chunk, o.block, o.blocks_per_cu = chunk, block, bpc
        o.w_min, o.lottery_chance = base_opts.w_min, base_opts.lottery_chance
        o.use_lottery = base_opts.use_lottery
        _, rep = ctx.run_batch(cal, lut, o, naccum)
        res.append((rep.kernel_ms, chunk, block, bpc))
    res.sort()
    return res


def tuned_opts(batch, packed, force: bool = False):
    arr, lut, naccum = packed
    ctx = batch.context
    base = batch.opts()
    key = '{}cu-{}'.format(ctx.compute_units, _signature(arr))
    cache = _load_cache()
    if not force and key in cache:
        chunk, block, bpc = cache[key]['best']
    else:
        res = sweep(ctx, arr, lut, naccum, base)
        _, chunk, block, bpc = res[0]
        cache[key] = {'best': [chunk, block, bpc],
                      'sweep_ms': [[r[0], r[1], r[2], r[3]] for r in res]}
        _store_cache(cache)
    return batch.opts(chunk, block, bpc)
