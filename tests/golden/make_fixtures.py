"""Write the golden fixtures under tests/golden/ from the reference tree.

Run in the development container only (the reference tree is not shipped):
    python tests/golden/make_fixtures.py [/root/reference]

Outputs (data only; no reference source is copied):
  cuda_singlelayer_linesource_radial.npz
      The CUDAMCML-derived radial reflectance profile and its simulation
      parameters, read from xopto/mcml/test/reference/
      cuda_singlelayer_linesource_radial.pkl. The pickle is decoded by a
      static scan of its opcode stream (pickletools.genops): no object is
      constructed from the file and no global it names is resolved. Only
      the primitive values (float, int, str, bytes) and containers are
      rebuilt; the ndarray payload is read from its raw bytes with the
      dtype string ('f8', little endian) and shape recorded next to it.
  cuda_singlelayer_uniformfiber.npz, cuda_doublelayer_uniformfiber.npz
      The fiber-collected reflectance of the reference's uniform-fiber
      tests (20 x 20 grid of mua, mus' of the varied layer, 5 source-
      detector separations) and the parameters of those runs, decoded the
      same way from the .pkl files of the same folder.
  seeds.npz
      Rows 0..4095 and a SHA-256 of the multiplier column of
      xopto/data/primes/safeprimes_base32_500k.npz (np.load without
      pickles), plus init_RNG outputs (x, a for 4096 streams) for three
      initializers, computed by the reference's own rng.cpp built into
      oracle/_ref/ (the oracle Makefile), and the rand_cc float sequences
      of that library for one stream.
"""
import ctypes
import hashlib
import os
import pickletools
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _static_pickle_values(path):
    """Rebuild dicts/lists/tuples of primitives and ndarray payloads from a
    pickle's opcode stream without executing it."""
    data = open(path, 'rb').read()
    stack, memo, marks = [], {}, []

    class Global:
        def __init__(self, mod, name):
            self.name = '{}.{}'.format(mod, name)

    class Reduced:
        def __init__(self, func, args):
            self.func, self.args, self.state = func, args, None

    for op, arg, pos in pickletools.genops(data):
        name = op.name
        if name in ('PROTO', 'FRAME', 'STOP'):
            continue
        if name in ('EMPTY_DICT',):
            stack.append({})
        elif name == 'EMPTY_LIST':
            stack.append([])
        elif name == 'EMPTY_TUPLE':
            stack.append(())
        elif name == 'MARK':
            marks.append(len(stack))
        elif name == 'MEMOIZE':
            memo[len(memo)] = stack[-1]
        elif name in ('BINGET', 'LONG_BINGET'):
            stack.append(memo[arg])
        elif name in ('SHORT_BINUNICODE', 'BINUNICODE', 'BINFLOAT',
                      'BININT', 'BININT1', 'BININT2', 'SHORT_BINBYTES',
                      'BINBYTES', 'BINBYTES8'):
            stack.append(arg)
        elif name == 'NONE':
            stack.append(None)
        elif name in ('NEWTRUE', 'NEWFALSE'):
            stack.append(name == 'NEWTRUE')
        elif name == 'SETITEM':
            v = stack.pop(); k = stack.pop(); stack[-1][k] = v
        elif name == 'SETITEMS':
            m = marks.pop(); items = stack[m:]; del stack[m:]
            d = stack[-1]
            for i in range(0, len(items), 2):
                d[items[i]] = items[i + 1]
        elif name == 'APPEND':
            v = stack.pop(); stack[-1].append(v)
        elif name == 'APPENDS':
            m = marks.pop(); items = stack[m:]; del stack[m:]
            stack[-1].extend(items)
        elif name == 'TUPLE':
            m = marks.pop(); items = tuple(stack[m:]); del stack[m:]
            stack.append(items)
        elif name in ('TUPLE1', 'TUPLE2', 'TUPLE3'):
            k = int(name[-1]); items = tuple(stack[-k:]); del stack[-k:]
            stack.append(items)
        elif name == 'STACK_GLOBAL':
            nm = stack.pop(); mod = stack.pop(); stack.append(Global(mod, nm))
        elif name == 'REDUCE':
            args = stack.pop(); func = stack.pop()
            stack.append(Reduced(func, args))
        elif name == 'BUILD':
            state = stack.pop(); stack[-1].state = state
        else:
            raise ValueError('unsupported opcode {} at {}'.format(name, pos))
    root = stack[-1]

    def convert(v):
        if isinstance(v, dict):
            return {k: convert(x) for k, x in v.items()}
        if isinstance(v, list):
            return [convert(x) for x in v]
        if isinstance(v, tuple):
            return tuple(convert(x) for x in v)
        if isinstance(v, Reduced) and v.func.name.endswith('_reconstruct'):
            # state = (version, shape, dtype-Reduced, fortran, raw bytes)
            version, shape, dt, fortran, raw = v.state
            dtype_str = dt.args[0]
            endian = dt.state[1]
            arr = np.frombuffer(raw, dtype=np.dtype(endian + dtype_str))
            return arr.reshape(shape, order='F' if fortran else 'C').copy()
        return v
    return convert(root)


def make_fixture_radial(ref_root):
    src = os.path.join(ref_root, 'xopto', 'mcml', 'test', 'reference',
                       'cuda_singlelayer_linesource_radial.pkl')
    d = _static_pickle_values(src)
    layers = d['layers']
    out = {
        'reflectance': np.asarray(d['reflectance'], dtype=np.float64),
        'layer_n': np.array([l['basic']['n'] for l in layers]),
        'layer_d': np.array([l['basic']['d'] for l in layers]),
        'layer_mua': np.array([l['basic']['mua'] for l in layers]),
        'layer_mus': np.array([l['basic']['mus'] for l in layers]),
        'layer_g': np.array([l['pf']['pfparams'][0] for l in layers]),
        'axis_start': d['detector']['axis']['start'],
        'axis_stop': d['detector']['axis']['stop'],
        'axis_n': d['detector']['axis']['n'],
        'cosmin': d['detector']['cosmin'],
        'rmax': d['mc']['rmax'],
    }
    np.savez(os.path.join(HERE, 'cuda_singlelayer_linesource_radial.npz'),
             **out)
    return out


def make_fixture_fiber(ref_root, name):
    """The CUDAMCML-derived fiber reflectance of the reference's uniform
    fiber tests (xopto/mcml/test/validate.py:348-511 single layer,
    :692-850 double layer): reflectance[mua][musr][sds] and the parameters
    of the run, decoded statically like the radial fixture."""
    src = os.path.join(ref_root, 'xopto', 'mcml', 'test', 'reference',
                       name + '.pkl')
    d = _static_pickle_values(src)
    layers = d['layers']
    out = {
        'reflectance': np.asarray(d['reflectance'], dtype=np.float64),
        'layer_n': np.array([l['basic']['n'] for l in layers]),
        'layer_d': np.array([l['basic']['d'] for l in layers]),
        'layer_g': np.array([l['pf']['pfparams'][0] for l in layers]),
        'source_na': d['source']['na'],
        'source_dcore': d['source']['dcore'],
        'source_dcladding': d['source']['dcladding'],
        'source_ncore': d['source']['ncore'],
        'axis_start': d['detector']['axis']['start'],
        'axis_stop': d['detector']['axis']['stop'],
        'axis_n': d['detector']['axis']['n'],
        'cosmin': d['detector']['cosmin'],
        'sds': np.array(d['detector']['fibers']['sds'], dtype=np.float64),
        'fiber_dcore': d['detector']['fibers']['dcore'],
        'rmax': d['mc']['rmax'],
    }
    lut = d['lut']
    if isinstance(lut, list):           # double layer: fixed top layer
        out['layer1_mua'] = float(np.asarray(lut[0]['mua']).ravel()[0])
        out['layer1_musr'] = float(np.asarray(lut[0]['musr']).ravel()[0])
        lut = lut[1]
    out['lut_mua'] = np.asarray(lut['mua'], dtype=np.float64)
    out['lut_musr'] = np.asarray(lut['musr'], dtype=np.float64)
    np.savez(os.path.join(HERE, name + '.npz'), **out)
    return out


def make_seeds(ref_root):
    table = np.load(os.path.join(ref_root, 'xopto', 'data', 'primes',
                                 'safeprimes_base32_500k.npz'))['data']
    fora = np.ascontiguousarray(table[:, 0].astype(np.uint32))
    sha = hashlib.sha256(fora.astype('<u4').tobytes()).hexdigest()
    lib_path = os.path.join(REPO, 'oracle', '_ref', 'librng_ref.so')
    if not os.path.isfile(lib_path):
        raise RuntimeError('build oracle/_ref first: make -C oracle ref')
    lib = ctypes.CDLL(lib_path)
    P = ctypes.POINTER
    lib.init_RNG.argtypes = [P(ctypes.c_uint64), P(ctypes.c_uint32),
                             P(ctypes.c_uint32), ctypes.c_uint32,
                             ctypes.c_uint64]
    lib.rand_cc.argtypes = [P(ctypes.c_float), ctypes.c_uint32,
                            P(ctypes.c_uint64), ctypes.c_uint32]
    n = 4096
    xinits = np.array([0x5EED000000000001, 123456789123, 0xABCDEF0123456789],
                      dtype=np.uint64)
    xs, as_ = [], []
    for xinit in xinits:
        x = np.zeros(n, np.uint64)
        a = np.zeros(n, np.uint32)
        rc = lib.init_RNG(x.ctypes.data_as(P(ctypes.c_uint64)),
                          a.ctypes.data_as(P(ctypes.c_uint32)),
                          fora.ctypes.data_as(P(ctypes.c_uint32)),
                          ctypes.c_uint32(n), ctypes.c_uint64(int(xinit)))
        assert rc == 0
        xs.append(x)
        as_.append(a)
    bad = np.uint64(0xFFFFFFF000000001)
    x = np.zeros(1, np.uint64); a = np.zeros(1, np.uint32)
    rc_bad = lib.init_RNG(x.ctypes.data_as(P(ctypes.c_uint64)),
                          a.ctypes.data_as(P(ctypes.c_uint32)),
                          fora.ctypes.data_as(P(ctypes.c_uint32)),
                          ctypes.c_uint32(1), ctypes.c_uint64(int(bad)))
    # rand_cc: ((uint32)x & 0x7FFFFF)/0x7FFFFF after each MWC step
    xs0 = ctypes.c_uint64(int(xs[0][0]))
    cc = np.zeros(1000, np.float32)
    lib.rand_cc(cc.ctypes.data_as(P(ctypes.c_float)), ctypes.c_uint32(1000),
                ctypes.byref(xs0), ctypes.c_uint32(int(as_[0][0])))
    np.savez(os.path.join(HERE, 'seeds.npz'), table_head=table[:n],
             table_rows=np.int64(table.shape[0]), fora_sha256=sha,
             xinits=xinits, x=np.stack(xs), a=np.stack(as_),
             bad_xinit=bad, bad_rc=np.int64(rc_bad), rand_cc=cc,
             rand_cc_x_after=np.uint64(xs0.value))


if __name__ == '__main__':
    ref = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
    fx = make_fixture_radial(ref)
    print('radial fixture:', fx['reflectance'].shape, fx['layer_mua'],
          fx['layer_mus'], fx['rmax'])
    for name in ('cuda_singlelayer_uniformfiber', 'cuda_doublelayer_uniformfiber'):
        fx = make_fixture_fiber(ref, name)
        print(name, fx['reflectance'].shape, fx['sds'], fx['layer_n'])
    make_seeds(ref)
    print('seeds written')
