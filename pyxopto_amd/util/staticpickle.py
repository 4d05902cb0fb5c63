"""Decode the object entries of .npz result files without running code.

numpy stores a dict, a None or any other Python object saved with
``np.savez_compressed`` as an object array whose payload is a pickle stream;
``np.load(..., allow_pickle=True)`` would execute whatever that stream
names. This decoder interprets the stream's opcodes itself
(``pickletools.genops``) and rebuilds only plain values: None, bool, int,
float, str, bytes, tuples, lists, dicts, sets, numpy scalars and numpy
arrays (from their raw bytes and recorded dtype/shape). Any other global a
stream names raises ValueError; nothing it names is imported or called.
"""
import pickletools

import numpy as np

_NUMPY_MODULES = ('numpy.core.multiarray', 'numpy._core.multiarray',
                  'numpy.core.numeric', 'numpy._core.numeric')


class _Global:
    __slots__ = ('module', 'name')

    def __init__(self, module, name):
        self.module, self.name = module, name

    def key(self):
        if self.module in _NUMPY_MODULES:
            return 'multiarray.' + self.name
        return '{}.{}'.format(self.module, self.name)


class _Dtype:
    """numpy.dtype(str, align, copy) with its BUILD state."""
    __slots__ = ('descr', 'state')

    def __init__(self, descr):
        self.descr, self.state = descr, None

    def dtype(self) -> np.dtype:
        descr = self.descr
        endian = '|'
        if self.state is not None and len(self.state) > 1:
            endian = self.state[1]
        if descr.startswith('O'):
            return np.dtype(object)
        if isinstance(self.state, tuple) and len(self.state) > 3 and \
                self.state[3] is not None:
            raise ValueError('structured dtypes are not supported')
        return np.dtype((endian if endian in '<>' else '') + descr)


class _Array:
    """numpy.core.multiarray._reconstruct(ndarray, (0,), b'b') + BUILD."""
    __slots__ = ('state',)

    def __init__(self):
        self.state = None

    def value(self):
        if self.state is None:
            raise ValueError('array without state')
        _version, shape, dt, fortran, payload = self.state
        dtype = dt.dtype()
        order = 'F' if fortran else 'C'
        if dtype == np.dtype(object):
            out = np.empty(shape, dtype=object)
            flat = list(payload)
            if len(flat) != out.size:
                raise ValueError('object array size mismatch')
            view = out.reshape(-1, order=order)
            for i, v in enumerate(flat):
                view[i] = _finish(v)
            return out
        if isinstance(payload, str):   # protocol 2 stores bytes as latin-1
            payload = payload.encode('latin-1')
        arr = np.frombuffer(bytes(payload), dtype=dtype)
        return arr.reshape(shape, order=order).copy()


def _finish(v):
    if isinstance(v, _Array):
        return v.value()
    if isinstance(v, dict):
        return {_finish(k): _finish(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_finish(x) for x in v]
    if isinstance(v, tuple):
        return tuple(_finish(x) for x in v)
    if isinstance(v, (_Global, _Dtype)):
        raise ValueError('unsupported object in stream')
    return v


def _reduce(func, args):
    if not isinstance(func, _Global):
        raise ValueError('call of a non-global')
    key = func.key()
    if key == 'numpy.dtype':
        return _Dtype(str(args[0]))
    if key == 'multiarray._reconstruct':
        return _Array()
    if key == 'multiarray.scalar':
        dt = args[0].dtype()
        raw = args[1]
        if isinstance(raw, str):
            raw = raw.encode('latin-1')
        return np.frombuffer(bytes(raw), dtype=dt)[0]
    if key == 'multiarray._frombuffer':   # protocol 5 arrays
        buf, dt, shape, order = args
        arr = np.frombuffer(bytes(buf), dtype=dt.dtype())
        return arr.reshape(shape, order=order).copy()
    if key == '_codecs.encode':
        return str(args[0]).encode(args[1] if len(args) > 1 else 'utf-8')
    if key == 'collections.OrderedDict':
        return dict(args[0]) if args else {}
    if key in ('builtins.set', '__builtin__.set'):
        return set(args[0]) if args else set()
    if key in ('builtins.frozenset', '__builtin__.frozenset'):
        return frozenset(args[0]) if args else frozenset()
    raise ValueError('global {} is not decoded'.format(key))


_ALLOWED_GLOBALS = {'numpy.dtype', 'numpy.ndarray', 'multiarray._reconstruct',
                    'multiarray.scalar', 'multiarray._frombuffer',
                    '_codecs.encode',
                    'collections.OrderedDict', 'builtins.set',
                    '__builtin__.set', 'builtins.frozenset',
                    '__builtin__.frozenset'}

_LITERALS = {'SHORT_BINUNICODE', 'BINUNICODE', 'BINUNICODE8', 'UNICODE',
             'SHORT_BINSTRING', 'BINSTRING', 'STRING', 'SHORT_BINBYTES',
             'BINBYTES', 'BINBYTES8', 'BININT', 'BININT1', 'BININT2', 'INT',
             'LONG', 'LONG1', 'LONG4', 'BINFLOAT', 'FLOAT', 'BYTEARRAY8'}


def loads(data: bytes):
    """Value of the pickle stream `data`, rebuilt from its opcodes."""
    stack, memo, marks = [], {}, []

    def pop_mark():
        m = marks.pop()
        items = stack[m:]
        del stack[m:]
        return items

    for op, arg, pos in pickletools.genops(data):
        name = op.name
        if name in ('PROTO', 'FRAME'):
            continue
        if name == 'STOP':
            break
        if name in _LITERALS:
            stack.append(arg)
        elif name == 'NONE':
            stack.append(None)
        elif name in ('NEWTRUE', 'NEWFALSE'):
            stack.append(name == 'NEWTRUE')
        elif name == 'EMPTY_DICT':
            stack.append({})
        elif name == 'EMPTY_LIST':
            stack.append([])
        elif name == 'EMPTY_TUPLE':
            stack.append(())
        elif name == 'EMPTY_SET':
            stack.append(set())
        elif name == 'MARK':
            marks.append(len(stack))
        elif name == 'MEMOIZE':
            memo[len(memo)] = stack[-1]
        elif name in ('PUT', 'BINPUT', 'LONG_BINPUT'):
            memo[arg] = stack[-1]
        elif name in ('GET', 'BINGET', 'LONG_BINGET'):
            stack.append(memo[arg])
        elif name == 'POP':
            stack.pop()
        elif name == 'POP_MARK':
            pop_mark()
        elif name == 'DUP':
            stack.append(stack[-1])
        elif name == 'SETITEM':
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif name == 'SETITEMS':
            items = pop_mark()
            d = stack[-1]
            for i in range(0, len(items), 2):
                d[items[i]] = items[i + 1]
        elif name == 'APPEND':
            v = stack.pop()
            stack[-1].append(v)
        elif name == 'APPENDS':
            items = pop_mark()
            stack[-1].extend(items)
        elif name == 'ADDITEMS':
            items = pop_mark()
            stack[-1].update(items)
        elif name == 'LIST':
            stack.append(list(pop_mark()))
        elif name == 'DICT':
            items = pop_mark()
            stack.append({items[i]: items[i + 1]
                          for i in range(0, len(items), 2)})
        elif name == 'TUPLE':
            stack.append(tuple(pop_mark()))
        elif name in ('TUPLE1', 'TUPLE2', 'TUPLE3'):
            k = int(name[-1])
            items = tuple(stack[-k:])
            del stack[-k:]
            stack.append(items)
        elif name == 'FROZENSET':
            stack.append(frozenset(pop_mark()))
        elif name in ('GLOBAL', 'STACK_GLOBAL'):
            if name == 'GLOBAL':
                module, qual = arg.split(' ', 1)
            else:
                qual = stack.pop()
                module = stack.pop()
            g = _Global(module, qual)
            if g.key() not in _ALLOWED_GLOBALS:
                raise ValueError('global {} is not decoded (opcode at '
                                 '{})'.format(g.key(), pos))
            stack.append(g)
        elif name == 'REDUCE':
            args = stack.pop()
            func = stack.pop()
            stack.append(_reduce(func, args))
        elif name == 'BUILD':
            state = stack.pop()
            target = stack[-1]
            if not isinstance(target, (_Array, _Dtype)):
                raise ValueError('BUILD of an unsupported object at '
                                 '{}'.format(pos))
            target.state = state
        else:
            raise ValueError('unsupported opcode {} at {}'.format(name, pos))
    if len(stack) != 1:
        raise ValueError('malformed stream')
    return _finish(stack[0])
