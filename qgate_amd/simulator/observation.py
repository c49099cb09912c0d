"""Measured-value containers (qgate/simulator/observation.py:1-128, utils.py:30-95).

An observation packs one bit per Reference (bit i = reflist[i]); `none_mask` marks
references that were never written (they read as None).
"""
import collections
import numbers


def _bits(value, mask, width):
    v = '{{:0{}b}}'.format(width).format(value)
    m = '{{:0{}b}}'.format(width).format(mask)
    return ''.join(c if mc == '0' else '*' for c, mc in zip(v, m))


class Observation:
    def __init__(self, reflist, value, none_mask):
        self._reflist = reflist
        self._value = value
        self._none_mask = none_mask

    def __int__(self):
        return int(self._value)

    @property
    def int(self):
        return self._value

    def __call__(self, key):
        if key not in self._reflist:
            raise RuntimeError('unknown ref, {}.'.format(key))
        bit = 1 << self._reflist.index(key)
        if self._value & bit:
            return 1
        return None if self._none_mask & bit else 0

    def _check_operand(self, other):
        if isinstance(other, numbers.Integral):
            raise ValueError('To compare with integer, use int() or Observation.int property.')

    def __eq__(self, other):
        self._check_operand(other)
        if not isinstance(other, Observation):
            raise TypeError('comparison with Observation and {} not supported.'.format(type(other)))
        return (self._reflist == other._reflist and self._value == other._value
                and self._none_mask == other._none_mask)

    def __ne__(self, other):
        return not self.__eq__(other)

    def __lt__(self, other):
        self._check_operand(other)
        return NotImplemented

    __hash__ = None

    def __repr__(self):
        return _bits(self._value, self._none_mask, len(self._reflist))


class ObservationList:
    def __init__(self, reflist, values, mask):
        self._reflist = reflist
        self._values = values
        self._mask = mask

    @property
    def intarray(self):
        return self._values

    def __len__(self):
        return len(self._values)

    def __iter__(self):
        return (Observation(self._reflist, v, self._mask) for v in self._values)

    def __getitem__(self, key):
        if isinstance(key, slice):
            return ObservationList(self._reflist, self._values[key], self._mask)
        return Observation(self._reflist, self._values[key], self._mask)

    def __call__(self, ref):
        if ref not in self._reflist:
            raise RuntimeError('unknown ref, {}.'.format(ref))
        bit = 1 << self._reflist.index(ref)
        if self._mask & bit:
            return [None] * len(self._values)
        return [1 if (int(v) & bit) else 0 for v in self._values]

    def histgram(self):
        hist = collections.Counter(int(v) for v in self._values)
        return ObservationHistgram(hist, len(self._reflist), len(self))

    def __repr__(self):
        w = len(self._reflist)
        items = [_bits(int(v), self._mask, w) for v in self._values]
        if len(items) > 256:
            items = items[:3] + ['...'] + items[-3:]
        return '[' + ', '.join(items) + ']'


class ObservationHistgram:
    def __init__(self, hist, n_bits, n_samples):
        self._hist = hist
        self._n_bits = n_bits
        self._n_samples = n_samples

    @property
    def n_samples(self):
        return self._n_samples

    def __len__(self):
        return len(self._hist)

    def __iter__(self):
        return iter(self._hist)

    def __getitem__(self, key):
        return self._hist.get(key, 0)

    def keys(self):
        return self._hist.keys()

    def values(self):
        return self._hist.values()

    def items(self):
        return self._hist.items()

    def __repr__(self):
        fmt = '{{:0{}b}}: {{}}'.format(self._n_bits)
        keys = sorted(self._hist)
        items = [fmt.format(k, self._hist[k]) for k in keys]
        if len(items) > 256:
            items = items[:3] + ['...'] + items[-3:]
        return '{' + ', '.join(items) + '}'
