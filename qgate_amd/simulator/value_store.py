"""Values written by measure/prob, keyed by Reference id (qgate/simulator/value_store.py)."""
from .. import model


class ValueStore:

    def __init__(self):
        self.valuedict = {}

    def reset(self):
        self.valuedict = {}

    def sync_refs(self, refs):
        for ref in refs:
            if self.valuedict.get(ref.id) is None:
                self.valuedict[ref.id] = None

    def set(self, ref, value):
        self.valuedict[ref.id] = value

    def get(self, refs):
        if isinstance(refs, model.Reference):
            return self.valuedict.get(refs.id)
        return [self.valuedict.get(ref.id) for ref in refs]

    def get_packed_value(self, ref_array):
        """Bit i is set when ref_array[i] holds 1; unset refs count as 0 (OpenQASM rule)."""
        packed = 0
        for i, ref in enumerate(ref_array):
            if self.valuedict.get(ref.id, 0) == 1:
                packed |= 1 << i
        return packed

    def get_mask(self, ref_array):
        mask = 0
        for i, ref in enumerate(ref_array):
            if ref.id not in self.valuedict:
                mask |= 1 << i
        return mask
