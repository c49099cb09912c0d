"""qgate_amd: an MI355X-native state-vector backend with Qgate's circuit and simulator API."""
from . import model
from . import simulator
from .model import prefs
from . import script
from . import openqasm


def dump(obj, file=None, number_format=None):
    """Pretty-print a circuit, qubit states, values or observations (qgate/__init__.py:6-27)."""
    import sys
    from .simulator import qubits, value_store, observation
    file = sys.stdout if file is None else file
    if isinstance(obj, (model.GateList, list)):
        model.dump(obj, file)
    elif isinstance(obj, (qubits.Qubits, qubits.StateGetter)):
        getter = obj.states if isinstance(obj, qubits.Qubits) else obj
        if number_format is None:
            number_format = '{1: .3f}' if getter.mathop is qubits.null else '{1:g}'
        fmt = '|{{0:0{0:d}b}}> '.format(getter.n_qregs) + number_format
        for idx, value in enumerate(getter[:]):
            print(fmt.format(idx, value), file=file)
    elif isinstance(obj, value_store.ValueStore):
        for key, value in obj.valuedict.items():
            print('{:d}:'.format(key), value, file=file)
    elif isinstance(obj, (observation.ObservationList, observation.ObservationHistgram)):
        print(repr(obj), file=file)
    else:
        raise RuntimeError('unknown object, {}.'.format(type(obj)))
