"""Circuit model: registers, operators, gate types, preprocessing."""
from .ir import (Qreg, Reference, Operator, GateType, Gate, GatelistMacro, MultiQubitGate,
                 Measure, Prob, PauliObserver, PauliMeasure, PauliProb, Barrier, Reset,
                 IfClause, ClauseBegin, ClauseEnd, NewQreg, ReleaseQreg, Join, Separate,
                 GateList, GateListIterator, dump)
from .preprocessor import Preprocessor
from . import gate_type
from . import expand
from . import prefs
