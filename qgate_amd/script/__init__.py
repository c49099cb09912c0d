"""Circuit DSL (qgate/script/__init__.py)."""
from .script import (new_qreg, new_qregs, release_qreg, new_reference, new_references,
                     new_gatelist, measure, prob, barrier, reset, if_)
from .script import (I, H, S, T, X, Y, Z, Rx, Ry, Rz, U1, U2, U3, controlled, ctrl, Swap, SH,
                     Expii, Expiz, Expi)
