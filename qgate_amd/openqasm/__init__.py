"""OpenQASM 2.0 front-end: QASM text -> Python source over qgate_amd.script -> circuit.

The interface follows qgate/openqasm/importer.py:9-50: translate(qasm) and
translate_file(path) return the Python source of the program; load_circuit(qasm) and
load_circuit_from_file(path) run it and return a module-like object holding one attribute per
register (lists of qregs / references) and `circuit`, the gate list. Gate names map as in
qgate/openqasm/formatter.py:78-112 (U -> U3, rz -> U1, sdg -> S.Adj, cu1 -> ctrl(a).U1,
ccx -> ctrl(a, b).X, ...).

Differences from the reference, where its translation would not run or would not mean what
the QASM says (it maps the same constructs otherwise):
  * `a ^ b` is a power (the reference emits Python's xor); `ln` is math.log (the reference
    emits math.ln, which does not exist); `cu3` maps to ctrl(a).U3 (the reference emits an
    undefined name);
  * a multi-qubit gate over whole registers is applied element-wise (OpenQASM broadcasting);
    the reference would control on the whole register;
  * `measure q -> c` over registers is element-wise whatever the register names; the
    reference decides by the length of the name.
The lexer and parser are this package's own: the reference's need `ply`, which is not part
of this environment.
"""
import math as _math
import sys as _sys
import traceback as _traceback

from .parser import parse

# name -> (number of controls, script factory, parameter count or None, adjoint)
_GATES = {
    'U': (0, 'U3', 3, False), 'u3': (0, 'U3', 3, False), 'u2': (0, 'U2', 2, False),
    'u1': (0, 'U1', 1, False), 'rz': (0, 'U1', 1, False),
    'rx': (0, 'Rx', 1, False), 'ry': (0, 'Ry', 1, False),
    'x': (0, 'X', None, False), 'y': (0, 'Y', None, False), 'z': (0, 'Z', None, False),
    'h': (0, 'H', None, False), 's': (0, 'S', None, False), 't': (0, 'T', None, False),
    'sdg': (0, 'S', None, True), 'tdg': (0, 'T', None, True), 'id': (0, 'I', None, False),
    'CX': (1, 'X', None, False), 'cx': (1, 'X', None, False), 'cy': (1, 'Y', None, False),
    'cz': (1, 'Z', None, False), 'ch': (1, 'H', None, False),
    'cu1': (1, 'U1', 1, False), 'cu3': (1, 'U3', 3, False),
    'ccx': (2, 'X', None, False),
}


def _ref(arg):
    return arg[0] if len(arg) == 1 else '{}[{}]'.format(arg[0], arg[1])


class _Writer:

    def __init__(self):
        self.lines = ['from qgate_amd.script import *', 'import math', '']
        self.level = 0
        self.list_open = False
        self.circuit_defined = False

    def write(self, text):
        self.lines.append('    ' * self.level + text)

    def open_list(self):
        if self.list_open:
            return
        self.write('circuit += [' if self.circuit_defined else 'circuit = [')
        self.circuit_defined = True
        self.list_open = True
        self.level += 1

    def close_list(self):
        if self.list_open:
            self.level -= 1
            self.write(']')
            self.write('')
            self.list_open = False

    # ---- statements ----------------------------------------------------------------------
    def statement(self, st):
        kind = st[0]
        if kind == 'decl':
            self.close_list()
            factory = 'new_qregs' if st[1] == 'qreg' else 'new_references'
            self.write('{} = {}({})'.format(st[2], factory, st[3]))
            return
        self.open_list()
        self.op(st)

    def op(self, st):
        kind = st[0]
        if kind == 'gate':
            self.gate(st[1], st[2], st[3])
        elif kind == 'measure':
            q, c = st[1], st[2]
            if len(q) == 1 and len(c) == 1:
                self.write('[measure(_c, _q) for _c, _q in zip({}, {})],'.format(c[0], q[0]))
            elif len(q) == 2 and len(c) == 2:
                self.write('measure({}, {}),'.format(_ref(c), _ref(q)))
            else:
                raise RuntimeError('measure needs two registers or two indexed bits, {} -> {}.'
                                   .format(_ref(q), _ref(c)))
        elif kind == 'reset':
            self.write(self._each('reset({})', [st[1]]))
        elif kind == 'barrier':
            for arg in st[1]:
                self.write(self._each('barrier({})', [arg]))
        elif kind == 'if':
            self.write('if_({}, {}, ['.format(st[1], st[2]))
            self.level += 1
            self.op(st[3])
            self.level -= 1
            self.write(']),')
        else:
            raise AssertionError('unexpected statement {}'.format(kind))

    @staticmethod
    def _each(pattern, args):
        """pattern applied to the args; whole registers are taken element by element, in step."""
        regs = [i for i, a in enumerate(args) if len(a) == 1]
        if not regs:
            return pattern.format(*[_ref(a) for a in args]) + ','
        names = ['_a{}'.format(i) if len(a) == 1 else _ref(a) for i, a in enumerate(args)]
        loop_vars = ', '.join('_a{}'.format(i) for i in regs)
        if len(regs) == 1:
            source = args[regs[0]][0]
        else:
            source = 'zip({})'.format(', '.join(args[i][0] for i in regs))
        return '[{} for {} in {}],'.format(pattern.format(*names), loop_vars, source)

    def gate(self, name, params, args):
        if name == 'swap':
            if len(args) != 2:
                raise RuntimeError('swap takes 2 qubits, {} given.'.format(len(args)))
            self.write(self._each('Swap({}, {})', args))
            return
        spec = _GATES.get(name)
        if spec is None:
            raise RuntimeError('unknown gate, {}.'.format(name))
        n_ctrl, factory, n_params, adjoint = spec
        if (n_params or 0) != len(params or []):
            raise RuntimeError('{} takes {} parameter(s), {} given.'.format(name, n_params or 0, len(params or [])))
        call = factory + ('({})'.format(', '.join(params)) if params else '') + ('.Adj' if adjoint else '')
        if n_ctrl == 0:
            # single-qubit gate: one application per listed qubit or register
            for arg in args:
                self.write(self._each(call + '({})', [arg]))
            return
        if len(args) != n_ctrl + 1:
            raise RuntimeError('{} takes {} qubits, {} given.'.format(name, n_ctrl + 1, len(args)))
        ctrls = ', '.join(['{}'] * n_ctrl)
        self.write(self._each('ctrl(' + ctrls + ').' + call + '({})', args))

    def source(self, statements):
        for st in statements:
            self.statement(st)
        self.close_list()
        return '\n'.join(self.lines) + '\n'


def translate(qasm):
    """Python source (over qgate_amd.script) of an OpenQASM 2.0 program."""
    return _Writer().source(parse(qasm))


def translate_file(filename):
    with open(filename, 'r') as f:
        return translate(f.read())


class QasmModule:
    """Namespace of a loaded program: its registers and `circuit`."""


def _import(name, globals=None, locals=None, fromlist=(), level=0):
    if name == 'math':
        return _math
    if name == 'qgate_amd.script':
        from .. import script
        return script
    raise ImportError('{} cannot be imported from OpenQASM programs'.format(name))


def load_circuit(qasm):
    """Translate and run a program; errors from running it carry the offending source line
    (importer.py:27-43)."""
    code = translate(qasm)
    module = QasmModule()
    module.__dict__['__builtins__'] = {'__import__': _import, 'zip': zip}
    try:
        exec(compile(code, '<openqasm>', 'exec'), module.__dict__)
    except Exception as e:
        lineno = None
        for frame in _traceback.extract_tb(_sys.exc_info()[2]):
            if frame.filename == '<openqasm>':
                lineno = frame.lineno
        if lineno is None:
            raise
        detail = e.args[0] if e.args else str(e)
        line = code.splitlines()[lineno - 1].strip()
        raise e.__class__("{}, '{}'".format(detail, line)) from None
    del module.__dict__['__builtins__']
    return module


def load_circuit_from_file(filename):
    with open(filename, 'r') as f:
        return load_circuit(f.read())
