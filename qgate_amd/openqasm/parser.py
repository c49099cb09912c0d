"""Recursive-descent parser of OpenQASM 2.0 into statement tuples.

The accepted language is the grammar of qgate/openqasm/yacc.py:16-240: an optional
`OPENQASM <real>;` header, `include "qelib1.inc";` lines (any other file raises
NotImplementedError), then qreg/creg declarations, gate applications (U, CX and named gates
with optional parameters over indexed qubits or whole registers), measure, reset, barrier and
`if (creg == int) <op>`. Gate declarations and opaque gates are parsed and then rejected with
NotImplementedError, as the reference does. Syntax errors raise SyntaxError carrying the line
and offset of the offending token.

Statements (each a tuple):
  ('decl', 'qreg' | 'creg', name, size)
  ('gate', name, params, args)        params: list of Python expression strings or None
  ('measure', qarg, carg)             args are (name,) or (name, index)
  ('reset', arg)
  ('barrier', args)
  ('if', creg_name, value, statement)
"""
from .lexer import tokenize

# Expression precedence (yacc.py:8-13): + - < * / < ^ < unary minus.
_BINARY = {'+': 1, '-': 1, '*': 2, '/': 2, '^': 3}
_RIGHT_ASSOC = {'^'}
_PY_FUNC = {'sin': 'math.sin', 'cos': 'math.cos', 'tan': 'math.tan', 'exp': 'math.exp',
            'ln': 'math.log', 'sqrt': 'math.sqrt'}


class Parser:

    def __init__(self, text):
        self.tokens = tokenize(text)
        self.i = 0

    # ---- token helpers -------------------------------------------------------------------
    def peek(self, k=0):
        return self.tokens[min(self.i + k, len(self.tokens) - 1)]

    def take(self):
        tok = self.tokens[self.i]
        self.i += 1
        return tok

    def error(self, tok, what):
        err = SyntaxError('Syntax Error: {} at {!r}, line {}'.format(what, tok.text or 'end of input', tok.line))
        err.lineno, err.pos = tok.line, tok.pos
        return err

    def expect(self, kind):
        tok = self.peek()
        if tok.kind != kind:
            raise self.error(tok, 'expected {}'.format(kind))
        return self.take()

    def accept(self, kind):
        if self.peek().kind == kind:
            return self.take()
        return None

    # ---- program -------------------------------------------------------------------------
    def program(self):
        if self.accept('OPENQASM'):
            self.expect('REAL')
            self.expect(';')
        while self.accept('INCLUDE'):
            name = self.expect('FILENAME')
            self.expect(';')
            if name.text != '"qelib1.inc"':
                raise NotImplementedError('include is not implemented.')
        statements = []
        while self.peek().kind != 'EOF':
            statements.append(self.statement())
        return statements

    def statement(self):
        kind = self.peek().kind
        if kind in ('QREG', 'CREG'):
            kw = self.take().text
            name = self.expect('ID').text
            self.expect('[')
            size = int(self.expect('NMINTEGER').text)
            self.expect(']')
            self.expect(';')
            return ('decl', kw, name, size)
        if kind == 'GATE':
            self.gate_declaration()
            raise NotImplementedError('gatedecl not implemented')
        if kind == 'OPAQUE':
            self.take()
            self.expect('ID')
            if self.accept('('):
                if self.peek().kind != ')':
                    self.explist()
                self.expect(')')
            self.idlist()
            self.expect(';')
            raise NotImplementedError('opaque is not implemented.')
        if kind == 'IF':
            self.take()
            self.expect('(')
            creg = self.expect('ID').text
            self.expect('EQ')
            value = int(self.expect('NMINTEGER').text)
            self.expect(')')
            return ('if', creg, value, self.qop())
        if kind == 'BARRIER':
            self.take()
            args = self.anylist()
            self.expect(';')
            return ('barrier', args)
        return self.qop()

    def gate_declaration(self):
        self.expect('GATE')
        self.expect('ID')
        if self.accept('('):
            if self.peek().kind != ')':
                self.idlist()
            self.expect(')')
        self.idlist()
        self.expect('{')
        while not self.accept('}'):
            if self.accept('BARRIER'):
                self.idlist()
                self.expect(';')
            else:
                self.uop()

    def qop(self):
        kind = self.peek().kind
        if kind == 'MEASURE':
            self.take()
            q = self.argument()
            self.expect('-')
            self.expect('>')
            c = self.argument()
            self.expect(';')
            return ('measure', q, c)
        if kind == 'RESET':
            self.take()
            arg = self.argument()
            self.expect(';')
            return ('reset', arg)
        return self.uop()

    def uop(self):
        tok = self.peek()
        if tok.kind == 'U':
            self.take()
            self.expect('(')
            params = self.explist()
            self.expect(')')
            arg = self.argument()
            self.expect(';')
            return ('gate', 'U', params, [arg])
        if tok.kind == 'CX':
            self.take()
            a0 = self.argument()
            self.expect(',')
            a1 = self.argument()
            self.expect(';')
            return ('gate', 'CX', None, [a0, a1])
        if tok.kind == 'ID':
            name = self.take().text
            params = None
            if self.accept('('):
                params = [] if self.peek().kind == ')' else self.explist()
                self.expect(')')
                if not params:
                    params = None
            args = self.anylist()
            self.expect(';')
            return ('gate', name, params, args)
        raise self.error(tok, 'expected a statement')

    # ---- lists and arguments ---------------------------------------------------------------
    def argument(self):
        name = self.expect('ID').text
        if self.accept('['):
            index = int(self.expect('NMINTEGER').text)
            self.expect(']')
            return (name, index)
        return (name,)

    def anylist(self):
        args = [self.argument()]
        while self.accept(','):
            args.append(self.argument())
        return args

    def idlist(self):
        names = [self.expect('ID').text]
        while self.accept(','):
            names.append(self.expect('ID').text)
        return names

    def explist(self):
        exps = [self.expression()]
        while self.accept(','):
            exps.append(self.expression())
        return exps

    # ---- expressions: Python source text, fully parenthesised --------------------------------
    def expression(self, min_prec=1):
        lhs = self.unary()
        while True:
            op = self.peek().kind
            prec = _BINARY.get(op)
            if prec is None or prec < min_prec:
                return lhs
            self.take()
            rhs = self.expression(prec if op in _RIGHT_ASSOC else prec + 1)
            lhs = '({} {} {})'.format(lhs, '**' if op == '^' else op, rhs)

    def unary(self):
        if self.accept('-'):
            return '(-{})'.format(self.unary())
        return self.primary()

    def primary(self):
        tok = self.peek()
        if tok.kind in ('REAL', 'NMINTEGER', 'ID'):
            return self.take().text
        if tok.kind == 'PI':
            self.take()
            return 'math.pi'
        if tok.kind == 'UNARYOP':
            self.take()
            self.expect('(')
            inner = self.expression()
            self.expect(')')
            return '{}({})'.format(_PY_FUNC[tok.text], inner)
        if tok.kind == '(':
            self.take()
            inner = self.expression()
            self.expect(')')
            return inner
        raise self.error(tok, 'expected an expression')


def parse(text):
    """Statement tuples of an OpenQASM 2.0 program."""
    return Parser(text).program()
