"""Tokens of OpenQASM 2.0 as qgate reads it (qgate/openqasm/lex.py:5-75).

Token kinds: the keywords OPENQASM, include, gate, opaque, qreg, creg, barrier, reset,
measure, if, U and CX; ID (a lower-case letter, then letters, digits, '_'); PI ('pi');
UNARYOP (sin, cos, tan, exp, ln, sqrt); NMINTEGER (digits only) and REAL (digits with a
fraction and/or exponent); FILENAME (a quoted name); EQ ('=='); and the single characters
+ - * / ^ ( ) [ ] ; , { } >. Whitespace and // comments separate tokens. Any other
character raises SyntaxError with its line and offset, as the reference's lexer does.
"""
import re

KEYWORDS = {
    'OPENQASM': 'OPENQASM', 'include': 'INCLUDE', 'gate': 'GATE', 'opaque': 'OPAQUE',
    'qreg': 'QREG', 'creg': 'CREG', 'barrier': 'BARRIER', 'reset': 'RESET',
    'measure': 'MEASURE', 'if': 'IF', 'U': 'U', 'CX': 'CX',
}
UNARY_OPS = ('sin', 'cos', 'tan', 'exp', 'ln', 'sqrt')
SINGLE = set('+-*/^()[];,{}>')

_SCAN = re.compile(r'''
    (?P<ws>[ \t\r]+) |
    (?P<nl>\n) |
    (?P<comment>//[^\n]*) |
    (?P<number>[0-9]+(?:\.[0-9]*)?(?:[eE][-+]?[0-9]+)?) |
    (?P<word>[A-Za-z][A-Za-z0-9_]*) |
    (?P<filename>"[A-Za-z0-9_.]+") |
    (?P<eq>==) |
    (?P<single>[-+*/^()\[\];,{}>])
''', re.VERBOSE)


class Token:
    __slots__ = ('kind', 'text', 'line', 'pos')

    def __init__(self, kind, text, line, pos):
        self.kind, self.text, self.line, self.pos = kind, text, line, pos

    def __repr__(self):
        return 'Token({}, {!r}, line {})'.format(self.kind, self.text, self.line)


def _syntax_error(msg, line, pos):
    err = SyntaxError(msg)
    err.lineno, err.pos = line, pos
    return err


def tokenize(text):
    """List of Tokens of `text` (line ends normalised), ending with an 'EOF' token."""
    text = '\n'.join(text.splitlines())
    tokens, line, pos = [], 1, 0
    while pos < len(text):
        m = _SCAN.match(text, pos)
        if m is None:
            raise _syntax_error('Syntax Error: unexpected character {!r}'.format(text[pos]), line, pos)
        kind, value = m.lastgroup, m.group()
        if kind == 'nl':
            line += 1
        elif kind == 'number':
            tokens.append(Token('NMINTEGER' if value.isdigit() else 'REAL', value, line, pos))
        elif kind == 'word':
            if value in KEYWORDS:
                tokens.append(Token(KEYWORDS[value], value, line, pos))
            elif value == 'pi':
                tokens.append(Token('PI', value, line, pos))
            elif value in UNARY_OPS:
                tokens.append(Token('UNARYOP', value, line, pos))
            elif value[0].islower():
                tokens.append(Token('ID', value, line, pos))
            else:   # identifiers start with a lower-case letter (lex.py:52-53)
                raise _syntax_error('Syntax Error: invalid identifier {!r}'.format(value), line, pos)
        elif kind == 'filename':
            tokens.append(Token('FILENAME', value, line, pos))
        elif kind == 'eq':
            tokens.append(Token('EQ', value, line, pos))
        elif kind == 'single':
            tokens.append(Token(value, value, line, pos))
        pos = m.end()
    tokens.append(Token('EOF', '', line, pos))
    return tokens
