"""OpenQASM qelib1 aliases built from the script gate set (qgate/script/qelib1.py:1-46)."""
from .script import S, T, ctrl


def cx(c, t):
    return ctrl(c).X(t)


def sdg(a):
    return S.Adj(a)


def tdg(a):
    return T.Adj(a)


def cz(a, b):
    return ctrl(a).Z(b)


def cy(a, b):
    return ctrl(a).Y(b)


def ch(a, b):
    return ctrl(a).H(b)


def ccx(a, b, c):
    return ctrl(a, b).X(c)


def crz(_lambda, a, b):
    return ctrl(a).Rz(_lambda)(b)


def cu1(_lambda, a, b):
    return ctrl(a).U1(_lambda)(b)


def cu3(theta, phi, _lambda, c, t):
    return ctrl(c).U3(theta, phi, _lambda)(t)
