"""Simulator preference keys (qgate/model/prefs.py:1-10).

circuit_prep selects how qregs are grouped into state vectors:
  one_static: the whole circuit is one state vector, allocated before the run;
  static:     independent sub-circuits get their own state vectors, allocated before the run;
  dynamic:    state vectors are created, joined and separated as the circuit runs.
"""
circuit_prep = 'circuit_prep'
dynamic = 'dynamic'
static = 'static'
one_static = 'one_static'
