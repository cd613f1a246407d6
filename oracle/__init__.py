"""Test infrastructure only: the CPU oracle for ProCell's plain method.

Importable by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg; the product package never imports it.
"""
