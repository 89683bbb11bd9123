"""CPU oracle for the BAMF hot path — TEST INFRASTRUCTURE ONLY.

Plain-C restatements of the reference algorithms (liboracle.so, built from
oracle/*.c by oracle/Makefile) with a ctypes wrapper (oracle.pyoracle).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the product library never calls it.

Parity status of each function is recorded in its docstring in pyoracle.
"""
