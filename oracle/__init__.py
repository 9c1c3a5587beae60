"""Parity oracles for the full-binary-pursuit stage.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

* ``fbp_oracle`` -- ctypes wrapper of the C restatement (fbp_oracle.c), FP32
  with the exact operation order the HIP path reproduces.
* ``fbp_fp64``   -- independent NumPy FP64 restatement for tiny inputs.
"""
