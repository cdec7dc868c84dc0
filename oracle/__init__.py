"""CPU oracle of the iactrace ray-tracing path -- test infrastructure only.

iact_oracle: numpy restatement (fp64 and reference-fp32 modes).
cpu_ref:     C++/OpenMP restatement of the image path and its adjoint (CPU baseline, FP64 run).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
"""
