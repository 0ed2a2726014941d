"""physis_amd -- MI355X (gfx950) execution target for the Physis stencil DSL.

Layout:
  include/physis/        C ABI headers (physis_hip.h replaces physis_cuda.h)
  include/physis/hip/    device skeleton headers used by generated code
  physis_amd/csrc/       libphysis_rt_hip sources (runtime, buffers,
                         reductions, built-in stencil programs)
  physis_amd/runtime.py  ctypes binding of the C ABI
"""
