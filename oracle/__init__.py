"""CPU oracle for the ssmhip hot path — TEST INFRASTRUCTURE ONLY.

Restatements of the reference's algorithms used to check the GPU kernels.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline; the product
(package ssmhip) never imports or calls it.

Modules
  em_numpy      float32 NumPy transcription of the sequential EM loops
                (ddm_flexbound, race_model, lca) and LBA, replay or NumPy RNG
  philox        Philox4x32-10 (uint32 NumPy) + float64 Box–Muller
  ddm_analytic  closed forms for the constant-bound DDM (Navarro & Fuss 2009)
  kde_ref       float64 LogKDE restatement (fit / eval / sample)
  mlp_ref       float32/float64 NumPy MLP forward/backward, Huber, Adam
  aux_ref       auxiliary labels of the LAN data generator (counts, histograms)
  cpu_baseline  ctypes loader of the C++/OpenMP restatement (cpu_baseline.cpp)

Pinning status of each function is stated in its docstring.
"""
