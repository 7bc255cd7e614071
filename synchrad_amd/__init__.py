"""MI355X (gfx950) implementation of synchrad's spectral radiation integral.

Modules:
  calc      -- ``SynchRad``, the reference's calculator API (drop-in)
  utils     -- the reference's analysis helpers and constants
  engine    -- CSR track batches, device grids, the batched spectrum call
  workloads -- seeded synthetic track sets of the benchmark configurations
  _abi      -- ctypes binding of include/synchrad_amd.h
"""
__version__ = "0.1.0"
