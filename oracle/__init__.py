"""ORACLE — test infrastructure, not product code.

CPU restatement of NeCT's training and query hot path, used only as the
checker: by tests/, by ``__graft_entry__.smoke()`` and by bench.py's
``cpu_baseline`` leg.  The product package (``nect_amd``) never imports,
links or calls anything here.

Contents (each function cites the reference lines it restates):
  sampler.c / sampler.py   ray sampler, nect/sampling/ct_sampling.cu:117-648
                           + nect/sampling/projector.py:102-131 (plain C,
                           -ffp-contract=off)
  hashgrid.py              multiresolution hash-grid encoding (numpy, exact
                           uint32 index arithmetic, float64 interpolation)
  mlp.py                   bias-free MLP forward/backward (torch CPU, float64)
  train.py                 line integral + loss + autograd step + Adam
                           (nect/trainers/base_trainer.py:517-654)
  analytic.py              float64 analytic line integrals through
                           ellipsoid/sphere phantoms
  geometry.py              geometry bookkeeping (max distance, C structs),
                           nect/sampling/geometry.py:91-178, 322-407
  query.py                 export grids, nect/static_export.py:106-141,
                           nect/dynamic_export.py:519-537,
                           nect/trainers/base_trainer.py:402-450
  networks.py              composite models (QuadCubes, HyperCubes, grid +
                           Identity) and networks of any width,
                           nect/network/networks.py:36-64, 437-449, 496-508

Parity status of each part is stated in its module docstring.
"""
