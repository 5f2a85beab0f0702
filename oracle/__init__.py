"""CPU oracle for the HMC hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy/SciPy, the arithmetic of the reference
library's HMC path so that the HIP engine in ``hmclab_amd`` can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or as the timed CPU comparator).  The product path (``hmclab_amd``) never
imports it and has no CPU fallback.

Modules
-------
hmc_oracle   distributions, LinearMatrix backends, mass matrices, integrators,
             Metropolis test, step-size autotuning, the proposal loop.
tomography   straight-ray CSR builder restated with a different algorithm
             (incremental cell walk) from the product's vectorised builder.
ess          rank-normalised split-chain bulk ESS (Vehtari et al. 2021).

Pinning status (what checks each function; see tests/test_oracle_*.py):
every function is pinned by closed forms, invariants, brute force on tiny
inputs, or independent library routines.  The reference's own Python cannot
be imported in this image without stand-ins for h5py/ipywidgets/IPython
(absent), so no bitwise golden vectors from the reference exist; the
end-to-end seeded chain is pinned by the draw order read from the reference
source plus per-step closed forms.
"""
