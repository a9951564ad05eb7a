"""MI355X-native batched KPP Rosenbrock Rodas3 gas-phase chemistry solver.

The product is the C-ABI library ``lib/libgcchemgpu.so`` (declarations in
``include/gcchemgpu.h``), a drop-in for the library GEOS-Chem's fullchem
driver binds through ISO_C_BINDING.  This Python package is the host-side
mirror used by the standalone driver, the tests and the gradient path:

* ``mechanism``  mechanism IR loading and counts (flops, bytes)
* ``lib``        ctypes binding of the C ABI (fails loudly without the library),
                 including the cost order and hybrid CPU/GPU split entry points
                 (``GcAmdHybrid``, ``CPU_TAIL_FN``)
* ``solver``     device-resident batch solve on torch tensors
* ``adjoint``    discrete adjoint (recording solve, reverse sweep, autograd)
* ``inputs``     seeded synthetic cells for the benchmark configurations
* ``synthetic``  mechanism IRs from reaction lists (the synthetic stress mechanisms)
* ``tuning``     stored launch parameters and measured traffic
* ``autotune``   launch-parameter sweep of the persistent kernel
* ``build``      in-tree hipcc build of the library (``codegen/`` tables)

The directory name is not a Python identifier; it is imported as
``gcchem_amd`` (see ``load_package`` in the repository's entry points).
"""
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
