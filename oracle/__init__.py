"""CPU oracle for the Lagrange assemble-and-solve path -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy/SciPy FP64 restatement of the reference
algorithm (weihuayi/fealpy, numpy backend), function by function, each citing
the reference file:line it follows.  It exists to *check* the MI355X product
(``fealpy_amd``): only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  Product code must never
import, call or fall back to anything in here.

Modules
-------
``oracle.basis``   multi-index matrices, simplex shape functions, quadrature rules
``oracle.mesh``    ``from_box`` meshes, topology (``construct``/``flocc``), DOF numbering
``oracle.fem``     integrators, COO->CSR assembly, ``DirichletBC``, ``cg``, L2 error

Pinning (see tests/test_oracle_*.py): every function here is checked against the
reference's own golden vectors (``tests/golden/ref_*.npz``), against fixtures
written by running the reference where its source is present
(``tests/golden/fixtures_*.npz``) and against closed forms / brute force.
Functions without such a check are marked "parity unpinned" in their docstring.
"""
