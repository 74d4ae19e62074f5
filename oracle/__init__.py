"""CPU FP64 oracle for the AFQMC walker-propagation path.

TEST INFRASTRUCTURE ONLY.  This package is a from-scratch CPU restatement of
the reference algorithm (hshi/AFQMCLAB), written with numpy and the same
LAPACK routines the reference calls (zgetrf, zgetrs, zgeqrf, zungqr, zheevd,
dsyevd via scipy.linalg.lapack).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline -- never as part of the product path, which
lives in ``afqmclab_amd`` and runs on the GPU through the C-ABI library.

Modules and the reference files they follow:

* ``rng``       -- counter-based uniform stream (the reference's SPRNG stream is
                   not available; uniforms are injected in the reference's
                   consumption order), Box-Muller gaussians
                   (common/randomHao/source/random_hao.cpp:73-85).
* ``mathfun``   -- solveCoshxEqExpy, solveCosxEqExpy, exp2by2Matrix
                   (common/mathHao/source/simple_fun.cpp:7-91).
* ``linalg``    -- QRMatrix, LU, logDeterminant, solve_lineq
                   (common/tensorHao/source/tensor_2d_bl_cpu.cpp:145-315,
                   common/tensorHao/source/tensor_2d_common_fun.cpp:58-105).
* ``blocks``    -- SD / SD2s / SD2is walkers, one-body and HS operations,
                   SDSDOperation, SD2sSD2isOperation, NiupNidn, CholeskyReal,
                   Taylor exponential (afqmc/blocks/...).
* ``models``    -- HubbardSOC + its fixed mixed measurement, RealMaterialMolecule
                   + its fixed mixed measurement, lattice hopping generator
                   (afqmc/model/..., scripts/supercubic/setHoping.py).
* ``comb``      -- popCheck, popConfiguration, populationControl
                   (afqmc/utilities/populationControl/...).
* ``drivers``   -- constraint-path and phaseless step loops
                   (tutorials/afqmcConstraintPath, tutorials/afqmcPhaseless).
* ``lanczos``   -- exact diagonalisation of small Hubbard lattices, the
                   end-to-end energy check (lanczos/ in the reference).

Pinning status of each function is recorded in its docstring; the checks that
pin it live in ``tests/test_oracle_*.py``.
"""
