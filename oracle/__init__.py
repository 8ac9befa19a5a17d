"""Parity oracle for the phaseless-AFQMC walker step.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy FP64 restatement of
the reference's algorithm (``OmegaQMC/afqmc_gto.py``,
``OmegaQMC/observables/{greens,energy,greens_th4,energy_th4}.py``,
``OmegaQMC/integrals/cholesky.py``, ``OmegaQMC/utils.py``), written so that the
HIP kernels of ``omegaqmc_amd`` can be checked against it on identical
inputs.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the reported CPU baseline -- never as a product code path.

Modules
-------
``afqmc``   single-determinant step, Green's functions, energies, QR, comb,
            driver loop, binning analysis.
``multidet`` explicit multi-determinant Green's function / overlap / energy.
``heg``     electron-gas step and sign-weighted local energy
            (``OmegaQMC/afqmc_pw_heg.py``).
``qed``     cavity-QED step: photon half steps, dipole Taylor factor, QED
            local energy and comb (``OmegaQMC/qed_afqmc_gto.py``).
``philox``  Philox4x32-10 counter-based generator (Salmon et al., SC'11) and
            the Box-Muller normal stream the GPU generator also implements.

How each function is pinned (closed forms, brute force, invariants) is
listed in the header of its module; every function has at least one such
check in ``tests/test_oracle_*.py``.
"""
