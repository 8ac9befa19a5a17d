"""omegaqmc_amd -- MI355X-native phaseless-AFQMC walker propagator.

Hand-written gfx950 HIP kernels behind a C ABI (include/omega_afqmc.h), with
a Python mirror of the reference's ``OmegaQMC.afqmc_gto`` operator API (and of
``OmegaQMC.qed_afqmc_gto`` for cavity QED in :mod:`.qed`).
"""

from .afqmc_gto import (  # noqa: F401
    FBBOUND_DEFAULT, WEIGHT_CLIP_FRACTION, PhiloxKey, System, _apply_exp_vhs,
    _apply_exp_vhs_from_chol, _build_vhs, _update_weights_phaseless, apply_one_body, aux_fields,
    build_propagator, comb_offset, determinant_energy_from_integrals, do_binning_analysis,
    force_bias, greens_function, greens_function_force_bias, greens_function_overlap,
    half_rotate_cholesky, local_energy, local_energy_2body, local_energy_streamed,
    one_body_chol_correction, orthogonalize_walkers, population_control_comb,
    prepare_afqmc_integrals_from_arrays, propagate_walkers,
    MultiDetSystem, greens_function_multidet, greens_function_multidet_overlap_only,
    local_energy_multidet, multidet_charge_density, occupations_from_trials,
    propagate_walkers_multidet)
from ._lib import AFQError  # noqa: F401
from .afqmc_gto import (greens_function_multidet_force_bias,  # noqa: F401
                        half_rotate_cholesky_multidet, local_energy_multidet_streamed,
                        walker_status, decode_walker_flags, Walkers,
                        population_control_pair_branch, WickTrial, build_wick_trial,
                        propagate_walkers_wick)
from .driver import AFQMCDriver, get_afqmc_func_from_integrals  # noqa: F401
from .estream import EStreamDriver, get_afqmc_estream_func_from_integrals  # noqa: F401
from .heg import HEGDriver, build_3deg_system, get_afqmc_3deg_func  # noqa: F401
from .qed import (QEDDriver, QEDWalkers, get_qed_afqmc_func_from_integrals,  # noqa: F401
                  prepare_qed_integrals_from_arrays, propagate_qed_walkers, qed_local_energy)
from .edga import load_edga_result, make_omegaqmc_trial, select_top_determinants  # noqa: F401

__version__ = "0.1.0"
