"""sunnyhip: MI355X-native spin-dynamics hot path of Sunny.jl.

Host-side mirror of Sunny's `System` / `Langevin` / `ImplicitMidpoint` /
`step!` API over the C ABI of libsunnyhip.so (include/sunny_hip.h).
"""
from .crystal import (Bond, Crystal, bond_parity, cubic_crystal, diamond_crystal, lattice_vectors,
                      pyrochlore_crystal, pyrochlore_local_axes, triangular_crystal)
from .integrators import (ConvergenceError, ImplicitMidpoint, Langevin, ReplicaBatch, energy, energy_grad_coherents,
                          energy_grad_dipoles, energy_per_site, step_)
from .operators import (SCALAR_BIQUAD_METRIC, StevensExpansion, decompose_pair_coupling, hermitian_basis,
                        matrix_to_stevens_coefficients, rcs_factors, spin_matrices, stevens_matrices,
                        tensor_expansion)
from .system import (VACUUM_PERMEABILITY, ModelParam, Moment, System, clone_system, enable_dipole_dipole_,
                     expected_spin_host, get_param, get_params, ket_from_dipole, param_spec, polarize_spins_,
                     randomize_spins_, repopulate_couplings_from_params_, set_coherent_, set_dipole_, set_exchange_,
                     set_field_, set_onsite_coupling_, set_pair_biquad_, set_pair_coupling_, set_param_, set_params_,
                     set_spin_rescaling_, set_stevens_expansion_, to_product_space)
from .inhomogeneous import (is_homogeneous, set_exchange_at_, set_field_at_, set_onsite_coupling_at_,
                            set_pair_coupling_at_, set_vacancy_at_, sites_to_internal_bond, to_inhomogeneous)
from . import workloads  # noqa: F401
from .adjoint import midpoint_gradient, structure_factor, structure_factor_loss
from .tempering import ParallelTempering
from .correlations import SampledCorrelations, add_sample_
from .optimize import MinimizationResult, minimize_energy_, suggest_timestep
from .samplers import (LocalSampler, Proposal, local_coloring, mix_proposals, propose_delta, propose_flip,
                       propose_uniform)

__all__ = [n for n in dir() if not n.startswith("_")]
