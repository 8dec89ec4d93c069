"""torch_sim_amd -- MI355X-native batched classical force-and-step path.

A drop-in for TorchSim's classical pair-potential path: the periodic
cell-list neighbor list over concatenated systems, the Lennard-Jones / Morse /
soft-sphere energy / force / virial models behind ``model(state)``, and the
``integrate`` / ``optimize`` entry points, all running hand-written gfx950 HIP
kernels through the C ABI in ``include/tsx.h``.
"""

from . import _lib, neighbors, units, workloads
from .engine import BatchedMD
from .integrators import (
    INTEGRATOR_KWARG_UNITS,
    INTEGRATOR_REGISTRY,
    Integrator,
    MDState,
    NoseHooverChain,
    NoseHooverChainFns,
    NPTCRescaleState,
    NPTLangevinAnisotropicState,
    NPTLangevinIsotropicState,
    NPTNoseHooverIsotropicState,
    NPTState,
    NVTNoseHooverState,
    NVTVRescaleState,
    construct_nose_hoover_chain,
    initialize_momenta,
    momentum_step,
    npt_crescale_anisotropic_step,
    npt_crescale_init,
    npt_crescale_isotropic_step,
    npt_crescale_triclinic_average_step,
    npt_crescale_triclinic_step,
    npt_langevin_anisotropic_init,
    npt_langevin_anisotropic_step,
    npt_langevin_isotropic_init,
    npt_langevin_isotropic_step,
    npt_nose_hoover_isotropic_init,
    npt_nose_hoover_isotropic_invariant,
    npt_nose_hoover_isotropic_step,
    nve_init,
    nve_step,
    nvt_langevin_init,
    nvt_langevin_step,
    nvt_nose_hoover_init,
    nvt_nose_hoover_invariant,
    nvt_nose_hoover_step,
    nvt_vrescale_init,
    nvt_vrescale_step,
    position_step,
)
from .models import (
    LennardJonesModel,
    ModelInterface,
    MorseModel,
    MultiLennardJonesModel,
    MultiLennardJonesPairFn,
    MultiSoftSpherePairFn,
    PairForcesModel,
    PairPotentialModel,
    ParticleLifeModel,
    SerialSumModel,
    SoftSphereModel,
    SoftSphereMultiModel,
    SumModel,
    lennard_jones_pair,
    morse_pair,
    particle_life_pair_force,
    soft_sphere_pair,
)
from .neighbors import default_batched_nl, torch_nl_linked_cell, torch_nl_n2, torchsim_nl
from .quantities import calc_kinetic_energy, calc_kT, calc_temperature, system_wise_max_force
from .optimizers import (
    OPTIM_REGISTRY,
    CellFilter,
    CellFireState,
    FireState,
    Optimizer,
    fire_init,
    fire_step,
)
from .runners import (
    InFlightAutoBatcher,
    generate_energy_convergence_fn,
    generate_force_convergence_fn,
    integrate,
    optimize,
)
from .state import SimState, concatenate_states, initialize_state, select_systems
from .trajectory import TorchSimTrajectory, TrajectoryReporter
from .units import MetalUnits

__version__ = "0.1.0"
