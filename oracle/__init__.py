"""CPU float64 oracle for the batched classical force-and-step path.

TEST INFRASTRUCTURE ONLY.  This package is a plain-PyTorch (CPU, float64)
restatement of the TorchSim reference algorithms that the HIP product path
replaces.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or as the timed CPU baseline).  The product package ``torch_sim_amd`` never
imports it and has no CPU fallback.

Every function cites the reference file:line it restates.  How each function
is pinned (golden vectors from the reference tests, closed forms, scipy,
brute force, invariants) is listed in ``PINNING`` below; a function without an
entry there is "parity unpinned".

Modules:
    nl      -- wrap, image-count, brute-force and linked-cell neighbor lists,
               full->half reduction (torch_sim/transforms.py,
               torch_sim/neighbors/torch_nl.py, models/pair_potential.py).
    models  -- LJ / Morse / soft-sphere pair functions and the half-list
               pair-potential forward with autograd forces and virial stress
               (torch_sim/models/*.py).
    md      -- Maxwell-Boltzmann momenta, velocity Verlet, BAOAB Langevin
               (torch_sim/integrators/{md,nve,nvt}.py).
    fire    -- ASE-FIRE / VV-FIRE and the Frechet cell filter
               (torch_sim/optimizers/{fire,cell_filters}.py).
    mat33   -- 3x3 expm / logm / Frechet derivative (torch_sim/math.py),
               checked against scipy.linalg.
    units   -- MetalUnits constants (torch_sim/units.py).
    rng     -- Philox4x32-10 + Box-Muller Langevin normals of the fused engine
               (the reference draws torch.randn, integrators/nvt.py:74-83).
    thermo  -- Nose-Hoover chain, CSVR, anisotropic NPT Langevin, MTK
               Nose-Hoover NPT and stochastic cell rescaling
               (torch_sim/integrators/{md,nvt,npt}.py).
"""

PINNING = {
    "nl.wrap_positions": "reference KATs test_transforms.py:358-416 (golden fixture)",
    "nl.num_cell_repeats": "reference KATs test_transforms.py:993-1025; closed form",
    "nl.brute_force_nl": "exhaustive enumeration; closed-form fcc shell counts "
    "(Ar 48 pairs, test_transforms.py:1334-1368; 78-neighbour shells); "
    "lattice-image invariance (test_neighbors.py:232-304)",
    "nl.linked_cell_nl": "set-equal to brute_force_nl on random/triclinic inputs",
    "nl.full_to_half_list": "reference KAT test_pair_potential.py:102-124",
    "models.*_pair": "closed forms test_lennard_jones.py / test_morse.py / "
    "test_soft_sphere.py; finite differences",
    "models.pair_forward": "perfect-lattice closed sums (E/atom, F=0); F=-dE/dx and "
    "stress=-(1/V)dE/deps by finite differences; half==full; sum F=0",
    "models.MultiPairFn": "species lookup restated from test_soft_sphere.py:84-116 with exact "
    "values (0.125 / 0.28125 / 0.3472); row order follows the atomic_numbers list; eps and "
    "alpha matrices used; LJ branch V_ab(2^(1/6) sigma_ab) = -eps_ab with zero slope; dimers "
    "of every species pair through pair_forward against the written-out LJ energy and force",
    "models.pair_forward per-atom energies / stresses": "perfect fcc lattice: every atom's "
    "energy is the full shell sum, the stress is N/(6V) sum_s n_s r_s V'(r_s) I and each "
    "atom carries 1/N of it (at a = 3.61 and compressed 3.50); on a jittered triclinic batch "
    "sum energies = 2E and sum stresses = the system stress per system",
    "models.particle_life_force": "test_particle_life.py:13-35 (inside beta negative, zero "
    "from sigma on, linear in A) and the closed form: r/beta - 1, the tent's zero at beta "
    "and its peak A at (1 + beta) / 2",
    "models.pair_forces_forward": "non-periodic dimer written out (forces -2f r_hat / +2f "
    "r_hat on a full list, -f / +f on a half list, stress -(f r / V) x x on both); with "
    "f = -dV/dr of Morse on a jittered Cu batch the half-list forces, stress and per-atom "
    "stresses equal pair_forward's and the full list doubles the forces only",
    "md.*": "BAOAB one-step closed form with supplied noise; NVE energy drift; "
    "COM removal; batch==single",
    "md.gjf_particle_position / gjf_particle_velocity": "harmonic oscillator at a large dt: "
    "the GJF update samples <x^2> = kT / k and <v^2> = kT / m (the property that fixes its "
    "coefficients a, b)",
    "md.gjf_strain_position / gjf_strain_velocity": "harmonic strain potential: <eps^2> = "
    "kT / k and <eps_dot^2> = kT / Q",
    "md.strain_force": "perfect lattice: 3 V (P_virial + n kT / V - P_ext) with the "
    "lattice's closed-form virial pressure",
    "md.npt_langevin_isotropic_step": "without forces and noise the step is an isotropic "
    "scaling: positions scale by exp(eps' - eps), the cell is exp(eps') h0, the strain "
    "velocity follows the kinetic pressure alone",
    "md.kinetic_energy / md.calc_kT": "constructed momenta p = sqrt(m) u give KE = 0.5 sum "
    "|u|^2 and kT = sum |u|^2 / (3n) per system; Maxwell-Boltzmann momenta at kT with the COM "
    "removed measure kT (n-1)/n",
    "fire.ase_fire_step": "P<0 branch (test_optimizers.py:600-662); mixing known "
    "answers with per-system norms; parallel-velocity identity; dt_max and max_step "
    "clamps; first step uses raw forces; Frechet first step = scalar exp scaling",
    "fire.vv_fire_step": "P<0 zeroes velocities; n_pos ordering vs ASE flavour "
    "(fire.py:244-246 vs 342-345)",
    "fire.fire_init / compute_cell_forces": "isotropic-stress closed forms for the "
    "Frechet and unit filters; hydrostatic / constant-volume projections; Frechet "
    "adjoint identity vs the 9-direction sum",
    "gradients (models.pair_forward under torch.autograd, create_graph=True)": "central "
    "finite differences of the force-matching loss and of the energy",
    "NVE adjoint (autograd through md.nve_step)": "central finite differences of "
    "E(x_T) wrt the pair parameters; checkpoint-interval invariance",
    "mat33.*": "scipy.linalg.expm / logm / expm_frechet",
    "rng.philox4x32_10": "Random123 known-answer vectors (counter/key all zero, all ones, "
    "and the pi-digit pair)",
    "rng.langevin_normals": "Box-Muller identities (|z0, z1| radius from u1, angle from u2); "
    "moments over 2^18 draws",
    "thermo.SY_WEIGHTS": "each set sums to 1 and is symmetric; float32 as the reference's "
    "default-dtype tensors hold them",
    "thermo.nhc_* / nvt_nose_hoover_step / nvt_nose_hoover_invariant": "harmonic "
    "oscillators: the extended energy is conserved over 3000 steps and a cold start heats to "
    "dof kT / 2; one substep and 3 x 3 Suzuki-Yoshida substeps agree to O(dt^3)",
    "thermo.vrescale_update / nvt_vrescale_step": "ideal gas: the kinetic energy samples the "
    "canonical gamma distribution (mean dof kT / 2, variance dof kT^2 / 2)",
    "thermo.aniso_strain_force / npt_langevin_anisotropic_step": "the per-axis strain forces "
    "sum to md.strain_force; at balanced pressure without noise or friction the strains move "
    "ballistically and positions and cell rows scale per Cartesian axis",
    "thermo.npt_nose_hoover_isotropic_step / _invariant": "soft-sphere crystal near its "
    "virial pressure: the extended Hamiltonian is conserved over 300 steps",
    "thermo.crescale_barostat / npt_crescale_step / rotate_gram_schmidt": "balanced pressure "
    "without noise leaves the cell and drifts positions by p dt / m in every mode; the "
    "triclinic deformation preserves volume (det(cell') = det(cell) (sqrt V' / sqrt V)^2); "
    "Gram-Schmidt rows keep the columns' lengths and angles",
}
