/*
 * tsx.h -- C ABI of the MI355X-native batched classical force-and-step path.
 *
 * One shared library (libtsx.so, gfx950) exports every entry point below.
 * Conventions shared by all functions:
 *   - All pointers are device pointers allocated by the caller (PyTorch);
 *     the library never allocates or frees user memory.  Scratch space comes
 *     in through explicit workspace arguments sized by the *_ws_bytes calls.
 *   - ``stream`` is a hipStream_t passed as void*.  Work is enqueued
 *     asynchronously on it; nothing synchronizes the device, nothing reads
 *     back to the host, so every call is capturable in a hipGraph.
 *   - Return value: 0 on success, <0 on a bad argument (TSX_ERR_ARG) or a HIP
 *     error (TSX_ERR_HIP); tsx_last_error() returns a thread-local message.
 *     Capacity problems discovered on the device (neighbor-list or bin
 *     overflow) are reported through the int32 ``status`` word that the
 *     caller reads when it synchronizes (TSX_STATUS_* bits).
 *   - Cells are the column-vector SimState cells [S,3,3] (lattice vector k is
 *     column k), float64.  Positions are float64 [N,3] (stride 3) or the
 *     packed engine layout (stride 4).  Atoms of one system are contiguous;
 *     ``sys_start`` [S+1] (int32) gives the atom range of each system and
 *     ``atom_sys`` [N] (int32) the system of each atom.
 *
 * Reference interfaces replaced (TorchSim v0.6.1, paths relative to the
 * reference repository):
 *   tsx_nl_build / tsx_nl_count_pairs / tsx_nl_emit_pairs
 *       torch_sim/neighbors/__init__.py:46-98 (torchsim_nl) and
 *       torch_sim/neighbors/torch_nl.py:156-215 (torch_nl_linked_cell),
 *       torch_sim/transforms.py:160-210 (wrap), 435-467 (image counts),
 *       torch_sim/models/pair_potential.py:86-127 (full_to_half_list)
 *   tsx_pair_forces / tsx_system_reduce
 *       torch_sim/models/pair_potential.py:130-464 (PairPotentialModel.forward)
 *       with the pair functions of lennard_jones.py:23-47, morse.py:23-47,
 *       soft_sphere.py:35-59 and 90-167
 *   tsx_md_* / tsx_md_fused_step
 *       torch_sim/integrators/md.py:113-208, nve.py:74-130, nvt.py:27-240
 *   tsx_fire_pre / tsx_fire_post
 *       torch_sim/optimizers/fire.py:21-499,
 *       torch_sim/optimizers/cell_filters.py:26-60, 137-178, 180-284, 353-399,
 *       torch_sim/math.py:106-189, 614-778
 *   tsx_pair_grad / tsx_nve_adjoint_update
 *       autograd through pair_potential.py:351-464 with create_graph=True
 *       (the reference itself detaches dV/dr, pair_potential.py:419-425) and
 *       through the NVE loop of nve.py:74-130
 */
#ifndef TSX_H_
#define TSX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSX_OK 0
#define TSX_ERR_ARG (-1)
#define TSX_ERR_HIP (-2)

/* device status word bits */
#define TSX_STATUS_NL_OVERFLOW 1      /* an atom has more neighbors than kmax */
#define TSX_STATUS_BIN_OVERFLOW 2     /* total bins exceed bin_cap            */
#define TSX_STATUS_SHIFT_OVERFLOW 4   /* image shift outside the packed range */
#define TSX_STATUS_SINGULAR_CELL 8    /* periodic system with singular cell   */
#define TSX_STATUS_REBUILD 16         /* engine: list must be rebuilt         */
#define TSX_STATUS_BAD_SIZE 32        /* a system exceeds max_system_atoms    */

/* pair potential kinds */
#define TSX_LJ 0          /* 4 eps [(s/r)^12 - (s/r)^6]                  */
#define TSX_MORSE 1       /* eps (1 - exp(-a (r - s)))^2 - eps            */
#define TSX_SOFT_SPHERE 2 /* eps/a (1 - r/s)^a for r < s                  */
/* particle life: a force without an energy (tsx_pair_forces only), with
 * (sigma, epsilon, alpha) = (outer radius s, amplitude A, inner radius b):
 *   f(r) = r/b - 1 (r < b),  A (1 - |2r - 1 - b| / (1 - b)) (b <= r < s),  0
 * and the force on j due to i f(r) (r_j - r_i) / r
 * (torch_sim/models/particle_life.py:24-56, pair_potential.py:467-612) */
#define TSX_PARTICLE_LIFE 3

/* Library version (major*10000 + minor*100 + patch). */
int tsx_version(void);

/* Thread-local message of the last failing call (empty string if none). */
const char* tsx_last_error(void);

/* ------------------------------------------------------------------ NL --
 * Cell-list neighbor list over concatenated systems.
 *
 * Per system: fractional binning with nb_k = max(1, floor(q w_k / r_list))
 * bins along each periodic axis (w_k = distance between lattice planes,
 * q = bins_per_rlist), one bin along non-periodic axes; stencil half-width
 * m_k = ceil(r_list nb_k / w_k).  Atoms are binned by count (atomics),
 * prefix-scanned, scattered into bin order and sorted by atom index within a
 * bin (deterministic), then every atom walks its stencil and keeps the
 * images with FP64 |wr_i - wr_j - S.L|^2 < r_list^2 on wrapped coordinates.
 *
 * Output is an ELL list in "thread order": row t of the list belongs to
 * atom order[t] (atoms sorted by bin).  Entry n of row t lives at
 *     nbr[(n / tpa) * (ld * tpa) + t * tpa + (n % tpa)]
 * and packs the neighbor's system-local index in the low ``jbits`` bits and
 * the unwrapped-frame image shift S' = S + w_i - w_j (valid for the input
 * positions) as three ``sbits``-bit two's-complement fields above it
 * (sbits = (32 - jbits) / 3; a zero shift is a zero high part).
 * ``count[t]`` is the number of entries; overflowing rows set
 * TSX_STATUS_NL_OVERFLOW and are truncated at kmax.
 *
 * half != 0 keeps one copy of every unordered pair (i < j, or i == j with a
 * lexicographically positive S'), as full_to_half_list does.
 *
 * ``setup`` != 0 recomputes the per-system geometry (inverse cell, grid,
 * bin offsets) and clears the bin counters; a caller whose cells do not
 * change passes 1 on the first build only.  The workspace keeps the geometry
 * between builds and the bin counters are left cleared after every build.
 *
 * ``gate`` (device int32 [S], may be NULL = all systems): only systems with
 * gate[s] != 0 are rebuilt; their rows, order entries and bins are rewritten
 * and every other system's list is left untouched.  A captured graph can
 * therefore carry a rebuild that runs only for the systems whose atoms left
 * the Verlet skin.  ``counter`` (may be NULL) is incremented once per rebuilt
 * system.  ``xref`` (stride 4, may be NULL) receives a copy of the positions
 * of the rebuilt systems (the displacement reference of the skin) and
 * ``cell_build`` ([S,3,3], may be NULL) a copy of their cells.
 *
 * ``max_system_atoms`` (0 = unknown): when every system has at most 2048
 * atoms, one workgroup per rebuilt system bins, scans, scatters and sorts it
 * (one launch); otherwise these stages run as grid-wide kernels.
 *
 * ``fill_mode``: TSX_FILL_BIN stages each bin's stencil candidates in LDS
 * (one 256-thread workgroup per bin: dense lists, partial rebuilds of many
 * systems);
 * TSX_FILL_ATOM gives each atom a group of lanes reading its stencil columns
 * from L2 (sparse bins, full builds of large systems); TSX_FILL_WAVE (stencil
 * words only) gives each active bin one wave that tests every candidate of
 * the bin's stencil against all of the bin's atoms at once (the MD engine's
 * gated rebuilds); TSX_FILL_AUTO picks BIN for bins_per_rlist >= 2,
 * kmax >= 64 and more than one system, and WAVE for stencil words with
 * bins_per_rlist >= 2 and kmax >= 64 (one large system included).  WAVE
 * tests up to 8 centre atoms per pass over the candidates when n_atoms is at
 * least 3 bin_cap (dense bins), else 4; TSX_FILL_WAVE_SPARSE /
 * TSX_FILL_WAVE_DENSE fix that width.
 * All write every row in the same order (stencil image, then atom index).
 */
#define TSX_FILL_AUTO 0
#define TSX_FILL_BIN 1
#define TSX_FILL_ATOM 2
#define TSX_FILL_WAVE 3
#define TSX_FILL_WAVE_SPARSE 4
#define TSX_FILL_WAVE_DENSE 5

/* ``word_mode``: TSX_WORDS_SHIFT writes the image shift S' in the three word
 * fields (the layout above, read by every consumer of the reference triple);
 * TSX_WORDS_STENCIL writes instead the stencil offset d of the neighbour's
 * bin relative to the centre's bin (|d_k| <= m_k), for the fused MD step,
 * which then needs a bin-relative frame (full lists only).  For an entry
 * (i, j, d) with anchors A (below) the displacement is
 *     r_j - r_i + S'.L = (r_j - A_j) - (r_i - A_i) + sum_k d_k ustep_k,
 * every term of which is of the order of a bin, so it is accurate in FP32. */
#define TSX_WORDS_SHIFT 0
#define TSX_WORDS_STENCIL 1

/* Bin-relative frame written by a build (any pointer may be NULL but
 * e_atom says whether the next four are used):
 *   anchor [N,4] float64, by atom index: A_i = sum_k (w_ik + b_ik / nb_k) a_k,
 *       the corner of atom i's bin in its unwrapped image (w = wrap shift,
 *       b = bin, a_k = lattice vector k); along a non-periodic direction the
 *       coefficient is the fractional coordinate of the system's first atom,
 *       and a system with no periodic axis (or a singular cell) anchors at
 *       its first atom;
 *   rel [N,4] float32, by atom index: r_i - A_i rounded to float, and the low
 *       32 bits of the packed species word of a stride-4 position;
 *   ustep [S,12] float64: entries 3k..3k+2 = a_k / nb_k, entry 9 =
 *       sum_k |a_k / nb_k| (written by setup builds only).
 * Anchor and rel are written for the rebuilt systems.
 *
 * e_atom [N] float64, energy [S], gate_clear [S], step [1] (optional, NULL
 * e_atom: unused): the per-system reduction of the MD step before this
 * rebuild, done by extra workgroups of the build's one-workgroup-per-system
 * kernel (beside the builds) instead of a launch of its own.  Every system s, rebuilt or not, gets energy[s] = the
 * sum of e_atom over its atoms in the fixed order of tsx_md_reduce (bitwise
 * the same value), gate_clear[s] = 0, and step[0] advances by one.  Needs
 * 0 < max_system_atoms <= 2048 and n_atoms > 0 (the one-workgroup build). */
typedef struct tsx_nl_frame {
  double* anchor;
  float* rel;
  double* ustep;
  const double* e_atom;
  double* energy;
  int32_t* gate_clear;
  int64_t* step;
} tsx_nl_frame;

int64_t tsx_nl_ws_bytes(int64_t n_atoms, int32_t n_systems, int32_t bin_cap);

int tsx_nl_build(const double* pos, int32_t pos_stride, const double* cell_col,
                 const uint8_t* pbc, int32_t pbc_per_system, const int32_t* sys_start,
                 const int32_t* atom_sys, int64_t n_atoms, int32_t n_systems,
                 double r_list, int32_t bins_per_rlist, int32_t self_interaction,
                 int32_t half, int32_t jbits, int32_t tpa, int32_t kmax, int64_t ld,
                 int32_t bin_cap, int32_t setup, void* ws, uint32_t* nbr, int32_t* count,
                 int32_t* order, double* xref, int32_t* status, const int32_t* gate,
                 int32_t* counter, double* cell_build, int32_t fill_mode, int32_t word_mode,
                 const tsx_nl_frame* frame, int32_t max_system_atoms, void* stream);

/* tsx_nl_build without a workspace argument: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the workspace is a static device array of 36 MiB inside the
 * library, one per GPU, held until the process ends; the call itself
 * allocates nothing.  Limits: n_atoms <= TSX_NOALLOC_MAX_ATOMS, n_systems <=
 * TSX_NOALLOC_MAX_SYSTEMS, bin_cap <= TSX_NOALLOC_MAX_BINS; larger sizes
 * return TSX_ERR_ARG with a message.  Every call rebuilds the per-system
 * geometry (setup = 1 of tsx_nl_build), because the array keeps whatever the
 * previous call on the device left in it; outputs are those of tsx_nl_build
 * with setup = 1.  All calls on one GPU share the array, so calls that
 * overlap on that GPU (two streams, two threads) race: serialise them.  The
 * first call on each device looks the array's address up
 * (hipGetSymbolAddress); make that first call outside graph capture. */
#define TSX_NOALLOC_MAX_ATOMS 524288
#define TSX_NOALLOC_MAX_SYSTEMS 4096
#define TSX_NOALLOC_MAX_BINS 524288
int tsx_nl_build_noalloc(const double* pos, int32_t pos_stride, const double* cell_col,
                         const uint8_t* pbc, int32_t pbc_per_system, const int32_t* sys_start,
                         const int32_t* atom_sys, int64_t n_atoms, int32_t n_systems,
                         double r_list, int32_t bins_per_rlist, int32_t self_interaction,
                         int32_t half, int32_t jbits, int32_t tpa, int32_t kmax, int64_t ld,
                         int32_t bin_cap, uint32_t* nbr, int32_t* count, int32_t* order,
                         double* xref, int32_t* status, const int32_t* gate, int32_t* counter,
                         double* cell_build, int32_t fill_mode, int32_t word_mode,
                         const tsx_nl_frame* frame, int32_t max_system_atoms, void* stream);

/* Exclusive scan of ``count`` into ``row_ptr`` [n+1] (int64); the total is
 * row_ptr[n].  Used before tsx_nl_emit_pairs.  ``ws`` holds
 * tsx_scan_ws_bytes(n) bytes. */
int64_t tsx_scan_ws_bytes(int64_t n);
int tsx_nl_count_pairs(const int32_t* count, int64_t n, int64_t* row_ptr, void* ws,
                       void* stream);

/* Expand the ELL list into the reference triple: mapping [2,P] int64
 * (mapping[0] = i, mapping[1] = j), system_mapping [P] int64 and
 * shifts [P,3] float64, in row order. */
int tsx_nl_emit_pairs(const uint32_t* nbr, const int32_t* count, const int32_t* order,
                      const int64_t* row_ptr, const int32_t* sys_start,
                      const int32_t* atom_sys, int64_t n_atoms, int32_t jbits,
                      int32_t tpa, int64_t ld, int64_t* mapping, int64_t n_pairs,
                      int64_t* system_mapping, double* shifts, void* stream);

/* Verlet-skin validity of a list built at cell_build [S,3,3] (column) with
 * positions xref [N,4] and radius r_list, for the current positions and
 * cells: flags[s] = 1 when system s must be rebuilt (an atom pulled back to
 * the build cell moved more than (r_list - cutoff / sigma) / 2, sigma a lower
 * bound of the cell map's smallest singular value), else 0.  Replaces the
 * rebuild-every-call of the reference NL (neighbors/__init__.py:46-98) for
 * callers that keep a list with skin.  ws: tsx_nl_check_ws_bytes. */
int64_t tsx_nl_check_ws_bytes(int32_t n_systems);
int tsx_nl_check(const double* pos, int32_t pos_stride, const double* cell_col,
                 const double* cell_build, const double* xref, const int32_t* atom_sys,
                 int64_t n_atoms, int32_t n_systems, double r_list, double cutoff,
                 int32_t* flags, void* ws, void* stream);

/* Wrapped positions [N,3] and integer lattice shifts [N,3] (float64), as
 * pbc_wrap_batched_and_get_lattice_shifts (transforms.py:160-210).  ``ws``
 * holds tsx_wrap_ws_bytes(n_systems) bytes. */
int64_t tsx_wrap_ws_bytes(int32_t n_systems);
int tsx_wrap_positions(const double* pos, const double* cell_col, const uint8_t* pbc,
                       int32_t pbc_per_system, const int32_t* atom_sys, int64_t n_atoms,
                       int32_t n_systems, void* ws, int32_t* status, double* wrapped,
                       double* shifts, void* stream);

/* ---------------------------------------------------------------- pair --
 * Pair energies / forces / virials over the ELL list (full or half).
 * Pair arithmetic in FP32 on FP64 displacement vectors; the FP64 cutoff
 * test d^2 < rc^2 matches the reference's ``distances < cutoff`` mask.
 * Accumulation of forces, energies and virials is FP64.
 *   params: per-species-pair table [n_species^2][3] float64 of
 *           (sigma, epsilon, alpha); species [N] int32 (may be NULL when
 *           n_species == 1).
 *   forces  [N,3] float64 (written, not accumulated)
 *   e_atom  [N]   float64 per-atom energy (written)
 *   w_atom  [N,6] float64 per-atom virial sum_j 0.5 dr (x) f, (xx,yy,zz,yz,xz,xy)
 *           or NULL.
 * For a full list each ordered pair contributes to atom i only (no atomics);
 * for a half list (half=1) the j-side contributions use FP64 atomics.
 */
int tsx_pair_forces(int32_t kind, const double* params, int32_t n_species,
                    const int32_t* species, const double* pos, int32_t pos_stride,
                    const double* cell_col, const int32_t* sys_start,
                    const int32_t* atom_sys, const uint32_t* nbr, const int32_t* count,
                    const int32_t* order, int64_t n_atoms, int32_t jbits, int32_t tpa,
                    int64_t ld, double cutoff, int32_t half, double* forces,
                    double* e_atom, double* w_atom, void* stream);

/* Per-system reductions of per-atom energies and virials:
 *   energy [S] = sum e_atom;  stress [S,3,3] = -(sum w_atom) / |det cell|;
 *   stresses [N,3,3] (optional) = -w_atom / |det cell|.
 * Chunks of 2048 atoms counted from each system's first atom reduce in
 * parallel (one 1M-atom system or many small ones); a system's sum order
 * depends only on its size, so it is bitwise the same alone or in a batch. */
int tsx_system_reduce(const double* e_atom, const double* w_atom, const double* cell_col,
                      const int32_t* sys_start, int32_t max_system_atoms, int32_t n_systems,
                      double* energy, double* stress, double* stresses, void* ws,
                      void* stream);

/* Workspace bytes for tsx_system_reduce / tsx_md_reduce (max_system_atoms =
 * the largest system): chunk partials of systems above 2048 atoms. */
int64_t tsx_reduce_ws_bytes(int32_t max_system_atoms, int32_t n_systems);


/* ------------------------------------------------------------------ MD --
 * Per-step kernels (in place, float64 [N,3] momenta / positions):
 *   tsx_md_kick     p <- p + F dt                       (md.py:169-187)
 *   tsx_md_drift    x <- x + (p/m) dt                   (md.py:190-208)
 *   tsx_md_vv_half  B(dt/2) A(dt) of velocity Verlet    (nve.py:121-123)
 *   tsx_md_baoa     B(dt/2) A(dt/2) O(dt) A(dt/2)       (nvt.py:229-233, 27-86)
 *                   O: p <- c1 p + c2 sqrt(m) xi, c1 = exp(-gamma dt),
 *                   c2 = sqrt(kT (1 - c1^2)); kT and gamma per system [S];
 *                   xi = noise [N,3] when given, else Philox4x32-10 keyed by
 *                   (seed, step, atom index).
 */
int tsx_md_kick(double* p, const double* f, double dt, int64_t n_atoms, void* stream);
int tsx_md_drift(double* x, const double* p, const double* m, double dt, int64_t n_atoms,
                 void* stream);
int tsx_md_vv_half(double* x, double* p, const double* f, const double* m, double dt,
                   int64_t n_atoms, void* stream);
int tsx_md_baoa(double* x, double* p, const double* f, const double* m, const int32_t* atom_sys,
                const double* kT, const double* gamma, double dt, const double* noise,
                uint64_t seed, uint64_t step, int64_t n_atoms, void* stream);

/* Fused step over packed [N,4] positions (x, y, z, species bits): pair forces
 * on pos_in, the closing B(dt/2) of the current step when ``end_kick``, and,
 * when ``do_next``, the opening half of the next step (B A O A for Langevin,
 * B A(dt) for NVE) written to pos_out (which must differ from pos_in); else
 * pos_out = pos_in.  An atom of system s farther than ``half_skin`` from its
 * ``xref`` sets flag_set[s] = 1 (the rebuild gate of the next step).
 * ``step`` is a device counter giving the Philox step of the next step's
 * noise.  w_atom may be NULL.
 * The list must be a TSX_WORDS_STENCIL list with its frame (anchor, ustep and
 * rel_in = the rel of pos_in, i.e. (float)(pos_in - anchor) per component,
 * exactly as tsx_nl_build and the previous fused step write it; the centre
 * atom's FP32 position is read from it too): each neighbour costs one
 * 16-byte gather of rel; the FP32 displacement decides the cutoff outside a
 * proven rounding margin and the FP64 displacement (r_j - A_j) - (r_i - A_i)
 * + d.ustep decides inside it.  rel_out receives the rel of pos_out.
 * ld * tpa must stay below 2^27 and n_atoms below 2^28 (32-bit buffer
 * offsets of the word and record loads), else TSX_ERR_ARG.
 * ``orthorhombic`` nonzero promises that every system's cell is orthorhombic
 * (the off-diagonal entries of cell_col, hence of ustep, are exactly zero):
 * the stencil shift d.ustep then costs one product per component, with the
 * same bits as the general form.  Pass 0 when unsure.
 * ``stream_words`` nonzero loads the list words non-temporally.  Each word is
 * read once per step; when a system is so large that its rows reuse a
 * neighbour's position only after many megabytes of words (one 1M-atom box),
 * streaming them keeps those positions in L2 (8 % faster on the 1M-atom LJ
 * box); for batches of small systems the reuse is close and plain loads are
 * as fast.  It applies to single-species lists without the virial (the
 * other variants always load plainly).  The results do not depend on it. */
int tsx_md_fused_step(int32_t kind, const double* params, int32_t n_species,
                      const double* pos_in, double* pos_out, const double* xref,
                      const double* cell_col, const int32_t* sys_start,
                      const int32_t* atom_sys, const uint32_t* nbr, const int32_t* count,
                      const int32_t* order, int64_t n_atoms, int32_t jbits, int32_t tpa,
                      int64_t ld, double cutoff, double* p, const double* m, double* f,
                      double* e_atom, double* w_atom, double dt, int32_t langevin,
                      const double* kT, const double* gamma, const double* noise,
                      uint64_t seed, const int64_t* step, int32_t end_kick, int32_t do_next,
                      double half_skin, int32_t* flag_set, const float* rel_in, float* rel_out,
                      const double* anchor, const double* ustep, int32_t orthorhombic,
                      int32_t stream_words, void* stream);

/* Per-system energy [S] = sum of e_atom (the chunked fixed-order reduction
 * of tsx_system_reduce, same workspace); gate_clear[s] = 0 when given (the
 * step's rebuild gates, read by the list build before the fused step); when
 * ``advance``, *step += 1. */
int tsx_md_reduce(const double* e_atom, const int32_t* sys_start, int32_t max_system_atoms,
                  int32_t n_systems, double* energy, int64_t* step, int32_t advance,
                  int32_t* gate_clear, void* ws, void* stream);

/* Copy n_doubles of pos into xref when *gate != 0 (or gate is NULL) and count
 * the copies in *counter (may be NULL): the displacement reference of the
 * Verlet skin, refreshed together with a gated list rebuild. */
int tsx_copy_ref(const double* pos, double* xref, int64_t n_doubles, const int32_t* gate,
                 int32_t* counter, void* stream);


/* ------------------------------------------------------------ FIRE ------ */

#define TSX_FIRE_ASE 0      /* fire_flavor="ase_fire" (fire.py:288-499)    */
#define TSX_FIRE_VV 1       /* fire_flavor="vv_fire"  (fire.py:179-285)    */
#define TSX_CELL_NONE 0     /* positions only                              */
#define TSX_CELL_UNIT 1     /* CellFilter.unit    (cell_filters.py:180-227) */
#define TSX_CELL_FRECHET 2  /* CellFilter.frechet (cell_filters.py:230-284) */
#define TSX_MAX_LOG_DEFORM 2.0 /* cell_filters.py:23 */

/* fire_step keyword arguments (fire.py:117-128) plus the cell-filter options
 * (cell_filters.py:230-240).  eps is 1e-16 for float64 (fire.py:149).
 * first_step: every velocity is NaN (fire.py:313-315): raw forces, no mixing
 * and no dt / alpha update on this step (ASE flavour). */
typedef struct tsx_fire_params {
  double dt_max, f_inc, f_dec, alpha_start, f_alpha, max_step, eps;
  int32_t n_min, flavor, cell_filter, hydrostatic_strain, constant_volume, first_step;
} tsx_fire_params;

/* First half of a FIRE step, one workgroup per system (replaces the part of
 * _ase_fire_step before the model call, fire.py:306-481, or of _vv_fire_step,
 * fire.py:193-212).  Updates vel, pos (stride 3 or 4; column 3 untouched),
 * dt, alpha, n_pos and, under a cell filter, cell_col, cell_pos, cell_vel.
 * active (may be NULL): systems with active[s] == 0 are left untouched.
 * List check (flags non-NULL): flags[s] = 1 when an atom of s, pulled back to
 * the cell of the last list build (cell_build), moved more than the distance
 * that keeps a list of radius r_list valid for cutoff.  ws: tsx_fire_ws_bytes
 * of device scratch (per-system step data between the pre kernels). */
int64_t tsx_fire_ws_bytes(int32_t n_systems);
int tsx_fire_pre(const tsx_fire_params* prm, int32_t n_systems, const int32_t* sys_start,
                 const int32_t* atom_sys, int64_t n_atoms, double* pos, int32_t pos_stride,
                 double* vel, const double* forces, const double* masses, double* dt,
                 double* alpha, int32_t* n_pos, double* cell_col, const double* ref_cell,
                 const double* cell_factor, double* cell_pos, double* cell_vel,
                 const double* cell_forces, const int32_t* active, const double* cell_build,
                 const double* xref, double r_list, double cutoff, int32_t* flags, void* ws,
                 void* stream);

/* Second half of a FIRE step after the model call: cell forces from stress
 * (compute_cell_forces, cell_filters.py:353-399; with init != 0 the filter
 * init functions, cell_filters.py:180-284), the VV flavour's second half
 * kick / power / mixing (fire.py:227-285, skipped when init != 0), and
 * fmax[2*s] = max_i |F_i|, fmax[2*s+1] = max row norm of the cell forces
 * (quantities.py:327-342, runners.py:527-534).  cell_factor [S], pressure and
 * stress [S,3,3], cell_masses [S] (total mass).  flags (may be NULL): for
 * systems whose list was rebuilt, cell_build[s] = cell[s] and flags[s] = 0. */
int tsx_fire_post(const tsx_fire_params* prm, int32_t n_systems, const int32_t* sys_start,
                  double* vel, const double* forces, const double* masses, double* dt,
                  double* alpha, int32_t* n_pos, const double* cell_col, const double* ref_cell,
                  const double* cell_factor, const double* pressure, const double* stress,
                  double* cell_vel, double* cell_forces, const double* cell_masses,
                  const int32_t* active, double* fmax, double* cell_build, int32_t* flags,
                  int32_t init, void* stream);


/* ------------------------------------------------- backward (gradients) --- */

/* Workspace bytes for tsx_pair_grad (block partials of the parameter
 * gradient and per-atom cell gradients). */
int64_t tsx_grad_ws_bytes(int64_t n_atoms, int32_t tpa, int32_t n_species);

/* Gradients of the pair model over the full ELL list of tsx_nl_build
 * (replaces torch.autograd through PairPotentialModel.forward with
 * create_graph=True, torch_sim/models/pair_potential.py:351-464).
 * With w [N,3] (cotangent of the forces): the vector-Jacobian product of the
 * forces, dpos[i] = sum_j H(d_ij) (w_j - w_i) (a Hessian-vector product),
 * dcell = d<w,F>/d cell, dparams = d<w,F>/d params.  With w == NULL: the
 * gradient of sum_s ebar[s] E_s (ebar NULL = all ones): dpos = -ebar F,
 * dcell, dparams.  params / dparams are [n_species, n_species, 3]
 * (sigma, epsilon, alpha); dparams is returned symmetrised, so
 * dparams[a,b] + dparams[b,a] (a != b) is the derivative with respect to
 * the shared parameter of the species pair.  dcell is the gradient with
 * respect to the column-vector cell [S,3,3] at fixed Cartesian positions.
 * Any of dpos / dcell / dparams may be NULL; accumulate != 0 adds into dcell
 * and dparams instead of overwriting (dpos is always overwritten).  Parameter gradients support up to 4
 * species; sums are fixed-order (bitwise reproducible). */
int tsx_pair_grad(int32_t kind, const double* params, int32_t n_species, const int32_t* species,
                  const double* pos, int32_t pos_stride, const double* cell_col,
                  const int32_t* sys_start, const int32_t* atom_sys, int32_t n_systems,
                  const uint32_t* nbr, const int32_t* count, const int32_t* order,
                  int64_t n_atoms, int32_t jbits, int32_t tpa, int64_t ld, double cutoff,
                  const double* w, const double* ebar, double* dpos, double* dcell,
                  double* dparams, int32_t accumulate, void* ws, void* stream);

/* ------------------------------------------- isotropic NPT Langevin (GJF) --- */

/* Workspace bytes for the NPT step: per-system strain force, cell noise,
 * old strain and scale (4 doubles) and the per-atom noise beta (3 doubles),
 * carried from _pre to _post. */
int64_t tsx_npt_ws_bytes(int64_t n_atoms, int32_t n_systems);

/* First half of one isotropic NPT Langevin step (replaces
 * torch_sim/integrators/npt.py:987-1070, the cell and particle position
 * updates of npt_langevin_isotropic_step; state npt.py:653-701): per system,
 * cell_masses = (n + 1) kT b_tau^2, the strain force 3 V (P - p_ext) with
 * P = -tr(stress) / 3 + n kT / V and V = det(ref_cell) exp(3 eps), the GJF
 * strain update eps -> eps' and cell_col = exp(eps') ref_cell; per atom,
 * beta = sqrt(2 alpha kT dt) noise and the GJF position update scaled by
 * exp(eps' - eps).  cell_noise [S] and noise [N,3] are standard normals.
 * All arrays FP64, per-system ones [S] (cells [S,3,3] column vectors). */
int tsx_npt_langevin_iso_pre(int32_t n_systems, const int32_t* sys_start,
                             const int32_t* atom_sys, int64_t n_atoms, double* pos,
                             const double* mom, const double* forces, const double* masses,
                             const double* kT, const double* p_ext, const double* alpha,
                             const double* cell_alpha, const double* b_tau, double dt,
                             const double* stress, const double* ref_cell, double* eps,
                             const double* eps_dot, double* cell_masses, double* cell_col,
                             const double* cell_noise, const double* noise, void* ws,
                             void* stream);

/* Second half, after the model call at the new positions and cell
 * (npt.py:1072-1107): the strain velocity from the old and new strain
 * forces, and the GJF momentum update from f_old, f_new and the beta of
 * tsx_npt_langevin_iso_pre (same ws). */
int tsx_npt_langevin_iso_post(int32_t n_systems, const int32_t* sys_start,
                              const int32_t* atom_sys, int64_t n_atoms, double* mom,
                              const double* f_old, const double* f_new, const double* masses,
                              const double* kT, const double* p_ext, const double* alpha,
                              const double* cell_alpha, double dt, const double* stress,
                              const double* ref_cell, const double* eps, double* eps_dot,
                              const double* cell_masses, const void* ws, void* stream);

/* First half of one anisotropic NPT Langevin step (replaces
 * torch_sim/integrators/npt.py:513-620, the cell and particle position
 * updates of npt_langevin_anisotropic_step; state npt.py:62-131): the same
 * GJF scheme with one strain eps_k per Cartesian axis ([S,3]), strain forces
 * V (P_kk - p_ext[S,3]) with P_kk = -stress_kk + n kT / V and V = det(cell_col)
 * of the current cell, the cell rows scaled by exp(eps_k') (cell_col updated
 * in place), the particles' components by exp(eps_k' - eps_k).  cell_noise
 * [S,3] and noise [N,3] are standard normals.  ws as tsx_npt_ws_bytes. */
int tsx_npt_langevin_aniso_pre(int32_t n_systems, const int32_t* sys_start,
                               const int32_t* atom_sys, int64_t n_atoms, double* pos,
                               const double* mom, const double* forces, const double* masses,
                               const double* kT, const double* p_ext, const double* alpha,
                               const double* cell_alpha, const double* b_tau, double dt,
                               const double* stress, const double* ref_cell, double* cell_col,
                               double* eps, const double* eps_dot, double* cell_masses,
                               const double* cell_noise, const double* noise, void* ws,
                               void* stream);

/* Second half (npt.py:622-650): per-axis strain velocities from the old and
 * new strain forces (V of the new cell) and the GJF momentum update with the
 * beta of tsx_npt_langevin_aniso_pre (same ws). */
int tsx_npt_langevin_aniso_post(int32_t n_systems, const int32_t* sys_start,
                                const int32_t* atom_sys, int64_t n_atoms, double* mom,
                                const double* f_old, const double* f_new, const double* masses,
                                const double* kT, const double* p_ext, const double* alpha,
                                const double* cell_alpha, double dt, const double* stress,
                                const double* cell_col, double* eps_dot,
                                const double* cell_masses, const void* ws, void* stream);

/* Workspace (bytes) of tsx_md_scale_kick_reduce and tsx_npt_nh_post: chunk
 * partials of up to 6 doubles per (system, 2048-atom chunk). */
int64_t tsx_mom_reduce_ws_bytes(int32_t max_system_atoms, int32_t n_systems);

/* Per atom p <- scale[s] p + h F (scale and forces each optional, NULL =
 * skip), then per system out[s] = factor * sum p^2 / m (tensor = 0), or the
 * six components xx yy zz yz xz xy of sum p p^T / m ([S,6], tensor = 1), over
 * the updated momenta; out NULL = no reduction.  Replaces the momentum
 * rescales of torch_sim/integrators/md.py:488-493 and nvt.py:650-653, the
 * momentum_step of md.py:169-187, calc_kinetic_energy (quantities.py:104-139)
 * and the kinetic tensor of compute_instantaneous_pressure_tensor
 * (quantities.py:156-198).  Sums run in a fixed order per system (2048-atom
 * chunks from the system's first atom), independent of the batch. */
int tsx_md_scale_kick_reduce(int32_t n_systems, const int32_t* sys_start,
                             int32_t max_system_atoms, double* mom, const double* masses,
                             const double* scale, const double* forces, double h,
                             int32_t tensor, double factor, double* out, void* ws,
                             void* stream);

/* Per atom p <- scale[atom_sys] p + h F, x <- x + (p / m) dt: a per-system
 * momentum scale fused with the first half of velocity Verlet
 * (md.py:211-245 after nvt.py:455-459 / 650-653). scale NULL = 1. */
int tsx_md_scale_kick_drift(int64_t n_atoms, const int32_t* atom_sys, double* pos, double* mom,
                            const double* scale, const double* forces, double h,
                            const double* masses, double dt, void* stream);

/* Half a step of a Nose-Hoover chain per system (replaces the chain
 * functions of torch_sim/integrators/md.py:338-565: half_step_chain_fn with
 * substep_fn, and update_chain_mass_fn when update_mass): chain_steps x
 * sy_steps Suzuki-Yoshida substeps over chain positions xi, momenta pxi and
 * masses [S, chain_length]; ke [S] is the coupled kinetic energy (scaled in
 * place), scale_out [S] the product of the momentum scales to apply to the
 * particles, sys_mom [S] (optional) a per-system momentum scaled in place
 * (the barostat's).  tau [S] is needed only with update_mass. */
int tsx_nhc_half_step(int32_t n_systems, int32_t chain_length, int32_t chain_steps,
                      int32_t sy_steps, double dt, const double* kT, const double* tau,
                      const double* dof, double* xi, double* pxi, double* masses, double* ke,
                      int32_t update_mass, double* scale_out, double* sys_mom, void* stream);

/* CSVR momentum scale per system (replaces nvt.py:594-654
 * _vrescale_update): ke2 = sum p^2 / m [S], dof, kT, tau [S]; r1 standard
 * normal, r2 = 2 Gamma((dof - 1) / 2) draws [S]; lam [S] out. */
int tsx_vrescale_factor(int32_t n_systems, const double* ke2, const double* dof,
                        const double* kT, const double* tau, double dt, const double* r1,
                        const double* r2, double* lam, void* stream);

/* Workspace (bytes) of tsx_npt_nh_pre / _post: 8 doubles per system. */
int64_t tsx_npt_nh_ws_bytes(int32_t n_systems);

/* First half of the MTK inner step (replaces torch_sim/integrators/
 * npt.py:1532-1590 and the cell mass of npt.py:1267-1309): per system the
 * cell mass 3 (n + 1) kT b_tau^2, the cell force alpha 2K - tr(stress) V -
 * 3 p_ext V (alpha = 1 + 3 / dof, K = ke2 / 2 of the thermostatted momenta,
 * V = det(ref_cell) exp(3 eps)), cell_mom += dt/2 G, eps += (cell_mom / W)
 * dt and cell_col = (V'/V0)^(1/3) ref_cell; per atom the momentum map
 * exp(iL2 dt/2) with the current forces and the position map exp(iL1 dt). */
int tsx_npt_nh_pre(int32_t n_systems, const int32_t* sys_start, const int32_t* atom_sys,
                   int64_t n_atoms, double* pos, double* mom, const double* forces,
                   const double* masses, const double* kT, const double* b_tau,
                   const double* dof, const double* ke2, const double* stress,
                   const double* p_ext, const double* ref_cell, double* eps, double* cell_mom,
                   double* cell_mass, double* cell_col, double dt, void* ws, void* stream);

/* Second half (npt.py:1591-1640, 1940-1960): exp(iL2 dt/2) with the new
 * forces fused with ke2 = sum p^2 / m, the cell force at the new stress and
 * volume, cell_mom += dt/2 G, ke_therm = ke2 / 2, ke_baro = cell_mom^2 / (2 W).
 * ws from tsx_npt_nh_pre, red_ws as tsx_mom_reduce_ws_bytes. */
int tsx_npt_nh_post(int32_t n_systems, const int32_t* sys_start, int32_t max_system_atoms,
                    double* mom, const double* forces, const double* masses,
                    const double* stress, const double* p_ext, const double* ref_cell,
                    const double* eps, double* cell_mom, const double* cell_mass, double dt,
                    double* ke2, double* ke_therm, double* ke_baro, const void* ws,
                    void* red_ws, void* stream);

/* Workspace (bytes) of tsx_crescale_step: 27 doubles per system. */
int64_t tsx_crescale_ws_bytes(int32_t n_systems);

/* One stochastic cell-rescaling barostat step (replaces torch_sim/
 * integrators/npt.py:2186-2492): mode 0 isotropic, 1 anisotropic, 2
 * triclinic (optional deviatoric target p_ext_tensor [S,3,3] with the
 * initial cell inverse and volume), 3 triclinic with the average kinetic
 * pressure (dof [S] = 3n - 3).  kin [S,6] = sum p p^T / m of the kicked
 * momenta (modes 0-2).  noise [S,11] = (r_a, R[9] (mode 1: R[3]), r_b)
 * standard normals.  Per system: sqrt(V) propagated in two halves around
 * the traceless deformation (exp, Gram-Schmidt), the cell updated in place;
 * per atom: x <- R x + (Mx p) dt / (2 m), p <- Vp p (unchanged in mode 3). */
int tsx_crescale_step(int32_t n_systems, int32_t mode, const int32_t* atom_sys, int64_t n_atoms,
                      double* pos, double* mom, const double* masses, const double* kin,
                      const double* stress, double* cell_col, const double* kT,
                      const double* compressibility, const double* tau_p, double dt,
                      const double* p_ext, const double* p_ext_tensor, const double* cell0_inv,
                      const double* vol0, const double* dof, const double* noise, void* ws,
                      void* stream);

/* One reverse step of the discrete adjoint of velocity Verlet in momentum
 * form (torch_sim/integrators/nve.py:74-130): xbar += vjp;
 * pbar += (dt / m) xbar; w = wscale * pbar (the next force-VJP cotangent). */
int tsx_nve_adjoint_update(double* xbar, double* pbar, const double* vjp, const double* masses,
                           double dt, double wscale, double* w, int64_t n_atoms, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TSX_H_ */
