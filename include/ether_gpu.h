/*
 * ether_gpu.h - C-ABI of the MI355X (gfx950) Monte Carlo sweep engine.
 *
 * The reference (mukkelian/Ether) has no FFI: its hot path is a set of
 * implicit-interface Fortran externals that share state through module
 * `init` (src/init.f90:20-98).  Each entry point below replaces one of those
 * externals or one of their call sites in the driver, and is bound from
 * Fortran through `iso_c_binding` (ether_amd/fortran/ether_gpu_mod.f90) and
 * from Python through ctypes (ether_amd/engine.py).
 *
 * Conventions
 *   - Every call returns an int status (ETHER_OK == 0).  On failure the
 *     message is kept in the context and read with ether_last_error(); the
 *     Fortran driver maps any nonzero status to its terminate(remark), which
 *     keeps the reference's print-and-stop convention (src/terminate.f90:20-35).
 *   - The context owns every device buffer.  Host arrays belong to the caller
 *     and are copied in at set/commit time or out at get time.
 *   - Site indices are the reference's supercell order: cell k (z) outer,
 *     j (y), i (x), basis l innermost (src/generate_supercell.f90:115-132),
 *     0-based here.  Species ids are 1-based as in ion(4,:).
 *   - A "replica" is one Markov chain (one temperature x one repeat of the
 *     reference, or one member of a tempering ladder).  A "slot" is a
 *     temperature of the ladder; replica r sits at slot slot_of[r].  Tempering
 *     swaps permute slots, never spins.
 *   - One host thread per context; all kernels run in order on the context's
 *     stream.  Calls that return host data synchronise that stream.  Distinct
 *     contexts may be created, set up and committed from several host threads
 *     at once; the error text of a call without a context is kept per thread.
 *   - No MPI, no CUDA shims.  gfx950 only.
 */
#ifndef ETHER_GPU_H
#define ETHER_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ether_ctx ether_ctx;

/* status codes */
#define ETHER_OK          0
#define ETHER_ERR_ARG     1   /* bad argument (size, pointer, range)          */
#define ETHER_ERR_STATE   2   /* call not valid in the current state          */
#define ETHER_ERR_HIP     3   /* HIP runtime error                            */
#define ETHER_ERR_LIMIT   4   /* exceeds a documented capacity                */
#define ETHER_ERR_NODEV   5   /* no gfx950 device visible                     */

/* spin models: in.ether `model = ising | xyz` (src/read_input.f90:62-66) */
#define ETHER_MODEL_ISING 0
#define ETHER_MODEL_XYZ   1

/* ether_init_spins modes (the generate_supercell start state of
 * src/generate_supercell.f90:34-63,115-132 is written with ether_set_spins) */
#define ETHER_INIT_FRESH  1  /* fresh_spins: one unconditional proposal per site (src/fresh_spins.f90:31-38) */
#define ETHER_INIT_RANDOM 2  /* independent random state: Ising +-1, XYZ uniform on S^2                       */

/* compat flags (ether_set_compat) */
#define ETHER_COMPAT_PT_ETOT  1  /* tempering swap uses the double-counted E_tot (src/ether.f90:177,194)      */
#define ETHER_COMPAT_SPLIT_OVRR 2 /* ether_run: a whole OVRR pass before the sweep, as src/ether.f90:163-172;
                                    by default a pass of at most one reflection per site is fused into the
                                    sweep, each colour phase reflecting a site about its local field and then
                                    proposing from the reflected spin (the same stationary distribution)     */
#define ETHER_COMPAT_MEASURE_PASS 4 /* ether_run: every measurement as a pass of its own after the sweep, as
                                    get_tot_energy after Metropolis (src/ether.f90:172-177); by default an
                                    XYZ sweep that a measurement follows books it as its phases update the
                                    sites (the same energies to FP32 rounding)                            */

#define ETHER_COMPAT_CHAIN_REPEATS 8 /* the drivers: the REPEAT repeats run one after another on nscan
                                    replicas, each from fresh_spins of the previous repeat's final state, as the
                                    reference's rank runs them (src/ether.f90:152-155, src/fresh_spins.f90:31-38);
                                    by default every repeat is its own replica and starts from fresh_spins of the
                                    generate_supercell state.  The library itself accepts and ignores it.       */

/* per-replica moment record read by ether_read_moments (8 doubles) */
#define ETHER_MOM_N     0   /* number of samples                                  */
#define ETHER_MOM_E     1   /* sum of e    (e = E_tot/(2N), src/get_eng.f90:36)   */
#define ETHER_MOM_E2    2   /* sum of e^2                                         */
#define ETHER_MOM_E4    3   /* sum of e^4                                         */
#define ETHER_MOM_M     4   /* sum of m    (m = |M|/N, src/get_mag.f90:36-37)     */
#define ETHER_MOM_M2    5   /* sum of m^2                                         */
#define ETHER_MOM_M4    6   /* sum of m^4                                         */
#define ETHER_MOM_ACC   7   /* sum over sweeps of accepted/N (src/ether.f90:172-174) */
#define ETHER_MOM_COUNT 8

/* ---------------------------------------------------------------------------
 * Host-side lattice set-up (no GPU).  Replaces getting_nbd / search_nbd
 * (src/getting_nbd.f90:46-82, src/search_nbd.f90:34-58): same capture rule
 * |d_shell - |r_i - (r_j + image)|| <= tol over images {0,+1,-1} per closed
 * axis, j != i, each j stored once per shell, ascending j.  Uses the lattice's
 * translation symmetry instead of the O(N^2 * 27) scan.
 *
 *   abc[3*3]     lattice vectors as rows (Angstrom), abc[3*a+c]
 *   n_basis      sites per cell; basis_xyz[3*n_basis] Cartesian positions
 *   sc[3]        supercell repeats; open_bc[3] nonzero for 'o'
 *   shell_d[n_shells] bond lengths; tol = NBDFC
 *   capacity     nbd_capacity (src/defaults.f90:56); overflow -> ETHER_ERR_LIMIT
 *   nn_count[n_shells*N]            out: neighbours of site i in shell s at [s*N+i]
 *   nn_idx[n_shells*N*capacity]     out: [(s*N+i)*capacity + k], 0-based site ids
 * N = sc[0]*sc[1]*sc[2]*n_basis.  Returns ETHER_OK or an error code; a message
 * for the last failure of these context-free calls is read with
 * ether_last_error(NULL, ...).
 * ------------------------------------------------------------------------- */
int ether_build_neighbours(const double* abc, int32_t n_basis, const double* basis_xyz,
                           const int32_t* sc, const int32_t* open_bc,
                           int32_t n_shells, const double* shell_d, double tol,
                           int32_t capacity, int32_t* nn_count, int32_t* nn_idx);

/* The same neighbour lists built on the GPU (one thread per site and shell
 * applying the shell's offset template for the site's basis atom; identical
 * output, including the capacity error), plus a colouring of the neighbour
 * graph (all shells) that repeats with the lattice:
 *   colour(site) depends only on its basis atom and its cell index modulo a
 *   period m_k per axis (m_k in {1, 2, 3, 4, 6}, dividing L_k on closed axes),
 *   the fewest colours over the periods tried (2 for even simple-cubic
 *   lattices, 3 for triangular ones with L divisible by 3).
 * colour[N] (nullable) and n_colours (nullable) receive it; n_colours = 0
 * when no period gives a proper colouring (ether_commit then colours
 * greedily).  Pass colour to ether_set_lattice.  Needs a GPU
 * (ETHER_ERR_NODEV otherwise); no counterpart in the reference beyond the
 * lists of src/getting_nbd.f90:46-82, whose colouring is none (a typewriter
 * sweep, src/Metropolis.f90:39). */
int ether_build_neighbours_device(const double* abc, int32_t n_basis, const double* basis_xyz,
                                  const int32_t* sc, const int32_t* open_bc,
                                  int32_t n_shells, const double* shell_d, double tol,
                                  int32_t capacity, int32_t* nn_count, int32_t* nn_idx,
                                  int32_t* colour, int32_t* n_colours);

/* The periodic colouring alone, on the host (the same colours as
 * ether_build_neighbours_device gives). */
int ether_colour_lattice(const double* abc, int32_t n_basis, const double* basis_xyz,
                         const int32_t* sc, const int32_t* open_bc, int32_t n_shells,
                         const double* shell_d, double tol, int32_t* colour, int32_t* n_colours);

/* ---------------------------------------------------------------------------
 * Spiral-state partners: replaces get_spiral_state('ions_for_spiral')
 * (src/get_spiral_state.f90:47-82), which fills ifs(spiral_capacity, N) and
 * spiral_count.  pos[3N] Cartesian site positions (ion(6:8,:)), dir[3] =
 * ss_direc, dis = ss_dis, latency_deg = ss_lat (src/read_input.f90:128-139).
 * Site j != i is a partner of i when |r_j - r_i| <= dis and
 * acos((r_j - r_i).dir / |r_j - r_i|) in degrees <= latency_deg; no periodic
 * images.  partners[N*capacity] out: 0-based partner ids in ascending order,
 * -1 after the last one; count out (nullable) = spiral_count.  More than
 * `capacity` partners for a site -> ETHER_ERR_LIMIT.  Context-free: errors are
 * read with ether_last_error(NULL, ...).
 * ------------------------------------------------------------------------- */
int ether_build_spiral_partners(int32_t n_sites, const double* pos, const double* dir,
                                double dis, double latency_deg, int32_t capacity,
                                int32_t* partners, int64_t* count);

/* ---------------------------------------------------------------------------
 * Start state of the basis sites: replaces the draw of generate_supercell
 * (src/generate_supercell.f90:34-63), which the reference takes from the
 * Fortran runtime's random_number after random_seed(put=seed+rank)
 * (src/ether.f90:63), whose stream depends on the compiler.  Here it is a
 * pure function of `seed`, so every driver that calls it gets the same start
 * state: Ising S = (0, 0, (-1)^file_index[l]) with file_index the 1-based
 * atom number in structure.vasp; XYZ Marsaglia (angle = 0) a uniform point
 * of S^2 by George_Marsaglia's disc rule (src/George_Marsaglia.f90:29-39),
 * phi = 0; ANGLE (angle = 1) S_z ~ U(-1,1), phi ~ U(0, 2 pi).  Draws use
 * Philox4x32-10 with counter {l, 0, 0, 5 << 16 | attempt}.  S[3*n_basis]
 * (x, y, z per basis site), phi[n_basis].  Host-only, context-free; the
 * supercell copies basis site l to every cell (src/generate_supercell.f90:115-132).
 * ------------------------------------------------------------------------- */
int ether_start_state(uint64_t seed, int32_t model, int32_t angle, int32_t n_basis,
                      const int32_t* file_index, double* S, double* phi);

/* ---------------------------------------------------------------------------
 * Context life cycle
 * ------------------------------------------------------------------------- */

/* Replaces the module-`init` allocation done by generate_supercell/getting_nbd.
 * n_sites = total_ions (src/generate_supercell.f90:137), n_replicas chains,
 * seed = in.ether `seed` (src/read_input.f90:118-120). */
int ether_ctx_create(int32_t model, int32_t n_sites, int32_t n_replicas, uint64_t seed,
                     ether_ctx** out);
void ether_ctx_destroy(ether_ctx* ctx);

/* Copies the last error message (NUL-terminated, truncated to len). ctx may be NULL. */
int ether_last_error(const ether_ctx* ctx, char* buf, int32_t len);

/* Lattice: species[N] (1-based, ion(4,:)), stagger[N] (stgg_ion,
 * src/get_staggered_info.f90:78-82 and src/generate_supercell.f90:127),
 * neighbour lists as produced by ether_build_neighbours (the nn array of
 * src/getting_nbd.f90:34).  colour[N] may be NULL: the library then colours
 * the interaction graph greedily (bipartite lattices get 2 colours). */
int ether_set_lattice(ether_ctx* ctx, int32_t n_species, const int32_t* species,
                      const double* stagger, int32_t n_shells, int32_t capacity,
                      const int32_t* nn_count, const int32_t* nn_idx, const int32_t* colour);

/* Exchange: j_exc[((s*n_species + a)*n_species + b)*3 + c] in energy units
 * already scaled as the reference's j_exc after j_values + parameters
 * (s_a*s_b, /1000 or /J_para: src/j_values.f90:83-84, src/parameters.f90:26-41).
 * spin[n_species] = s (src/read_input.f90:74-75), used for M. */
int ether_set_exchange(ether_ctx* ctx, const double* j_exc, const double* spin);

/* Zeeman term -g*mu_B*s*(S.H) (src/Hamiltonian.f90:108-124). */
int ether_set_zeeman(ether_ctx* ctx, int32_t on, const double* h3, double g_factor, double mu_b);

/* Single-ion anisotropy D.(S*S) (src/Hamiltonian.f90:127-143), sia[n_species*3]
 * scaled as sia_vec (src/get_sia_values.f90:58, src/parameters.f90).  XYZ only. */
int ether_set_sia(ether_ctx* ctx, int32_t on, const double* sia);

/* Proposal: angle=0 Marsaglia (src/George_Marsaglia.f90:29-39); angle=1 ANGLE
 * moves with window dphi in radians (src/update_spin_details.f90:39-46). */
int ether_set_proposal(ether_ctx* ctx, int32_t angle, double dphi_rad);

int ether_set_compat(ether_ctx* ctx, int32_t flags);

/* Builds the colour-sorted device layout and uploads it.  Must follow the
 * set_* calls and precede any sweep. */
int ether_commit(ether_ctx* ctx);

/* Number of colour classes and, optionally, colour[N] and the device slot of
 * every site (rng_id[N]); the counter-based RNG is keyed by the slot. */
int ether_get_layout(ether_ctx* ctx, int32_t* n_colours, int32_t* colour, int32_t* rng_id);

/* ---------------------------------------------------------------------------
 * Temperatures and ladders
 * ------------------------------------------------------------------------- */

/* beta per slot (beta = 1/(kB*T), src/ether.f90:139-140).  Resets slot_of to
 * the identity. */
int ether_set_betas(ether_ctx* ctx, const double* beta);

/* Replaces the temperature of every slot but keeps which replica sits in
 * which slot (slot_of), e.g. to move an adaptive ladder during equilibration. */
int ether_retune_betas(ether_ctx* ctx, const double* beta);

/* Tempering groups: slots [group_start[g], group_start[g+1]) form one ladder,
 * ordered as the reference orders its ranks (hottest first, src/ether.f90:79-84).
 * n_groups = 0 disables tempering. group_start has n_groups+1 entries. */
int ether_set_ladders(ether_ctx* ctx, int32_t n_groups, const int32_t* group_start);

/* ---------------------------------------------------------------------------
 * Hot path
 * ------------------------------------------------------------------------- */

/* fresh_spins (src/ether.f90:155) and the generate_supercell start state. */
int ether_init_spins(ether_ctx* ctx, int32_t mode);

/* n_mcs Metropolis sweeps of every replica (src/Metropolis.f90:20-70), one
 * colour class after another.  Per-replica acceptance ratios are added to the
 * ETHER_MOM_ACC moment.  accept[R] (nullable) receives the ratio of the last
 * sweep. */
int ether_sweep(ether_ctx* ctx, int32_t n_mcs, double* accept);

/* Overrelaxation (src/overrelaxation.f90:20-63): reflection of each spin about
 * its local exchange (+ Zeeman) field, colour by colour.  `passes` is the
 * expected number of reflections per site (ovrr_para): floor(passes) full
 * passes plus a Bernoulli(frac) pass.  XYZ only; refused with SIA on. */
int ether_overrelax(ether_ctx* ctx, double passes);

/* get_tot_energy + get_tot_magnetisation (src/get_tot_energy.f90:20-49,
 * src/get_tot_magnetisation.f90:20-42) for every replica.  With sample != 0 the
 * observable() moments are accumulated (src/observable.f90:30-43).
 * E[R] (nullable) = E_tot as the reference defines it (exchange double
 * counted, Zeeman/SIA once), M[3R] (nullable) = staggered magnetisation. */
int ether_measure(ether_ctx* ctx, int32_t sample, double* E, double* M);

/* Split energies per replica: parts[4R] = {E_exchange(double counted),
 * E_zeeman, E_sia, E_physical}.  Ising: also the exact integer bond sums
 * bonds[R*n_classes] (nullable) and n_classes. Uses the last ether_measure. */
int ether_get_energy_parts(ether_ctx* ctx, double* parts, int64_t* bonds, int32_t* n_classes);

int ether_read_moments(ether_ctx* ctx, double* moments /* [R*ETHER_MOM_COUNT] */);
int ether_reset_moments(ether_ctx* ctx);   /* zeroes('avg') (src/zeroes.f90:51-58) */

/* Spiral-state parameter (ssp, XYZ only; src/get_spiral_state.f90:84-120).
 * ether_set_spiral, before ether_commit: partners[N*capacity] as produced by
 * ether_build_spiral_partners (0-based, -1 after a row's last partner);
 * spiral_count = the number of partner entries; capacity 0 switches it off.
 * With partners set, every measure pass also forms psi = |sum_i sum_j
 * S_i x S_j| / spiral_count per replica, and sampling passes add psi, psi^2,
 * psi^4 to the replica's temperature slot ('spiral_state' case).
 * ether_read_spiral: moments[3R] (nullable) = those sums per slot (the sample
 * count is the ETHER_MOM_N moment); psi[R] (nullable) = psi of each replica's
 * last measured state.  ether_reset_moments clears the sums too. */
int ether_set_spiral(ether_ctx* ctx, int32_t capacity, const int32_t* partners);
int ether_read_spiral(ether_ctx* ctx, double* moments, double* psi);

/* parallel_tempering (src/parallel_tempering.f90:20-110) for every ladder:
 * energies from a fresh measure pass, pairs (k,k+1) with k%2 == parity,
 * accept with min(1, exp((beta_k - beta_k+1)(E_k - E_k+1))).  Swaps slots.
 * n_accepted (nullable) = accepted swaps in this call. */
int ether_pt_swap(ether_ctx* ctx, int32_t parity, int32_t* n_accepted);

/* slot_of[R] (nullable), per-slot accepted/attempted swap counts with the
 * slot above (acc[R], att[R], nullable) and completed cold->hot->cold round
 * trips per replica (trips[R], nullable). */
int ether_pt_state(ether_ctx* ctx, int32_t* slot_of, int64_t* acc, int64_t* att, int64_t* trips);

/* One whole MCS schedule of src/ether.f90:159-203 for steps
 * first_step..first_step+n_steps-1 (1-based like stepi):
 *   OVRR   if ovrr_interval>0 and stepi%ovrr_interval==0 and stepi%coa!=0
 *   sweep
 *   PT     if pt_interval>0 and stepi%pt_interval==0 and stepi%coa!=0
 *   sample if stepi>mcs_eq and stepi%coa==0
 * swap_count (in/out) is the reference's per-repeat PT counter. */
int ether_run(ether_ctx* ctx, int32_t first_step, int32_t n_steps, int32_t mcs_eq,
              int32_t coa, int32_t ovrr_interval, double ovrr_passes,
              int32_t pt_interval, int32_t* swap_count);

/* ether_run replays a captured HIP graph for every whole period
 * P = lcm(coa, ovrr_interval, pt_interval) of the schedule that lies entirely
 * in equilibration or entirely in production (P <= 1024); other steps are
 * launched one by one.  Both paths give bit-identical results.  on = 0 turns
 * graphs off (default on). */
int ether_set_graphs(ether_ctx* ctx, int32_t on);

/* Ising sweeps with the lattice resident in LDS (no counterpart in the
 * reference, whose sweep is one loop over sites, src/Metropolis.f90:35-62):
 * one workgroup per word of four replicas holds that word of every site in
 * LDS and runs all the sweeps between two measurements (a swap or a sample)
 * in one launch, with a workgroup barrier after each colour phase, booking
 * the measurement the next swap or sample reads.  Used by ether_sweep and
 * ether_run when the context is Ising, has at most 16 colours and
 * 4*N (+ 16 * its threshold table) bytes fit 60 KiB; batches keep the phase
 * kernels.  While the replicas take fewer than twice the rounds of
 * workgroups over the compute units that the words take (fewer replicas than
 * compute units), each replica gets a workgroup of its own instead (its int8
 * spins, N bytes, and its threshold table in LDS), which draws the word's
 * Philox block and uses its own word of it.  Results are bit-identical to
 * launching the colour phases one by one (the same Philox counters, counts
 * and integer thresholds).  on: 0 forces the phase kernels, 1 chooses
 * (default), 2 forces a workgroup per word, 3 a workgroup per replica;
 * anything else is ETHER_ERR_ARG. */
int ether_set_resident(ether_ctx* ctx, int32_t on);

/* The Ising sweep path a committed context takes now (no counterpart in the
 * reference): 0 the colour phases launched one by one (an XYZ context, a
 * lattice that does not fit, or ether_set_resident(ctx, 0)), 2 the resident
 * kernel with a workgroup per word of four replicas, 3 with a workgroup per
 * replica.  For reports and profiles. */
int ether_get_resident(ether_ctx* ctx, int32_t* kind);

/* The grid forms of the structure factor (ether_structure_factor and
 * ether_structure_factor_basis) take every lattice axis of at most 96 cells
 * with a kernel that keeps the axis's DFT matrices resident in LDS and reads
 * each input row once per axis; longer axes take the tiled kernel, which
 * stages both operands per K step.  The two make the same products in the
 * same order, so S(q) is bit-identical either way.  on = 0 forces the tiled
 * kernel on every axis (default on).  A tuning and testing switch of this
 * library; no routine of the reference corresponds to it. */
int ether_set_dft_resident(ether_ctx* ctx, int32_t on);

/* Captures, instantiates and uploads every graph that ether_run with the same
 * arguments would replay, and runs nothing, so that the ether_run that
 * follows spends no time on capture or instantiation (a timed production
 * loop, a serving loop).  No counterpart in the reference.  A no-op with
 * graphs off or kernel timing on. */
int ether_run_prepare(ether_ctx* ctx, int32_t first_step, int32_t n_steps, int32_t mcs_eq,
                      int32_t coa, int32_t ovrr_interval, double ovrr_passes, int32_t pt_interval);

/* ---------------------------------------------------------------------------
 * Batches of lattices (no counterpart in the reference, which runs one
 * lattice per process: src/ether.f90:110-214 over MPI ranks).  Each lattice
 * is an ordinary committed context with its own sites, couplings, replicas,
 * temperatures, ladders and seed; a batch advances all of them through the
 * schedule of ether_run with one launch per kernel kind and template
 * instance (instead of one per lattice), whole periods replayed as one
 * captured graph.  Every lattice's result is bit-identical to ether_run on
 * its context alone.  Requirements: the same spin model and SIA switch,
 * either all or none with ladders, kernel timing off.  The contexts stay
 * usable (moments, spins, tempering state); destroy the batch before any of
 * its contexts.  Errors are read with ether_last_error(NULL, ...).
 * ------------------------------------------------------------------------- */
typedef struct ether_batch ether_batch;
int ether_batch_create(ether_ctx* const* ctxs, int32_t n, ether_batch** out);
int ether_batch_run(ether_batch* batch, int32_t first_step, int32_t n_steps, int32_t mcs_eq,
                    int32_t coa, int32_t ovrr_interval, double ovrr_passes, int32_t pt_interval,
                    int32_t* swap_count);
int ether_batch_set_graphs(ether_batch* batch, int32_t on);
/* Whether the last ether_batch_run swept with the spins resident in LDS:
 * Ising batches whose every context allows it (ether_set_resident mode 1 or
 * 3) and fits a workgroup's LDS (int8 spins of every site plus the
 * threshold table, at most 150 KiB: up to about 150,000 sites) run
 * the sweeps between two measurements as one launch with a workgroup per
 * replica of every lattice, largest lattices first; otherwise the colour
 * phases are batched launches of their own.  Results are identical. */
int ether_batch_get_resident(ether_batch* batch, int32_t* on);
void ether_batch_destroy(ether_batch* batch);

/* Launch shape of the sweep colour phases (no counterpart in the reference,
 * whose sweep is one OpenMP loop over sites, src/Metropolis.f90:35-62):
 * sites_per_wave (1..16) consecutive sites of a colour class per wavefront,
 * xcd_map 1 = workgroups remapped so that each XCD (own L2) sweeps one
 * contiguous site range with every group of 64 replicas, 2 = the same with
 * the replica group outermost (an XCD sweeps more sites of fewer groups, whose
 * neighbour spins then stay in its L2 between uses).  Defaults 4 and 2.
 * Results do not depend on the shape: every update draws from its own (site,
 * replica, sweep) counter. */
int ether_set_launch(ether_ctx* ctx, int32_t sites_per_wave, int32_t xcd_map);
/* Lanes of a wavefront given to one site in the sweep when a context has few
 * replicas: 64 (one site per wave), or fewer so that a wave updates
 * 64 / lanes whole sites at once (lanes past the last whole site idle).
 * XYZ: any count from the replica count to 64 (default: 16 or 32, the fewer
 * that hold all replicas).  Ising, whose lanes hold 4 replicas each: 8, 16,
 * 32 or 64 (default: the fewest that hold all replicas).  0 restores the
 * default.  Results do not depend on it.  In
 * ether_set_launch, sites_per_wave then counts steps of 64 / lanes sites. */
int ether_set_lanes(ether_ctx* ctx, int32_t lanes_per_site);
/* Autotuner: times `reps` sweeps for each of sites_per_wave {1,2,4,8,16} x
 * xcd_map {0,1,2} on the context's own lattice and replicas, in two rounds
 * (each candidate keeps its faster one), keeps the fastest, or among the
 * candidates within 1% of it the one with the most sites per wave (returned
 * through the two nullable pointers), and restores the spins, the
 * acceptance counters and the sweep counter.  table_ms[15] (nullable)
 * receives the time per sweep of candidate i = xcd_map*5 + index of
 * sites_per_wave.  Allocates a temporary copy of the spin planes. */
int ether_autotune(ether_ctx* ctx, int32_t reps, int32_t* sites_per_wave, int32_t* xcd_map,
                   double* table_ms);

/* Kernel timing with HIP events recorded on the context's stream around the
 * launches of a kind: 0 Metropolis colour phases (one event pair around the
 * back-to-back phases of a sweep), 1 overrelaxation colour phase,
 * 2 measurement (partials + finalize, or the finalize alone after a sweep
 * that booked it), 3 tempering swap, 4 structure factor (device part of one
 * ether_structure_factor call), 5 the colour phases of a sweep that also
 * books the measurement that follows it (XYZ, inside ether_run), 6 the
 * resident Ising launches (ether_set_resident; counted in sweeps).  Adds no
 * synchronisation; ether_get_timing synchronises and reports the summed
 * event time and the number of launches it covers since timing was switched
 * on. */
#define ETHER_TIME_SWEEP      0
#define ETHER_TIME_OVRR       1
#define ETHER_TIME_MEASURE    2
#define ETHER_TIME_PT         3
#define ETHER_TIME_SQ         4
#define ETHER_TIME_SWEEP_MEAS 5
#define ETHER_TIME_RESIDENT   6  /* resident Ising sweeps: one event pair per launch, counted in sweeps */
int ether_set_timing(ether_ctx* ctx, int32_t on);
int ether_get_timing(ether_ctx* ctx, int32_t kind, double* total_ms, int64_t* launches);
/* Waits for all work queued on the context's stream. */
int ether_synchronize(ether_ctx* ctx);

/* Spins of one replica in site order: S[3N] (x,y,z per site), phi[N] nullable.
 * Ising spins are returned as (0,0,+-1). */
int ether_get_spins(ether_ctx* ctx, int32_t replica, double* S, double* phi);
int ether_set_spins(ether_ctx* ctx, int32_t replica, const double* S, const double* phi);

/* Local exchange field h_i = sum_j J_ij o S_j (src/get_ovrr_vec.f90:41-85) and
 * site energies e_i (src/Hamiltonian.f90:20-54, site mode) of one replica,
 * computed by the device kernels; h[3N], e[N] nullable. */
int ether_local_fields(ether_ctx* ctx, int32_t replica, double* h, double* e);

/* ---------------------------------------------------------------------------
 * Structure factor (utility/Structure_Factor/structure_factor.f90:124-183)
 *   S_a(q) = sum_r S_a(r) exp(i q.r) on the commensurate grid
 *   q = 2*pi*(m1*b1/L1 + m2*b2/L2 + m3*b3/L3), m_k in [0, L_k), a_i.b_j = delta_ij,
 * evaluated as per-axis DFT contractions on the matrix cores (split-BF16
 * operands, FP32 accumulation).  out[R*Q] = sum_a |S_a(q)|^2 / N^2 per
 * replica, Q = N = L1*L2*L3 (one site per cell), q index m1 fastest.
 * Keeps its intermediate planes and tables between calls.
 * ------------------------------------------------------------------------- */
int ether_structure_factor(ether_ctx* ctx, const int32_t* sc, float* out);

/* The same grid for a supercell of L1 x L2 x L3 cells with n_basis sites per
 * cell (N = L1*L2*L3*n_basis, sites in the supercell order, basis site
 * innermost): frac[3*n_basis] are the basis sites' fractional coordinates
 * (r_l = sum_k frac[3l+k] a_k), so that q.r = 2*pi*sum_k m_k (n_k + frac_lk)/L_k.
 * out[R*Q], Q = L1*L2*L3 grid points, m1 fastest: sum_a |S_a(q)|^2 / N^2. */
int ether_structure_factor_basis(ether_ctx* ctx, const int32_t* sc, int32_t n_basis,
                                 const double* frac, float* out);

/* The reference's own S(q) (utility/Structure_Factor/structure_factor.f90:116-177):
 * q_m = q_start + m*dq for m = 0..n_q-1, with n_q = |nint((q_stop - q_start)/dq)| + 1
 * (checked), and S_a(q_m) = sum_i S_a(r_i) exp(i q_m k_vec.r_i) over every site,
 * where k_vec = b1 + b2 + b3 with b the reciprocal vectors (2 pi included) of
 * the lattice vectors as the utility scales them (:95-116) and pos[3N] the
 * Cartesian site positions in site order (gss.dat's ion(4:6)).
 * out[R*n_q*4], per replica and q: {sum_a |S_a|^2 / N^2, |S_x|/N, |S_y|/N,
 * |S_z|/N} (:175-177; Ising spins are (0, 0, +-1)).  The phases are formed in
 * FP64 and reduced to [-pi, pi] before their FP32 sine and cosine; the
 * contraction is split-BF16 on the matrix cores with FP32 accumulation.
 * Keeps its device buffers (positions, k_vec.r, output) between calls. */
int ether_structure_factor_line(ether_ctx* ctx, const double* pos, const double* k_vec,
                                double q_start, double q_stop, double dq, int32_t n_q, float* out);

/* ether_structure_factor_line without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as ether_structure_factor_line; its
 * device buffers (positions, k.r, output staging: 32*N + 16*R*n_q bytes, each
 * part rounded up to 256 bytes) come from one static 16 MiB device array per
 * GPU instead of the context, and larger requests are refused with
 * ETHER_ERR_LIMIT.  Calls on one GPU share that array: each call writes what
 * it reads before reading it and synchronises the context's stream before it
 * returns, so calls that overlap on one GPU (from two host threads) race. */
int ether_structure_factor_line_noalloc(ether_ctx* ctx, const double* pos, const double* k_vec,
                                        double q_start, double q_stop, double dq, int32_t n_q,
                                        float* out);

/* ---------------------------------------------------------------------------
 * Learned tempering ladder.  No reference counterpart: the nearest is the
 * commented Hukushima stub at src/startup.f90:115-122.
 *   A 1-64-64-2 tanh MLP maps x = (T - t_mid)/t_half to (<E>(T), log C_v(T))
 *   normalised; weights live on the device; training is full-batch Adam with
 *   hand-written forward/backward kernels.
 * ------------------------------------------------------------------------- */
#define ETHER_MLP_NPARAM (64 + 64 + 64*64 + 64 + 2*64 + 2)
int ether_mlp_init(ether_ctx* ctx, uint64_t seed);
int ether_mlp_set_params(ether_ctx* ctx, const float* params);
int ether_mlp_get_params(ether_ctx* ctx, float* params);
/* loss/grad on a batch: x[n], y[2n], w[2n] (per-output weights). grad[NPARAM]. */
int ether_mlp_loss_grad(ether_ctx* ctx, int32_t n, const float* x, const float* y,
                        const float* w, float* loss, float* grad);
int ether_mlp_forward(ether_ctx* ctx, int32_t n, const float* x, float* y);
int ether_mlp_train(ether_ctx* ctx, int32_t n, const float* x, const float* y, const float* w,
                    int32_t iters, float lr, float* final_loss);
/* Ladder placement from the trained surrogate: with x = (T - x_mid)/x_half,
 * or x = (ln T - x_mid)/x_half with mode bit ETHER_PLACE_LOG_T, and
 * log C_v(T) = const + logcv_scale * y1(x), places n_temps temperatures on
 * [t_lo, t_hi] (hottest first, like src/ether.f90:79-84) so that the
 * thermodynamic length s = int sqrt(C_v)/T dT is equal between neighbours,
 * which makes the Gaussian estimate erfc(dT sqrt(C_v)/(2T)) of the swap
 * acceptance uniform.  With mode bit ETHER_PLACE_LENGTH y1 is s(T) itself
 * (any affine scale; its running maximum is used). */
#define ETHER_PLACE_LOG_T   1
#define ETHER_PLACE_LENGTH  2
int ether_ladder_place(ether_ctx* ctx, double x_mid, double x_half, double logcv_scale,
                       double t_lo, double t_hi, int32_t n_temps, int32_t mode, double* T_out);

/* Training protocol of the learned ladder, in the library so that every
 * driver runs the same one.  No reference counterpart (the nearest is the
 * stub at src/startup.f90:115-122); the plain ladder it starts from and falls
 * back to is the reference's linear one (src/ether.f90:79-84).
 * The context holds R = L x n_temps replicas; every one of its L ladders is
 * slots [g*n_temps, (g+1)*n_temps), and the slots already hold `temps`
 * (hottest first) through ether_set_betas.  From first_step it runs, with the
 * ether_run schedule (coa, OVRR, swaps every pt_interval MCS):
 *   - `windows` windows over train_mcs MCS, the first on the plain ladder;
 *     each measures the pair swap acceptance a_k, which gives the length
 *     density sqrt(C)/T = 2 erfcinv(a_k)/|dT_k| on a 64-point map of
 *     [T_lo, T_hi] (an interval with no accepted swap bounds it from below),
 *     and <e> per slot; with learn != 0 the ladder is then re-placed at equal
 *     steps of int sqrt(C)/T dT, by the MLP (placement 0: ether_mlp_init with
 *     mlp_seed, then ether_mlp_train on <e> and log C of the map, 4000 Adam
 *     steps, and ether_ladder_place) or straight from the map (placement 1),
 *     and the slots are retuned in place (ether_retune_betas);
 *   - with validate_mcs > 0, three windows of validate_mcs MCS: plain,
 *     learned, plain.  The learned ladder is kept only if its round-trip cost
 *     Lambda = sum_k (1 - a_k)/a_k is below the plain ladder's by more than
 *     twice the standard error below and its lowest pair acceptance above
 *     the plain ladder's (plain: the mean of its two windows); otherwise the
 *     slots return to the plain ladder.
 * learn = 0 runs the same windows on the plain ladder only.  Replicas keep
 * their state and slot throughout; the swap counters are reset at the end.
 * T_out[n_temps] = the ladder the slots hold for production, hottest first;
 * swap_count in/out as in ether_run; report[8] (nullable) = {Lambda plain,
 * Lambda learned, lowest acceptance plain, lowest acceptance learned (all
 * from validation), learned kept (0/1), last MLP loss (NaN if none), MCS run,
 * standard error of Lambda plain - Lambda learned from the binomial spread
 * of every pair's acceptance over its attempts: sum_k (1 - a_k)/(a_k^3 n_k)
 * per window with a = (accepted + 1)/(n + 2)}.
 * The production run continues at first_step + report[6]. */
int ether_ladder_train(ether_ctx* ctx, int32_t n_temps, const double* temps, double kb,
                       int32_t first_step, int32_t train_mcs, int32_t windows,
                       int32_t validate_mcs, int32_t coa, int32_t ovrr_interval,
                       double ovrr_passes, int32_t pt_interval, int32_t learn,
                       int32_t placement, uint64_t mlp_seed, int32_t* swap_count,
                       double* T_out, double* report);

#ifdef __cplusplus
}
#endif
#endif /* ETHER_GPU_H */
