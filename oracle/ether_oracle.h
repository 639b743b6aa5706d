/*
 * ether_oracle.h - CPU FP64 oracle for the Monte Carlo sweep engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline), never as a product path.
 *
 * Each function restates one routine of the reference (mukkelian/Ether) in
 * plain C, citing the file:line it follows.  The reference draws its random
 * numbers from the Fortran runtime's random_number, whose stream depends on
 * the compiler and cannot be reproduced; the oracle instead uses the same
 * counter-based Philox4x32-10 contract as the device kernels (below), so that
 * one sweep of the oracle and one sweep of the GPU can be compared site by
 * site.  With a single colour class in site order the oracle sweep is the
 * reference's typewriter sweep (src/Metropolis.f90:39).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   philox          Random123 known-answer vectors (Salmon et al., SC'11)
 *   neighbours      neighbour counts of the reference's example inputs
 *                   (the examples directory), and an independent O(N^2) search
 *   energies        closed forms (FM/Neel SC, 120-degree triangular,
 *                   pyrochlore tetrahedra) and an O(N^2) pair sum from positions
 *   sweep           exact enumeration (4x4 Ising), Langevin single spin,
 *                   single-spin SIA integral, detailed balance on 2 sites
 *   overrelaxation  involution, |S|=1, energy conservation
 *   estimators      closed forms on synthetic sample series
 *   tempering       two-level analytic acceptance
 *   spiral state    partner counts on a cubic grid (one +z partner per site
 *                   below the top layer), psi of rotated planar spirals
 *                   (= sin of the pitch angle), estimators on constant series
 *   mlp             central finite differences
 *   structure fac.  FM S(0)=1, Neel peak, direct FP64 sum
 */
#ifndef ETHER_ORACLE_H
#define ETHER_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG contract shared with the device kernels -------------------------
 * key = {seed & 0xffffffff, seed >> 32}
 * ctr = {id, replica, step, (purpose << 16) | attempt}
 * purposes: */
#define O_RNG_METROPOLIS 1
#define O_RNG_INIT       2
#define O_RNG_OVRR       3
#define O_RNG_PT         4
#define O_RNG_START      5

void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void o_rng(uint64_t seed, uint32_t id, uint32_t replica, uint32_t step, uint32_t purpose,
           uint32_t attempt, uint32_t out[4]);

/* ---- model description (module init, src/init.f90:42-77) ---------------- */
typedef struct {
    int32_t model;          /* 0 Ising, 1 XYZ                                   */
    int32_t n_sites;        /* total_ions                                       */
    int32_t n_species;
    int32_t n_shells;       /* no_of_nbd                                        */
    int32_t capacity;       /* nbd_capacity                                     */
    const int32_t* species; /* [N] 1-based (ion(4,:))                           */
    const double* stagger;  /* [N] stgg_ion                                     */
    const int32_t* nn_count;/* [n_shells*N]                                     */
    const int32_t* nn_idx;  /* [n_shells*N*capacity], 0-based                   */
    const double* j_exc;    /* [n_shells][n_species][n_species][3]              */
    const double* spin;     /* [n_species] s                                    */
    int32_t zeeman;
    double h[3], g_factor, mu_b;
    int32_t sia;
    const double* sia_vec;  /* [n_species][3]                                   */
    int32_t angle;
    double dphi;            /* radians                                          */
} o_model;

/* neighbours: getting_nbd + search_nbd, O(N^2*27) exactly as the reference */
int o_neighbours_brute(const double* abc, int32_t n_basis, const double* basis_xyz,
                       const int32_t* sc, const int32_t* open_bc, int32_t n_shells,
                       const double* shell_d, double tol, int32_t capacity,
                       int32_t* nn_count, int32_t* nn_idx);

/* Hamiltonian (src/Hamiltonian.f90:20-54).  S is [3N]. */
double o_site_energy(const o_model* m, const double* S, int32_t i);
double o_delta_energy(const o_model* m, const double* S, int32_t i, const double* s_new);
void   o_exchange_field(const o_model* m, const double* S, int32_t i, double* h3);
/* get_tot_energy split: parts = {E_exch (double counted), E_zeeman, E_sia} */
void   o_total_energy(const o_model* m, const double* S, double* parts3);
/* Ising integer bond sums B_k = sum_i sigma_i sum_{j in class k} sigma_j for
 * the classes given by class_of_edge (same layout as nn_idx) */
void   o_ising_bond_sums(const o_model* m, const double* S, const int32_t* class_of_edge,
                         int32_t n_classes, int64_t* bonds);
void   o_magnetisation(const o_model* m, const double* S, double* M3);

/* proposal (src/update_spin_details.f90:20-53, src/George_Marsaglia.f90:29-39)
 * drawn with the Philox contract for (seed, id, replica, step, purpose).
 * eta receives the acceptance uniform in (0,1] (word 2 of attempt 0). */
void o_propose(const o_model* m, uint64_t seed, uint32_t id, uint32_t replica, uint32_t step,
               uint32_t purpose, const double* s_old, double phi_old, double* s_new,
               double* phi_new, double* eta);

/* Proposals of one site for replicas 0..R-1, as the device draws them: XYZ
 * Marsaglia candidates are pooled over each group of 64 consecutive replicas
 * (a replica whose first pair misses the disc takes, in replica order, the
 * second pair of the replicas whose second pair hits it; if those run out it
 * draws attempts 1, 2, ... itself).  s_old/s_new [R*3], phi_old (nullable),
 * phi_new, eta [R]. */
void o_propose_site(const o_model* m, uint64_t seed, uint32_t id, int32_t R, uint32_t step,
                    const double* s_old, const double* phi_old, double* s_new, double* phi_new,
                    double* eta);

/* Start state of the basis sites (src/generate_supercell.f90:34-63), drawn
 * with the Philox contract for (seed, id = basis index l, replica 0, step 0,
 * O_RNG_START): Ising (-1)^file_index[l]; Marsaglia: the point o_propose
 * draws; ANGLE: S_z ~ U(-1,1), phi ~ U(0, 2 pi) from words 0 and 1.
 * S [3*nb], phi [nb]. */
void o_start_state(uint64_t seed, int32_t model, int32_t angle, int32_t nb,
                   const int32_t* file_index, double* S, double* phi);

/* Ising acceptance threshold: accept iff w < thr or thr == 0xffffffff.
 * The Ising word of replica r at a site is word r&3 of the Philox block with
 * ctr {rng_id, r>>2, step, purpose}. */
uint32_t o_ising_threshold(double beta, double dE);

/* One sweep of R replicas (S[R][N][3], phi[R][N] nullable, beta[R]) in the
 * device's order and with its proposal rule; acc[R] (nullable) += accepted. */
int64_t o_sweep_batch(const o_model* m, double* S, double* phi, int32_t R, const int32_t* order,
                      int32_t n, const int32_t* rng_id, const double* beta, uint64_t seed,
                      uint32_t step, int64_t* acc);

/* One Metropolis sweep (src/Metropolis.f90:35-62) visiting sites in `order`
 * (n entries); rng_id[i] is the RNG id of site i.  S [3N], phi [N] (nullable
 * unless angle).  Returns the number of accepted moves. */
int64_t o_sweep(const o_model* m, double* S, double* phi, const int32_t* order, int32_t n,
                const int32_t* rng_id, double beta, uint64_t seed, uint32_t replica,
                uint32_t step);

/* Overrelaxation reflection of one site (src/overrelaxation.f90:30-49) about
 * the field A (exchange, plus Zeeman when on). */
void o_reflect(const o_model* m, double* S, int32_t i);
/* A pass over `order` with Bernoulli(frac) selection (frac >= 1: every site). */
void o_overrelax_pass(const o_model* m, double* S, double* phi, const int32_t* order,
                      int32_t n, const int32_t* rng_id, double frac, uint64_t seed,
                      uint32_t replica, uint32_t step, uint32_t pass);

/* One chain through the MCS loop of src/ether.f90:159-203 without tempering:
 * OVRR gate, sweep, samples; mom[8] accumulates the device's moment record. */
void o_run_chain(const o_model* m, double* S, double* phi, const int32_t* order, int32_t n,
                 const int32_t* rng_id, double beta, uint64_t seed, uint32_t replica,
                 int32_t mcs, int32_t eq, int32_t coa, int32_t ovrr_iv, double ovrr_p,
                 double* mom);

/* Estimators.  Per repeat (src/get_eng.f90:43-56, src/get_mag.f90:44-55):
 * mom = {n, sum e, e2, e4, m, m2, m4}; out = {e_avg, U_e, cv, m_avg, U_m, chi} */
void o_repeat_observables(const double* mom, double beta, int32_t n_sites, double* out6);
/* Over repeats (src/observables_per_repeat.f90:29-69): per[repeat*6] -> mean[6], sem[6] */
void o_over_repeats(const double* per, int32_t n_repeat, double* mean6, double* sem6);

/* Spiral state (src/get_spiral_state.f90).  Partners, case 'ions_for_spiral'
 * (:47-82): the reference's O(N^2) scan over j != i in ascending order; pos
 * [3N] Cartesian.  partners[N*cap] gets 0-based ids, -1 after a row's last;
 * returns spiral_count, or -1 if a site has more than cap partners. */
int64_t o_spiral_partners(int32_t n, const double* pos, const double* dir, double dis,
                          double latency_deg, int32_t cap, int32_t* partners);
/* Case 'spiral_state' (:84-120): psi = |sum_i sum_j S_i x S_j| / spiral_count,
 * S[3N] one replica's spins, in the reference's summation order. */
double o_spiral_psi(int32_t n, const double* S, const int32_t* partners, int32_t cap,
                    int64_t spiral_count);
/* Case 'avg_spiral_state' (:122-140): sums = {sum psi, psi^2, psi^4} over
 * n_samples -> out3 = {<psi>, U = 1.5 - 0.5 <psi^4>/<psi^2>^2,
 * chi = beta (<psi^2> - <psi>^2) spiral_count}. */
void o_spiral_observables(const double* sums, double n_samples, double beta, int64_t spiral_count,
                          double* out3);

/* PT decision (src/parallel_tempering.f90:77-89) */
int o_pt_accept(double beta, double beta_p, double E, double E_p, double u);

/* MLP 1-64-64-2 tanh, FP64.  params layout as ETHER_MLP_NPARAM:
 * W1[64] b1[64] W2[64*64 row=out] b2[64] W3[2*64] b3[2] */
#define O_MLP_NPARAM (64 + 64 + 64 * 64 + 64 + 2 * 64 + 2)
void o_mlp_forward(const double* p, int32_t n, const double* x, double* y);
double o_mlp_loss_grad(const double* p, int32_t n, const double* x, const double* y,
                       const double* w, double* grad);

/* S(q) direct sum (utility/Structure_Factor/structure_factor.f90:150-177) at
 * arbitrary q (Cartesian, 1/Angstrom): returns sum_a |S_a(q)|^2 / N^2 */
double o_structure_factor(int32_t n, const double* pos, const double* S, const double* q3);
/* the reference's whole output line (:175-177): out4 = {sum_a |S_a|^2/N^2,
 * |S_x|/N, |S_y|/N, |S_z|/N} */
void o_structure_factor_parts(int32_t n, const double* pos, const double* S, const double* q3, double* out4);

#ifdef __cplusplus
}
#endif
#endif
