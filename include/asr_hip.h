/* C ABI of the MI355X chain engine (libasr_hip.so): the augmented-space operator, its real and complex-symmetric Lanczos
 * chains, and the KPM moments of a supercell realisation.
 *
 * Conventions: 0-based indices, -1 = no link, row-major arrays, complex numbers as interleaved (re, im) doubles.  Host arrays
 * are read only for the duration of a call.  Every call returns 0 on success; on failure it returns non-zero and
 * asr_gpu_last_error() describes the error (the Fortran side turns it into `error stop`, as the reference does).
 * Calls are made from one host thread; every handle owns its device memory and a HIP stream.
 *
 * The species parameter array of a chain, `params`, holds c, d, o, e (each nsp x 9, [species][orbital]), as augcls_op%c,
 * %d, %o, %e of the reference (src/asr_augcls.f90:50, set by augcls_set_params :121-126).  A chain's start component is
 * 9 * node 0 + start_orb (the start node is node 0, as iv of src/asr_edc.f90:73-74).
 */
#ifndef ASR_HIP_H
#define ASR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct asr_gpu_op asr_gpu_op;
typedef struct asr_gpu_real asr_gpu_real;

const char* asr_gpu_last_error(void);
int asr_gpu_device_count(int* n);

/* ---- the site-class augmented-space operator (augcls_op, src/asr_augcls.f90:45-68) */
/* Dimensions: nn nodes, nslot bond slots, nop group operations (<= 64), nsmax states per site, ncls classes, nsp species,
 * ntype site types.  Replaces the allocation part of build_augcls_operator (src/asr_augcls.f90:709-727). */
int asr_gpu_op_create(int nn, int nslot, int nop, int nsmax, int ncls, int nsp, int ntype, asr_gpu_op** out);
/* Node tables (src/asr_augcls.f90:55-63): hopr/hopg nn x nslot, flipr/flipg nn x nsmax, lst/ncl/ntyp/fz/sq nn,
 * stab_ptr nn + 1, stab nstab.  fz: one byte per node (signed char, Fortran's integer(c_signed_char)): 1 the node sits on
 * the start site (augcls_freeze_site applies there, :129-141), 2 on a frozen pair partner (augcls_freeze_any, :147-161;
 * the EPI pair operator), 0 elsewhere.
 * Frozen sites must have been non-flippable when the node set was built. */
int asr_gpu_op_set_tables(asr_gpu_op* op, const int* hopr, const int* hopg, const int* flipr, const int* flipg, const int* lst,
                          const int* ncl, const int* ntyp, const signed char* fz, const double* sq, const int* stab_ptr,
                          const int* stab, int nstab);
/* Orbital operation tables pi, sg (nop x 9): D(g)[pi(l), l] = sg(l) (src/asr_augcls.f90:727-736). */
int asr_gpu_op_set_orbital_ops(asr_gpu_op* op, const int* pi, const double* sg);
/* Bond blocks block (ntype x nslot x 9 x 9) and on-site blocks block0 (ntype x 9 x 9): mdl%block, mdl%block0
 * (augcls_set_blocks, src/asr_augcls.f90:99-104). */
int asr_gpu_op_set_blocks(asr_gpu_op* op, const double* block, const double* block0);
/* Classes: nsc (ncls), spc (ncls x nsmax), mt (ncls x nsmax x nsmax x nsmax) = class_mt (src/asr_augcls.f90:190-205). */
int asr_gpu_op_set_classes(asr_gpu_op* op, const int* nsc, const int* spc, const double* mt);
/* Member tile of the batched chains (members per wave; 64 / mt nodes per wave): a power of two <= 64, 65 = tile 64 with
 * complex members run as pairs of real lanes, or 0 = chosen by the engine's autotuner per member count (default).  Replaces the reference's chain_teams / blk_nbmax rules
 * (src/asr_chain.f90:140-155, src/asr_augblk.f90:32-36). */
int asr_gpu_op_set_tile(asr_gpu_op* op, int mt);
/* Adaptive depth of the complex-symmetric chains (ed_depth_tol and ed_tail, edc_green, src/asr_edc.f90:121-127): at
 * levels 8, 12, 16, ... a chain that has not reached its depth is terminated with the given tail (0 hull, 1 bp, 2 avg2,
 * 3 none; cf_species, src/asr_chain.f90:162-262) and leaves when |G_k - G_{k-4}| <= depth_tol |G_k| (k >= 12).
 * depth_tol = 0 (the default): every chain runs to its depth.  With the tails hull, avg2 and none the test runs on the
 * device (the chain matrix's eigenvalues by a complex-symmetric QL iteration, one thread per chain, chains of at most 128
 * levels), so a batch stays one stream of launches or one HIP graph; with bp (or deeper chains) the host evaluates it
 * (one synchronisation per tested level; mode 2 of the batched calls then runs as mode 0). */
int asr_gpu_op_set_depth_tol(asr_gpu_op* op, double depth_tol, int tail);
/* Target states of the projected chains: tnode / tgop of build_augcls_operator with tsite (src/asr_augcls.f90:805-819,
 * asr_host_tables_targets): ntarget x nsmax representative nodes (-1: none) and the operations that map the target
 * states onto them. */
int asr_gpu_op_set_targets(asr_gpu_op* op, int ntarget, const int* tnode, const int* tgop);
/* Projected chains of route edc for spectral functions: replaces edc_green_proj (src/asr_edc.f90:174-232).  The chains
 * of asr_gpu_chains_edc (same arguments) run all nlev levels (no depth test); at every level the vector is projected on the
 * target states with the dressing weights tw (nchain x ntarget x nsmax x 9, spectral_dressing, src/asr_spectral.f90:
 * 208-228): cpr_k(t, l) = sum_ks tw(t, ks, l) psi_{u_k}(tnode(t, ks), l) (augcls_amp_c).  On the host g = G(z) of the
 * chain with the operator's tail, gp(t, l) = sum_k cpr_k(t, l) pv_k (pv the first column of the terminated resolvent) and
 * gp4 the same for the chain cut four levels earlier.  g: nchain complex; gp, gp4: nchain x ntarget x 9 complex; a, b:
 * nchain x ldab complex.  mode 0 batched, 1 plain, 2 one HIP graph. */
int asr_gpu_chains_edc_proj(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                            const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab,
                            const double* tw, double* a, double* b, int* mused, double* g, double* gp, double* gp4, int mode,
                            double* elapsed_ms);
/* Full reorthogonalisation of the real chains (lanczos with reorth = .true., src/reference/asr_recursion.f90:17-58): every
 * level projects the new vector twice against all stored Lanczos vectors of its chain, w = w - Q (Q^T w), before b_n is
 * taken.  The stored vectors cost depth x (nn + 1) x 9 x members doubles (at most 64 GiB): for tests and small cases. */
int asr_gpu_op_set_reorth(asr_gpu_op* op, int reorth);
/* Form of one level of the complex-symmetric edc chains (edc_green, src/asr_edc.f90:95-155, whose loop applies the
 * operator, takes a_n = sum(u w), then updates and projects w): form 0 (the default) takes a_n from the first of the
 * level's two operator passes, a_n = sum_i sq_i^2 [u~ (E + X(z)) u~ + u~ h u~ - (h u~) o (h u~)]_i, equal to sum(u w)
 * on the symmetric subspace, and folds the three-term update, the projection and the norm into the second pass; form 1
 * runs the update as a pass of its own after the operator, in the reference's order.  The coefficients of the two forms
 * agree to rounding.  Form 0 applies to chains on the member-tile operator pass (member tile 8 or 16, one d for every
 * species, the projection on); other chains take form 1.  The chain blocks take the same forms: route ed3 (chain,
 * src/asr_ed3.f90:169-184) with a_n = sum_i sq_i^2 [u~ (z - et) u~ + u~ h1 - h1 (u u~) - h1 t]_i, h1 = h' u~,
 * t = u u~ - Qt h1, and the EPI pair blocks; in form 0 the second pass also takes the partial sums of <t, w> for the
 * projection c_n+1 of the block's second vector.  For asr_gpu_lanczos_adjoint form 0 (on the member-tile pass)
 * computes the recomputed level u_k+1 and the combination u_bar_k on the operator's last pass instead of in passes of
 * their own; the gradients agree to rounding. */
int asr_gpu_op_set_level_form(asr_gpu_op* op, int form);
/* Member compaction of the batched chains with the adaptive depth (on != 0, the default): edc_green_blk
 * (src/asr_edc.f90:237-316) lets every member of a batch leave at its own depth; here, between host-issued levels (mode 0),
 * after each level of the depth test the members still running move to the front of the batch so that whole member
 * tiles stop, and the batch stops when none runs.  The coefficients are those of the uncompacted run bit for bit (every
 * member's arithmetic is unchanged); off keeps every member in its slot. */
int asr_gpu_op_set_compaction(asr_gpu_op* op, int on);
/* Window runs (nslot > 0; 0, the default, turns them off): route-edc chains with the adaptive depth tested on the device
 * (tails hull, avg2, none), issued as host levels (mode 0), with more chains than nslot hold nslot member slots; the other
 * chains wait in a queue (deepest first) and take the slot of a member as soon as it leaves, at its own level.  This is
 * the dynamic schedule of chain_worker (src/asr_class_scf.f90:1093-1105) over the members of one batched run: the member
 * tiles stay full while chains remain, and the vectors of nslot members, not of every chain, are allocated.  Each chain's
 * coefficients equal those of the batched run to rounding (the lagged normalisation rescales at the run's levels, not
 * the chain's). */
int asr_gpu_op_set_window(asr_gpu_op* op, int nslot);
/* Where a window run's queue lives (on_device != 0, the default: on the device): a kernel after every level hands the
 * slots of members that have left to the next chains (their results to the chains' rows, their tables and start vectors
 * to the slots) and the host only reads the queue's counters every 8 levels; once the queue is empty the run continues on
 * the host path (slot hand-over, then member compaction).  0: the host follows every slot after each level.  The
 * coefficients are the same either way; no counterpart in the reference (chain_worker's schedule runs on the host). */
int asr_gpu_op_set_window_queue(asr_gpu_op* op, int on_device);
/* The number of member compactions of the operator's last batched run (diagnostics). */
int asr_gpu_op_get_compactions(asr_gpu_op* op, int* n);
/* The tile the autotuner picked for nchain members of the given kind (0: none tuned yet). */
int asr_gpu_op_get_tile(asr_gpu_op* op, int nchain, int is_complex, int* mt);
/* Release the level-vector blocks the operator keeps cached between batched runs (the batches of a contour pass reuse
 * them instead of allocating tens of GB per call).  No counterpart in the reference (its work arrays are per thread and
 * allocated once, chain_worker src/asr_class_scf.f90:1126-1136); asr_gpu_op_destroy releases them too. */
int asr_gpu_op_trim(asr_gpu_op* op);
void asr_gpu_op_destroy(asr_gpu_op* op);

/* y = H x, H = E_nu + h - h o h in normalised coordinates (augcls_apply, src/asr_augcls.f90:275-297) for nb chains at once
 * (blk_hamx without the energy term, src/asr_augblk.f90:219-250).  x, y: nb x 9 nn.  frozen_sp: species of the frozen
 * start site per chain (-1 none) or NULL; conc: nb x ncls x nsmax class concentrations (augcls_set_class_conc) or NULL. */
int asr_gpu_apply_batched(asr_gpu_op* op, int nb, const int* start_orb, const double* params, const int* frozen_sp,
                          const double* conc, const double* x, double* y);
/* y = H x + X(z) x for complex x (augcls_apply on Re and Im plus augcls_rand_c, src/asr_edc.f90:106-116); X(z) = z pa + pb,
 * pab: nb x 2 x nsp x 9 (asr_edc.f90:253-258); z: nb complex.  x, y: nb x 9 nn complex. */
int asr_gpu_apply_complex(asr_gpu_op* op, int nb, const int* start_orb, const double* params, const double* pab,
                          const int* frozen_sp, const double* conc, const double* z, const double* x, double* y);

/* Real Lanczos chains (lanczos, src/reference/asr_recursion.f90:17-58) from the unit vector of each chain's start
 * component, nlev[c] levels each.  a, b: nchain x ldab; b of the last level is the norm of the last residual.  mused: levels
 * done.  mode 0: the batched engine (all chains stacked as one block vector); mode 1: the plain form, one chain per launch
 * sequence with a host synchronisation per level.  elapsed_ms: device time of the recursion (HIP events). */
int asr_gpu_lanczos_real(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const int* frozen_sp,
                         const double* conc, const int* nlev, double tiny, int ldab, double* a, double* b, int* mused, int mode,
                         double* elapsed_ms);
/* Complex-symmetric chains of route edc (edc_green, src/asr_edc.f90:75-164; block form edc_green_blk :237-316) without the
 * adaptive depth: chain c at energy z[c], depth nlev[c], projection onto the symmetric subspace when project != 0.
 * a, b: nchain x ldab complex.  Replaces the chain_worker task loop (src/asr_class_scf.f90:1093-1157) for one pass. */
int asr_gpu_chains_edc(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                       const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab,
                       double* a, double* b, int* mused, int mode, double* elapsed_ms);

/* Chains on several operators at once (one group per operator, e.g. the start sublattices of an ordered or partially
 * disordered compound: the per-class operator copies of chain_worker, src/asr_class_scf.f90:1093-1130).  Group g holds
 * nchain[g] chains; every per-chain array is concatenated over the groups in order (params and pab in each operator's
 * nsp, conc in its ncls x nsmax).  Members are sorted by depth inside a group and member tiles whose chains have all left
 * are skipped.  mode 0: levels issued from the host onto one stream per group (groups overlap on the GPU); mode 2: the whole
 * run captured into one HIP graph and launched once; mode 1: the plain form (one chain per launch sequence). */
int asr_gpu_chains_edc_multi(int ngroup, asr_gpu_op* const* ops, const int* nchain, const int* start_orb,
                             const double* params, const double* pab, const int* frozen_sp, const double* conc, const double* z,
                             const int* nlev, int project, int ldab, double* a, double* b, int* mused, int mode,
                             double* elapsed_ms);
int asr_gpu_lanczos_real_multi(int ngroup, asr_gpu_op* const* ops, const int* nchain, const int* start_orb,
                               const double* params, const int* frozen_sp, const double* conc, const int* nlev, double tiny,
                               int ldab, double* a, double* b, int* mused, int mode, double* elapsed_ms);

/* Route ed3, the full-ASA energy-dependent recursion: replaces ed3_blocks / ed3_blocks_blk (src/asr_ed3.f90:247-340) with
 * their start vectors (ed3_start :131-144) and chains (chain / chain_blk :148-211, 307-392).  Per input chain j (start
 * orbital l0, energy z, start site frozen to the central species) two complex-symmetric chains on z - K'(z),
 * K' = et - h' + u h' + h' u - h' Qt h' (ed3_kmul :99-129): from |a> = |u> + o_c h'|u> (normalised) projected on |v> = h'|u>,
 * and from |v> (normalised) projected on |a>.  params: nchain x 4 x nsp x 9 with c = g, d = D_c^1/2 and o = e = 0
 * (ed3_setup :53-84); cf: nchain x 6 x nsp x 9, et = z cf1 + cf2, u = z cf3 + cf4, Qt = z cf5 + cf6 (the cf array of
 * ed3_blocks_blk); oc: o_c of the start orbital per chain; z: nchain complex; nlev: depth per chain (<= ldab).
 * bk: nchain x 4 complex = (G_11, (hG)_11, (hGh)_11, q1a - q1b) as ed3_blocks; a, b, c: nchain x 2 x ldab complex (the two
 * chains' coefficients and projections c_k = <t|u_k>); mused: nchain x 2.  The continued fractions use the operator's tail
 * (asr_gpu_op_set_depth_tol) or, for a chain that closed (b_k < 1e-12), the finite fraction.  mode 0: batched, host-issued
 * levels; 1: plain (one chain per launch sequence); 2: one HIP graph (mode 0 when the adaptive depth is on). */
int asr_gpu_ed3_blocks(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* cf,
                       const double* oc, const int* frozen_sp, const double* conc, const double* z, const int* nlev,
                       int project, int ldab, double* bk, double* a, double* b, double* c, int* mused, int mode,
                       double* elapsed_ms);

/* Effective pair interactions: replaces epi_blocks / epi_blocks_blk (src/asr_epi.f90:628-647,740-759) with their start
 * vectors (epi_start :578-587) and chains (chain / chain_blk :590-626,650-737) on the route-edc operator
 * H(z) = E_nu + h - h o h + X(z) (augcls_apply + augcls_rand_c).  Per block j: the symmetry-adapted start combination
 * v9 (9 values on the start node, unit norm) with its characters chi (nop values +-1 under the operator's group,
 * augcls_set_chi), the chain from it projected on |v> = h'|v9> and the chain from |v> / |v| projected on it, both to full
 * depth nlev.  params: nchain x 4 x nsp x 9 and pab: nchain x 2 x nsp x 9 as epi_setup (c, d = D_c^1/2, o, e = 0 and
 * e_s(z) = z pa + pb, species nsp - 1 the path point); frozen_sp: species of the start site (fz slot 1, the path point);
 * frozen_sp2: species of the pair partner (fz slot 2) or NULL; z: nchain complex.  x: nchain x 4 complex = (X0 = G_00,
 * X1 = (hG)_00, X2 = (hGh)_00, X1a - X1b) for the trace of epi_integrand; a, b, c: nchain x 2 x ldab complex; mused:
 * nchain x 2.  A chain that closed (b_k < 1e-12) gets no tail, the others the operator's tail.  mode as
 * asr_gpu_ed3_blocks. */
int asr_gpu_epi_blocks(asr_gpu_op* op, int nchain, const double* v9, const double* chi, const double* params,
                       const double* pab, const int* frozen_sp, const int* frozen_sp2, const double* conc, const double* z,
                       const int* nlev, int project, int ldab, double* x, double* a, double* b, double* c, int* mused,
                       int mode, double* elapsed_ms);

/* Gradient of a loss L(a, b) of real Lanczos chains with respect to the chain parameters (asr_gpu_adjoint of the
 * replacement interface).  The reference has no adjoint: its nearest analogue is the Hellmann-Feynman trace of EPI,
 * dOmega/dlambda = -(1/pi) Im sum_j w_j Tr[G(z_j) dH/dlambda] (src/asr_epi.f90:14-28), for site-0 parameters only.  Here the
 * forward is asr_gpu_lanczos_real (lanczos, src/reference/asr_recursion.f90:17-58) for nchain chains of depth nlev on
 * H = c + d S d (the o and e parameter rows must be zero), with the Lanczos vectors checkpointed every ckpt levels
 * (0 = 8); the reverse sweep recomputes each segment from its checkpoint and runs the adjoint three-term recursion with
 * one operator application per level.  dl_da, dl_db: nchain x nlev (dL/da_n, dL/db_n of every chain).  a, b: nchain x
 * nlev, the forward's coefficients.  grad: nchain x 2 x nsp x 3, dL/dc and dL/dd of each species summed over the l shell
 * (s, p, d) of the nine orbitals, per chain (sum over chains for parameters shared between chains).  fwd_ms, adj_ms:
 * device time of the forward and of the reverse sweep. */
int asr_gpu_lanczos_adjoint(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const int* frozen_sp,
                            const double* conc, int nlev, int ckpt, const double* dl_da, const double* dl_db, double* a,
                            double* b, double* grad, double* fwd_ms, double* adj_ms);

/* Per-kernel durations of the chain level step (HIP events around every launch, each averaged over the levels in which
 * it ran): the chains of asr_gpu_chains_edc (pab != NULL) or asr_gpu_lanczos_real (pab == NULL).  On return *nk (<= 16)
 * launches per level, kernel_ms[q] their mean duration and kind[q] their kind: 0 operator pass (S pass + flip terms),
 * 1 flip-gather terms only, 2 a_n reduction, 3 Lanczos update + projection, 4 b_n reduction, 5 scaling (every 16 levels
 * for the complex chains), 6 the terms of the nodes with flip partners, 7 operator pass with the three-term update (the
 * fused edc level).  Measurement support; no reference counterpart. */
int asr_gpu_profile_chains(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                           const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int* nk,
                           double* kernel_ms, int* kind);

/* ---- KPM on a supercell realisation (lmto_real_op, src/reference/asr_lmto_realization.f90:14-23) */
/* nbr: nsite x nslot neighbour sites (nslot <= 18), species: nsite bytes (signed char, 0 .. nsp - 1), par: 4 x nsp x 9
 * (c, d, o, e; o = e = 0 required: the kernel applies h = c + d S d), block: nslot x 9 x 9, block0: 9 x 9. */
int asr_gpu_real_create(int nsite, int nslot, const int* nbr, int nsp, const signed char* species, const double* par,
                        const double* block, const double* block0, asr_gpu_real** out);
void asr_gpu_real_destroy(asr_gpu_real* r);
/* Form of the Chebyshev steps: 1 or 2 vectors per lane on 64-vector rows (2: 16-byte loads, two sites per wave), 3 = 16-vector
 * slices, 4 = the block contraction on the FP64 matrix cores (one wave per site, 18 neighbour slots; other slot counts take
 * form 2), or 0 = automatic (2). */
int asr_gpu_real_set_tile(asr_gpu_real* r, int vt);
/* Moment doubling for asr_gpu_kpm_moments (on != 0): the Chebyshev step that makes t_m = T_m(H~) r also gives
 * mu_2m = 2 <t_m|t_m> / dim - mu_0 and mu_2m-1 = 2 <t_m|t_m-1> / dim - mu_1 (T_2m = 2 T_m^2 - T_0, T_2m-1 = 2 T_m T_m-1 - T_1),
 * so nmom moments take nmom / 2 steps instead of nmom - 1.  The moments equal those of kpm_moments
 * (src/reference/asr_kpm.f90:39-65, mu_n = <r|t_n> / dim) up to rounding; the default (off) computes them as it does. */
int asr_gpu_real_set_doubling(asr_gpu_real* r, int on);
/* y = h x for nvec vectors (h_apply, src/reference/asr_lmto_realization.f90:120-147); x, y: nvec x 9 nsite. */
int asr_gpu_real_apply(asr_gpu_real* r, int nvec, const double* x, double* y);
/* Chebyshev moments with nvec Rademacher vectors as one batched SpMM per moment (kpm_moments,
 * src/reference/asr_kpm.f90:39-65).  The vectors are r(v, row) = +-1 from splitmix64((seed << 40) ^ (v << 28) ^ row).
 * mu: nmom; mu_per: nvec x nmom or NULL; step_ms: mean duration of one Chebyshev step; total_ms: whole recurrence. */
int asr_gpu_kpm_moments(asr_gpu_real* r, double c, double s, int nmom, int nvec, uint64_t seed, double* mu, double* mu_per,
                        double* step_ms, double* total_ms);
/* asr_gpu_kpm_moments for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): no counterpart in the reference.  The same recurrence, vectors and moments (mu
 * bitwise equal to asr_gpu_kpm_moments'), with its work arrays (three Chebyshev vectors of 9 nsite x nvec rounded up to
 * 64, the block partials and the per-vector moments) in a fixed device scratch of 64 MiB that the library holds on each
 * device until the process ends; a size beyond it returns an error.  mu_dev: nmom doubles of device memory (output).  No
 * timing, no per-vector output; the call returns when the moments are in mu_dev.  Every call on one GPU shares the
 * scratch: calls that overlap on one GPU (several host threads, several handles) race. */
int asr_gpu_kpm_moments_noalloc(asr_gpu_real* r, double c, double s, int nmom, int nvec, uint64_t seed, double* mu_dev);
/* Extremal eigenvalues of h = c + d S d for the KPM scaling (spectral_bounds, src/reference/asr_kpm.f90:14-36): nstep
 * Lanczos steps from nvec (<= 64) random start vectors at once (vector v: uniform [-1/2, 1/2) entries from
 * splitmix64(((seed + v) << 40) ^ row)), the eigenvalues of each chain's tridiagonal matrix (the role of dsterf);
 * emin = smallest - margin, emax = largest + margin over the chains.  a, b: nvec x nstep coefficients or NULL. */
int asr_gpu_spectral_bounds(asr_gpu_real* r, int nstep, int nvec, uint64_t seed, double margin, double* emin, double* emax,
                            double* a, double* b);

#ifdef __cplusplus
}
#endif
#endif
