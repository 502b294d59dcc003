/* Host-side C ABI of the augmented-space chain engine (libasr_host.so): the node-set builder and the routines that stay on
 * the host around the GPU chains (terminated continued fractions, tails, KPM DOS, contours).
 *
 * Conventions: 0-based indices, -1 = no link, row-major arrays, complex numbers as interleaved (re, im) doubles.  Every
 * function returns 0 on success; a non-zero status is an error (the Fortran side turns it into `error stop`).
 */
#ifndef ASR_HOST_H
#define ASR_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct asr_augcls_tables asr_augcls_tables;

/* Node set and node tables of the site-class, point-group reduced augmented-space operator.
 * Replaces build_augcls_operator (src/asr_augcls.f90:596-823) for the table part: the caller keeps the bond blocks,
 * species parameters and orbital tables.  perm: nop x nsite (operation 0 = identity), flippable: nsite bytes or NULL,
 * nsc: states per class, maxrank <= 0: unlimited.  err receives a message on failure. */
int asr_host_build_augcls(int nsite, const int* nbr_ptr, const int* nbr, const int* cls, int nop, const int* perm, int start,
                          int maxlevel, int maxrank, const unsigned char* flippable, const int* nsc, int ncls, int nthreads,
                          asr_augcls_tables** out, char* err, int errlen);
/* dims = {nn, nslot, nop, nsmax, nstab} */
int asr_host_tables_dims(const asr_augcls_tables* t, int* dims);
/* hopr/hopg: nn x nslot, flipr/flipg: nn x nsmax, lst/site/level/orb: nn, stab_ptr: nn + 1, stab: nstab */
int asr_host_tables_get(const asr_augcls_tables* t, int* hopr, int* hopg, int* flipr, int* flipg, int* lst, int* site,
                        int* level, int* orb, int* stab_ptr, int* stab);
/* Target nodes (build_augcls_operator with tsite, src/asr_augcls.f90:805-819): for target site tsite[t] and local state k
 * (the site alone in state k, every other site in its reference state) tnode[t * nsmax + k] = representative node,
 * tgop[t * nsmax + k] = the operation that maps the node onto it (augcls_amp_c); -1 = not in the node set. */
int asr_host_tables_targets(const asr_augcls_tables* t, int nt, const int* tsite, int* tnode, int* tgop);
void asr_host_tables_free(asr_augcls_tables* t);

/* Terminated continued fraction of a complex-symmetric chain: cf_species (src/asr_chain.f90:67-167).
 * tail: 0 hull, 1 bp, 2 avg2, 3 none.  Returns the number of bp fallbacks. */
int asr_host_cf_species(const double* a, const double* b, const double* c, int m, const double* z, int tail, double* g11,
                        double* f, double* edges, double* pv);
/* The same for a chain that closed (b_m < tiny, Krylov space exhausted): the finite fraction without a tail, cf_closed
 * (src/asr_ed3.f90:214-230). */
int asr_host_cf_closed(const double* a, const double* b, const double* c, int m, const double* z, double* g11, double* f);
/* G(z) of n chains at once (cf_species per chain with the same tail, src/asr_chain.f90:162-262): a, b n x ld complex,
   m[j] levels of chain j, z n complex, g11 n complex out.  The host end of a contour pass (chain_greens,
   src/asr_class_scf.f90:1127-1150).  Returns 0 (a tail that cannot be formed falls back as in cf_species). */
int asr_host_cf_species_batch(int n, const double* a, const double* b, int ld, const int* m, const double* z, int tail,
                              double* g11);
/* Beer-Pettifor asymptote: bp_asymptote (src/reference/asr_terminator.f90:39-67).  1 = not converged. */
int asr_host_bp_asymptote(const double* a, const double* b, int n, double* ainf, double* binf, int* niter);
/* G(z) of a real chain with a square-root tail: terminated_green (src/reference/asr_terminator.f90:117-126). */
int asr_host_terminated_green(const double* a, const double* b, int n, double ainf, double binf, int ne, const double* z,
                              double* g);
/* Jackson-kernel DOS from Chebyshev moments: kpm_dos (src/reference/asr_kpm.f90:68-87). */
int asr_host_kpm_dos(const double* mu, int nmom, double c, double s, int ne, const double* e, double* rho);
/* Eigenvalues of a real symmetric tridiagonal matrix (ascending, in place in d): the role of dsterf. */
int asr_host_tridiag_eigvals(int n, double* d, const double* e);
/* Eigenvalues of the complex-symmetric tridiagonal chain matrix: the role of zgeev in cf_species. */
int asr_host_chain_eigvals(int n, const double* a, const double* b, double* ev);
/* Gauss-Legendre rule (src/asr_contour.f90:132-151) and the closed occupation contour (src/asr_contour.f90:39-58). */
int asr_host_gauss_legendre(int n, double* x, double* w);
int asr_host_contour_closed(double ebot, double ef, double eta, int n_arc, int n_vert, double* z, double* w);
/* Semicircle E_bot -> E_F (contour_nodes, src/asr_contour.f90:17-34; eta_end <= 0: no eta_end).  1 = bad arguments. */
int asr_host_contour_nodes(double ebot, double ef, int n, double eta_end, double* z, double* w);
/* Fermi-smeared contour: fermi_npole, fermi_nodes_count, contour_nodes_fermi (src/asr_contour.f90:61-121).  z, w hold
   fermi_nodes_count(npole, n_main, n_step) (re, im) pairs.  1 = bad arguments. */
int asr_host_fermi_npole(double kt);
int asr_host_fermi_nodes_count(int npole, int n_main, int n_step);
int asr_host_contour_fermi(double ebot, double mu, double kt, int npole, int n_main, int n_step, double* z, double* w);

#ifdef __cplusplus
}
#endif
#endif
