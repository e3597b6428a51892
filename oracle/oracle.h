/*
 * oracle.h -- FP64 CPU restatement of monofonIC's nLPT field pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker.
 *
 * Every function restates a reference routine literally (same loop order,
 * same index maps, same scalar factors) and cites the reference file:line it
 * follows.  Arrays use the reference's in-place r2c layout
 * (src/grid_fft.cc:43-112):
 *   real space : relem(i,j,k) = data[(i*n + j)*(n+2) + k]
 *   k space    : kelem(i,j,k) = cdata[(i*n + j)*(n/2+1) + k]   (complex)
 * for cubic grids of side n.  Complex values are interleaved (re,im) doubles.
 *
 * Parity unpinned (restated by reading the reference; no golden vector,
 * closed form or invariant fixes them, so a mistake here would go unseen):
 *   - orc_ngenic_seedtable: the order in which the four blocks of each octant
 *     draw their seeds (random_ngenic.cc:47-75); its frequency mapping is
 *     pinned by the N / N/2 mode agreement;
 *   - the Threefry counter packing of a Panphasia cell and the conversion of
 *     its words to uniforms with tail replacement
 *     (uniform_rand_threefry4x64.c:254-493); the Threefry4x64-20 function
 *     itself is pinned by the reference's three known-answer tests;
 *   - the Panphasia descriptor check digit (uniform_rand_threefry4x64.c:495,
 *     restated in the tests) and the assignment of a cell's deviates to
 *     parity blocks (high_order_panphasia_routines.c:73-176); the cell solve
 *     and the synthesis are pinned by moment propagation and a direct sum.
 */
#ifndef MFX_ORACLE_H
#define MFX_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------- 1D / 3D FFT (unitary, FFTW sign conventions) ---------------- */
/* out-of-place complex DFT of length n, sign -1 forward, +1 backward, unnormalised */
void orc_dft1d(const double *in, double *out, int n, int sign);
/* in-place 3D r2c on the padded real layout, then multiply by 1/sqrt(n^3)
 * (grid_fft.cc:209-238 fftw_execute + ApplyNorm) */
void orc_fft_r2c(double *data, int n);
/* in-place 3D c2r, Hermitian semantics of FFTW c2r (imag parts of the kz=0 and
 * kz=n/2 coefficients ignored), then multiply by 1/sqrt(n^3) (grid_fft.cc:242-270) */
void orc_fft_c2r(double *data, int n);

/* ---------- k-space helpers (include/grid_fft.hh:282-390) --------------- */
/* get_k component: (i - [i>n/2] n) * kfac */
double orc_get_k(size_t i, size_t n, double kfac);
/* gradient component value (imaginary part): 0 at i == n/2 */
double orc_grad_k(size_t i, size_t n, double kfac);

/* ---------- multipliers used by BaseConvolver (convolution.hh:74-292) ---- */
/* A multiplier is a sum of up to two terms  coef * prod_d f_d(i_d)^p_d, where
 * f_d is get_k (nyq=0) or the Nyquist-zeroed gradient (nyq=1). */
typedef struct {
  int nterms;
  double coef_re[2], coef_im[2];
  int pow[2][3];
  int nyq[2];
} orc_mult_t;

enum { ORC_OP_ASSIGN = 0, ORC_OP_ADD = 1, ORC_OP_SUB = 2, ORC_OP_MULADD = 3 };

/* OrszagConvolver::convolve2 (convolution.hh:512-535): pad_insert both inputs
 * (569-592), c2r at 3n/2, pointwise product, r2c, unpad + op (617-653). */
void orc_convolve2(const double *fa, const orc_mult_t *ma, const double *fb, const orc_mult_t *mb,
                   double *out, int op, double opfac, int n, double boxlen);
/* OrszagConvolver::convolve3 (convolution.hh:542-549) */
void orc_convolve3(const double *fa, const orc_mult_t *ma, const double *fb, const orc_mult_t *mb,
                   const double *fc, const orc_mult_t *mc, double *out, int op, double opfac, int n,
                   double boxlen);
/* convenience: Hessian multiplier built from gradient() (convolution.hh:136-160) */
void orc_mult_hessian(orc_mult_t *m, int a, int b);
void orc_mult_sum_hessians(orc_mult_t *m, int a, int b, int c, int d);
void orc_mult_diff_hessians(orc_mult_t *m, int a, int b, int c, int d);
void orc_mult_gradient(orc_mult_t *m, int a);          /* gradient(): +i k, Nyquist zeroed */
void orc_mult_getk_gradient(orc_mult_t *m, int a);     /* (0,-k) from get_k, no zeroing */
void orc_mult_getk_hessian(orc_mult_t *m, int a, int b); /* -k_a k_b from get_k */
void orc_mult_identity(orc_mult_t *m);

/* ---------- k-space operators ------------------------------------------- */
/* ic_generator.cc:341-346 */
void orc_noise_normalise(double *wn, int n, double volfac, int fixing, int inversion);
/* ic_generator.cc:450-455: phi = -wn*A(|k|)/|k|^2, then zero_DC_mode.  The
 * amplitude is looked up in amp_table[n2] with n2 = ix^2+iy^2+iz^2 (signed
 * wave numbers), evaluated by the caller at k = kfac*sqrt(n2). */
void orc_phi1(double *phi, const double *wn, int n, double boxlen, const double *amp_table);
/* grid_fft.hh:440-449 apply_InverseLaplacian + zero_DC_mode */
void orc_inverse_laplacian(double *f, int n, double boxlen);
/* operator*= in k space */
void orc_scale(double *f, int n, double s);

/* ---------- LPT sequence (ic_generator.cc:464-563) ---------------------- */
/* Given phi (k space, before scaling by g1) computes phi2, phi3, A3[0..2]
 * (k space, inverse Laplacian applied, NOT yet scaled by growth factors).
 * order = 2 or 3.  phi3/A3 may be NULL when order == 2. */
void orc_lpt_terms(const double *phi, double *phi2, double *phi3, double *A3x, double *A3y,
                   double *A3z, int n, double boxlen, int order);

/* ---------- assembly (ic_generator.cc:821-944, particle_generator.hh:499-638) */
/* Builds the k-space field for displacement (kind 0) or velocity (kind 1)
 * component idim from the already growth-scaled potentials, applies unit
 * factor, zero_DC, c2r; writes the real field into out (padded real layout). */
void orc_assemble_component(const double *phi, const double *phi2, const double *phi3,
                            const double *A3x, const double *A3y, const double *A3z,
                            double *out, int n, double boxlen, int order, int kind, int idim,
                            double vfac, double unit);
/* pos[p] = q_d*lunit + field (sc lattice, particle order (i,j,k)) */
void orc_set_positions(const double *field, double *pos, int n, int idim, double lunit);
/* Grid_FFT::shift_field (grid_fft.hh:832-842) of a real padded field, in place */
extern int orc_lattice_mutation;
void orc_shift_field(double *f, int n, double boxlen, double sx, double sy, double sz);
/* Grid_FFT::FourierInterpolateCopyTo (grid_fft.cc:274-333, the single-process
 * branch) between k-space buffers of sides f[3] and t[3] (complex
 * interleaved, n0 x n1 x (n2/2+1)): the destination is zeroed and receives the
 * modes of equal signed wave number below both grids' Nyquist modes.  Index
 * bounds in signed arithmetic (the reference's size_t bounds wrap when a side
 * more than doubles). */
void orc_fourier_copy(const double *from, const int f[3], double *to, const int t[3]);
/* Grid_FFT<ccomplex_t> transform of an n^3 complex grid, unitary, in place */
void orc_fft_c2c(double *z, int n, int sign);
/* semiclassical (PPT) Eulerian density and velocities (ic_generator.cc:674-790) */
extern int orc_ppt_mutation;
void orc_ppt_fields(const double *phi, const double *phi2, int n, double boxlen, double D, double vunit,
                    double *delta, double *v, double *hbar_out);
/* set_positions with a sublattice offset s (cells) along idim */
void orc_set_positions_shifted(const double *field, double *pos, int n, int idim, double lunit, double s);

/* ---------- reductions / diagnostics (grid_fft.cc:1133-1210) ------------- */
int orc_power_spectrum(const double *f, int n, double boxlen, double *bin_k, double *bin_P,
                       double *bin_eP, double *bin_count, int maxbins);

/* ---------- random numbers ---------------------------------------------- */
/* MT19937 (Matsumoto & Nishimura 1998) with GSL seeding (0 -> 4357) */
typedef struct { uint32_t mt[624]; int mti; } orc_mt_t;
void orc_mt_seed(orc_mt_t *s, unsigned long seed);
uint32_t orc_mt_next(orc_mt_t *s);
/* Leva (1992) ratio-of-uniforms normal deviate, as used through
 * gsl_ran_ugaussian_ratio_method (random_music_wnoise_generator.cc:613) */
double orc_gauss_ratio(orc_mt_t *s);
/* the same with the constants (s, t, a, b, r1, r2) supplied */
double orc_gauss_ratio_c(orc_mt_t *s, const double *c6);
/* count deviates of the stream MT(seed); c6 = NULL: the published constants */
void orc_gauss_ratio_bulk(unsigned long seed, size_t count, const double *c6, double *out);
/* the first count words of MT19937(seed) */
void orc_mt_words(unsigned long seed, size_t count, uint32_t *out);
/* MUSIC1 white noise fill (random_music_wnoise_generator.cc:35-66,588-627,747-796;
 * random_music.cc:92-116) into the real-space padded layout, FP64 */
void orc_music1_fill(double *data, int n, long baseseed, int cubesize);
/* same, but with the cube values rounded to float before the mean removal
 * (the FLOAT build stores real_t=float) */
void orc_music1_fill_f32(float *data, int n, long baseseed, int cubesize);

/* Threefry-4x64-20 (Salmon et al. 2011; threefry.h) */
void orc_threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);
/* Panphasia-HO uniform stream for one cell, first `count` values
 * (uniform_rand_threefry4x64.c:254-493) */
void orc_panphasia_uniforms(size_t l, size_t j1, size_t j2, size_t j3, size_t seed, double *u,
                            int count);
/* Box-Muller in double (high_order_panphasia_routines.c:196-237) */
void orc_box_muller(const double *u, double *g, int count);
/* cell white noise: first Box-Muller deviate of the cell's stream, level l = log2 n */
void orc_threefry_fill(double *data, int n, size_t seed);
/* the same fill evaluated at m cells (i, j, k) = ijk[3c..3c+2] of an n^3 grid */
void orc_threefry_cells(const long long *ijk, int m, int n, size_t seed, double *out);

/* NGenIC (random_ngenic.cc): RANLUX ranlxd1 + column fill */
typedef struct { double xdbl[12]; double carry; unsigned int ir, jr, is, is_old, pr; } orc_ranlxd_t;
void orc_ranlxd_seed(orc_ranlxd_t *s, unsigned long seed, int luxury);
double orc_ranlxd_uniform(orc_ranlxd_t *s);
void orc_ngenic_seedtable(unsigned int *table, int n, long seed);
void orc_ngenic_fill(double *cdata, int n, const unsigned int *table);

/* ---------- cosmology (include/cosmology_calculator.hh) ------------------ */
typedef struct {
  double h, Omega_m, Omega_b, Omega_DE, Omega_r, Omega_k, w_0, w_a, n_s, sigma_8, Tcmb;
  /* derived */
  double pnorm, sqrtpnorm, Dnow;
  /* E&H */
  double eh[24];
  /* growth table (log a, log D, log f) + natural cubic spline second derivatives */
  int ntab;
  double *la, *lD, *lf, *d2D, *d2f;
} orc_cosmo_t;
void orc_cosmo_init(orc_cosmo_t *c, double H0, double Omega_m, double Omega_b, double n_s,
                    double sigma_8, double Tcmb, int zero_radiation);
void orc_cosmo_free(orc_cosmo_t *c);
double orc_cosmo_transfer(const orc_cosmo_t *c, double k);
/* the baryon and CDM parts T_b, T_c of TFfit_onek at k in Mpc^-1 */
void orc_cosmo_eh_parts(const orc_cosmo_t *c, double k_mpc, double *tb, double *tc);
double orc_cosmo_Dplus(const orc_cosmo_t *c, double a);
double orc_cosmo_f(const orc_cosmo_t *c, double a);
double orc_cosmo_vfact(const orc_cosmo_t *c, double a);
double orc_cosmo_H(const orc_cosmo_t *c, double a);
double orc_cosmo_amplitude(const orc_cosmo_t *c, double k);
double orc_cosmo_sigma8_unnorm(const orc_cosmo_t *c);

/* ---------- PLT corrections, sc lattice (include/particle_plt.hh:48-585) -- */
/* real-space dynamical matrix on ng^3 sites incl. the r = 0 background
 * (:161-284); out = 6 padded real grids xx, xy, xz, yy, yz, zz */
void orc_plt_dynamical_matrix(int ng, double *out);
/* the two halves of the Ewald split of init_D at the wrapped position ar,
 * added into D (row-major 3x3), for split parameter eta: the real-space sum
 * over images |n_i| <= nimg (add_greensftide_sr, particle_plt.hh:175-191) and
 * the k-space sum over |m_i| <= nk, k = 2 pi m / ng (add_greensftide_lr,
 * :194-204, with the fft_norm 1/ng^3) */
void orc_plt_ewald_real(const double ar[3], double eta, int nimg, double D[9]);
/* planted mistakes for mutation checks: 1 = sign of K1, 2 = k-space exponent */
extern int orc_plt_mutation;
void orc_plt_ewald_kspace(const double ar[3], double eta, int nk, int ng, double D[9]);
/* largest eigenpair of a symmetric 3x3 matrix; eval ascending */
void orc_sym3_eigen(const double A[3][3], double eval[3], double evec_max[3]);
/* the same, with the eigenvector of a degenerate largest pair defined as the
 * projection of kdir onto its eigenspace */
void orc_plt_top_eigen(const double A[3][3], const double kdir[3], double eval[3], double v[3]);
/* align_with_k with a fixed sign for transverse eigenvectors */
void orc_plt_align(double v[3], const double kv[3]);
/* init_D (:68-515): tabs = 4 tables over ng x ng x (ng/2+1): D_r, D_theta,
 * D_phi, growth-rate correction; mu1 (optional) largest eigenvalue */
void orc_plt_tables(int ng, double *tabs, double *mu1);
/* lattice_gradient::gradient (imaginary parts) and vfac_corr (:551-585) */
void orc_plt_gradient(const double *tabs, int ng, int n, double boxlen, size_t i, size_t j, size_t k, double g[3],
                      double *vfac);
/* assembly with the PLT gradient (ic_generator.cc:821-919, ENABLE_PLT) */
void orc_assemble_component_plt(const double *phi, const double *phi2, const double *phi3, const double *A3x,
                                const double *A3y, const double *A3z, const double *tabs, int ng, double *out, int n,
                                double boxlen, int order, int kind, int idim, double vfac, double unit);

#ifdef __cplusplus
}
#endif

/* Panphasia-HO (panphasia.c) */
extern int orc_pan_mutation;
void orc_pan_tables(double *C, int *deg, int *rowidx);
void orc_pan_solve(const double *parent, const double *children_in, double *out, int constrained);
void orc_pan_cell_gauss(size_t level, size_t j1, size_t j2, size_t j3, double *g);
void orc_pan_root(double *root);
void orc_pan_coefficients(uint64_t ox, uint64_t oy, uint64_t oz, int base_level, int rel, size_t np, double *coef);
double orc_sph_bessel(int l, double x);
void orc_panphasia_noise(uint64_t ox, uint64_t oy, uint64_t oz, int base_level, uint64_t size, uint64_t kk, int n,
                         double *out);

#endif
