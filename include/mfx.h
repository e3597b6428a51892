/*
 * mfx.h -- C ABI of the MI355X (gfx950) nLPT field pipeline.
 *
 * Plain C, plain pointers and sizes.  Each entry point names the monofonIC
 * interface it replaces (file:line in the monofonIC source tree).  Every call
 * returns 0 on success and a nonzero status otherwise; mfx_last_error()
 * returns the message of the last failure on the calling thread.
 *
 * Device memory of a grid belongs to its mfx_grid handle (this replaces the
 * fftw_malloc ownership of Grid_FFT, src/grid_fft.cc:50,97).  All work of a
 * context is issued in order on one HIP stream; calls return after enqueueing
 * unless documented otherwise (downloads, reductions and mfx_ctx_synchronize
 * block).  Grids are cubic or rectangular with even last dimension; the memory
 * layout is the reference's in-place r2c layout:
 *   real space relem(i,j,k) = data[(i*n1 + j)*(n2+2) + k]
 *   k space    kelem(i,j,k) = cdata[(i*n1 + j)*(n2/2+1) + k]
 */
#ifndef MFX_H
#define MFX_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFX_OK 0
#define MFX_EINVAL 1 /* invalid argument or unsupported size */
#define MFX_EHIP 2   /* HIP runtime failure */
#define MFX_ESTATE 3 /* call not valid in the grid's current space */

#define MFX_FLOAT 4
#define MFX_DOUBLE 8

/* space_t (include/grid_fft.hh:29) */
#define MFX_KSPACE 0
#define MFX_RSPACE 1

/* output operators op::assign_to/add_to/subtract_from/multiply_add_to
 * (include/operators.hh:25-35) */
#define MFX_OP_ASSIGN 0
#define MFX_OP_ADD 1
#define MFX_OP_SUB 2
#define MFX_OP_MULADD 3

typedef struct mfx_ctx mfx_ctx;
typedef struct mfx_grid mfx_grid;
typedef struct mfx_conv mfx_conv;
typedef struct mfx_cosmo mfx_cosmo;

/* ------------------------------------------------------------ context */
const char *mfx_last_error(void);
int mfx_version(void);
/* one HIP stream on `device`; replaces the FFTW/OpenMP set-up of src/main.cc:156-182 */
int mfx_ctx_create(int device, mfx_ctx **out);
int mfx_ctx_destroy(mfx_ctx *c);
int mfx_ctx_synchronize(mfx_ctx *c);
/* stream-ordered timers: record event `slot` (0..15), elapsed ms between two
 * recorded slots (blocks until the later one completed) */
int mfx_event_record(mfx_ctx *c, int slot);
int mfx_event_elapsed(mfx_ctx *c, int slot0, int slot1, float *ms);
/* per-kernel profiling with HIP events around every launch: name, launches,
 * total ms, total algorithmic bytes */
int mfx_profile_enable(mfx_ctx *c, int on);
int mfx_profile_reset(mfx_ctx *c);
int mfx_profile_count(mfx_ctx *c, int *count);
int mfx_profile_get(mfx_ctx *c, int idx, char *name, size_t namelen, int *launches, double *ms, double *bytes);
/* device memory owned by the context (bytes), e.g. FP64 particle arrays */
int mfx_device_alloc(mfx_ctx *c, size_t bytes, void **ptr);
int mfx_device_free(mfx_ctx *c, void *ptr);
int mfx_memcpy_d2h(mfx_ctx *c, void *host, const void *dev, size_t bytes);
int mfx_memcpy_h2d(mfx_ctx *c, void *dev, const void *host, size_t bytes);

/* ------------------------------------------------------------ autotuning
 * No counterpart in the reference (FFTW plans are made with FFTW_ESTIMATE,
 * src/grid_fft.cc:53-54).  Times the column-pass variants of lengths n and
 * 3n/2 in `precision` -- tile width (64 KiB tile, half or twice), plain or
 * persistent pipelined kernel, and for lengths 384, 768, 1024 and 1536 the
 * second radix plan -- and keeps the fastest per length and direction in
 * the context.  Tile width and kernel form never change results; the radix
 * plan changes them at the level of rounding. */
int mfx_autotune(mfx_ctx *c, size_t n, int precision);
/* set the column-pass variant of length L explicitly: the tile width (the
 * widest tile whose exchange area fits 64 KiB of LDS, half of it, or -- where
 * one workgroup can hold it -- twice it), or'ed with MFX_COL_PIPELINED for the
 * persistent pipelined kernel and with MFX_COL_PLAN_ALT for the length's
 * second radix plan (384 = 12.8.4, 768 = 12.8.8, 1024 = 16.16.4,
 * 1536 = 12.8.8.2 instead of 12.4.4.2, 12.4.4.4, 8.8.4.4, 12.4.4.4.2) */
#define MFX_COL_PIPELINED 0x100
#define MFX_COL_PLAN_ALT 0x200
int mfx_set_column_tile(mfx_ctx *c, int L, int precision, int inverse, int tile);
/* text table of the decisions taken so far */
int mfx_autotune_report(mfx_ctx *c, char *buf, size_t len);

/* ------------------------------------------------------------ Grid_FFT
 * Grid_FFT<real_t,false> ctor/allocate/reset (include/grid_fft.hh:77-101,
 * src/grid_fft.cc:43-112).  precision MFX_FLOAT or MFX_DOUBLE. */
int mfx_grid_create(mfx_ctx *c, size_t n0, size_t n1, size_t n2, double L0, double L1, double L2, int precision,
                    int space, mfx_grid **out);
int mfx_grid_destroy(mfx_grid *g);
/* ntot_ (number of reals), space_, precision */
int mfx_grid_info(const mfx_grid *g, size_t *ntot, int *space, int *precision);
int mfx_grid_data(mfx_grid *g, void **dev_ptr);
/* whole ntot_ buffer; nbytes must equal ntot*precision */
int mfx_grid_upload(mfx_grid *g, const void *host, size_t nbytes);
int mfx_grid_download(const mfx_grid *g, void *host, size_t nbytes);
/* Grid_FFT::zero (grid_fft.hh:150-156) */
int mfx_grid_zero(mfx_grid *g);
/* Grid_FFT::copy_from (grid_fft.hh:158-178) */
int mfx_grid_copy_from(mfx_grid *dst, const mfx_grid *src);
/* FourierTransformForward/Backward (src/grid_fft.cc:209-270): unitary
 * 1/sqrt(N) normalisation, no-op when already in the target space;
 * do_transform = 0 only flips the space flag */
int mfx_grid_fft_forward(mfx_grid *g, int do_transform);
int mfx_grid_fft_backward(mfx_grid *g, int do_transform);
/* operator*= (grid_fft.hh:392-403), applied in the current space */
int mfx_grid_scale(mfx_grid *g, double s);
/* apply_InverseLaplacian (grid_fft.hh:440-449): forward FFT if needed,
 * -x/|k|^2, zero DC */
int mfx_grid_inverse_laplacian(mfx_grid *g);
/* zero_DC_mode (grid_fft.hh:844-887) */
int mfx_grid_zero_dc(mfx_grid *g);
/* mean() and std() (grid_fft.hh:507-595).  In k space std() runs over the
 * n0 n1 (n2/2+1) modes (real parts and |z|^2) and mean() over the raw buffer
 * relem(i,j,k) with k < n2/2+1, as the reference's loops over sizes_ do */
int mfx_grid_stats(mfx_grid *g, double *mean, double *std);
/* Grid_FFT::shift_field (include/grid_fft.hh:832-842): the grid is taken to k
 * space (FourierTransformForward, a no-op there), every mode is multiplied by
 * exp(i s.k dx) with k from get_k, dx = L/n per axis and (sx, sy, sz) in
 * cells, and with transform_back the grid returns to real space.  The phase
 * is formed in FP64. */
int mfx_grid_shift_field(mfx_grid *g, double sx, double sy, double sz, int transform_back);
/* Grid_FFT::FourierInterpolateCopyTo (src/grid_fft.cc:274-333, the single-process
 * branch): `from` is taken to k space, `to` is marked k space without a
 * transform; equal sides copy the buffer (copy_from), other sides zero `to` and
 * copy the modes of equal signed wave number below both grids' Nyquist modes,
 * unchanged (unitary amplitudes).  Both grids belong to one context and have one
 * precision.  The receive bounds are signed, so a side may also more than double
 * (where the reference's size_t bounds wrap). */
int mfx_grid_fourier_interpolate_copy(mfx_grid *from, mfx_grid *to);
/* Compute_PowerSpectrum (src/grid_fft.cc:1133-1210) */
int mfx_grid_power_spectrum(mfx_grid *g, double *bin_k, double *bin_P, double *bin_eP, double *bin_count,
                            int maxbins, int *nbins);

/* ------------------------------------------------------------ k-space operators of ic_generator::run */
/* noise normalisation, fixing and inversion (src/ic_generator.cc:341-346) */
int mfx_noise_normalise(mfx_grid *wn, double volfac, int fixing, int inversion);
/* phi = -wn * A(|k|)/|k|^2 and zero_DC_mode (src/ic_generator.cc:450-457);
 * amp_table[n2] = A(kfac*sqrt(n2)) in FP64 for n2 = ix^2+iy^2+iz^2 <= 3(n/2)^2 */
int mfx_phi1(mfx_grid *phi, const mfx_grid *wn, const double *amp_table, size_t ntable);

/* ------------------------------------------------------------ OrszagConvolver
 * OrszagConvolver<real_t>(N,L) (include/convolution.hh:463-470).  A Fourier
 * multiplier is sum_t coef_t * prod_d f_d(i_d)^pow_t[d], with f_d the
 * Nyquist-zeroed gradient wave number (nyq = 1, Grid_FFT::gradient) or get_k
 * (nyq = 0). */
typedef struct {
  int nterms;
  double coef_re[2], coef_im[2];
  int pow[2][3];
  int nyq[2];
} mfx_mult;

int mfx_conv_create(mfx_ctx *c, size_t n, double L, int precision, mfx_conv **out);
int mfx_conv_destroy(mfx_conv *cv);
/* generic OrszagConvolver::convolve2 (convolution.hh:512-535) with the
 * multipliers evaluated inside pad_insert, op applied by unpad (617-653) */
int mfx_conv2(mfx_conv *cv, mfx_grid *a, const mfx_mult *ma, mfx_grid *b, const mfx_mult *mb, mfx_grid *out,
              int op, double opfac);
/* OrszagConvolver::convolve3 (convolution.hh:542-549) */
int mfx_conv3(mfx_conv *cv, mfx_grid *a, const mfx_mult *ma, mfx_grid *b, const mfx_mult *mb, mfx_grid *c,
              const mfx_mult *mc, mfx_grid *out, int op, double opfac);
/* BaseConvolver named variants (convolution.hh:74-292) */
int mfx_conv_gradients(mfx_conv *cv, mfx_grid *inl, int d1l, mfx_grid *inr, int d1r, mfx_grid *out, int op,
                       double opfac);
int mfx_conv_gradient_and_hessian(mfx_conv *cv, mfx_grid *inl, int d1l, mfx_grid *inr, const int d2r[2],
                                  mfx_grid *out, int op, double opfac);
int mfx_conv_hessians(mfx_conv *cv, mfx_grid *inl, const int d2l[2], mfx_grid *inr, const int d2r[2],
                      mfx_grid *out, int op, double opfac);
int mfx_conv_hessians3(mfx_conv *cv, mfx_grid *inl, const int d2l[2], mfx_grid *inm, const int d2m[2],
                       mfx_grid *inr, const int d2r[2], mfx_grid *out, int op, double opfac);
int mfx_conv_sum_of_hessians(mfx_conv *cv, mfx_grid *inl, const int d2l[2], mfx_grid *inr, const int d2r1[2],
                             const int d2r2[2], mfx_grid *out, int op, double opfac);
int mfx_conv_difference_of_hessians(mfx_conv *cv, mfx_grid *inl, const int d2l[2], mfx_grid *inr,
                                    const int d2r1[2], const int d2r2[2], mfx_grid *out, int op, double opfac);
int mfx_conv_multiply_field(mfx_conv *cv, mfx_grid *inl, mfx_grid *inr, mfx_grid *out, int op, double opfac);

/* All nLPT source terms of src/ic_generator.cc:464-547 in one call: phi2,
 * phi3 = 3a + 3b and A3 with the inverse Laplacian applied (growth factors
 * not applied).  Same results as the 5 (2LPT) / 33 (3LPT) convolve calls of
 * the reference, computed from the Hessian tensors of phi and phi2 so that
 * each padded transform is done once.  order 2: phi3/A3 may be NULL. */
int mfx_lpt_sources(mfx_conv *cv, mfx_grid *phi, mfx_grid *phi2, mfx_grid *phi3, mfx_grid *A3x, mfx_grid *A3y,
                    mfx_grid *A3z, int order);

/* 3LPT evaluation path of mfx_lpt_sources (no counterpart in the reference,
 * whose convolve calls have no such choice):
 *   MFX_LPT_STREAMED  every stage transforms the Hessian rows it needs
 *                     (workspace: 3 m*n*hp + 16 m*m*hp complex, m = 3n/2,
 *                     hp = n/2+1 rounded up to a multiple of 16);
 *   MFX_LPT_STORED    stage A also stores phi's real-space Hessian rows and
 *                     the phi3 and A3 sources are formed in one pass over them
 *                     (half the padded transforms; workspace 11 instead of 16
 *                     m*m*hp slots plus 6 m^3 reals);
 *   MFX_LPT_AUTO      stored when its workspace fits in free device memory
 *                     with headroom, streamed otherwise (the default).
 * Both paths give the same fields up to rounding.  The choice is made when the
 * 3LPT workspace is first reserved and kept until the path is set again.
 * mfx_conv_lpt_path reports it (MFX_LPT_AUTO while not yet made). */
#define MFX_LPT_AUTO 0
#define MFX_LPT_STREAMED 1
#define MFX_LPT_STORED 2
int mfx_conv_set_lpt_path(mfx_conv *cv, int path);
int mfx_conv_lpt_path(mfx_conv *cv, int *path);

/* ------------------------------------------------------------ assembly
 * k-space displacement (kind 0) or velocity (kind 1) component idim built from
 * the growth-scaled potentials, times unit/L, zero DC
 * (src/ic_generator.cc:821-860, 877-919); tmp is left in k space */
int mfx_assemble_kspace(mfx_grid *tmp, mfx_grid *phi, mfx_grid *phi2, mfx_grid *phi3, mfx_grid *A3x,
                        mfx_grid *A3y, mfx_grid *A3z, int order, int kind, int idim, double vfac, double unit);
/* ------------------------------------------------------------ PLT corrections
 * particle::lattice_gradient for the sc lattice (include/particle_plt.hh:48-585).
 * mfx_plt_create runs init_D (:68-515) on an ngrid^3 mesh of the unit box (the
 * reference uses 64) on the device: Ewald-summed dynamical matrix, FFT, largest
 * eigenpair per mode, spherical projection and growth-rate factor. */
typedef struct mfx_plt mfx_plt;
int mfx_plt_create(mfx_ctx *c, int ngrid, mfx_plt **out);
int mfx_plt_destroy(mfx_plt *p);
/* tables, 4 x ngrid x ngrid x (ngrid/2+1) doubles: D_r (= 1), D_theta, D_phi,
 * growth-rate factor (the D_xx_, D_yy_, D_zz_, D_xy_ grids after init_D) */
int mfx_plt_tables(const mfx_plt *p, double *host, size_t count);
/* real parts of the k-space dynamical matrix, 6 x ngrid x ngrid x (ngrid/2+1):
 * xx, xy, xz, yy, yz, zz, unnormalised transform (:286-313) -- test hook */
int mfx_plt_dmatrix(const mfx_plt *p, double *host, size_t count);
/* lattice_gradient::gradient (imaginary parts, 3 per mode) and vfac_corr
 * (:551-585) at `count` k-space indices ijk (3 per mode) of an n-grid of side
 * boxlen -- test hook */
int mfx_plt_gradient(const mfx_plt *p, size_t n, double boxlen, const long long *ijk, int count, double *g,
                     double *vfac);
/* mfx_assemble_kspace with the PLT gradient and, for velocities, its growth-
 * rate factor (ic_generator.cc:821-919 in ENABLE_PLT builds) */
int mfx_assemble_kspace_plt(mfx_grid *tmp, mfx_grid *phi, mfx_grid *phi2, mfx_grid *phi3, mfx_grid *A3x,
                            mfx_grid *A3y, mfx_grid *A3z, int order, int kind, int idim, double vfac, double unit,
                            const mfx_plt *plt);

/* set_positions / set_velocities for the sc lattice, 64-bit reals
 * (include/particle_generator.hh:499-543, 597-638): out_dev is a device
 * array of n^3 doubles, particle order (i,j,k); kind 0 positions, 1 velocities */
int mfx_set_particles(mfx_grid *field, double *out_dev, int idim, double lunit, int kind);

/* ------------------------------------------------------------ RNG plugins
 * RNG_plugin::Fill_Grid (include/random_plugin.hh:26-39) */
/* MUSIC1 (src/plugins/random_music.cc:92-116, random_music_wnoise_generator.cc
 * :35-66,588-627,747-796): real-space N(0,1), leaves the grid in real space.
 * The grid size is a power of two (random_music.cc:66-74) and the cube size
 * (clipped to the grid size) at most 1024, else an error is returned.  A cube
 * size that does not divide the grid repeats cubes as the reference's cell
 * lookup does (random_music_wnoise_generator.hh:195-222): grid point i reads
 * cube (i / cubesize) mod floor(n / cubesize), cell i mod cubesize, and a cell
 * read by m grid points has the mean subtracted m times, as the reference's
 * zero-mean loop over grid points does (random_music_wnoise_generator.cc:44-65). */
int mfx_rng_music1_fill(mfx_grid *g, long seed, int cubesize);
/* No counterpart in the reference.  The MUSIC1 fill above for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc/hipFree and creates
 * no events; its seed and scratch (mean, per-cube sums) live in a 1 MiB
 * file-scope device array, one per GPU, held until the process ends.  At most
 * 65520 cubes ((n/cubesize)^3, i.e. up to 1280^3 at cube size 32), else an
 * error is returned.  Asynchronous on the grid's context stream (synchronise
 * before reading the grid on the host).  Every call on one GPU shares that
 * GPU's array: calls on different streams of one GPU that overlap race. */
int mfx_rng_music1_fill_noalloc(mfx_grid *g, long seed, int cubesize);
/* NGENIC (src/plugins/random_ngenic.cc:40-75 seed table, :85-152
 * Fill_Grid): k-space noise, leaves the grid in k space */
int mfx_rng_ngenic_fill(mfx_grid *g, long seed);
/* Panphasia-HO Threefry-4x64-20 cell stream (external/panphasia_ho/
 * uniform_rand_threefry4x64.c:254-493): first Box-Muller deviate of each
 * cell (l = log2 n, j = (i,j,k)), real space */
int mfx_rng_threefry_fill(mfx_grid *g, unsigned long long seed);
/* PANPHASIA with a Panphasia-HO descriptor ([Panph6,L..,(x,y,z),S..,(KK..,)CH..,name]):
 * RNG_panphasia::Fill_Grid (src/plugins/random_panphasia.cc:131-180) -- the octree
 * Legendre-block field (external/panphasia_ho/high_order_panphasia_routines.c:73-150,
 * 888-1240), its spherical-Bessel k-space synthesis on N0 = 2 * size * 2^rel points
 * (pan_mpi_routines.c:23-460) and Grid_FFT::FourierInterpolateCopyTo into the grid
 * (grid_fft.cc:274-333); leaves the grid in k space.  Panph1 (Fortran Panphasia)
 * descriptors and an odd size * 2^rel are errors; a check digit other than -999 must
 * match (compute_check_digit_, uniform_rand_threefry4x64.c:495-531). */
int mfx_rng_panphasia_fill(mfx_grid *g, const char *descriptor);
/* descriptor fields as parsed (parse_and_validate_descriptor_,
 * high_order_panphasia_routines.c:1242-1342): out[0..6] order, level, x, y, z,
 * size, KK limit; out[7] the computed check digit; for n > 0 out[8..10] the
 * relative level, coefficient cube size * 2^rel and Fourier grid N0 used for an
 * n-grid (random_panphasia.cc:143-150) */
int mfx_panphasia_info(const char *descriptor, int n, long long out[11]);
/* test hook: the order-6 basis tables as this library builds them from closed forms
 * (pan_matrices_order6.h: constraint_matrices, irank_p, list_by_parity); any
 * pointer may be null.  C is 84 x 84, degrees 3 x 84, row_index 84. */
int mfx_panphasia_tables(double *C, int *degrees, int *row_index);
/* raw streams (test hooks): MT19937 words, ranlxd1 uniforms, Threefry cell
 * uniforms (cells = 4 x ncells of (l, j1, j2, j3)); outputs are host arrays */
int mfx_rng_mt19937_raw(mfx_ctx *c, const unsigned long *seeds, int nseeds, int count, unsigned *out);
int mfx_rng_ranlxd1_raw(mfx_ctx *c, const unsigned *seeds, int nseeds, int count, double *out);
int mfx_rng_threefry_uniforms(mfx_ctx *c, const unsigned long long *cells, int ncells, unsigned long long seed,
                              int count, double *out);

/* ------------------------------------------------------------ cosmology (host, FP64)
 * cosmology::calculator (include/cosmology_calculator.hh:166-210) with the
 * Eisenstein & Hu transfer function (src/plugins/transfer_eisenstein.cc) */
#define MFX_COSMO_TRANSFER 0
#define MFX_COSMO_DPLUS 1
#define MFX_COSMO_F 2
#define MFX_COSMO_VFACT 3
#define MFX_COSMO_H 4
#define MFX_COSMO_AMPLITUDE 5
#define MFX_COSMO_PNORM 6
#define MFX_COSMO_OMEGA_DE 7 /* cosmo_param_["Omega_DE"] (cosmology_parameters.hh) */
int mfx_cosmo_create(double H0, double Omega_m, double Omega_b, double n_s, double sigma_8, double Tcmb,
                     int zero_radiation, mfx_cosmo **out);
int mfx_cosmo_destroy(mfx_cosmo *cc);
int mfx_cosmo_query(const mfx_cosmo *cc, int what, double x, double *out);
int mfx_cosmo_amplitude_table(const mfx_cosmo *cc, size_t n, double boxlen, double *out, size_t ntable);

/* ------------------------------------------------------------ LPT driver
 * ic_generator::run for the sc, bcc, fcc and rsc lattices and generic (field_lagrangian) or
 * particle output (src/ic_generator.cc:90-971): noise, phi(1..3), A(3),
 * growth scaling and FP64 position/velocity assembly, all on the device. */
typedef struct {
  size_t n;
  double boxlen, zstart;
  int lpt_order;        /* 1, 2 or 3 */
  int precision;        /* bulk FFT/convolution precision: MFX_FLOAT or MFX_DOUBLE */
  int generator;        /* 0 MUSIC1, 1 NGENIC, 2 Threefry, 3 PANPHASIA (descriptor) */
  long seed;
  int fixing, inversion;
  double H0, Omega_m, Omega_b, n_s, sigma_8, Tcmb;
  int zero_radiation;
  int use_fused_lpt;    /* 1: mfx_lpt_sources, 0: the reference's convolve call sequence */
  int use_plt;          /* 1: PLT-corrected assembly (ENABLE_PLT builds, ic_generator.cc:366-370) */
  int plt_ngrid;        /* PLT mesh; 0 selects the reference's 64 */
  int autotune;         /* 1: mfx_autotune for this grid size at creation (once per context) */
  int assemble_each;    /* 1: assemble and transform one component at a time in one grid
                           (least memory); 0: all six components from one read of the
                           potentials when device memory allows */
  int field_output;     /* 1: generic (field_lagrangian) output (output_generic.cc:62-67):
                           arrays 0..2 hold the displacements Psi without the lattice
                           coordinate (DM_dx..DM_dz) and array 6 the 1st-order density
                           (DM_delta, ic_generator.cc:945-955); 0: particle positions q + Psi */
  int own_particles;    /* 1: the output arrays get their own allocation, so that a download
                           of one realisation (mfx_icgen_download_async) overlaps the next
                           realisation's noise, LPT and assembly stages; 0: they share the
                           LPT workspace when it is large enough (least memory) */
  int lattice;          /* [setup] ParticleLoad (ic_generator.cc:131-139): 0 sc, 1 bcc, 2 fcc,
                           3 rsc, i.e. 1, 2, 4 or 8 particles per cell
                           (particle_generator.hh:47-54,499-543,597-638); not with PLT,
                           field output or assemble_each */
  int baryons;          /* [setup] DoBaryons (ic_generator.cc:153-168,608-671): a second species
                           on the secondary lattice (particle_generator.hh:56-63), arrays 7..12;
                           with the Eisenstein-Hu transfer function its difference modes
                           delta_bc, theta_bc are zero (transfer_eisenstein.cc:240), so both
                           species follow the total-matter fields; particle output only, not
                           with PLT or assemble_each */
  int eulerian;         /* with field_output: [output] generic_out_eulerian (output_generic.cc:
                           43,63-68), the semiclassical (PPT) Eulerian fields of one species
                           (ic_generator.cc:674-790) instead of the Lagrangian ones: array 6
                           the density |psi|^2 - 1 (DM_delta), arrays 3..5 the velocities
                           (DM_vx..DM_vz); arrays 0..2 are not written.  Not with PLT; grid
                           sizes 16..1024 of the form 2^k or 3*2^k */
  const char *descriptor; /* generator 3 (PANPHASIA): the Panphasia-HO descriptor; the
                             field is computed once at creation (it depends on the
                             descriptor only; seeds do not apply) */
  int cubesize;         /* generator 0 (MUSIC1): [random] cubesize (random_music.cc:120);
                           0 selects the reference's default 32.  A size that does not
                           divide n repeats cubes (see mfx_rng_music1_fill) */
  const double *amplitude_table; /* NULL: the amplitude A(k) of the Eisenstein-Hu transfer
                           function and the cosmology above (mfx_cosmo_amplitude_table);
                           else the caller's A(kfac sqrt(q)) for q = 0 .. 3(n/2)^2, as
                           ic_generator.cc:450-455 evaluates get_amplitude(|k|) from any
                           transfer plugin (CLASS, CAMB file, ...).  Read once, at creation;
                           with DoBaryons it applies to both species.  The growth factors
                           still follow from the cosmology parameters */
  size_t amplitude_ntable; /* entries of amplitude_table, at least 3(n/2)^2 + 1 */
} mfx_icgen_params;

typedef struct mfx_icgen mfx_icgen;
int mfx_icgen_create(mfx_ctx *c, const mfx_icgen_params *p, mfx_icgen **out);
int mfx_icgen_destroy(mfx_icgen *ic);
/* one full realisation with the seed of the parameters; results stay on the
 * device */
int mfx_icgen_run(mfx_icgen *ic);
/* Graph replay of realisations (no counterpart in the reference, which runs
 * one realisation per process).  mfx_icgen_set_seed sets the white-noise seed
 * of the next realisation and starts its host part (for NGENIC the seed table,
 * computed by a worker thread); mfx_icgen_run or mfx_icgen_launch copies it to
 * the device.  mfx_icgen_capture runs one realisation and captures the device
 * sequence into a HIP graph; mfx_icgen_launch replays it on the context's
 * stream.  Replays give the same results as mfx_icgen_run with that seed. */
int mfx_icgen_set_seed(mfx_icgen *ic, long seed);
int mfx_icgen_capture(mfx_icgen *ic);
int mfx_icgen_launch(mfx_icgen *ic);
/* hbar of the last realisation's semiclassical fields (ic_generator.cc:718-725:
 * max |phi(neighbour) - phi| / pi / D+ * 1.01, computed on the device); waits
 * for the realisation.  Eulerian field output only. */
int mfx_icgen_hbar(mfx_icgen *ic, double *hbar);
/* device pointers of the FP64 particle arrays: what 0..2 positions x,y,z
 * (displacements with field_output), 3..5 velocities x,y,z, 6 the density
 * (field_output only) (n^3 doubles each, order (i,j,k); for a bcc, fcc or
 * rsc load arrays 0..5 hold 2, 4 or 8 such blocks, one per sublattice, as
 * particle_generator.hh stores them); with baryons, 7..12 the baryons'
 * positions and velocities on the secondary lattice.  With the fused
 * LPT engine they may share device memory with its workspace: valid until the
 * next mfx_icgen_run */
int mfx_icgen_particles(mfx_icgen *ic, int what, double **dev_ptr);
/* Asynchronous download of the output arrays of the last realisation: bit w
 * of `what` copies array w (0..12 as mfx_icgen_particles: 6 the density of
 * field output, 7..12 the baryons'; n^3 doubles each, times the particles
 * per cell except for 6; dst has an entry per bit up to the highest set)
 * into host buffer dst[w] on the
 * generator's copy stream, after the realisation's last kernel.  Host
 * buffers from mfx_host_alloc (pinned) make the copy asynchronous.  The next
 * mfx_icgen_run waits for these copies before its first kernel that
 * overwrites the arrays (the assembly, with own_particles; its first kernel
 * otherwise); mfx_icgen_launch waits before the replay.  No counterpart in
 * the reference, whose output plugins read host memory (output_plugin.hh). */
int mfx_icgen_download_async(mfx_icgen *ic, unsigned what, void *const dst[]);
/* blocks until the copies of mfx_icgen_download_async are complete */
int mfx_icgen_download_wait(mfx_icgen *ic);
/* Batched realisations (no counterpart in the reference, which generates
 * one realisation per run): the device sequences of `count` generators, of
 * any mix of grid sizes, captured into ONE HIP graph whose branches run on
 * min(nstreams, count) concurrent streams (generator i on branch i mod
 * nstreams), forked from and joined to the batch's own stream.  Creation
 * runs each generator once.  mfx_batch_launch copies every generator's
 * staged seed (mfx_icgen_set_seed) and replays the graph: one launch for all
 * realisations; each generator's results equal mfx_icgen_run with its seed.
 * mfx_batch_synchronize waits for it.  A generator in a batch cannot be
 * destroyed before the batch. */
typedef struct mfx_batch mfx_batch;
int mfx_batch_create(mfx_icgen *const *gens, int count, int nstreams, mfx_batch **out);
int mfx_batch_launch(mfx_batch *b);
int mfx_batch_synchronize(mfx_batch *b);
int mfx_batch_destroy(mfx_batch *b);
/* pinned host memory (hipHostMalloc) for the downloads */
int mfx_host_alloc(size_t bytes, void **ptr);
int mfx_host_free(void *ptr);
/* grids of the realisation: 0 wnoise, 1 phi, 2 phi2, 3 phi3, 4..6 A3, 7 tmp.
 * After a run the noise grid holds, for NGENIC, the k-space field before its
 * normalisation (ic_generator.cc:341-346 normalises in place; here the
 * normalisation is applied as phi(1) reads it) and, for MUSIC1 and Threefry
 * noise, whose forward transform writes phi(1) from its last pass, an
 * intermediate (scratch); the potentials hold the fields before the
 * growth factors (ic_generator.cc:550-563 scales them in place; here the
 * factors are applied as the assembly reads them) and tmp is scratch. */
int mfx_icgen_grid(mfx_icgen *ic, int which, mfx_grid **g);
/* scalars: Dplus(astart), vfact(astart) */
int mfx_icgen_scalars(mfx_icgen *ic, double *dplus, double *vfact);

#ifdef __cplusplus
}
#endif
#endif
