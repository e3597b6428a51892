/* mfx_io.h -- host output boundary: the HDF5 output plugins of monofonIC for
 * the fields and particles the device pipeline produces.
 *
 * libmfx_io.so is a host-only library (it links the HDF5 C library and no
 * HIP code).  Every call returns 0 on success or a nonzero status, with the
 * message in mfx_io_last_error().  Arrays are host memory; the IC driver's
 * mfx_icgen_download_async stages device results into pinned host buffers
 * that can be passed here directly.
 */
#ifndef MFX_IO_H
#define MFX_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfx_out mfx_out;

const char *mfx_io_last_error(void);

/* generic_output_plugin ctor (src/plugins/output_generic.cc:33-56): creates
 * (truncates) fname and writes the Header group with the double attributes
 * Boxsize, astart, Omega_b, Omega_m, Omega_L */
int mfx_out_generic_create(const char *fname, double boxsize, double astart, double omega_b, double omega_m,
                           double omega_l, mfx_out **out);

/* Grid_FFT::Write_to_HDF5 for a real-space grid (src/grid_fft.cc:669-760),
 * as generic_output_plugin::write_grid_data calls it (output_generic.cc
 * :119-124): dataset `name` (DM_dx ... DM_vz, DM_delta) of n0 x n1 x n2
 * values of `precision` bytes (4 float, 8 double) taken from rows of
 * `pitch` >= n2 elements (pitch n2 + 2: the in-place r2c layout) */
int mfx_out_write_field(mfx_out *o, const char *name, const void *data, int precision, size_t n0, size_t n1,
                        size_t n2, size_t pitch);

/* gadget_hdf5_output_plugin (src/plugins/output_gadget_hdf5.cc:73-177):
 * double_prec selects gadget_hdf5_double (double Coordinates/Velocities)
 * over gadget_hdf5 (float); long_ids = [output] UseLongids;
 * gadget2_compat = [output] Gadget2Compatibility (32-bit particle counts) */
typedef struct {
  double zstart, boxlength, omega_m, omega_de, omega_b, h;
  int lpt_order, grid_res;
  int double_prec, long_ids, gadget2_compat;
  int do_fixing, do_inversion, do_baryons, do_baryons_vrel;
  const char *particle_load; /* "sc", "bcc", "fcc" or "rsc" */
  const char *transfer;      /* "eisenstein" */
  const char *param_set;     /* "none" */
  const char *rng;           /* "MUSIC1", "NGENIC", ... */
  const char *precision;     /* code precision string */
  const char *convolutions;  /* convolver string */
  const char *plt;           /* "yes" / "no" */
  const char *git_rev, *git_tag, *git_branch;
} mfx_gadget_params;

int mfx_out_gadget_create(const char *fname, const mfx_gadget_params *p, mfx_out **out);

/* write_particle_data (output_gadget_hdf5.cc:268-308) for the particles of
 * an n^3 grid and the particle load of mfx_out_gadget_create (particle_load
 * sc, bcc, fcc or rsc: N = 1, 2, 4 or 8 n^3 particles, sublattice-major, each
 * sublattice in (i,j,k) order): species 1 (dm) / 0 (baryon) / 3 (neutrino);
 * pos[d] and vel[d] are the three FP64 component arrays of N values;
 * Coordinates = pos * lscale and Velocities = vel * vscale (interleaved to
 * N x 3); ParticleIDs = overload * ((i n + j) n + k) + sublattice
 * (particle_generator.hh:229-250); the mass table entry
 * omega_species * rhoc L^3 / N (rhoc = 27.7519737 h^2 1e10 Msun/Mpc^3).
 * With do_baryons (DoBaryons) the masses go to a Masses dataset instead
 * (the mass table entry is 0), every particle (1/N) (omega_species rhoc L^3)
 * (the Eisenstein-Hu transfer function has no baryon-CDM difference mode),
 * and species 0 (baryons) numbers its particles from N on (IDoffset 1). */
int mfx_out_gadget_write_particles(mfx_out *o, int species, const double *const pos[3], const double *const vel[3],
                                   size_t n, double omega_species, double lscale, double vscale);

/* closes the file; a gadget file gets its Header group here, as the
 * plugin's destructor writes it (output_gadget_hdf5.cc:180-214) */
int mfx_out_close(mfx_out *o);

#ifdef __cplusplus
}
#endif

#endif
