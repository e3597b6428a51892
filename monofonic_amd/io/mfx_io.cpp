// mfx_io.cpp -- the HDF5 output plugins behind include/mfx_io.h.
//
// Files are laid out as the reference's plugins lay them out (dataset and
// attribute names, shapes and types): the generic field file of
// src/plugins/output_generic.cc and the Gadget-HDF5 particle file of
// src/plugins/output_gadget_hdf5.cc, using the HDF5 C API the way
// include/HDF_IO.hh does (scalar attributes for the generic header, one-
// element and six-element vectors for the Gadget header, fixed-length
// null-padded strings).  The file stays open for the writer's lifetime;
// fields are written one x-plane at a time from the caller's padded rows,
// particles in slabs that are interleaved to N x 3 as they are written.
#include "mfx_io.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct IoError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define IO_CHECK(cond, msg)                 \
  do {                                      \
    if (!(cond)) throw IoError(msg);        \
  } while (0)

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// attribute writers (HDF_IO.hh:1084-1160, 1229-1250)
void attr_scalar(hid_t loc, const char* name, hid_t type, const void* v) {
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
  IO_CHECK(a >= 0, std::string("cannot create attribute ") + name);
  herr_t st = H5Awrite(a, type, v);
  H5Aclose(a);
  H5Sclose(sp);
  IO_CHECK(st >= 0, std::string("cannot write attribute ") + name);
}

void attr_vector(hid_t loc, const char* name, hid_t type, const void* v, size_t n) {
  hsize_t d = n;
  hid_t sp = H5Screate_simple(1, &d, nullptr);
  hid_t a = H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
  IO_CHECK(a >= 0, std::string("cannot create attribute ") + name);
  herr_t st = H5Awrite(a, type, v);
  H5Aclose(a);
  H5Sclose(sp);
  IO_CHECK(st >= 0, std::string("cannot write attribute ") + name);
}

void attr_string(hid_t loc, const char* name, const std::string& s) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, std::max<size_t>(s.size(), 1));
  H5Tset_strpad(t, H5T_STR_NULLPAD);
  attr_scalar(loc, name, t, s.c_str());
  H5Tclose(t);
}

hid_t make_group(hid_t file, const char* name) {
  hid_t g = H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  IO_CHECK(g >= 0, std::string("cannot create group ") + name);
  return g;
}

// Gadget header (output_gadget_hdf5.cc:38-60)
struct GadgetHeader {
  unsigned npart[6] = {};
  unsigned long long npart64[6] = {};
  double mass[6] = {};
  double time = 0, redshift = 0;
  unsigned npartTotal[6] = {};
  unsigned long long npartTotal64[6] = {};
  unsigned npartTotalHighWord[6] = {};
  int num_files = 1;
  double BoxSize = 0, Omega0 = 0, OmegaLambda = 0, HubbleParam = 0;
};

}  // namespace

struct mfx_out {
  hid_t file = -1;
  bool gadget = false;
  bool double_prec = true, long_ids = false, gadget2 = false;
  double boxlength = 0;
  int overload = 1;  // particles per cell of the particle load
  bool masses = false;  // individual particle masses (DoBaryons)
  GadgetHeader hdr;
};

extern "C" {

const char* mfx_io_last_error(void) { return g_err.c_str(); }

int mfx_out_generic_create(const char* fname, double boxsize, double astart, double omega_b, double omega_m,
                           double omega_l, mfx_out** out) {
  return guard([&] {
    IO_CHECK(fname && out, "null argument");
    auto* o = new mfx_out();
    o->file = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (o->file < 0) {
      delete o;
      throw IoError(std::string("cannot create ") + fname);
    }
    hid_t g = make_group(o->file, "Header");
    attr_scalar(g, "Boxsize", H5T_NATIVE_DOUBLE, &boxsize);
    attr_scalar(g, "astart", H5T_NATIVE_DOUBLE, &astart);
    attr_scalar(g, "Omega_b", H5T_NATIVE_DOUBLE, &omega_b);
    attr_scalar(g, "Omega_m", H5T_NATIVE_DOUBLE, &omega_m);
    attr_scalar(g, "Omega_L", H5T_NATIVE_DOUBLE, &omega_l);
    H5Gclose(g);
    *out = o;
  });
}

int mfx_out_write_field(mfx_out* o, const char* name, const void* data, int precision, size_t n0, size_t n1,
                        size_t n2, size_t pitch) {
  return guard([&] {
    IO_CHECK(o && o->file >= 0 && name && data, "null argument");
    IO_CHECK(precision == 4 || precision == 8, "precision must be 4 or 8 bytes");
    IO_CHECK(pitch >= n2 && n0 && n1 && n2, "invalid field extent");
    const hid_t type = precision == 8 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
    hsize_t dims[3] = {n0, n1, n2};
    hid_t fs = H5Screate_simple(3, dims, nullptr);
    hid_t ds = H5Dcreate2(o->file, name, type, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    IO_CHECK(ds >= 0, std::string("cannot create dataset ") + name);
    // one x-plane per write: the memory space strides over the row padding
    hsize_t mdims[3] = {1, n1, pitch};
    hid_t ms = H5Screate_simple(3, mdims, nullptr);
    hsize_t mstart[3] = {0, 0, 0}, count[3] = {1, n1, n2};
    H5Sselect_hyperslab(ms, H5S_SELECT_SET, mstart, nullptr, count, nullptr);
    const char* base = static_cast<const char*>(data);
    herr_t st = 0;
    for (size_t i = 0; i < n0 && st >= 0; ++i) {
      hsize_t start[3] = {i, 0, 0};
      H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
      st = H5Dwrite(ds, type, ms, fs, H5P_DEFAULT, base + i * n1 * pitch * (size_t)precision);
    }
    H5Sclose(ms);
    H5Sclose(fs);
    H5Dclose(ds);
    IO_CHECK(st >= 0, std::string("cannot write dataset ") + name);
  });
}

int mfx_out_gadget_create(const char* fname, const mfx_gadget_params* p, mfx_out** out) {
  return guard([&] {
    IO_CHECK(fname && p && out, "null argument");
    auto* o = new mfx_out();
    o->gadget = true;
    o->double_prec = p->double_prec != 0;
    o->long_ids = p->long_ids != 0;
    o->gadget2 = p->gadget2_compat != 0;
    o->boxlength = p->boxlength;
    // particles per cell of [setup] ParticleLoad (particle_generator.hh:223)
    const std::string load = p->particle_load ? p->particle_load : "sc";
    o->overload = load == "sc" ? 1 : load == "bcc" ? 2 : load == "fcc" ? 4 : load == "rsc" ? 8 : 0;
    if (o->overload == 0) {
      delete o;
      throw IoError("particle load must be sc, bcc, fcc or rsc, not " + load);
    }
    // with baryons every species is written with individual masses
    // (ic_generator.cc:637, particle_generator.hh:380-420, output_gadget_hdf5.cc:269-300)
    o->masses = p->do_baryons != 0;
    const double astart = 1.0 / (1.0 + p->zstart);
    GadgetHeader& h = o->hdr;
    h.time = astart;
    h.redshift = 1.0 / astart - 1.0;
    h.BoxSize = p->boxlength;
    h.Omega0 = p->omega_m;
    h.OmegaLambda = p->omega_de;
    h.HubbleParam = p->h;
    o->file = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (o->file < 0) {
      delete o;
      throw IoError(std::string("cannot create ") + fname);
    }
    // ICs_parameters (output_gadget_hdf5.cc:140-170)
    auto s = [](const char* v, const char* dflt) { return std::string(v ? v : dflt); };
    hid_t g = make_group(o->file, "ICs_parameters");
    attr_string(g, "Code", "MUSIC2 - monofonIC");
    attr_string(g, "Git Revision", s(p->git_rev, "unknown"));
    attr_string(g, "Git Tag", s(p->git_tag, "unknown"));
    attr_string(g, "Git Branch", s(p->git_branch, "unknown"));
    attr_string(g, "Precision", s(p->precision, "DOUBLE"));
    attr_string(g, "Convolutions", s(p->convolutions, "ORSZAG"));
    attr_string(g, "PLT", s(p->plt, "no"));
    attr_scalar(g, "LPT Order", H5T_NATIVE_INT, &p->lpt_order);
    attr_string(g, "Particle Load", s(p->particle_load, "sc"));
    attr_string(g, "Transfer Function", s(p->transfer, "eisenstein"));
    attr_string(g, "Cosmology Parameter Set", s(p->param_set, "none"));
    attr_string(g, "Random Generator", s(p->rng, "MUSIC1"));
    attr_scalar(g, "Mode Fixing", H5T_NATIVE_INT, &p->do_fixing);
    attr_scalar(g, "Mode inversion", H5T_NATIVE_INT, &p->do_inversion);
    attr_scalar(g, "Baryons", H5T_NATIVE_INT, &p->do_baryons);
    attr_scalar(g, "Baryons Relative Velocity", H5T_NATIVE_INT, &p->do_baryons_vrel);
    attr_scalar(g, "Grid Resolution", H5T_NATIVE_INT, &p->grid_res);
    H5Gclose(g);
    *out = o;
  });
}

int mfx_out_gadget_write_particles(mfx_out* o, int species, const double* const pos[3], const double* const vel[3],
                                   size_t n, double omega_species, double lscale, double vscale) {
  return guard([&] {
    IO_CHECK(o && o->gadget && o->file >= 0, "not a Gadget-HDF5 writer");
    IO_CHECK(species >= 0 && species < 6, "species index must be 0..5");
    for (int d = 0; d < 3; ++d) IO_CHECK(pos && vel && pos[d] && vel[d], "null particle array");
    const unsigned long long ncell = (unsigned long long)n * n * n, np = ncell * (unsigned long long)o->overload;
    GadgetHeader& h = o->hdr;
    h.npart[species] = (unsigned)np;
    h.npartTotal[species] = (unsigned)np;
    h.npartTotalHighWord[species] = (unsigned)(np >> 32);
    h.npart64[species] = np;
    h.npartTotal64[species] = np;
    const double rhoc = 27.7519737;  // h^2 1e10 Msun / Mpc^3
    const double munit = rhoc * std::pow(o->boxlength, 3);
    h.mass[species] = o->masses ? 0.0 : omega_species * munit / (double)np;
    // baryons (species 0) number their particles after the dark matter's:
    // IDoffset 1 (ic_generator.cc:629, particle_generator.hh:231)
    const unsigned long long idoff = (o->masses && species == 0) ? np : 0;

    const std::string grp = "PartType" + std::to_string(species);
    hid_t g = make_group(o->file, grp.c_str());
    const hid_t rtype = o->double_prec ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
    const size_t slab = (size_t)1 << 22;  // particles per write
    auto write_vec = [&](const char* name, const double* const* src, double scale) {
      hsize_t dims[2] = {np, 3};
      hid_t fs = H5Screate_simple(2, dims, nullptr);
      hid_t ds = H5Dcreate2(g, name, rtype, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      IO_CHECK(ds >= 0, std::string("cannot create dataset ") + name);
      std::vector<double> b64;
      std::vector<float> b32;
      herr_t st = 0;
      for (unsigned long long p0 = 0; p0 < np && st >= 0; p0 += slab) {
        const size_t m = (size_t)std::min<unsigned long long>(slab, np - p0);
        if (o->double_prec) b64.resize(3 * m); else b32.resize(3 * m);
        for (size_t q = 0; q < m; ++q)
          for (int d = 0; d < 3; ++d) {
            const double v = src[d][p0 + q] * scale;
            if (o->double_prec) b64[3 * q + d] = v; else b32[3 * q + d] = (float)v;
          }
        hsize_t start[2] = {p0, 0}, count[2] = {m, 3};
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
        hid_t ms = H5Screate_simple(2, count, nullptr);
        st = H5Dwrite(ds, rtype, ms, fs, H5P_DEFAULT,
                      o->double_prec ? (const void*)b64.data() : (const void*)b32.data());
        H5Sclose(ms);
      }
      H5Dclose(ds);
      H5Sclose(fs);
      IO_CHECK(st >= 0, std::string("cannot write dataset ") + name);
    };
    write_vec("Coordinates", pos, lscale);
    write_vec("Velocities", vel, vscale);
    // ParticleIDs: particle p = l ncell + cell of sublattice l gets
    // overload * cell + l (particle_generator.hh:229-250, IDoffset 0); the
    // 1D cell index for the sc load
    {
      const hid_t itype = o->long_ids ? H5T_NATIVE_ULLONG : H5T_NATIVE_UINT;
      hsize_t dims[1] = {np};
      hid_t fs = H5Screate_simple(1, dims, nullptr);
      hid_t ds = H5Dcreate2(g, "ParticleIDs", itype, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      IO_CHECK(ds >= 0, "cannot create dataset ParticleIDs");
      std::vector<unsigned long long> i64;
      std::vector<unsigned> i32;
      herr_t st = 0;
      for (unsigned long long p0 = 0; p0 < np && st >= 0; p0 += slab) {
        const size_t m = (size_t)std::min<unsigned long long>(slab, np - p0);
        const unsigned long long ov = (unsigned long long)o->overload;
        auto id = [&](unsigned long long q) { return idoff + ov * (q % ncell) + q / ncell; };
        if (o->long_ids) {
          i64.resize(m);
          for (size_t q = 0; q < m; ++q) i64[q] = id(p0 + q);
        } else {
          i32.resize(m);
          for (size_t q = 0; q < m; ++q) i32[q] = (unsigned)id(p0 + q);
        }
        hsize_t start[1] = {p0}, count[1] = {m};
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
        hid_t ms = H5Screate_simple(1, count, nullptr);
        st = H5Dwrite(ds, itype, ms, fs, H5P_DEFAULT, o->long_ids ? (const void*)i64.data() : (const void*)i32.data());
        H5Sclose(ms);
      }
      H5Dclose(ds);
      H5Sclose(fs);
      IO_CHECK(st >= 0, "cannot write dataset ParticleIDs");
    }
    // Masses: pmeanmass * rho with pmeanmass = 1/N and rho = (1 + C delta_bc)
    // Omega munit, delta_bc = 0 for the Eisenstein-Hu transfer function
    // (ic_generator.cc:650-667, particle_generator.hh:383-411)
    if (o->masses) {
      const double pm = (1.0 / (double)np) * (omega_species * munit);
      hsize_t dims[1] = {np};
      hid_t fs = H5Screate_simple(1, dims, nullptr);
      hid_t ds = H5Dcreate2(g, "Masses", rtype, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      IO_CHECK(ds >= 0, "cannot create dataset Masses");
      herr_t st = 0;
      for (unsigned long long p0 = 0; p0 < np && st >= 0; p0 += slab) {
        const size_t m = (size_t)std::min<unsigned long long>(slab, np - p0);
        std::vector<double> b64;
        std::vector<float> b32;
        if (o->double_prec) b64.assign(m, pm); else b32.assign(m, (float)pm);
        hsize_t start[1] = {p0}, count[1] = {m};
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
        hid_t ms = H5Screate_simple(1, count, nullptr);
        st = H5Dwrite(ds, rtype, ms, fs, H5P_DEFAULT, o->double_prec ? (const void*)b64.data() : (const void*)b32.data());
        H5Sclose(ms);
      }
      H5Dclose(ds);
      H5Sclose(fs);
      IO_CHECK(st >= 0, "cannot write dataset Masses");
    }
    H5Gclose(g);
  });
}

int mfx_out_close(mfx_out* o) {
  return guard([&] {
    if (!o) return;
    struct Del {
      mfx_out* o;
      ~Del() {
        if (o->file >= 0) H5Fclose(o->file);
        delete o;
      }
    } del{o};
    if (!o->gadget) return;
    // Header (output_gadget_hdf5.cc:180-204)
    const GadgetHeader& h = o->hdr;
    hid_t g = make_group(o->file, "Header");
    if (o->gadget2) {
      attr_vector(g, "NumPart_ThisFile", H5T_NATIVE_UINT, h.npart, 6);
      attr_vector(g, "NumPart_Total", H5T_NATIVE_UINT, h.npartTotal, 6);
      attr_vector(g, "NumPart_Total_HighWord", H5T_NATIVE_UINT, h.npartTotalHighWord, 6);
    } else {
      attr_vector(g, "NumPart_ThisFile", H5T_NATIVE_ULLONG, h.npart64, 6);
      attr_vector(g, "NumPart_Total", H5T_NATIVE_ULLONG, h.npartTotal64, 6);
    }
    attr_vector(g, "MassTable", H5T_NATIVE_DOUBLE, h.mass, 6);
    const int zero = 0;
    attr_vector(g, "Time", H5T_NATIVE_DOUBLE, &h.time, 1);
    attr_vector(g, "Redshift", H5T_NATIVE_DOUBLE, &h.redshift, 1);
    attr_vector(g, "Flag_Sfr", H5T_NATIVE_INT, &zero, 1);
    attr_vector(g, "Flag_Feedback", H5T_NATIVE_INT, &zero, 1);
    attr_vector(g, "Flag_Cooling", H5T_NATIVE_INT, &zero, 1);
    attr_vector(g, "NumFilesPerSnapshot", H5T_NATIVE_INT, &h.num_files, 1);
    attr_vector(g, "BoxSize", H5T_NATIVE_DOUBLE, &h.BoxSize, 1);
    attr_vector(g, "Omega0", H5T_NATIVE_DOUBLE, &h.Omega0, 1);
    attr_vector(g, "OmegaLambda", H5T_NATIVE_DOUBLE, &h.OmegaLambda, 1);
    attr_vector(g, "HubbleParam", H5T_NATIVE_DOUBLE, &h.HubbleParam, 1);
    attr_vector(g, "Flag_StellarAge", H5T_NATIVE_INT, &zero, 1);
    attr_vector(g, "Flag_Metals", H5T_NATIVE_INT, &zero, 1);
    attr_vector(g, "Flag_Entropy_ICs", H5T_NATIVE_INT, &zero, 1);
    H5Gclose(g);
  });
}

}  // extern "C"
