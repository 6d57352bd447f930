/*
 * adga_h5.h - C ABI of the HDF5 readers and writers of the reference's file formats (host code,
 * libadga_h5.so, linked to the HDF5 C library): the 1P file (DMFT self-energy, Green's function, double
 * counting, mu, beta, H(k)), the symmetrised 2P file the per-omega vertex is read from, the V(q) file and
 * the output file (susceptibilities, self-energies, occupations, the run's inputs).
 *
 * Conventions: complex arrays are interleaved (re, im) doubles in the reference's Fortran shapes (first
 * index fastest), exactly the arrays the reference's routines fill or take; every file dataset has the
 * reference's name, type (a compound of doubles "r" and "i" for complex data), shape and index order;
 * every entry point returns ADGA_H5_OK or an error code, adga_h5_last_error gives the message.  A file
 * stays open between calls (the reference opens and closes the 2P file for every omega).
 */
#ifndef ADGA_H5_H
#define ADGA_H5_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ADGA_H5_OK = 0, ADGA_H5_E_ARG = 1, ADGA_H5_E_IO = 2, ADGA_H5_E_FORMAT = 3 };

typedef struct adga_h5 adga_h5;

/* Opens an existing file (writable != 0: read-write) or creates one (adga_h5_create, truncating). */
int adga_h5_open(const char *path, int writable, adga_h5 **f);
int adga_h5_create(const char *path, adga_h5 **f);
int adga_h5_close(adga_h5 *f);
int adga_h5_last_error(adga_h5 *f, char *buf, int len);

/* Correlated-subspace layout of the orbitals (the reference's ndims(nineq, 2), Fortran order: ndims[ineq]
   = correlated orbitals of atom ineq, ndims[nineq + ineq] = its uncorrelated ones). */
typedef struct adga_h5_atoms {
  int ndim;          /* orbitals in total */
  int nineq;         /* inequivalent atoms */
  const int *ndims;  /* (nineq, 2) */
} adga_h5_atoms;

/*
 * 1P file.  siw/giw(2*iwmax, ndim): the spin average of <dmft_iter>/ineq-XXX/siw|giw/value (bands x spins x
 * 2*iwmax in the file) on each atom's correlated orbitals, zero elsewhere, orbital-averaged per atom when
 * orb_sym != 0.  dc(2, ndim) (real values): <dmft_iter>/ineq-XXX/dc/value.  mu: <dmft_iter>/mu/value.
 * beta: attribute general.beta of /.config.  hk(ndim, ndim, nkp) from start/hk/value (nkp and the
 * band count from the file; hk may be NULL to query them).
 * Replaces: read_siw, read_giw, read_dc, read_mu, read_beta, read_hk_w2dyn (src/hdf5_module.f90:380-627).
 */
int adga_h5_read_siw(adga_h5 *f, const char *dmft_iter, const adga_h5_atoms *at, int orb_sym, int iwmax,
                     double *siw);
int adga_h5_read_giw(adga_h5 *f, const char *dmft_iter, const adga_h5_atoms *at, int orb_sym, int iwmax,
                     double *giw);
int adga_h5_read_dc(adga_h5 *f, const char *dmft_iter, const adga_h5_atoms *at, double *dc);
int adga_h5_read_mu(adga_h5 *f, const char *dmft_iter, double *mu);
int adga_h5_read_beta(adga_h5 *f, double *beta);
int adga_h5_read_hk(adga_h5 *f, int *ndim, int *nkp, double *hk);

/*
 * Symmetrised 2P file (the vertex_chann_sym output): chi_loc_dens/magn(maxdim, maxdim), maxdim =
 * ndim^2 * 2 * iwfmax_small, for bosonic frequency iwb, from ineq-XXX/dens|magn/<iwb+iwbmax>/<band
 * compound index>/value (2*iwfmax x 2*iwfmax each, times beta), in the compound index
 * i1 = {nu1, i, j}, i2 = {nu2, l, k} of the box |nu| < iwfmax_small, with the disconnected terms of
 * vertex_type (0 full G4, 1 connected G4, 2 chi) formed from giw(2*iwmax, ndim).
 * Replaces: read_vertex (src/hdf5_module.f90:629-773).
 */
typedef struct adga_h5_vertex {
  adga_h5_atoms atoms;
  int iwfmax;        /* fermionic box of the file */
  int iwbmax;        /* bosonic box of the file */
  int iwfmax_small;  /* box of the BSE */
  int iwmax;         /* frequencies of giw: giw holds 2*iwmax rows */
  int vertex_type;
  double beta;
  const double *giw; /* (2*iwmax, ndim) */
} adga_h5_vertex;
int adga_h5_read_vertex(adga_h5 *f, int iwb, const adga_h5_vertex *v, double *chi_loc_dens, double *chi_loc_magn);

/*
 * Writes one bosonic frequency of atom ineq into a symmetrised 2P file in the layout read_vertex reads:
 * ineq-XXX/dens|magn/<iwb+iwbmax>/<band compound index>/value for every band pattern of the atom's nbands
 * correlated orbitals, from g4_dens/g4_magn(nbands, nbands, 2*iwfmax, nbands, nbands, 2*iwfmax) (the
 * values as stored, i.e. before read_vertex's factor beta).  The format of the vertex_chann_sym output
 * (src/vertex_chann_sym.f90:203-231); a frequency already written is an error.
 */
int adga_h5_write_vertex(adga_h5 *f, int ineq, int iwb, int iwbmax, int nbands, int iwfmax, const double *g4_dens,
                         const double *g4_magn);

/*
 * V(q) file: v_q(ndim^2, ndim^2, nqp) in read_vq's compound index (v(i1 = {k i}, i2 = {j l}) =
 * V_ijkl) for every q of the list at once: position iq of dataset <band compound index> when the q list is
 * the file's q volume (q_vol != 0, the file must hold nqp points), else position q_data(iq) of the whole k
 * grid.  Replaces: read_vq for every iq (src/interaction_module.f90:39-123).
 */
int adga_h5_read_vq(adga_h5 *f, int ndim, int nqp, int q_vol, const int *q_data, double *v_q);

/*
 * Output file.  adga_h5_init_output creates the groups and writes the run's inputs into /input
 * (mu, beta, dc, giw, siw, hk, n_dmft, n_dmft_k, the frequency boxes, nkp, nqp, nkpxyz, nqpxyz and, for
 * a q path or an EOM k path, the q / k coordinates).
 * Replaces: init_h5_output (src/hdf5_module.f90:997-1225).
 */
typedef struct adga_h5_run {
  int ndim, iwmax, iwfmax, iwbmax, iwfmax_small, iwbmax_small;
  int nkp, nkpx, nkpy, nkpz, nqp, nqpx, nqpy, nqpz, nkp_eom;
  int do_chi, do_eom, nonlocal, gamma_group, q_path_susc, k_path_eom;
  double mu, beta;
  const double *dc;         /* (2, ndim) complex */
  const double *giw, *siw;  /* (2*iwmax, ndim) complex */
  const double *hk;         /* (ndim, ndim, nkp) complex */
  const double *n_dmft;     /* (ndim) complex */
  const double *n_fock;     /* (nkp, ndim, ndim) complex, may be NULL (then not written) */
  const double *k_data;     /* (3, nkp) k coordinates, for the q / k paths (may be NULL otherwise) */
  const int *q_data;        /* (nqp) 1-based k indices */
  const int *k_data_eom;    /* (nkp_eom) 1-based k indices */
} adga_h5_run;
int adga_h5_init_output(adga_h5 *f, const adga_h5_run *r);

/*
 * susceptibility/nonloc/<channel>: chi_qw(ndim^2, ndim^2, nqp*(2*iwbmax_small+1)) (points omega-major,
 * q on the nqpx x nqpy x nqpz volume), as the full 8-index array or its reduced (i i j j) part, gzip level
 * `gzip` with shuffle and Fletcher32 in chunks of one omega; on a q path the full or reduced array
 * without filters.  susceptibility/loc/<channel>: chi_loc(ndim^2, ndim^2, 2*iwbmax_small+1), full or
 * reduced, filtered in chunks of one omega.
 * Replaces: output_chi_qw_full_h5, output_chi_qw_reduced_h5, output_chi_qpath_full_h5,
 * output_chi_qpath_reduced_h5, output_chi_loc_full_h5, output_chi_loc_reduced_h5
 * (src/hdf5_module.f90:1229-1644).
 */
int adga_h5_write_chi_qw(adga_h5 *f, const char *channel, int full, int ndim, int iwbmax_small, const int nq[3],
                         const double *chi_qw, int gzip);
int adga_h5_write_chi_qpath(adga_h5 *f, const char *channel, int full, int ndim, int iwbmax_small, int nqp,
                            const double *chi);
int adga_h5_write_chi_loc(adga_h5 *f, const char *channel, int full, int ndim, int iwbmax_small,
                          const double *chi_loc, int gzip);

/*
 * Self-energies: with nonlocal, selfenergy/nonloc/dga from sigma_sum(ndim, ndim, 2*iwfmax_small, nkp) and
 * selfenergy/nonloc/hartree_fock from sigma_sum_hf(ndim, ndim, nkp) (repeated over nu), both on the
 * nkpx x nkpy x nkpz grid, filtered, and selfenergy/loc/dga_ksum from sigma_loc(ndim, ndim, 2*iwfmax_small);
 * always selfenergy/loc/dmft from sigma_sum_dmft.  occupation/n_dga(ndim) and n_dga_k from
 * n_dga_k(nkp, ndim, ndim) (real parts).
 * Replaces: output_eom_h5, output_occ_h5 (src/hdf5_module.f90:1648-1797, 1912-1966).
 */
int adga_h5_write_eom(adga_h5 *f, int nonlocal, int ndim, int iwfmax_small, const int nk[3], const double *sigma_sum,
                      const double *sigma_sum_hf, const double *sigma_loc, const double *sigma_sum_dmft, int gzip);
int adga_h5_write_occ(adga_h5 *f, int ndim, const int nk[3], const double *n_dga, const double *n_dga_k);

/*
 * Generic access for checks and tools: the dims (C order, at most 8) of a dataset, a compound (r, i) or
 * double dataset read whole into interleaved doubles in file (C) order.
 */
int adga_h5_dims(adga_h5 *f, const char *path, int *rank, long long *dims);
int adga_h5_read_complex(adga_h5 *f, const char *path, double *out);
int adga_h5_read_real(adga_h5 *f, const char *path, double *out);
/* ... and written: a compound (r, i) or double dataset of the given C-order dims (rank 0: a scalar
   double), creating the groups of the path as needed; a double attribute on a group. */
int adga_h5_write_complex(adga_h5 *f, const char *path, int rank, const long long *dims, const double *data);
int adga_h5_write_real(adga_h5 *f, const char *path, int rank, const long long *dims, const double *data);
int adga_h5_write_attr_real(adga_h5 *f, const char *group, const char *name, double value);

#ifdef __cplusplus
}
#endif

#endif /* ADGA_H5_H */
