// HDF5 readers and writers of the reference's file formats (include/adga_h5.h), on the HDF5 C library.
// Host code only: the data go to the GPU through the hot path's C ABI (include/adga_amd.h).
//
// A Fortran array a(n1, ..., nr) written or read through HDF5's Fortran interface is, in the file, the
// C-order dataset of dims (nr, ..., n1) whose memory image is the Fortran array's: so the reference's
// Fortran dims are reversed here, and a dataset read whole is the Fortran array as the reference had it.
#include <hdf5.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "adga_h5.h"

struct adga_h5 {
  hid_t file = -1;
  std::string err;
};

namespace {

using cd = std::complex<double>;
using dims_t = std::vector<hsize_t>;

struct Fail {
  int code;
  std::string msg;
};

// owned HDF5 identifier, closed with the matching H5*close
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  operator hid_t() const { return id; }
};

hid_t ck(hid_t id, const std::string& what) {
  if (id < 0) throw Fail{ADGA_H5_E_IO, "HDF5 failure: " + what};
  return id;
}
void ckh(herr_t e, const std::string& what) {
  if (e < 0) throw Fail{ADGA_H5_E_IO, "HDF5 failure: " + what};
}

template <typename F>
int guarded(adga_h5* f, F&& fn) {
  try {
    fn();
    return ADGA_H5_OK;
  } catch (const Fail& e) {
    if (f) f->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (f) f->err = std::string("host exception: ") + e.what();
    return ADGA_H5_E_IO;
  }
}

// compound {"r", "i"} of doubles: the reference's compound_id
hid_t complex_type() {
  hid_t t = ck(H5Tcreate(H5T_COMPOUND, 2 * sizeof(double)), "H5Tcreate");
  H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "i", sizeof(double), H5T_NATIVE_DOUBLE);
  return t;
}

dims_t dataset_dims(hid_t d, const std::string& path) {
  Id sp(ck(H5Dget_space(d), "H5Dget_space " + path), H5Sclose);
  const int r = H5Sget_simple_extent_ndims(sp);
  if (r < 0) throw Fail{ADGA_H5_E_IO, "rank of " + path};
  dims_t dims(r);
  H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
  return dims;
}

size_t count_of(const dims_t& d) {
  size_t n = 1;
  for (hsize_t x : d) n *= (size_t)x;
  return n;
}

// whole compound (r, i) dataset into interleaved doubles (file order)
std::vector<double> read_cplx(hid_t loc, const std::string& path, dims_t* dims_out = nullptr) {
  Id d(ck(H5Dopen2(loc, path.c_str(), H5P_DEFAULT), "open " + path), H5Dclose);
  const dims_t dims = dataset_dims(d, path);
  std::vector<double> v(2 * count_of(dims));
  Id t(complex_type(), H5Tclose);
  ckh(H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read " + path);
  if (dims_out) *dims_out = dims;
  return v;
}

std::vector<double> read_real(hid_t loc, const std::string& path, dims_t* dims_out = nullptr) {
  Id d(ck(H5Dopen2(loc, path.c_str(), H5P_DEFAULT), "open " + path), H5Dclose);
  const dims_t dims = dataset_dims(d, path);
  std::vector<double> v(count_of(dims));
  ckh(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read " + path);
  if (dims_out) *dims_out = dims;
  return v;
}

// dataset of complex values (interleaved doubles, C order of dims); chunk non-empty: chunked with the
// reference's filters in its order (deflate level gzip, shuffle, Fletcher32)
void write_cplx(hid_t loc, const std::string& name, const dims_t& dims, const double* data,
                const dims_t& chunk = {}, int gzip = 0) {
  Id sp(ck(H5Screate_simple((int)dims.size(), dims.data(), nullptr), "dataspace " + name), H5Sclose);
  Id pl(ck(H5Pcreate(H5P_DATASET_CREATE), "plist"), H5Pclose);
  if (!chunk.empty()) {
    ckh(H5Pset_chunk(pl, (int)chunk.size(), chunk.data()), "chunk " + name);
    ckh(H5Pset_deflate(pl, (unsigned)gzip), "deflate " + name);
    ckh(H5Pset_shuffle(pl), "shuffle " + name);
    ckh(H5Pset_fletcher32(pl), "fletcher32 " + name);
  }
  Id t(complex_type(), H5Tclose);
  Id d(ck(H5Dcreate2(loc, name.c_str(), t, sp, H5P_DEFAULT, pl, H5P_DEFAULT), "create " + name), H5Dclose);
  ckh(H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write " + name);
}

void write_num(hid_t loc, const std::string& name, hid_t type, const dims_t& dims, const void* data) {
  Id sp(ck(dims.empty() ? H5Screate(H5S_SCALAR) : H5Screate_simple((int)dims.size(), dims.data(), nullptr),
           "dataspace " + name),
        H5Sclose);
  Id d(ck(H5Dcreate2(loc, name.c_str(), type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "create " + name),
       H5Dclose);
  ckh(H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write " + name);
}

std::string ineq_name(int ineq) {  // "ineq-001" for ineq = 1
  char b[32];
  snprintf(b, sizeof b, "ineq-%03d", ineq);
  return b;
}
std::string five(int x) {  // the reference's (I5.5) group names
  char b[16];
  snprintf(b, sizeof b, "%05d", x);
  return b;
}

// first correlated orbital (1-based) of atom ineq (1-based) and their number
void atom_range(const adga_h5_atoms* at, int ineq, int* start, int* count) {
  int s = 1;
  for (int i = 2; i <= ineq; ++i) s += at->ndims[i - 2] + at->ndims[at->nineq + i - 2];
  *start = s;
  *count = at->ndims[ineq - 1];
}

void check_atoms(const adga_h5_atoms* at) {
  if (!at || at->ndim < 1 || at->nineq < 1 || !at->ndims) throw Fail{ADGA_H5_E_ARG, "bad atom layout"};
  int s, c;
  atom_range(at, at->nineq, &s, &c);
  if (s + c - 1 > at->ndim) throw Fail{ADGA_H5_E_ARG, "correlated orbitals beyond ndim"};
}

// all four legs on the same atom's correlated orbitals (aux.f90:189-221)
bool same_correlated(const adga_h5_atoms* at, int m, int n, int o, int p) {
  int a = 0, b = 0, c = 0, d = 0;
  for (int ineq = 1; ineq <= at->nineq; ++ineq) {
    int s, cnt;
    atom_range(at, ineq, &s, &cnt);
    const int e = s + cnt - 1;
    if (m >= s && m <= e) a = ineq;
    if (n >= s && n <= e) b = ineq;
    if (o >= s && o <= e) c = ineq;
    if (p >= s && p <= e) d = ineq;
  }
  return a == b && c == d && a == d && a != 0;
}

// spin-averaged siw / giw (read_siw, read_giw: hdf5_module.f90:380-500)
void read_1p_freq(adga_h5* f, const char* iter, const adga_h5_atoms* at, int orb_sym, int iwmax, const char* what,
                  double* out) {
  check_atoms(at);
  if (!iter || iwmax < 1 || !out) throw Fail{ADGA_H5_E_ARG, "bad argument"};
  const int n = at->ndim, nw = 2 * iwmax;
  cd* o = reinterpret_cast<cd*>(out);
  for (int i = 0; i < nw * n; ++i) o[i] = 0.0;
  for (int ineq = 1; ineq <= at->nineq; ++ineq) {
    int s, cnt;
    atom_range(at, ineq, &s, &cnt);
    const std::string path = std::string(iter) + "/" + ineq_name(ineq) + "/" + what + "/value";
    dims_t dims;
    const std::vector<double> v = read_cplx(f->file, path, &dims);
    // file dims (bands, spins, 2*iwmax): the reference's buffer (2, -iwmax:iwmax-1, spins, bands)
    if (dims.size() != 3 || (int)dims[2] != nw || dims[1] < 2 || (int)dims[0] < cnt)
      throw Fail{ADGA_H5_E_FORMAT, path + ": expected (bands, 2 spins, 2*iwmax) values"};
    const size_t ns = dims[1];
    for (int b = 0; b < cnt; ++b)
      for (int w = 0; w < nw; ++w) {
        const double* u = &v[2 * (((size_t)b * ns + 0) * nw + w)];
        const double* d = &v[2 * (((size_t)b * ns + 1) * nw + w)];
        o[w + (size_t)(s - 1 + b) * nw] = cd(u[0] + d[0], u[1] + d[1]) / 2.0;
      }
    if (orb_sym && cnt > 1) {
      for (int w = 0; w < nw; ++w) {
        cd acc = o[w + (size_t)(s - 1) * nw];
        for (int b = 1; b < cnt; ++b) acc += o[w + (size_t)(s - 1 + b) * nw];
        acc /= (double)cnt;
        for (int b = 0; b < cnt; ++b) o[w + (size_t)(s - 1 + b) * nw] = acc;
      }
    }
  }
}

// intermediate groups of a path "a/b/c" (a and a/b), created when missing
void make_parents(hid_t file, const std::string& path) {
  for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1)) {
    const std::string g = path.substr(0, p);
    if (H5Lexists(file, g.c_str(), H5P_DEFAULT) <= 0)
      H5Gclose(ck(H5Gcreate2(file, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group " + g));
  }
}

hid_t open_group(hid_t file, const char* name) { return ck(H5Gopen2(file, name, H5P_DEFAULT), std::string("group ") + name); }

}  // namespace

extern "C" {

int adga_h5_open(const char* path, int writable, adga_h5** f) {
  if (!path || !f) return ADGA_H5_E_ARG;
  *f = nullptr;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as codes and messages
  const hid_t id = H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
  if (id < 0) return ADGA_H5_E_IO;
  *f = new adga_h5();
  (*f)->file = id;
  return ADGA_H5_OK;
}

int adga_h5_create(const char* path, adga_h5** f) {
  if (!path || !f) return ADGA_H5_E_ARG;
  *f = nullptr;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const hid_t id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (id < 0) return ADGA_H5_E_IO;
  *f = new adga_h5();
  (*f)->file = id;
  return ADGA_H5_OK;
}

int adga_h5_close(adga_h5* f) {
  if (!f) return ADGA_H5_OK;
  const herr_t e = f->file >= 0 ? H5Fclose(f->file) : 0;
  delete f;
  return e < 0 ? ADGA_H5_E_IO : ADGA_H5_OK;
}

int adga_h5_last_error(adga_h5* f, char* buf, int len) {
  if (!f || !buf || len <= 0) return ADGA_H5_E_ARG;
  snprintf(buf, (size_t)len, "%s", f->err.c_str());
  return ADGA_H5_OK;
}

int adga_h5_read_siw(adga_h5* f, const char* iter, const adga_h5_atoms* at, int orb_sym, int iwmax, double* siw) {
  if (!f) return ADGA_H5_E_ARG;
  return guarded(f, [&] { read_1p_freq(f, iter, at, orb_sym, iwmax, "siw", siw); });
}

int adga_h5_read_giw(adga_h5* f, const char* iter, const adga_h5_atoms* at, int orb_sym, int iwmax, double* giw) {
  if (!f) return ADGA_H5_E_ARG;
  return guarded(f, [&] { read_1p_freq(f, iter, at, orb_sym, iwmax, "giw", giw); });
}

int adga_h5_read_dc(adga_h5* f, const char* iter, const adga_h5_atoms* at, double* dc) {
  if (!f || !iter || !dc) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    check_atoms(at);
    cd* o = reinterpret_cast<cd*>(dc);
    for (int i = 0; i < 2 * at->ndim; ++i) o[i] = 0.0;
    for (int ineq = 1; ineq <= at->nineq; ++ineq) {
      int s, cnt;
      atom_range(at, ineq, &s, &cnt);
      const std::string path = std::string(iter) + "/" + ineq_name(ineq) + "/dc/value";
      dims_t dims;
      const std::vector<double> v = read_real(f->file, path, &dims);  // (bands, spins): Fortran dc_data(spin, band)
      if (dims.size() != 2 || dims[1] != 2 || (int)dims[0] < cnt)
        throw Fail{ADGA_H5_E_FORMAT, path + ": expected (bands, 2 spins) values"};
      for (int b = 0; b < cnt; ++b)
        for (int sp = 0; sp < 2; ++sp) o[sp + 2 * (size_t)(s - 1 + b)] = v[(size_t)b * 2 + sp];
    }
  });
}

int adga_h5_read_mu(adga_h5* f, const char* iter, double* mu) {
  if (!f || !iter || !mu) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    dims_t dims;
    const std::vector<double> v = read_real(f->file, std::string(iter) + "/mu/value", &dims);
    if (v.size() != 1) throw Fail{ADGA_H5_E_FORMAT, "mu/value is not a scalar"};
    *mu = v[0];
  });
}

int adga_h5_read_beta(adga_h5* f, double* beta) {
  if (!f || !beta) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    Id g(open_group(f->file, ".config"), H5Gclose);
    Id a(ck(H5Aopen(g, "general.beta", H5P_DEFAULT), "attribute general.beta"), H5Aclose);
    ckh(H5Aread(a, H5T_NATIVE_DOUBLE, beta), "read general.beta");
  });
}

int adga_h5_read_hk(adga_h5* f, int* ndim, int* nkp, double* hk) {
  if (!f || !ndim || !nkp) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    Id d(ck(H5Dopen2(f->file, "start/hk/value", H5P_DEFAULT), "open start/hk/value"), H5Dclose);
    const dims_t dims = dataset_dims(d, "start/hk/value");  // (nkp, bands, bands)
    if (dims.size() != 3 || dims[1] != dims[2]) throw Fail{ADGA_H5_E_FORMAT, "start/hk/value: expected (nkp, n, n)"};
    *nkp = (int)dims[0];
    *ndim = (int)dims[1];
    if (hk) {
      Id t(complex_type(), H5Tclose);
      ckh(H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, hk), "read start/hk/value");
    }
  });
}

int adga_h5_read_vertex(adga_h5* f, int iwb, const adga_h5_vertex* vx, double* chi_d, double* chi_m) {
  if (!f || !vx || !chi_d || !chi_m) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const adga_h5_atoms* at = &vx->atoms;
    check_atoms(at);
    const int n = at->ndim, n2 = n * n, iwf = vx->iwfmax_small, nfb = 2 * vx->iwfmax;
    if (iwf < 1 || iwf > vx->iwfmax || iwb < -vx->iwbmax || iwb > vx->iwbmax || !vx->giw ||
        vx->iwmax < iwf + std::abs(iwb) || vx->vertex_type < 0 || vx->vertex_type > 2)
      throw Fail{ADGA_H5_E_ARG, "bad vertex box, frequency or vertex_type"};
    // g4(b1, b2, nu1, b3, b4, nu2) of both channels over the file's box (hdf5_module.f90:645-705)
    const size_t nb = (size_t)nfb * nfb;
    std::vector<cd> g4m((size_t)n2 * n2 * nb, 0.0), g4d((size_t)n2 * n2 * nb, 0.0);
    auto g4i = [=](int b1, int b2, int u1, int b3, int b4, int u2) {  // 0-based bands, 0-based frequency rows
      return (((((size_t)u2 * n + b4) * n + b3) * nfb + u1) * n + b2) * n + b1;
    };
    Id t(complex_type(), H5Tclose);
    std::vector<double> buf(2 * nb);
    for (int ineq = 1; ineq <= at->nineq; ++ineq) {
      int s, cnt;
      atom_range(at, ineq, &s, &cnt);
      const std::string gm = ineq_name(ineq) + "/magn/" + five(iwb + vx->iwbmax);
      const std::string gd = ineq_name(ineq) + "/dens/" + five(iwb + vx->iwbmax);
      Id grp(open_group(f->file, gm.c_str()), H5Gclose);
      H5G_info_t info;
      ckh(H5Gget_info(grp, &info), "members of " + gm);
      for (hsize_t m = 0; m < info.nlinks; ++m) {
        char name[64];
        if (H5Lget_name_by_idx(grp, ".", H5_INDEX_NAME, H5_ITER_INC, m, name, sizeof name, H5P_DEFAULT) < 0)
          throw Fail{ADGA_H5_E_IO, "member name in " + gm};
        const int ind = atoi(name);
        // band pattern of the compound index (aux.f90:169-186), within the atom's correlated orbitals
        const int c3 = cnt * cnt * cnt, c2 = cnt * cnt, r = ind - 1;
        const int b1 = r / c3, b2 = (r - c3 * b1) / c2, b3 = (r - c3 * b1 - c2 * b2) / cnt,
                  b4 = r - c3 * b1 - c2 * b2 - cnt * b3;
        if (ind < 1 || b1 >= cnt) throw Fail{ADGA_H5_E_FORMAT, "band index " + std::string(name) + " in " + gm};
        for (int ch = 0; ch < 2; ++ch) {
          const std::string path = (ch ? gd : gm) + "/" + name + "/value";
          Id d(ck(H5Dopen2(f->file, path.c_str(), H5P_DEFAULT), "open " + path), H5Dclose);
          const dims_t dims = dataset_dims(d, path);
          if (dims.size() != 2 || (int)dims[0] != nfb || (int)dims[1] != nfb)
            throw Fail{ADGA_H5_E_FORMAT, path + ": expected (2*iwfmax, 2*iwfmax)"};
          ckh(H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "read " + path);
          std::vector<cd>& g4 = ch ? g4d : g4m;
          // Fortran tmp(nu1, nu2): nu1 fastest
          for (int u2 = 0; u2 < nfb; ++u2)
            for (int u1 = 0; u1 < nfb; ++u1) {
              const double* p = &buf[2 * ((size_t)u2 * nfb + u1)];
              g4[g4i(s - 1 + b1, s - 1 + b2, u1, s - 1 + b3, s - 1 + b4, u2)] = cd(p[0], p[1]) * vx->beta;
            }
        }
      }
    }
    // compound index of the BSE box and the disconnected terms (hdf5_module.f90:707-770)
    const int N = n2 * 2 * iwf;
    cd* xm = reinterpret_cast<cd*>(chi_m);
    cd* xd = reinterpret_cast<cd*>(chi_d);
    const cd* giw = reinterpret_cast<const cd*>(vx->giw);
    auto g = [&](int w, int orb) { return giw[(w + vx->iwmax) + (size_t)orb * 2 * vx->iwmax]; };  // 0-based orbital
    const double beta = vx->beta;
    for (int w2 = -iwf; w2 < iwf; ++w2)
      for (int l = 0; l < n; ++l)
        for (int k = 0; k < n; ++k) {
          const size_t i2 = (size_t)(w2 + iwf) * n2 + l * n + k;
          for (int w1 = -iwf; w1 < iwf; ++w1)
            for (int i = 0; i < n; ++i)
              for (int j = 0; j < n; ++j) {
                const size_t i1 = (size_t)(w1 + iwf) * n2 + i * n + j;
                const size_t e = g4i(i, j, w1 + vx->iwfmax, k, l, w2 + vx->iwfmax);
                cd vm = g4m[e], vd = g4d[e];
                const bool cor = same_correlated(at, i + 1, j + 1, k + 1, l + 1);
                const bool bubble = w2 == w1 && i == l && j == k;
                if (vx->vertex_type == 0) {
                  if (iwb == 0 && i == j && k == l && cor) vd -= 2.0 * beta * g(w1, i) * g(w2, l);
                  if (bubble && !cor) {
                    const cd b = beta * g(w1, i) * g(w2 - iwb, j);
                    vd -= b;
                    vm -= b;
                  }
                } else if (vx->vertex_type == 1) {
                  if (bubble) {
                    const cd b = beta * g(w1, i) * g(w2 - iwb, j);
                    vd -= b;
                    vm -= b;
                  }
                } else if (bubble && !cor) {
                  const cd b = beta * g(w1, i) * g(w2 - iwb, j);
                  vd -= b;
                  vm -= b;
                }
                xm[i1 + i2 * N] = vm;
                xd[i1 + i2 * N] = vd;
              }
        }
  });
}

int adga_h5_write_vertex(adga_h5* f, int ineq, int iwb, int iwbmax, int nbands, int iwfmax, const double* g4_dens,
                         const double* g4_magn) {
  if (!f || ineq < 1 || nbands < 1 || iwfmax < 1 || iwb < -iwbmax || iwb > iwbmax || !g4_dens || !g4_magn)
    return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int nb = nbands, nfb = 2 * iwfmax;
    const size_t plane = (size_t)nfb * nfb;
    std::vector<double> buf(2 * plane);
    for (int ch = 0; ch < 2; ++ch) {
      const std::string base = ineq_name(ineq) + (ch ? "/dens" : "/magn");
      // intermediate groups as needed
      for (const std::string& g : {ineq_name(ineq), base, base + "/" + five(iwb + iwbmax)})
        if (H5Lexists(f->file, g.c_str(), H5P_DEFAULT) <= 0)
          H5Gclose(ck(H5Gcreate2(f->file, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group " + g));
      const double* g4 = ch ? g4_dens : g4_magn;
      for (int b1 = 0; b1 < nb; ++b1)
        for (int b2 = 0; b2 < nb; ++b2)
          for (int b3 = 0; b3 < nb; ++b3)
            for (int b4 = 0; b4 < nb; ++b4) {
              const int ind = nb * nb * nb * b1 + nb * nb * b2 + nb * b3 + b4 + 1;  // aux.f90:157-164
              const std::string grp = base + "/" + five(iwb + iwbmax) + "/" + five(ind);
              H5Gclose(ck(H5Gcreate2(f->file, grp.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group " + grp));
              // value(nu1, nu2) = g4(b1, b2, nu1, b3, b4, nu2), nu1 fastest
              for (int u2 = 0; u2 < nfb; ++u2)
                for (int u1 = 0; u1 < nfb; ++u1) {
                  const size_t e = (((((size_t)u2 * nb + b4) * nb + b3) * nfb + u1) * nb + b2) * nb + b1;
                  buf[2 * ((size_t)u2 * nfb + u1)] = g4[2 * e];
                  buf[2 * ((size_t)u2 * nfb + u1) + 1] = g4[2 * e + 1];
                }
              Id gg(open_group(f->file, grp.c_str()), H5Gclose);
              write_cplx(gg, "value", {(hsize_t)nfb, (hsize_t)nfb}, buf.data());
            }
    }
  });
}

int adga_h5_read_vq(adga_h5* f, int ndim, int nqp, int q_vol, const int* q_data, double* v_q) {
  if (!f || ndim < 1 || nqp < 1 || !v_q || (!q_vol && !q_data)) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int n = ndim, n2 = n * n;
    cd* v = reinterpret_cast<cd*>(v_q);
    for (size_t i = 0; i < (size_t)n2 * n2 * nqp; ++i) v[i] = 0.0;
    dims_t qd;
    (void)read_real(f->file, ".axes/Q-points", &qd);  // (points, 3)
    if (qd.size() != 2) throw Fail{ADGA_H5_E_FORMAT, ".axes/Q-points: expected (points, 3)"};
    if (q_vol && (int)qd[0] != nqp) throw Fail{ADGA_H5_E_FORMAT, "Inconsistent number of q-points in V^q"};
    H5G_info_t info;
    ckh(H5Gget_info(f->file, &info), "members of /");
    for (hsize_t m = 0; m < info.nlinks; ++m) {
      char name[64];
      if (H5Lget_name_by_idx(f->file, ".", H5_INDEX_NAME, H5_ITER_INC, m, name, sizeof name, H5P_DEFAULT) < 0)
        throw Fail{ADGA_H5_E_IO, "member name of /"};
      if (name[0] == '.') continue;  // .axes
      const int ind = atoi(name), r = ind - 1;
      const int i = r / (n * n * n), j = (r / (n * n)) % n, k = (r / n) % n, l = r % n;
      if (ind < 1 || i >= n) throw Fail{ADGA_H5_E_FORMAT, std::string("band index ") + name};
      dims_t dd;
      const std::vector<double> x = read_cplx(f->file, name, &dd);
      // v(i1 = {k i}, i2 = {j l}) = vq(i, j, k, l) with i1 = i*n + k, i2 = l*n + j (0-based)
      const size_t i1 = (size_t)i * n + k, i2 = (size_t)l * n + j;
      for (int iq = 0; iq < nqp; ++iq) {
        const size_t p = q_vol ? (size_t)iq : (size_t)(q_data[iq] - 1);
        if (p >= x.size() / 2) throw Fail{ADGA_H5_E_FORMAT, std::string(name) + ": fewer points than needed"};
        v[i1 + i2 * n2 + (size_t)iq * n2 * n2] = cd(x[2 * p], x[2 * p + 1]);
      }
    }
  });
}

int adga_h5_init_output(adga_h5* f, const adga_h5_run* r) {
  if (!f || !r || r->ndim < 1) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const hid_t file = f->file;
    auto group = [&](hid_t loc, const char* name) {
      return ck(H5Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), std::string("group ") + name);
    };
    if (r->do_chi) {
      Id s(group(file, "susceptibility"), H5Gclose);
      if (r->nonlocal) H5Gclose(group(s, "nonloc"));
      H5Gclose(group(s, "loc"));
    }
    if (r->do_eom) {
      if (r->nonlocal) H5Gclose(group(file, "occupation"));
      Id s(group(file, "selfenergy"), H5Gclose);
      if (r->nonlocal) H5Gclose(group(s, "nonloc"));
      H5Gclose(group(s, "loc"));
    }
    if (!r->nonlocal && r->gamma_group) H5Gclose(group(file, "gamma"));
    Id in(group(file, "input"), H5Gclose);
    const int n = r->ndim;
    write_num(in, "mu", H5T_NATIVE_DOUBLE, {}, &r->mu);
    write_num(in, "beta", H5T_NATIVE_DOUBLE, {}, &r->beta);
    if (r->dc) write_cplx(in, "dc", {(hsize_t)n, 2}, r->dc);
    if (r->giw) write_cplx(in, "giw", {(hsize_t)n, (hsize_t)(2 * r->iwmax)}, r->giw);
    if (r->siw) write_cplx(in, "siw", {(hsize_t)n, (hsize_t)(2 * r->iwmax)}, r->siw);
    if (r->hk) {
      // hk_arr(kz, ky, kx, i2, i1) = hk(i1, i2, k): C order [i1][i2][kx][ky][kz] (hdf5_module.f90:1097-1114)
      std::vector<double> a(2 * (size_t)n * n * r->nkp);
      size_t o = 0;
      for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2)
          for (int kx = 0; kx < r->nkpx; ++kx)
            for (int ky = 0; ky < r->nkpy; ++ky)
              for (int kz = 0; kz < r->nkpz; ++kz) {
                const size_t k = kz + (size_t)ky * r->nkpz + (size_t)kx * r->nkpy * r->nkpz;
                const double* p = &r->hk[2 * (i1 + (size_t)i2 * n + k * n * n)];
                a[o++] = p[0];
                a[o++] = p[1];
              }
      write_cplx(in, "hk", {(hsize_t)n, (hsize_t)n, (hsize_t)r->nkpx, (hsize_t)r->nkpy, (hsize_t)r->nkpz}, a.data());
    }
    if (r->n_dmft) {
      std::vector<double> nd(n);
      for (int i = 0; i < n; ++i) nd[i] = r->n_dmft[2 * i];
      write_num(in, "n_dmft", H5T_NATIVE_DOUBLE, {(hsize_t)n}, nd.data());
    }
    if (r->n_fock) {
      // the reference writes real(n_fock) in n_fock's own (nkp, n, n) order under the dims
      // (kz, ky, kx, n, n) (hdf5_module.f90:1128-1145)
      std::vector<double> nf((size_t)r->nkp * n * n);
      for (size_t i = 0; i < nf.size(); ++i) nf[i] = r->n_fock[2 * i];
      write_num(in, "n_dmft_k", H5T_NATIVE_DOUBLE,
                {(hsize_t)n, (hsize_t)n, (hsize_t)r->nkpx, (hsize_t)r->nkpy, (hsize_t)r->nkpz}, nf.data());
    }
    const int ints[] = {r->iwmax, r->iwfmax, r->iwbmax, r->iwfmax_small, r->iwbmax_small, r->nkp, r->nqp};
    const char* names[] = {"iwmax", "iwfmax", "iwbmax", "iwfmax_small", "iwbmax_small", "nkp", "nqp"};
    for (int i = 0; i < 7; ++i) write_num(in, names[i], H5T_NATIVE_INT, {}, &ints[i]);
    const int nk[3] = {r->nkpx, r->nkpy, r->nkpz}, nq[3] = {r->nqpx, r->nqpy, r->nqpz};
    write_num(in, "nkpxyz", H5T_NATIVE_INT, {3}, nk);
    write_num(in, "nqpxyz", H5T_NATIVE_INT, {3}, nq);
    auto coords = [&](const char* name, int np, const int* idx) {
      if (!r->k_data || !idx) throw Fail{ADGA_H5_E_ARG, std::string(name) + " needs k_data and the index list"};
      std::vector<double> c(3 * (size_t)np);
      for (int i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a) c[3 * (size_t)i + a] = r->k_data[3 * (size_t)(idx[i] - 1) + a];
      write_num(in, name, H5T_NATIVE_DOUBLE, {(hsize_t)np, 3}, c.data());
    };
    if (r->q_path_susc) coords("qpath", r->nqp, r->q_data);
    if (r->k_path_eom) coords("kpath_eom", r->nkp_eom, r->k_data_eom);
  });
}

int adga_h5_write_chi_qw(adga_h5* f, const char* channel, int full, int n, int iwbmax_small, const int nq[3],
                         const double* chi, int gzip) {
  if (!f || !channel || n < 1 || iwbmax_small < 0 || !nq || !chi) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int nw = 2 * iwbmax_small + 1, nqx = nq[0], nqy = nq[1], nqz = nq[2], n2 = n * n;
    const size_t nqp = (size_t)nqx * nqy * nqz;
    Id g(open_group(f->file, "susceptibility/nonloc"), H5Gclose);
    std::vector<double> a;
    dims_t dims, chunk;
    auto put = [&](int r1, int c1, int qx, int qy, int qz, int w) {
      const size_t p = (size_t)w * nqp + qz + (size_t)qy * nqz + (size_t)qx * nqy * nqz;
      const double* s = &chi[2 * (r1 + (size_t)c1 * n2 + p * n2 * n2)];
      a.push_back(s[0]);
      a.push_back(s[1]);
    };
    if (full) {
      // chi_slice(qz, qy, qx, i4, i3, i2, i1) per omega = chi({i1 i2}, {i4 i3}, point) (hdf5_module.f90:1411-1427)
      for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2)
          for (int i3 = 0; i3 < n; ++i3)
            for (int i4 = 0; i4 < n; ++i4)
              for (int qx = 0; qx < nqx; ++qx)
                for (int qy = 0; qy < nqy; ++qy)
                  for (int qz = 0; qz < nqz; ++qz)
                    for (int w = 0; w < nw; ++w) put(i1 * n + i2, i4 * n + i3, qx, qy, qz, w);
      dims = {(hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)nqx, (hsize_t)nqy, (hsize_t)nqz, (hsize_t)nw};
      chunk = dims;
      chunk.back() = 1;
    } else {
      // (w, qz, qy, qx, i3, i1) = chi({i1 i1}, {i3 i3}, point) (hdf5_module.f90:1477-1491)
      for (int i1 = 0; i1 < n; ++i1)
        for (int i3 = 0; i3 < n; ++i3)
          for (int qx = 0; qx < nqx; ++qx)
            for (int qy = 0; qy < nqy; ++qy)
              for (int qz = 0; qz < nqz; ++qz)
                for (int w = 0; w < nw; ++w) put(i1 * n + i1, i3 * n + i3, qx, qy, qz, w);
      dims = {(hsize_t)n, (hsize_t)n, (hsize_t)nqx, (hsize_t)nqy, (hsize_t)nqz, (hsize_t)nw};
      chunk = dims;
      chunk.back() = 1;
    }
    write_cplx(g, channel, dims, a.data(), chunk, gzip);
  });
}

int adga_h5_write_chi_qpath(adga_h5* f, const char* channel, int full, int n, int iwbmax_small, int nqp,
                            const double* chi) {
  if (!f || !channel || n < 1 || iwbmax_small < 0 || nqp < 1 || !chi) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int nw = 2 * iwbmax_small + 1, n2 = n * n;
    Id g(open_group(f->file, "susceptibility/nonloc"), H5Gclose);
    std::vector<double> a;
    auto put = [&](int r1, int c1, int q, int w) {
      const double* s = &chi[2 * (r1 + (size_t)c1 * n2 + ((size_t)w * nqp + q) * n2 * n2)];
      a.push_back(s[0]);
      a.push_back(s[1]);
    };
    if (full) {
      // (w, q, i4, i3, i2, i1) = chi({i1 i2}, {i4 i3}, w nqp + q) (hdf5_module.f90:1546-1557)
      for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2)
          for (int i3 = 0; i3 < n; ++i3)
            for (int i4 = 0; i4 < n; ++i4)
              for (int q = 0; q < nqp; ++q)
                for (int w = 0; w < nw; ++w) put(i1 * n + i2, i4 * n + i3, q, w);
      write_cplx(g, channel, {(hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)nqp, (hsize_t)nw}, a.data());
    } else {
      // (w, q, i1, i3) = chi({i1 i1}, {i3 i3}, w nqp + q) (hdf5_module.f90:1610-1619)
      for (int i3 = 0; i3 < n; ++i3)
        for (int i1 = 0; i1 < n; ++i1)
          for (int q = 0; q < nqp; ++q)
            for (int w = 0; w < nw; ++w) put(i1 * n + i1, i3 * n + i3, q, w);
      write_cplx(g, channel, {(hsize_t)n, (hsize_t)n, (hsize_t)nqp, (hsize_t)nw}, a.data());
    }
  });
}

int adga_h5_write_chi_loc(adga_h5* f, const char* channel, int full, int n, int iwbmax_small, const double* chi,
                          int gzip) {
  if (!f || !channel || n < 1 || iwbmax_small < 0 || !chi) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int nw = 2 * iwbmax_small + 1, n2 = n * n;
    Id g(open_group(f->file, "susceptibility/loc"), H5Gclose);
    std::vector<double> a;
    auto put = [&](int r1, int c1, int w) {
      const double* s = &chi[2 * (r1 + (size_t)c1 * n2 + (size_t)w * n2 * n2)];
      a.push_back(s[0]);
      a.push_back(s[1]);
    };
    if (full) {
      // (w, i4, i3, i2, i1) = chi({i1 i2}, {i4 i3}, w) (hdf5_module.f90:1249-1259)
      for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2)
          for (int i3 = 0; i3 < n; ++i3)
            for (int i4 = 0; i4 < n; ++i4)
              for (int w = 0; w < nw; ++w) put(i1 * n + i2, i4 * n + i3, w);
      write_cplx(g, channel, {(hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)nw}, a.data(),
                 {(hsize_t)n, (hsize_t)n, (hsize_t)n, (hsize_t)n, 1}, gzip);
    } else {
      // (w, i3, i1) = chi({i1 i1}, {i3 i3}, w) (hdf5_module.f90:1320-1328)
      for (int i1 = 0; i1 < n; ++i1)
        for (int i3 = 0; i3 < n; ++i3)
          for (int w = 0; w < nw; ++w) put(i1 * n + i1, i3 * n + i3, w);
      write_cplx(g, channel, {(hsize_t)n, (hsize_t)n, (hsize_t)nw}, a.data(), {(hsize_t)n, (hsize_t)n, 1}, gzip);
    }
  });
}

int adga_h5_write_eom(adga_h5* f, int nonlocal, int n, int iwf, const int nk[3], const double* ss, const double* shf,
                      const double* sloc, const double* sdmft, int gzip) {
  if (!f || n < 1 || iwf < 1 || !nk || !sdmft || (nonlocal && (!ss || !shf || !sloc))) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int nf = 2 * iwf, nkx = nk[0], nky = nk[1], nkz = nk[2], n2 = n * n;
    const size_t nkp = (size_t)nkx * nky * nkz;
    auto local = [&](const double* s) {  // (iw, i2, i1) = s(i1, i2, iw): C order [i1][i2][iw]
      std::vector<double> a;
      for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2)
          for (int w = 0; w < nf; ++w) {
            const double* p = &s[2 * (i1 + (size_t)i2 * n + (size_t)w * n2)];
            a.push_back(p[0]);
            a.push_back(p[1]);
          }
      return a;
    };
    if (nonlocal) {
      Id g(open_group(f->file, "selfenergy/nonloc"), H5Gclose);
      const dims_t dims = {(hsize_t)n, (hsize_t)n, (hsize_t)nkx, (hsize_t)nky, (hsize_t)nkz, (hsize_t)nf};
      const dims_t chunk = {1, 1, (hsize_t)nkx, (hsize_t)nky, (hsize_t)nkz, (hsize_t)nf};
      // siwk(iw, kz, ky, kx, i2, i1) = sigma_sum(i1, i2, iw, k) (hdf5_module.f90:1670-1684); the Hartree-Fock
      // term repeated over nu (1707-1718)
      for (int which = 0; which < 2; ++which) {
        std::vector<double> a;
        a.reserve(2 * (size_t)n2 * nkp * nf);
        for (int i1 = 0; i1 < n; ++i1)
          for (int i2 = 0; i2 < n; ++i2)
            for (int kx = 0; kx < nkx; ++kx)
              for (int ky = 0; ky < nky; ++ky)
                for (int kz = 0; kz < nkz; ++kz) {
                  const size_t k = kz + (size_t)ky * nkz + (size_t)kx * nky * nkz;
                  for (int w = 0; w < nf; ++w) {
                    const double* p = which == 0 ? &ss[2 * (i1 + (size_t)i2 * n + (size_t)w * n2 + k * n2 * nf)]
                                                 : &shf[2 * (i1 + (size_t)i2 * n + k * n2)];
                    a.push_back(p[0]);
                    a.push_back(p[1]);
                  }
                }
        write_cplx(g, which == 0 ? "dga" : "hartree_fock", dims, a.data(), chunk, gzip);
      }
      Id gl(open_group(f->file, "selfenergy/loc"), H5Gclose);
      write_cplx(gl, "dga_ksum", {(hsize_t)n, (hsize_t)n, (hsize_t)nf}, local(sloc).data());
    }
    Id gl(open_group(f->file, "selfenergy/loc"), H5Gclose);
    write_cplx(gl, "dmft", {(hsize_t)n, (hsize_t)n, (hsize_t)nf}, local(sdmft).data());
  });
}

int adga_h5_write_occ(adga_h5* f, int n, const int nk[3], const double* n_dga, const double* n_dga_k) {
  if (!f || n < 1 || !nk || !n_dga || !n_dga_k) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const int nkx = nk[0], nky = nk[1], nkz = nk[2];
    const size_t nkp = (size_t)nkx * nky * nkz;
    Id g(open_group(f->file, "occupation"), H5Gclose);
    std::vector<double> nd(n);
    for (int i = 0; i < n; ++i) nd[i] = n_dga[2 * i];
    write_num(g, "n_dga", H5T_NATIVE_DOUBLE, {(hsize_t)n}, nd.data());
    // ndga_k_arr(kz, ky, kx, i2, i1) = n_dga_k(k, i1, i2) (hdf5_module.f90:1938-1950), real parts
    std::vector<double> a;
    a.reserve((size_t)n * n * nkp);
    for (int i1 = 0; i1 < n; ++i1)
      for (int i2 = 0; i2 < n; ++i2)
        for (int kx = 0; kx < nkx; ++kx)
          for (int ky = 0; ky < nky; ++ky)
            for (int kz = 0; kz < nkz; ++kz) {
              const size_t k = kz + (size_t)ky * nkz + (size_t)kx * nky * nkz;
              a.push_back(n_dga_k[2 * (k + (size_t)i1 * nkp + (size_t)i2 * nkp * n)]);
            }
    write_num(g, "n_dga_k", H5T_NATIVE_DOUBLE, {(hsize_t)n, (hsize_t)n, (hsize_t)nkx, (hsize_t)nky, (hsize_t)nkz},
              a.data());
  });
}

int adga_h5_dims(adga_h5* f, const char* path, int* rank, long long* dims) {
  if (!f || !path || !rank || !dims) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    Id d(ck(H5Dopen2(f->file, path, H5P_DEFAULT), std::string("open ") + path), H5Dclose);
    const dims_t dd = dataset_dims(d, path);
    if (dd.size() > 8) throw Fail{ADGA_H5_E_FORMAT, "rank above 8"};
    *rank = (int)dd.size();
    for (size_t i = 0; i < dd.size(); ++i) dims[i] = (long long)dd[i];
  });
}

int adga_h5_write_complex(adga_h5* f, const char* path, int rank, const long long* dims, const double* data) {
  if (!f || !path || rank < 1 || rank > 8 || !dims || !data) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    make_parents(f->file, path);
    write_cplx(f->file, path, dims_t(dims, dims + rank), data);
  });
}

int adga_h5_write_real(adga_h5* f, const char* path, int rank, const long long* dims, const double* data) {
  if (!f || !path || rank < 0 || rank > 8 || (rank > 0 && !dims) || !data) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    make_parents(f->file, path);
    write_num(f->file, path, H5T_NATIVE_DOUBLE, rank ? dims_t(dims, dims + rank) : dims_t{}, data);
  });
}

int adga_h5_write_attr_real(adga_h5* f, const char* group, const char* name, double value) {
  if (!f || !group || !name) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    if (H5Lexists(f->file, group, H5P_DEFAULT) <= 0) {
      make_parents(f->file, std::string(group) + "/x");
    }
    Id g(open_group(f->file, group), H5Gclose);
    Id sp(ck(H5Screate(H5S_SCALAR), "dataspace"), H5Sclose);
    Id a(ck(H5Acreate2(g, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT), std::string("attribute ") + name),
         H5Aclose);
    ckh(H5Awrite(a, H5T_NATIVE_DOUBLE, &value), std::string("write attribute ") + name);
  });
}

int adga_h5_read_complex(adga_h5* f, const char* path, double* out) {
  if (!f || !path || !out) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const std::vector<double> v = read_cplx(f->file, path);
    memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

int adga_h5_read_real(adga_h5* f, const char* path, double* out) {
  if (!f || !path || !out) return ADGA_H5_E_ARG;
  return guarded(f, [&] {
    const std::vector<double> v = read_real(f->file, path);
    memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

}  // extern "C"
