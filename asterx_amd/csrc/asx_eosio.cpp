// Host reader of StellarCollapse equation-of-state tables (HDF5) into the
// asx_eos_table layout (include/asterx_eosio.h).
#include "asterx_eosio.h"

#include <hdf5.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// EOSX/src/utils/eos_constants.hxx:7-12: CGS -> code units (G = c = Msun = 1)
constexpr double LENGTHGF = 6.77269222552442e-06;
constexpr double TIMEGF = 2.03040204956746e05;
constexpr double RHOGF = 1.61887093132742e-18;
constexpr double PRESSGF = 1.80123683248503e-39;
constexpr double EPSGF = 1.11265005605362e-21;

// One dataset of `count` elements of the given memory type; false if it is
// missing or has another number of elements.
bool read_dset(hid_t file, const char *name, hid_t mtype, size_t count, void *buf) {
  if (H5Lexists(file, name, H5P_DEFAULT) <= 0) return false;
  const hid_t d = H5Dopen2(file, name, H5P_DEFAULT);
  if (d < 0) return false;
  bool ok = false;
  const hid_t sp = H5Dget_space(d);
  if (sp >= 0) {
    const hssize_t np = H5Sget_simple_extent_npoints(sp);
    if (np == (hssize_t)count) ok = H5Dread(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0;
    H5Sclose(sp);
  }
  H5Dclose(d);
  return ok;
}

bool uniform(const std::vector<double> &a) {
  const double h = a[1] - a[0];
  if (!(h > 0)) return false;
  for (size_t i = 1; i < a.size(); ++i)
    if (std::fabs((a[i] - a[i - 1]) - h) > 1e-10 * h) return false;
  return true;
}

} // namespace

extern "C" {

int32_t asx_eos_read_stellarcollapse(const char *path, asx_eos_host_table *out) {
  if (!path || !out) return 1;
  std::memset(out, 0, sizeof(*out));
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); // report by return code only
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return 2;
  int st = 0;
  int nrho = 0, ntemp = 0, nye = 0;
  std::vector<double> lr, lt, ye, tab[4];
  double shift = 0.0;
  if (!read_dset(f, "pointsrho", H5T_NATIVE_INT, 1, &nrho) ||
      !read_dset(f, "pointstemp", H5T_NATIVE_INT, 1, &ntemp) ||
      !read_dset(f, "pointsye", H5T_NATIVE_INT, 1, &nye)) {
    st = 2;
  } else if (nrho < 2 || ntemp < 2 || nye < 2) {
    st = 3;
  } else {
    const size_t np = (size_t)nrho * ntemp * nye;
    lr.resize(nrho);
    lt.resize(ntemp);
    ye.resize(nye);
    static const char *names[4] = {"logpress", "logenergy", "cs2", "entropy"};
    bool ok = read_dset(f, "logrho", H5T_NATIVE_DOUBLE, nrho, lr.data()) &&
              read_dset(f, "logtemp", H5T_NATIVE_DOUBLE, ntemp, lt.data()) &&
              read_dset(f, "ye", H5T_NATIVE_DOUBLE, nye, ye.data()) &&
              read_dset(f, "energy_shift", H5T_NATIVE_DOUBLE, 1, &shift);
    for (int v = 0; ok && v < 4; ++v) {
      tab[v].resize(np);
      ok = read_dset(f, names[v], H5T_NATIVE_DOUBLE, np, tab[v].data());
    }
    if (!ok) st = 2;
    else if (!uniform(lr) || !uniform(lt) || !uniform(ye)) st = 3;
  }
  H5Fclose(f);
  if (st) return st;

  // eos_3p_tabulated3d.hxx:151-165: natural logs in code units; the table
  // reordered with the variable fastest (:117-123)
  const double ln10 = std::log(10.0);
  const size_t np = (size_t)nrho * ntemp * nye;
  out->n[0] = nrho;
  out->n[1] = ntemp;
  out->n[2] = nye;
  out->energy_shift = shift * EPSGF;
  out->axis[0] = (double *)std::malloc(nrho * sizeof(double));
  out->axis[1] = (double *)std::malloc(ntemp * sizeof(double));
  out->axis[2] = (double *)std::malloc(nye * sizeof(double));
  out->data = (double *)std::malloc(4 * np * sizeof(double));
  if (!out->axis[0] || !out->axis[1] || !out->axis[2] || !out->data) {
    asx_eos_host_table_free(out);
    return 2;
  }
  for (int i = 0; i < nrho; ++i) out->axis[0][i] = lr[i] * ln10 + std::log(RHOGF);
  for (int i = 0; i < ntemp; ++i) out->axis[1][i] = lt[i] * ln10;
  for (int i = 0; i < nye; ++i) out->axis[2][i] = ye[i];
  const double inv_time2 = 1 / (TIMEGF * TIMEGF);
  double emin = HUGE_VAL, emax = -HUGE_VAL;
  for (size_t p = 0; p < np; ++p) {
    double *d = out->data + 4 * p;
    d[0] = tab[0][p] * ln10 + std::log(PRESSGF);
    d[1] = tab[1][p] * ln10 + std::log(EPSGF);
    d[2] = tab[2][p] * LENGTHGF * LENGTHGF * inv_time2;
    d[3] = tab[3][p];
    const double e = std::exp(d[1]) - out->energy_shift;
    emin = e < emin ? e : emin;
    emax = e > emax ? e : emax;
  }
  // EOSX's eps range over the nodes, the minimum floored
  // (eos_3p_tabulated3d.hxx:341-357; the true min / max over the eps nodes,
  // where the reference indexes its interleaved table as variable-major)
  out->eps_min = emin > 1e-15 ? emin : 1e-15;
  out->eps_max = emax;
  return 0;
}

void asx_eos_host_table_free(asx_eos_host_table *t) {
  if (!t) return;
  std::free(t->data);
  for (int k = 0; k < 3; ++k) std::free(t->axis[k]);
  std::memset(t, 0, sizeof(*t));
}

} // extern "C"
