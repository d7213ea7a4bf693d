/*
 * asterx_eosio.h -- host-side reader of tabulated equations of state.
 *
 * EOSX reads a StellarCollapse table (HDF5) once at start-up, converts it to
 * code units and reorders it with the variable index fastest
 * (EOSX/src/eos_3p_tabulated3d/eos_3p_tabulated3d.hxx:53-193, helpers in
 * eos_readtable_scollapse.hxx).  This library does the same into host
 * memory, in the layout of asx_eos_table (include/asterx.h) for the four
 * variables the hot path reads; the caller copies the arrays to the device
 * and fills an asx_eos_table.  Separate from libasterx_hip so that only
 * callers that read a table file link the HDF5 library.
 */
#ifndef ASTERX_EOSIO_H
#define ASTERX_EOSIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A table in host memory: n[0] x n[1] x n[2] nodes in (log rho, log T, Ye),
 * data[v + 4 (i + n[0] (j + n[1] k))] with v = log p, log(eps + energy_shift),
 * c_s^2 (code units) and the entropy per baryon; axis[0] = log rho and
 * axis[1] = log T (natural logs, code units, T in MeV), axis[2] = Ye;
 * eps_min / eps_max = min and max of exp(log(eps + shift)) - shift over the
 * nodes, the minimum floored at 1e-15.  Arrays owned by the library until
 * asx_eos_host_table_free. */
typedef struct {
  int32_t n[3];
  double *data;
  double *axis[3];
  double energy_shift;
  double eps_min, eps_max;
} asx_eos_host_table;

/* EOSX's StellarCollapse reader (eos_3p_tabulated3d.hxx:53-193): the
 * integer datasets pointsrho, pointstemp, pointsye, the axes logrho,
 * logtemp (log10, g/cm^3 and MeV) and ye, energy_shift (erg/g), and the
 * [nye][ntemp][nrho] datasets logpress (log10 dyn/cm^2), logenergy
 * (log10 erg/g, shifted), cs2 (cm^2/s^2) and entropy (k_B per baryon).
 * Returns 0 on success, 1 for a NULL argument, 2 if the file or a dataset
 * cannot be read or has the wrong size, 3 for fewer than 2 points on an
 * axis or an axis that is not uniform to 1e-10 of its step (the
 * interpolation assumes uniform axes, eos_linear_interp_ND.hxx:50-64). */
int32_t asx_eos_read_stellarcollapse(const char *path, asx_eos_host_table *out);
void asx_eos_host_table_free(asx_eos_host_table *t);

#ifdef __cplusplus
}
#endif

#endif /* ASTERX_EOSIO_H */
