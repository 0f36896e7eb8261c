/*
 * wrfmi_history.h -- history output of a domain in the reference's wrfout layout.
 *
 * Replaces the history stream of the reference (share/output_wrf.F:3 output_wrf, called
 * from med_hist_out for io_form_history = 2) for the dynamics state: a netCDF classic file
 * (64-bit offset format, CDF-2) with the reference's dimension names (Time, DateStrLen,
 * west_east, south_north, bottom_top and their _stag forms), variable names (the Registry
 * 'h' stream of Registry/Registry.EM_COMMON: U, V, W, PH, PHB, T, THM, MU, MUB, P, PB,
 * P_TOP, ZNU, ZNW, HGT, QVAPOR, ...), per-variable attributes (FieldType, MemoryOrder,
 * description, units, stagger; external/io_netcdf/wrf_io.F90:2693-2702) and the global
 * attributes output_wrf.F writes for the options the step implements.
 *
 * The fields are read from the state's device arrays (or registered host arrays); each
 * record is packed on the GPU into the file's (k, j, i) big-endian layout, copied to the
 * host once and appended.  T is the dry potential temperature perturbation
 * (t + t0)/(1 + (Rv/Rd) qv) - t0 when the state carries moist theta (use_theta_m 1),
 * as module_big_step_utilities_em.F:5739 forms it; THM is the prognostic t itself.
 */
#ifndef WRFMI_HISTORY_H
#define WRFMI_HISTORY_H

#include "wrfmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wrfmi_history wrfmi_history;

/* Creates `path` and writes the header for the domain of `st` / `cfg` / `dims`.
 * start_date: "YYYY-MM-DD_hh:mm:ss" (START_DATE, SIMULATION_START_DATE and the XTIME units).
 * species_names: n_species names for st->q (NULL, or a NULL entry, for the defaults:
 * QVAPOR for q[0] of a moist state, Q<m+1> for other moist species, SCALAR<n> for passive
 * scalars).  Replaces the open_w_dataset + output_wrf header pass of med_hist_out. */
int wrfmi_history_open(const char *path, const wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *dims,
                       const char *start_date, const char *const *species_names, wrfmi_history **out);

/* Appends one time record: `date` ("YYYY-MM-DD_hh:mm:ss", the Times entry) and `xtime`
 * (minutes since start_date), with every variable read from `st` after the work already
 * queued on `stream` (NULL: the null stream).  The arrays of `st` must have the layout given
 * at open.  Replaces output_wrf's per-time write (share/output_wrf.F:3). */
int wrfmi_history_write(wrfmi_history *h, const wrfmi_state *st, const char *date, float xtime, void *stream);

/* Number of records written so far, or -1 for a null handle. */
int wrfmi_history_records(const wrfmi_history *h);

/* Closes the file and frees the handle (close_dataset). */
int wrfmi_history_close(wrfmi_history *h);

#ifdef __cplusplus
}
#endif
#endif
