// Host-side services of wrfmi_api.hip that the per-routine entry points with the reference's
// argument lists (wrfmi_routines.hip) share: error reporting, the library lock, the
// host-array registry, option and dims checks, a domain's work arrays and the stream those
// entry points enqueue on.  Internal to libwrfmi.so.
#pragma once
#include <mutex>
#include "kernels.h"

namespace wrfmi {
namespace host {
// sets the message wrfmi_last_error returns and returns `code`
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
std::mutex &lock();
// the device address of p: its registered mirror, or p itself (lock held)
void *translate(void *p);
// the option set the kernels implement, and the index sextets (tile = domain, halo >= 4)
int check_cfg(const wrfmi_config *c);
int make_dm(const wrfmi_dims *w, Dm *d);
// work arrays (dvdxi, rhs, FP64 low part) of the domain whose first 3-D array is at device
// address `key` (lock held)
int domain_scratch(const void *key, const Dm &d, float **dv, float **rhs, float **lo);
// WRFMI_E_POINTER unless p is device memory (a registered host array's mirror included)
int device_array(const void *p, const char *name);
// the stream wrfmi_set_stream selected for the calling thread (default: the null stream)
hipStream_t routine_stream();
// one state kernel over a tnx x tny x tnz tile grid, launched 1-D in the XCD-aware order
void launch_tiles(void (*kern)(KArgs), dim3 g, dim3 b, KArgs a, hipStream_t st);
}  // namespace host
}  // namespace wrfmi
