// qcd_internal.h -- helpers shared by the translation units of libqcdhip.so
// (not part of the C ABI).
#ifndef QCD_INTERNAL_H
#define QCD_INTERNAL_H

#include "../../include/qcd_hip.h"

namespace qcdhip {

// Records `msg` as the text qcd_last_error() returns; returns `code`.
int set_error(int code, const char* msg);

// Stores `count` F_munu records (12 doubles each, the layout qcd_read_fmunu
// returns) as records 0..count-1; records past the capacity are dropped.
int write_fmunu(qcd_lattice* lat, const double* f, int count);

}  // namespace qcdhip

#endif
