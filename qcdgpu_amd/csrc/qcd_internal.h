// qcd_internal.h -- helpers shared by the translation units of libqcdhip.so
// (not part of the C ABI).
#ifndef QCD_INTERNAL_H
#define QCD_INTERNAL_H

#include <hip/hip_runtime.h>

#include "../../include/qcd_hip.h"

namespace qcdhip {

// Records `msg` as the text qcd_last_error() returns; returns `code`.
int set_error(int code, const char* msg);

// Stores `count` F_munu records (12 doubles each, the layout qcd_read_fmunu
// returns) as records 0..count-1; records past the capacity are dropped.
int write_fmunu(qcd_lattice* lat, const double* f, int count);

// SU(2) FP32 lattices: enqueues on the lattice's stream a copy of every link
// into unit quaternions q[mu * V + site] (to_quats) or back into the table;
// *stream (optional) receives that stream.
int su2_quats(qcd_lattice* lat, float4* q, bool to_quats, hipStream_t* stream);

// N1..N4 of the lattice.
int lattice_dims(const qcd_lattice* lat, int n[4]);

}  // namespace qcdhip

#endif
