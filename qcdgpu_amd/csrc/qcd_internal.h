// qcd_internal.h -- helpers shared by the translation units of libqcdhip.so
// (not part of the C ABI).
#ifndef QCD_INTERNAL_H
#define QCD_INTERNAL_H

namespace qcdhip {

// Records `msg` as the text qcd_last_error() returns; returns `code`.
int set_error(int code, const char* msg);

}  // namespace qcdhip

#endif
