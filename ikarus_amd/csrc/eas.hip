// eas.hip — EAS (Green-Lagrange, 21 modes) hex8 element kernels: placeholder until the condensation
// kernel lands; the C ABI reports IK_ENOTIMPL for EAS assembly.
#include "ik_internal.hh"

namespace ik {
void easUpdate(ik_handle* h, const double* duDev) {
  (void)h;
  (void)duDev;
  throw Error{IK_ENOTIMPL, "EAS kernels not built yet"};
}
}  // namespace ik
