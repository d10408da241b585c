// surrogate.h -- device entry points of the learned energy model
// (surrogate.hip) used by the C-ABI driver.
#pragma once
#include "engine.h"

struct EmDeeSurrogate;

namespace emdee {
// Adds the model forces to F(layer) of the system (and its atom energies to
// the atom-energy array when that is on) and returns the model energy; W
// receives its pair virial.
double dev_surrogate_apply(Device* d, EmDeeSurrogate* s, int layer, double L, double* W);
// Stores the descriptors of the current configuration with per-atom labels
// (original atom order) as training rows.
void dev_surrogate_add_sample(Device* d, EmDeeSurrogate* s, double L, const double* labels);
}  // namespace emdee
