// surrogate.h -- device entry points of the learned energy model
// (surrogate.hip) used by the C-ABI driver.
#pragma once
#include "engine.h"
#include "kspace.h"

struct EmDeeSurrogate;

namespace emdee {
// What the model needs of the system it runs on: the box, the list cutoff
// and skin (its compact list follows the engine's Verlet list), and for the
// long-range channel the system's Ewald set-up, splitting parameter and the
// layer whose kCoul weights the channel uses.
struct SurrogateContext {
  double L = 0, Rc = 0, skin = 0;
  const EwaldSetup* ewald = nullptr;
  double alpha = 0;
  int lambdaLayer = 1;
};
// Adds the model forces to F(layer) of the system (and its atom energies to
// the atom-energy array when that is on) and returns the model energy; W
// receives its virial.
double dev_surrogate_apply(Device* d, EmDeeSurrogate* s, int layer, const SurrogateContext& c, double* W);
// The same in two halves: queue the network and the low-k sum, then, after
// dev_wait_posts, read the model energy and its virial.
void dev_surrogate_launch(Device* d, EmDeeSurrogate* s, int layer, const SurrogateContext& c);
double dev_surrogate_result(Device* d, EmDeeSurrogate* s, double* W);
// The model's descriptor cutoff.
double surrogate_cutoff(const EmDeeSurrogate* s);
// Stores the descriptors of the current configuration with per-atom labels
// (original atom order) as training rows.
void dev_surrogate_add_sample(Device* d, EmDeeSurrogate* s, const SurrogateContext& c, const double* labels);
}  // namespace emdee
