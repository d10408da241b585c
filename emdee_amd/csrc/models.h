// models.h -- host-side model set-up shared by the C-ABI driver.
#pragma once
#include "engine.h"

namespace emdee {

// pairContainer (modelClass_pair.f90:66-74)
struct PairEntry {
  Model model;
  bool coulomb = false;
  double kCoul = 0.0;
  PairEntry() { model.kind = PAIR_NONE; }
};

bool is_pair(const Model& m);
bool is_coul(const Model& m);
void setup_lj(Model& m, double eps, double sig);
void setup_softcore(Model& m, double eps, double sig, double lam);
void modifier_setup(Model& m, double cutoff);
void coul_cutoff_setup(Model& m, double cutoff);
void coul_kspace_setup(Model& m, double alpha);
PairEntry mix_pairs(const PairEntry& a, const PairEntry& b);
DevPair pack_pair(const PairEntry& p);
DevCoul pack_coul(const Model& m);

// KISS generator + ziggurat normals (math.f90:42-177).
struct Kiss {
  bool seeding_required = true;
  int32_t kn[128];
  double wn[128], fn[128];
  int32_t x = 0, y = 0, z = 0, w = 0;
  void setup(int32_t seed);
  int32_t i32();
  double normal();
};

}  // namespace emdee
