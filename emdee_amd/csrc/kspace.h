// kspace.h -- host-side Ewald set-up (modelClass_kspace.f90, kspace_ewald.f90)
// and the device inputs it produces.
#pragma once
#include <functional>
#include <vector>

#include "engine.h"

namespace emdee {

// Everything the k-space kernels need, in the device's order.
//   * k-vectors in the reference order (kspace_ewald.f90:115-138): n3 runs
//     fastest, so consecutive vectors that differ only by n3 += 1 form a
//     "row" along which e^{ik.r} advances by one factor of e^{2 pi i z/L}.
//   * charged atoms in "slots": grouped by charged type (ascending type,
//     ascending atom index, modelClass_kspace.f90:130-150), each type padded
//     with empty slots (item -1, charge 0) to a multiple of EWALD_GROUP so
//     that every wavefront group holds one type only.
constexpr int EWALD_GROUP = 256;

struct EwaldSetup {
  int nm = 0, nvecs = 0, nt = 0, nl = 0, nslots = 0;
  double alpha = 0;            // Ewald splitting parameter
  std::vector<int> kn;         // [nvecs][4]: n1, n2, n3, row-start flag
  std::vector<double> kv;      // [nvecs][4]: 2 prefac k (x, y, z), prefac
  std::vector<int> slotItem;   // [nslots]: original atom index (0-based) or -1
  std::vector<double> slotQ;   // [nslots]
  std::vector<int> groupType;  // [nslots / EWALD_GROUP]: local charged type (0-based)
  std::vector<int> typeGroups; // [nt + 1]: group range of each charged type
  std::vector<double> lambda;  // [nl][nt][nt]: kCoul of the type pair in each layer
};

struct EwaldHost {
  double alpha = 0, beta = 0, kmax = 0, volume = 0;
  std::vector<double> rigidEnergy;   // per layer: sum over type pairs of lambda * Erigid
  EwaldSetup dev;
};

// modelClass_kspace.f90:111-261 with kspace_ewald.f90:82-184.
//   R: coordinates (3,N) in original order, used for the intrabody pairs;
//   kCoul(i, j, l): pair(i,j,l)%kCoul with 1-based types and layer.
EwaldHost ewald_initialize(double accuracy, double Rc, double L, int nlayers, const std::vector<int>& atomType,
                           const std::vector<double>& charge, const std::vector<int>& bStart,
                           const std::vector<int>& bOrig, const double* R,
                           const std::function<double(int, int, int)>& kCoul);

// math.f90:642-657
double inverse_of_x_plus_ln_x(double y);

// Device side (ewald.hip): set-up, and one evaluation that adds the k-space
// forces to F(layer) and returns the reciprocal-space energy plus the self
// energy sum (kspace_ewald.f90:188-314).
void dev_ewald_setup(Device* d, const EwaldSetup& s);
double dev_ewald_compute(Device* d, int layer, double L);
// The same in two halves: queue the sum (false: nothing to sum), then, once
// dev_wait_posts has returned, its energy.
bool dev_ewald_launch(Device* d, int layer, double L);
double dev_ewald_energy(const Device* d);

// A partial k set: the vectors of the same half space with |k| <= kcut, the
// charges and slots of `full` and its kCoul weights of `layer` (1-based).
EwaldSetup lowk_setup(const EwaldSetup& full, int layer, double alpha, double L, double kcut);
// Its device state (ewald.hip): forces added to F (slot order), atom shares
// to eatom (may be null), energy returned and its virial in *W.
struct EwaldState;
EwaldState* dev_lowk_create(const EwaldSetup& s);
double dev_lowk_compute(Device* d, EwaldState* e, double* F, double* eatom, double L, double* W);
// The same in two halves: queue (false: nothing to sum), then read after
// dev_wait_posts.
bool dev_lowk_launch(Device* d, EwaldState* e, double* F, double* eatom, double L);
double dev_lowk_result(const Device* d, double* W);
int dev_lowk_vectors(const EwaldState* e);
void dev_lowk_release(EwaldState* e);

// Bonded terms (bonded.hip): harmonic bonds and angles, EmDeeData.f90:443-550.
struct DevStruct {
  int i, j, k, kind;          // 0-based original atoms; kind of the model
  double half_k, minus_k, x0, pad;
};
struct DevDihedral {
  int i, j, k, l, kind, pad;  // 0-based original atoms; kind of the model
  double half_k, minus_k, x0, pad2;
};
void dev_structs_setup(Device* d, const std::vector<DevStruct>& bonds, const std::vector<DevStruct>& angles,
                       const std::vector<DevDihedral>& dihedrals);
// Adds the bonded forces (and, when kspace is on, the erf discounts of the
// excluded 1-2 and 1-3 pairs) to F(layer); out = {Ebond, Wbond, Eangle,
// Wangle, Ecoul discount, Edihedral, Wdihedral}.
void dev_structs_compute(Device* d, int layer, double L, int ntypes, bool bonded, bool kspace, double alpha,
                         double beta, double* out);

}  // namespace emdee
