// engine.h -- internal interface between the host-side engine (api.cpp,
// models.cpp, rng.cpp; plain C++) and the device layer (device.hip; HIP).
//
// The host side owns the system description (types, models, exclusions,
// rigid bodies) and follows the reference driver logic; the device layer owns
// every per-atom array in HBM and every kernel.  No HIP type crosses this
// header.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace emdee {

// Model kinds: one per reference model module (file stem in the comment).
enum ModelKind : int {
  PAIR_NONE = 1,          // modelClass_pair.f90 pair_none
  PAIR_LJ_CUT,            // pair_lj_cut.f90
  PAIR_SOFTCORE_CUT,      // pair_softcore_cut.f90
  COUL_NONE,              // modelClass_coul.f90 coul_none
  COUL_CUT,               // coul_cut.f90
  COUL_SF,                // coul_sf.f90
  COUL_DAMPED,            // coul_damped.f90
  COUL_LONG,              // coul_long.f90
  COUL_DAMPED_SMOOTHED,   // coul_damped_smoothed.f90
  COUL_DAMPED_SQUARE_SMOOTHED,   // coul_damped_square_smoothed.f90
  COUL_SQUARE_SMOOTHED,          // coul_square_smoothed.f90
  COUL_SHIFTED_SQUARE_SMOOTHED,  // coul_shifted_square_smoothed.f90
  BOND_NONE, BOND_HARMONIC,      // modelClass_bond.f90, bond_harmonic.f90
  ANGLE_NONE, ANGLE_HARMONIC,    // modelClass_angle.f90, angle_harmonic.f90
  DIHEDRAL_NONE,                 // modelClass_dihedral.f90
  DIHEDRAL_HARMONIC,             // harmonic dihedral (no reference model file; see models.cpp)
  KSPACE_EWALD                   // kspace_ewald.f90
};

// Modifier codes of modelClass_nonbonded.f90:27-33.
enum Modifier : int { MOD_NONE = 0, MOD_SHIFTED, MOD_SHIFTED_FORCE, MOD_SMOOTHED,
                      MOD_SHIFTED_SMOOTHED, MOD_SQUARE_SMOOTHED, MOD_SHIFTED_SQUARE_SMOOTHED };

// One model instance as handed out by the constructors (the opaque void*
// of the C ABI).  Plain data: copying it is "allocate(source=...)".
struct Model {
  int kind = 0;
  int modifier = MOD_NONE;
  // non-bonded common fields (modelClass_nonbonded.f90:36-50)
  double eshift = 0, fshift = 0, skin = 0, Rm = 0, RmSq = 0, factor = 0, Rm2fac = 0;
  // pair parameters
  double epsilon = 0, sigma = 0, lambda = 0;
  double c0 = 0, c1 = 0, c2 = 0, c3 = 0;   // lj: eps4, eps24, sigsq; softcore: prefactor, prefactor6, invSigSq, shift
  // coulomb
  bool shifted = false, shifted_force = false, requires_kspace = false;
  double alpha = 0, beta = 0, damp = 0, skinWidth = 0, Rm2 = 0, invRm = 0;
  // bonded
  double k = 0, x0 = 0, minus_k = 0, half_k = 0;
  // kspace
  double accuracy = 0;
};

// Device records of the pair loop: one per (itype, jtype) and one for the
// layer's Coulomb model.  Mirrors pairContainer (modelClass_pair.f90:66-74).
struct DevPair {
  int kind;        // PAIR_* kind
  int modifier;
  int coulomb;     // pair%coulomb
  int pad;
  double kCoul;
  double c0, c1, c2, c3;
  double eshift, fshift, factor, Rm2fac;
};
struct DevCoul {
  int kind;
  int modifier;
  int pad0, pad1;
  double alpha, beta, eshift, fshift, Rm, Rm2, invRm, factor, Rm2fac;
};

// Per-call pair-loop configuration.
struct PairCall {
  int layer;          // 0-based
  int ntypes;
  int energy;         // Options.Compute
  int mixed;          // FP32 pair arithmetic with FP64 accumulation
  double L, invL2, Rc2;   // box and scaled squared cutoff
  int scaled;         // scaled positions are current (written by the trigger or the list build)
  int defer = 0;      // return once the kernels are queued; dev_pairs_result reads the sums
};

// Sums returned by one force evaluation (indices of EmDeeCode.f90:1219).
struct ForceSums { double Epair, Ecoul, Wpair, Wcoul; };

// Opaque device state of one system.
struct Device;

struct DeviceSetup {
  int N, ntypes, nlayers;
  std::vector<int> type;       // 0-based types, original order
  std::vector<int> body;       // body id per atom (free atoms have unique ids)
  std::vector<double> invMass;
  std::vector<double> charge;
  std::vector<int> freeFlag;   // 1 for atoms that are not in a rigid body
};

Device* dev_create(const DeviceSetup& s);
void dev_destroy(Device* d);
void dev_set_charges(Device* d, const std::vector<double>& q);
void dev_set_exclusions(Device* d, const std::vector<int>& offset, const std::vector<int>& items);
void dev_set_interact(Device* d, const std::vector<uint8_t>& table);
void dev_set_models(Device* d, int layer, const std::vector<DevPair>& pairs, const DevCoul& coul);

// Host <-> device transfers of (3,N) arrays in original atom order.
void dev_upload_positions(Device* d, const double* R);
void dev_download_positions(Device* d, double* R);
void dev_upload_momenta(Device* d, const double* P);      // free atoms; bodies via dev_bodies_assign_momenta
void dev_download_momenta(Device* d, double* P);
void dev_upload_forces(Device* d, int layer, const double* F);
void dev_download_forces(Device* d, int layer, double* F);

// Neighbor lists.  The maximum_approach_sq rebuild trigger in two parts:
// queue its kernels, which also write the scaled positions of the current
// coordinates (and, when mixed, their fixed-point copy) for the pair kernels
// of the same evaluation; then wait for the answer (true: rebuild).
void dev_trigger_launch(Device* d, double L, bool mixed, int layer);
bool dev_trigger_result(Device* d, double skinSq);
// Whether the last steps' trigger values make a rebuild at this step unlikely.
bool dev_trigger_unlikely(const Device* d, double skinSq);
void dev_rebuild(Device* d, double L, double xRc, double xInRc);  // sort + full Verlet list
int dev_neighbor_pairs(Device* d, int max, int* out);             // half set, 1-based i<j

// Forces of one layer: fills F(layer) and returns the reduced sums.
void dev_compute_pairs(Device* d, const PairCall& c, ForceSums* out);
// Sums of the last dev_compute_pairs call made with defer set (waits for them).
void dev_pairs_result(Device* d, ForceSums* out);
// Adds bonded forces (computed on the host, original order) into F(layer).
void dev_add_forces(Device* d, int layer, const double* F);

// Integrator (free atoms).  twoKE != nullptr requests the kinetic sums.
void dev_boost(Device* d, int layer, double CP, double CF, bool translate, double* twoKE);
void dev_displace(Device* d, double CR, double CP);
void dev_kinetic(Device* d, double* twoKE);

// Rigid bodies (members in CSR order, original atom indices, 0-based).
enum BodyField { BF_MASS, BF_MOI, BF_INVMOI, BF_RCM, BF_PCM, BF_Q, BF_PI, BF_OMEGA, BF_F, BF_TAU };
void dev_bodies_setup(Device* d, const std::vector<int>& start, const std::vector<int>& orig,
                      const std::vector<double>& mMass, const std::vector<double>& bodyMass);
void dev_bodies_update(Device* d, double L);          // update_rigid_bodies (unwraps R, principal axes)
void dev_bodies_boost(Device* d, int layer, double CP, double CF, bool translate, bool rotate, double* twoKE6);
void dev_bodies_move(Device* d, double CR, double CP, double dt, bool translate, bool rotate, int mode);
double dev_bodies_virial(Device* d, int layer);
// The same in two halves: queue it, then read it after dev_wait_posts.
void dev_bodies_virial_launch(Device* d, int layer);
double dev_bodies_virial_read(const Device* d);
// Waits until every result posted so far has arrived in host memory.
void dev_wait_posts(Device* d);
void dev_bodies_assign_momenta(Device* d, double* twoKE6);   // from momenta staged by dev_upload_momenta
void dev_bodies_particle_momenta(Device* d);                 // into the staging buffer
void dev_bodies_place(Device* d);                            // member positions from rcm, q
void dev_bodies_get(Device* d, int field, double* host);     // host layout [nb][ncomp]
void dev_bodies_set(Device* d, int field, const double* host);

// Shadow terms of EmDee_verlet_step (EmDeeCode.f90:1107-1211): pre stores
// s0, r0, q0 from the current state; post returns {Ks_t, Ks_r, Us}.
void dev_shadow_pre(Device* d, int layer, double dt, int mode);
void dev_shadow_post(Device* d, int layer, double dt, int mode, double* out3);

// EmDee_rdf pair counts over the current lists: count[bin + bins*(symm1D-1)];
// pairOn is the (nt+1)^2 table of requested type pairs (1-based types).
void dev_rdf(Device* d, int bins, double Rc, double L, const std::vector<uint8_t>& pairOn, int npt,
             std::vector<long long>& count);
// EmDee_rdf into g (bins x pairs) with no allocation; limits checked by the caller.
constexpr int RDF_NOALLOC_MAX_COUNTS = 65536, RDF_NOALLOC_MAX_VALUES = 65536, RDF_NOALLOC_MAX_PAIRS = 64,
              RDF_NOALLOC_MAX_TYPE = 31;
void dev_rdf_noalloc(Device* d, int bins, double Rc, double L, const int* itype, const int* jtype, int pairs,
                     const double* NiNj, int npt, double norm, double* g);

// Atom energies: while on, force evaluations also write each atom's share of
// the potential energy (half of each of its pair terms, its k-space energy
// and self energy); get returns them in original atom order.
void dev_set_atom_energies(Device* d, bool on);
void dev_get_atom_energies(Device* d, double* host);

// Profiling (HIP events around the pair kernel).
void dev_set_profiling(Device* d, bool on);
void dev_kernel_stats(Device* d, double* out, int n);
void dev_sync(Device* d);
// Phase-space sharing: copies coordinates and momenta of every atom (each
// system in its own slot order) and the rigid-body state from src to dst.
void dev_copy_phase(Device* dst, Device* src);
void dev_bodies_copy(Device* dst, Device* src);

// Runtime errors: prints "Error in <task>: <msg>." and exits (global.f90:51-56).
[[noreturn]] void fatal(const std::string& task, const std::string& msg);
void warning(const std::string& msg);

}  // namespace emdee
