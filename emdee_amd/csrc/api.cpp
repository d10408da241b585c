// api.cpp -- the EmDee C ABI (include/emdee.h) over the HIP device layer.
//
// Each entry point follows the reference procedure cited beside it
// (src/EmDeeCode.f90): same argument meaning, 1-based indices, same error
// messages and exit(1) on misuse, same energy/virial bookkeeping in the host
// struct.  Per-atom state lives on the GPU (device.hip); this file keeps the
// system description and drives the kernels.
#include <algorithm>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/emdee.h"
#include "../../include/emdee_gpu.h"
#include "engine.h"
#include "kspace.h"
#include "models.h"
#include "surrogate.h"

using namespace emdee;

namespace {

const char* VERSION = "MI355X 1.0";
const double EPS = 2.220446049250313e-16;   // epsilon(1.0_rb)

double wtime() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct EnergyRec { double Potential, Dispersion, Coulomb, Bond, Angle, Dihedral, ShadowPotential; bool UpToDate; };
struct VirialRec { double Total, Body; };

struct System;

// Phase space shared between systems (EmDee_share_phase_space,
// EmDeeCode.f90:239-269).  The reference makes the systems alias one set of
// coordinates, momenta, rigid bodies and box.  Each system here keeps its
// atoms on the GPU in its own cell order, so the sharing is kept coherent by
// version: the last member to change the phase space is its writer, and any
// other member copies the writer's state into its own arrays before it next
// reads or changes it.  Forces, lists and energies stay with each system, as
// in the reference.
struct SharedPhase {
  System* writer = nullptr;
  unsigned long long version = 0;
  std::vector<System*> members;
};

// tData (EmDeeData.f90:66-151), host part.
struct System {
  int natoms = 0, ntypes = 1, nlayers = 1, layer = 1, nthreads = 1;
  int nbodies = 0, nfree = 0;
  bool pairsQueued = false;   // forces_begin queued the pair kernels for forces_end to read
  double Rc = 0, skin = 0, RcSq = 0, xRc = 0, xRcSq = 0, skinSq = 0, InRc = 0, InRcSq = 0, xInRcSq = 0;
  double totalMass = 0, startTime = 0;
  bool hasL = false, hasR = false, initialized = false;
  double L = 0;
  std::vector<int> atomType;            // 1-based types, [N]
  std::vector<int> atomBody;            // reference body ids, [N]
  std::vector<int> freeList;            // 1-based free atoms
  std::vector<int> bStart, bOrig;       // rigid bodies: member CSR, 0-based original indices
  std::vector<double> bMemberMass, bodyMass;
  std::vector<double> mass, invMass, charge;
  std::vector<char> charged;
  Kiss rng;
  std::vector<std::vector<int>> excluded;   // per atom, sorted, 1-based
  std::vector<PairEntry> pair;              // pair(ntypes, ntypes, nlayers)
  std::vector<Model> coul;                  // coul(nlayers)
  bool kspace_active = false;
  Model kspace;
  std::vector<char> multilayer, overridable, interact;
  std::vector<char> pairs_exist, bonded, useInRc, forcesUpToDate;   // 1-based layers
  std::vector<EnergyRec> layerEnergy;
  std::vector<VirialRec> layerVirial;
  bool multilayer_coulomb = false;
  int nbonds = 0, nangles = 0, ndihedrals = 0;
  std::vector<DevStruct> bonds, angles;     // structs.f90 lists, 0-based atoms
  std::vector<DevDihedral> dihedrals;
  EwaldHost ewald;
  EmDeeSurrogate* surrogate = nullptr;       // learned model added to one layer
  int surrogateLayer = 0;
  bool atomEnergies = false;                 // next force evaluation writes atom energies
  int precision = 0;
  Device* dev = nullptr;
  // EmDee_memory_address views: host arrays (coordinates, momenta, forces,
  // layerForces) kept equal to the engine's state around every call
  struct View { std::vector<double> data, shadow; bool armed = false; };
  View views[4];
  std::shared_ptr<SharedPhase> phase;        // set once the system shares its phase space
  unsigned long long phaseSeen = 0;          // version of the shared state this system holds

  PairEntry& P(int i, int j, int l) { return pair[(i - 1) + ntypes * ((j - 1) + ntypes * (l - 1))]; }
  char& T2(std::vector<char>& a, int i, int j) { return a[(i - 1) + ntypes * (j - 1)]; }
};

System* S(const tEmDee& md) {
  if (!md.Data) fatal("EmDee", "invalid system handle");
  return static_cast<System*>(md.Data);
}

bool ranged(std::initializer_list<int> v, int imax) {
  for (int i : v) if (!(i > 0 && i <= imax)) return false;
  return true;
}

double phi_fn(double x) {   // math.f90:230-237
  if (std::fabs(x) > 1e-4) return (1.0 - std::exp(-x)) / x;
  return 1.0 + 0.5 * x * ((1.0 / 3.0) * x * (1.0 - 0.25 * x) - 1.0);
}

// Brings the system's copy of a shared phase space up to date.
void phase_acquire(System* me) {
  if (!me->phase) return;
  SharedPhase& h = *me->phase;
  if (h.writer && h.writer != me && me->phaseSeen != h.version) {
    dev_copy_phase(me->dev, h.writer->dev);
    me->L = h.writer->L;
    me->phaseSeen = h.version;
  }
}

// Records that the system has changed the shared phase space.
void phase_publish(System* me) {
  if (!me->phase) return;
  me->phase->writer = me;
  me->phaseSeen = ++me->phase->version;
}

// Leaves a shared phase space; the others first take the latest state.
void phase_leave(System* me) {
  if (!me->phase) return;
  std::shared_ptr<SharedPhase> h = me->phase;
  auto& m = h->members;
  m.erase(std::remove(m.begin(), m.end(), me), m.end());
  if (h->writer == me) {
    for (System* o : m) phase_acquire(o);
    h->writer = m.empty() ? nullptr : m.front();
  }
  me->phase.reset();
  me->phaseSeen = 0;
}

// Keeps a shared phase space coherent around an entry point: the state is
// brought up to date on entry and, for a call that changes it, published on
// exit.
// EmDee_memory_address (EmDeeCode.f90:212-235) hands the host pointers into
// the reference's own arrays.  Here the arrays live on the GPU in cell order,
// so the pointers are host views: on entry to a call, values the host wrote
// through a view go to the engine (as writes through the reference's pointers
// do: nothing is marked stale); on exit, every view is refreshed.
enum { VIEW_R, VIEW_P, VIEW_F, VIEW_LAYER_F };
void view_fill(System* me, int k) {
  System::View& v = me->views[k];
  size_t n = 3 * (size_t)me->natoms;
  v.data.resize(k == VIEW_LAYER_F ? n * me->nlayers : n);
  if (k == VIEW_R) dev_download_positions(me->dev, v.data.data());
  else if (k == VIEW_P) dev_download_momenta(me->dev, v.data.data());
  else if (k == VIEW_F) dev_download_forces(me->dev, me->layer - 1, v.data.data());
  else for (int l = 0; l < me->nlayers; l++) dev_download_forces(me->dev, l, v.data.data() + n * l);
  v.shadow = v.data;
}
void views_push(System* me) {
  size_t n = 3 * (size_t)me->natoms;
  for (int k = 0; k < 4; k++) {
    System::View& v = me->views[k];
    if (!v.armed || v.data == v.shadow) continue;
    if (k == VIEW_R) dev_upload_positions(me->dev, v.data.data());
    else if (k == VIEW_P) dev_upload_momenta(me->dev, v.data.data());
    else if (k == VIEW_F) dev_upload_forces(me->dev, me->layer - 1, v.data.data());
    else for (int l = 0; l < me->nlayers; l++) dev_upload_forces(me->dev, l, v.data.data() + n * l);
    v.shadow = v.data;
  }
}
void views_pull(System* me) {
  for (int k = 0; k < 4; k++) if (me->views[k].armed) view_fill(me, k);
}

struct PhaseScope {
  System* me;
  bool writes;
  PhaseScope(System* s, bool w) : me(s), writes(w) {
    phase_acquire(me);
    if (me->initialized) views_push(me);
  }
  ~PhaseScope() {
    if (writes) phase_publish(me);
    if (me->initialized) views_pull(me);
  }
};

void mark_stale(System* me, tEmDee* md) {
  for (int l = 1; l <= me->nlayers; l++) { me->forcesUpToDate[l] = 0; me->layerEnergy[l - 1].UpToDate = false; }
  md->Energy.UpToDate = false;
}

// EmDeeData.f90:194-223
void check_actual_interactions(System* me) {
  int nt = me->ntypes, nl = me->nlayers;
  std::vector<char> neutral(nt + 1);
  std::vector<char> inter((size_t)nt * nt * nl, 0);
  for (int i = 1; i <= nt; i++) {
    int c = 0;
    for (int a = 0; a < me->natoms; a++) if (me->atomType[a] == i && me->charged[a]) c++;
    neutral[i] = c == 0;
  }
  for (int i = 1; i <= nt; i++)
    for (int j = 1; j <= i; j++) {
      bool any = false;
      for (int k = 1; k <= nl; k++) {
        const PairEntry& p = me->P(i, j, k);
        bool no_pair = p.model.kind == PAIR_NONE;
        bool no_coul = me->coul[k - 1].kind == COUL_NONE || !p.coulomb;
        bool coul_only = no_pair && !no_coul;
        bool inert = (no_pair && no_coul) || (coul_only && neutral[i] && neutral[j]);
        inter[(i - 1) + nt * ((j - 1) + nt * (k - 1))] = !inert;
        any = any || !inert;
      }
      me->T2(me->interact, i, j) = any;
      me->T2(me->interact, j, i) = any;
    }
  for (int k = 1; k <= nl; k++) {
    bool e = false;
    for (int i = 1; i <= nt; i++) for (int j = 1; j <= i; j++) e = e || inter[(i - 1) + nt * ((j - 1) + nt * (k - 1))];
    me->pairs_exist[k] = e;
  }
}

void upload_models(System* me) {
  int nt = me->ntypes;
  std::vector<uint8_t> inter((size_t)nt * nt);
  for (int i = 1; i <= nt; i++) for (int j = 1; j <= nt; j++) inter[(i - 1) * nt + (j - 1)] = me->T2(me->interact, i, j);
  dev_set_interact(me->dev, inter);
  for (int l = 1; l <= me->nlayers; l++) {
    std::vector<DevPair> t((size_t)nt * nt);
    for (int i = 1; i <= nt; i++) for (int j = 1; j <= nt; j++) t[(i - 1) * nt + (j - 1)] = pack_pair(me->P(i, j, l));
    dev_set_models(me->dev, l - 1, t, pack_coul(me->coul[l - 1]));
  }
  std::vector<int> off(me->natoms + 1, 0), items;
  for (int i = 0; i < me->natoms; i++) {
    for (int j : me->excluded[i]) items.push_back(j - 1);
    off[i + 1] = (int)items.size();
  }
  dev_set_exclusions(me->dev, off, items);
  dev_set_charges(me->dev, me->charge);
}

// EmDeeData.f90:359-416
void perform_initialization(System* me, tEmDee* md) {
  const char* task = "system initialization";
  if (me->nbodies > 0) dev_bodies_update(me->dev, me->L);
  int bodyDoF = 6 * me->nbodies;
  md->RotDoF = bodyDoF - 3 * me->nbodies;
  md->DoF = 3 * me->nfree + bodyDoF - 3;
  check_actual_interactions(me);
  bool any_required = false;
  double kspaceRc = 0;
  for (int l = 1; l <= me->nlayers; l++)
    if (me->coul[l - 1].requires_kspace) {
      double lrc = me->useInRc[l] ? me->InRc : me->Rc;
      if (!any_required) kspaceRc = lrc;
      else if (lrc != kspaceRc) fatal(task, "all layers with ewald-like coulomb models must have the same cutoff");
      any_required = true;
    }
  if (any_required != me->kspace_active) {
    if (me->kspace_active) me->kspace_active = false;
    else fatal(task, "a kspace solver is required, but has not been defined");
  }
  if (me->kspace_active) {
    std::vector<double> R(3 * (size_t)me->natoms);
    dev_download_positions(me->dev, R.data());
    auto kCoul = [me](int i, int j, int l) { return me->P(i, j, l).kCoul; };
    me->ewald = ewald_initialize(me->kspace.accuracy, kspaceRc, me->L, me->nlayers, me->atomType, me->charge,
                                 me->bStart, me->bOrig, R.data(), kCoul);
    dev_ewald_setup(me->dev, me->ewald.dev);
    for (int l = 1; l <= me->nlayers; l++)
      if (me->coul[l - 1].requires_kspace) coul_kspace_setup(me->coul[l - 1], me->ewald.alpha);
  }
  if (me->nbonds + me->nangles + me->ndihedrals > 0) dev_structs_setup(me->dev, me->bonds, me->angles, me->dihedrals);
  upload_models(me);
  me->initialized = true;
}

void switch_layer(System* me, tEmDee* md, int layer) {
  if (layer != me->layer) {
    if (layer < 1 || layer > me->nlayers) fatal("model layer switch", "selected layer is out of range");
    me->layer = layer;
    std::memcpy(&md->Energy, &me->layerEnergy[layer - 1], sizeof(EnergyRec));
    md->Virial = me->layerVirial[layer - 1].Total;
    md->BodyVirial = me->layerVirial[layer - 1].Body;
  }
}

void compute_forces(tEmDee* md);

// EmDeeCode.f90:914-923
void initialize_system(System* me, tEmDee* md) {
  perform_initialization(me, md);
  for (int layer = me->nlayers; layer >= 1; layer--) {
    switch_layer(me, md, layer);
    compute_forces(md);
  }
}

// EmDeeCode.f90:1215-1277, in two parts: forces_begin queues the list
// update and the pair kernels; forces_end adds the other terms and fills the
// energy and virial records.  With `defer`, and a layer of pair terms only,
// forces_begin returns true without waiting for the pair sums: the caller
// queues more work (the kick of EmDee_boost) and reads them later with one
// wait for both.
// What an attached learned model needs of its system (surrogate.h).
static SurrogateContext surrogate_context(System* me) {
  SurrogateContext c;
  c.L = me->L; c.Rc = me->Rc; c.skin = me->skin;
  if (me->kspace_active) {
    c.ewald = &me->ewald.dev;
    c.alpha = me->ewald.alpha;
    for (int l = me->nlayers; l >= 1; l--) if (me->coul[l - 1].requires_kspace) c.lambdaLayer = l;
  }
  return c;
}

static bool forces_begin(tEmDee* md, bool defer, ForceSums& fs) {
  System* me = S(*md);
  if (!me->initialized) fatal("force computation", "box and coordinates have not been defined");
  phase_acquire(me);
  dev_set_atom_energies(me->dev, me->atomEnergies);
  bool compute = md->Options.Compute || me->atomEnergies;
  int l = me->layer;
  PairCall c;
  c.layer = l - 1;
  c.ntypes = me->ntypes;
  c.energy = compute ? 1 : 0;
  c.mixed = me->precision;
  c.L = me->L;
  c.invL2 = 1.0 / (me->L * me->L);
  c.Rc2 = (me->useInRc[l] ? me->InRcSq : me->RcSq) * c.invL2;
  c.scaled = 1;
  fs = ForceSums{ 0, 0, 0, 0 };
  bool pairsOnly = me->nbonds + me->nangles + me->ndihedrals == 0 && !me->coul[l - 1].requires_kspace &&
                   !(me->surrogate && me->surrogateLayer == l) && me->nbodies == 0;
  // pairs only: the caller may read the pair sums later (after its kick).
  // Pairs with Ewald, a learned model or rigid bodies (no bonded terms):
  // forces_end queues their kernels behind the pair kernels and reads every
  // result after one wait.
  bool toCaller = defer && pairsOnly && me->pairs_exist[l];
  me->pairsQueued = !pairsOnly && me->pairs_exist[l] && me->nbonds + me->nangles + me->ndihedrals == 0;
  c.defer = toCaller || me->pairsQueued;
  // handle_neighbor_lists: neighbor_lists.f90:175-195.  When the recent
  // trigger values make a rebuild unlikely, the pair kernels are queued
  // behind the trigger kernels before the answer is read, so the GPU does
  // not wait for the host in between; after a rebuild they run again on the
  // new list, and their sums replace the first ones.
  md->Time.Neighbor -= wtime();
  dev_trigger_launch(me->dev, me->L, me->precision != 0, l - 1);
  bool ahead = me->pairs_exist[l] && dev_trigger_unlikely(me->dev, me->skinSq);
  if (ahead) {
    PairCall ca = c;
    ca.defer = 1;
    dev_compute_pairs(me->dev, ca, &fs);
  }
  bool rebuild = dev_trigger_result(me->dev, me->skinSq);
  if (rebuild) {
    dev_rebuild(me->dev, me->L, me->xRc, std::sqrt(me->xInRcSq));
    md->Builds += 1;
  }
  md->Time.Neighbor += wtime();
  md->Time.Pair -= wtime();
  if (me->pairs_exist[l]) {
    if (!ahead || rebuild) dev_compute_pairs(me->dev, c, &fs);
    else if (!c.defer) dev_pairs_result(me->dev, &fs);
  } else {
    std::vector<double> zero(3 * (size_t)me->natoms, 0.0);
    dev_upload_forces(me->dev, l - 1, zero.data());
  }
  return toCaller;
}

static void forces_end(tEmDee* md, const ForceSums& fs0) {
  System* me = S(*md);
  double E[6] = { 0, 0, 0, 0, 0, 0 }, W[6] = { 0, 0, 0, 0, 0, 0 };
  enum { PAIR_ = 0, COUL_ = 1, LONG_ = 2, BOND_ = 3, ANGLE_ = 4, DIHEDRAL_ = 5 };
  bool compute = md->Options.Compute || me->atomEnergies;
  int l = me->layer;
  bool kspace = me->coul[l - 1].requires_kspace;
  ForceSums fs = fs0;
  const bool queued = me->pairsQueued;
  bool ewaldQueued = false;
  if (queued) {
    // the pair kernels are queued: the k-space sum and the body virial go
    // behind them, and the three results are read after one wait
    me->pairsQueued = false;
    if (kspace) ewaldQueued = dev_ewald_launch(me->dev, l - 1, me->L);
    if (me->surrogate && me->surrogateLayer == l)
      dev_surrogate_launch(me->dev, me->surrogate, l - 1, surrogate_context(me));
    if (me->nbodies != 0) dev_bodies_virial_launch(me->dev, l - 1);
    dev_wait_posts(me->dev);
    dev_pairs_result(me->dev, &fs);
  }
  E[PAIR_] = fs.Epair; E[COUL_] = fs.Ecoul; W[PAIR_] = fs.Wpair; W[COUL_] = fs.Wcoul;
  // compute_bonds / compute_angles, EmDeeData.f90:443-550, and the dihedrals
  // (EmDeeData.f90:554-640, which the reference does not call)
  if (me->nbonds + me->nangles + me->ndihedrals > 0 && (me->bonded[l] || kspace)) {
    double o[7];
    dev_structs_compute(me->dev, l - 1, me->L, me->ntypes, me->bonded[l], kspace, me->ewald.alpha, me->ewald.beta, o);
    E[BOND_] = o[0]; W[BOND_] = o[1]; E[ANGLE_] = o[2]; W[ANGLE_] = o[3];
    E[COUL_] += o[4];
    E[DIHEDRAL_] = o[5]; W[DIHEDRAL_] = o[6];
  }
  // compute_kspace, EmDeeData.f90:689-700; virial EmDeeCode.f90:1248-1252
  if (kspace) {
    const double ek = queued ? (ewaldQueued ? dev_ewald_energy(me->dev) : 0.0) : dev_ewald_compute(me->dev, l - 1, me->L);
    E[LONG_] = ek + me->ewald.rigidEnergy[l - 1];
    W[LONG_] = E[COUL_] + E[LONG_] - W[COUL_];
  }
  // learned model standing in for a costly term of this layer (no reference counterpart)
  if (me->surrogate && me->surrogateLayer == l) {
    double Ws = 0, Es = queued ? dev_surrogate_result(me->dev, me->surrogate, &Ws)
                               : dev_surrogate_apply(me->dev, me->surrogate, l - 1, surrogate_context(me), &Ws);
    E[COUL_] += Es;
    W[COUL_] += Ws;
  }
  md->Virial = W[0] + W[1] + W[2] + W[3] + W[4] + W[5];
  if (me->nbodies != 0) {
    md->BodyVirial = queued ? dev_bodies_virial_read(me->dev) : dev_bodies_virial(me->dev, l - 1);
    md->Virial = md->Virial + md->BodyVirial;
  }
  if (compute) {
    md->Energy.Dispersion = E[PAIR_];
    md->Energy.Coulomb = E[COUL_] + E[LONG_];
    md->Energy.Bond = E[BOND_];
    md->Energy.Angle = E[ANGLE_];
    md->Energy.Dihedral = E[DIHEDRAL_];
    md->Energy.Potential = E[0] + E[1] + E[2] + E[3] + E[4] + E[5];
    md->Energy.ShadowPotential = md->Energy.Potential;
  }
  md->Energy.UpToDate = compute;
  me->forcesUpToDate[l] = 1;
  std::memcpy(&me->layerEnergy[l - 1], &md->Energy, sizeof(EnergyRec));
  me->layerVirial[l - 1].Total = md->Virial;
  me->layerVirial[l - 1].Body = md->BodyVirial;
  double t = wtime();
  md->Time.Pair += t;
  md->Time.Total = t - me->startTime;
}

void compute_forces(tEmDee* md) {
  ForceSums fs;
  forces_begin(md, false, fs);
  forces_end(md, fs);
}

void set_kinetic(tEmDee* md, const double* twoKEt, const double* twoKEr) {
  for (int x = 0; x < 3; x++) { md->Kinetic.RotPart[x] = 0.5 * twoKEr[x]; md->Kinetic.TransPart[x] = 0.5 * twoKEt[x]; }
  md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
  md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
  md->Kinetic.ShadowKinetic = md->Kinetic.Kinetic;
  md->Kinetic.ShadowRotational = md->Kinetic.Rotational;
  md->Kinetic.UpToDate = true;
}

// EmDeeData.f90:268-355 (body bookkeeping).  Bodies are numbered as
// clean_body_indices numbers them: ids that occur more than once, in order of
// first appearance; ids that occur once (and ids <= 0) mark free atoms.  The
// reference finds the rank of an id by a linear search over the ids seen so
// far; a hash map gives the same rank in O(1).
void allocate_rigid_bodies(System* me, const int* bodies) {
  int N = me->natoms;
  me->atomBody.assign(N, 0);
  if (bodies) {
    std::vector<int> index(N + 1, 0), saved(N + 1, 0), amount(N + 1, 0), first(N + 1, 0);
    std::unordered_map<int, int> rank;
    rank.reserve((size_t)N);
    int n = 0;
    for (int i = 1; i <= N; i++) {
      int ib = bodies[i - 1];
      if (ib > 0) {
        auto it = rank.find(ib);
        if (it == rank.end()) { n++; rank.emplace(ib, n); amount[n] = 1; saved[n] = ib; first[n] = i; index[i] = 0; }
        else { int j = it->second; amount[j]++; index[i] = j; index[first[j]] = j; }
      }
    }
    int i = 0;
    for (int j = 1; j <= n; j++) if (amount[j] > 1) { i++; saved[j] = i; }
    int nb = 0;
    for (int k = 1; k <= N; k++) { if (index[k] > 0) index[k] = saved[index[k]]; me->atomBody[k - 1] = index[k]; nb = std::max(nb, index[k]); }
    me->nbodies = nb;
    me->freeList.clear();
    for (int k = 1; k <= N; k++) if (me->atomBody[k - 1] == 0) me->freeList.push_back(k);
    // members of body b in increasing atom index (EmDeeData.f90:300-303)
    std::vector<std::vector<int>> members(nb + 1);
    for (int k = 1; k <= N; k++) if (me->atomBody[k - 1] > 0) members[me->atomBody[k - 1]].push_back(k - 1);
    me->bStart.assign(1, 0);
    for (int b = 1; b <= nb; b++) {
      double m = 0;
      for (int a : members[b]) { me->bOrig.push_back(a); me->bMemberMass.push_back(me->mass[a]); m += me->mass[a]; }
      me->bodyMass.push_back(m);
      me->bStart.push_back((int)me->bOrig.size());
    }
    i = nb;
    for (int k = 0; k < N; k++) if (me->atomBody[k] == 0) me->atomBody[k] = ++i;
  } else {
    me->nbodies = 0;
    me->freeList.resize(N);
    for (int k = 0; k < N; k++) { me->freeList[k] = k + 1; me->atomBody[k] = k + 1; }
  }
  me->nfree = (int)me->freeList.size();
}

void add_exclusion(System* me, int i, int j) {
  std::vector<int>& v = me->excluded[i - 1];
  auto it = std::lower_bound(v.begin(), v.end(), j);
  if (it != v.end() && *it == j) return;
  v.insert(it, j);
}

}  // namespace

extern "C" {

// EmDeeCode.f90:69-208
tEmDee EmDee_system(int threads, int layers, double rc, double skin, int N, int* types, double* masses, int* bodies) {
  if (!getenv("EMDEE_QUIET")) printf("EmDee (version: %s)\n", VERSION);
  tEmDee md;
  std::memset(&md, 0, sizeof md);
  System* me = new System();
  me->nthreads = threads < 1 ? 1 : threads;
  me->nlayers = layers;
  if (layers < 1) fatal("system setup", "invalid number of layers");
  me->Rc = rc; me->skin = skin; me->RcSq = rc * rc;
  me->xRc = rc + skin; me->xRcSq = me->xRc * me->xRc;
  me->InRc = me->Rc; me->InRcSq = me->RcSq; me->xInRcSq = me->xRcSq;
  me->skinSq = skin * skin;
  me->natoms = N;
  me->atomType.assign(N, 1);
  if (types) {
    int mn = N ? types[0] : 1, mx = N ? types[0] : 1;
    for (int i = 0; i < N; i++) { mn = std::min(mn, types[i]); mx = std::max(mx, types[i]); }
    if (mn != 1) fatal("system setup", "wrong specification of atom types");
    me->ntypes = mx;
    for (int i = 0; i < N; i++) me->atomType[i] = types[i];
  }
  me->mass.assign(N, 1.0);
  me->invMass.assign(N, 1.0);
  if (masses) {
    double tm = 0;
    for (int i = 0; i < N; i++) { double m = masses[me->atomType[i] - 1]; me->mass[i] = m; me->invMass[i] = 1.0 / m; tm += m; }
    me->totalMass = tm;
  } else {
    me->totalMass = (double)N;
  }
  me->startTime = wtime();
  me->charge.assign(N, 0.0);
  me->charged.assign(N, 0);
  allocate_rigid_bodies(me, bodies);
  me->excluded.assign(N, std::vector<int>());
  int nt = me->ntypes;
  me->pair.assign((size_t)nt * nt * layers, PairEntry());
  me->multilayer.assign(nt * nt, 0);
  me->overridable.assign(nt * nt, 1);
  me->interact.assign(nt * nt, 0);
  me->coul.assign(layers, Model());
  for (auto& c : me->coul) c.kind = COUL_NONE;
  me->pairs_exist.assign(layers + 1, 0);
  me->useInRc.assign(layers + 1, 0);
  me->bonded.assign(layers + 1, 1);
  me->forcesUpToDate.assign(layers + 1, 0);
  me->layerEnergy.assign(layers, EnergyRec{ 0, 0, 0, 0, 0, 0, 0, false });
  me->layerVirial.assign(layers, VirialRec{ 0, 0 });
  const char* prec = getenv("EMDEE_PRECISION");
  me->precision = (prec && std::string(prec) == "mixed") ? 1 : 0;
  // device state
  DeviceSetup ds;
  ds.N = N; ds.ntypes = nt; ds.nlayers = layers;
  ds.type.resize(N); ds.body.resize(N); ds.invMass = me->invMass; ds.charge = me->charge; ds.freeFlag.resize(N);
  for (int i = 0; i < N; i++) { ds.type[i] = me->atomType[i] - 1; ds.body[i] = me->atomBody[i]; ds.freeFlag[i] = 0; }
  for (int k : me->freeList) ds.freeFlag[k - 1] = 1;
  me->dev = dev_create(ds);
  if (me->nbodies > 0) dev_bodies_setup(me->dev, me->bStart, me->bOrig, me->bMemberMass, me->bodyMass);
  md.Builds = 0;
  md.DoF = 3 * (N - 1);
  md.RotDoF = 0;
  md.Data = me;
  md.Options.Translate = true;
  md.Options.Rotate = true;
  md.Options.RotationMode = 0;
  md.Options.AutoBodyUpdate = true;
  md.Options.Compute = true;
  return md;
}

// EmDeeCode.f90:212-235.  The address is a host view of the engine's array
// (see views_push/views_pull): it stays valid and current for the life of the
// system, and values written through it reach the engine at the next call.
void* EmDee_memory_address(tEmDee md, char* option) {
  System* me = S(md);
  std::string item(option);
  int k = item == "coordinates" ? VIEW_R : item == "momenta" ? VIEW_P : item == "forces" ? VIEW_F
        : item == "layerForces" ? VIEW_LAYER_F : -1;
  if (k < 0) fatal("memory address retrieving", "invalid option " + item);
  PhaseScope ps(me, false);
  if (!me->views[k].armed) {
    me->views[k].armed = true;
    if (me->initialized) view_fill(me, k);
    else me->views[k].data.assign(3 * (size_t)me->natoms * (k == VIEW_LAYER_F ? me->nlayers : 1), 0.0);
  }
  return me->views[k].data.data();
}

// EmDeeCode.f90:239-269
void EmDee_share_phase_space(tEmDee mdkeep, tEmDee* mdlose) {
  const char* task = "phase space sharing";
  System *keep = S(mdkeep), *lose = S(*mdlose);
  if (!(keep->initialized && lose->initialized)) fatal(task, "EmDee system 1 has not been initialized");
  if (keep->natoms != lose->natoms) fatal(task, "different numbers of atoms");
  if (keep->atomType != lose->atomType) fatal(task, "atom types do not match");
  if (keep->mass != lose->mass) fatal(task, "atom masses do not match");
  if (keep->atomBody != lose->atomBody) fatal(task, "rigid bodies do not match");
  if (keep == lose) return;
  phase_acquire(keep);
  if (!keep->phase) {
    keep->phase = std::make_shared<SharedPhase>();
    keep->phase->members.push_back(keep);
    phase_publish(keep);
  }
  if (lose->phase != keep->phase) {
    phase_leave(lose);
    lose->phase = keep->phase;
    lose->phase->members.push_back(lose);
  }
  // lose now holds the shared state: copy it in place of the pointer swap
  lose->phaseSeen = 0;
  keep->phase->writer = keep;
  phase_acquire(lose);
  mdlose->Kinetic.Kinetic = mdkeep.Kinetic.Kinetic;
  for (int k = 0; k < 3; k++) {
    mdlose->Kinetic.TransPart[k] = mdkeep.Kinetic.TransPart[k];
    mdlose->Kinetic.RotPart[k] = mdkeep.Kinetic.RotPart[k];
  }
  mdlose->Kinetic.Rotational = mdkeep.Kinetic.Rotational;
}

// EmDeeCode.f90:273-305
void EmDee_layer_based_parameters(tEmDee md, double InternalRc, int* Apply, int* Bonded) {
  const char* task = "layer-based parameter setting";
  System* me = S(md);
  if (me->initialized) fatal(task, "system has already been initialized");
  bool any = false;
  for (int l = 1; l <= me->nlayers; l++) { me->useInRc[l] = Apply[l - 1] != 0; any = any || me->useInRc[l]; }
  if (any && (InternalRc <= 0.0 || InternalRc > me->Rc)) fatal(task, "invalid internal cutoff specification");
  me->InRc = InternalRc;
  me->InRcSq = InternalRc * InternalRc;
  me->xInRcSq = (InternalRc + me->skin) * (InternalRc + me->skin);
  for (int l = 1; l <= me->nlayers; l++) me->bonded[l] = Bonded[l - 1] != 0;
  for (int l = 1; l <= me->nlayers; l++) coul_cutoff_setup(me->coul[l - 1], me->useInRc[l] ? me->InRc : me->Rc);
}

}  // extern "C"

// EmDeeData.f90:227-264
static void set_pair_type(System* me, int itype, int jtype, int layer, const Model& model, double kCoul) {
  double layerRc = me->useInRc[layer] ? me->InRc : me->Rc;
  if (itype == jtype) {
    PairEntry& p = me->P(itype, itype, layer);
    p.model = model;
    p.coulomb = kCoul != 0.0;
    if (p.coulomb) p.kCoul = kCoul;
    modifier_setup(p.model, layerRc);
    for (int k = 1; k <= me->ntypes; k++)
      if (k != itype && me->T2(me->overridable, itype, k)) {
        me->P(itype, k, layer) = mix_pairs(me->P(k, k, layer), me->P(itype, itype, layer));
        modifier_setup(me->P(itype, k, layer).model, layerRc);
        me->P(k, itype, layer) = me->P(itype, k, layer);
      }
  } else {
    PairEntry& p = me->P(itype, jtype, layer);
    p.model = model;
    p.coulomb = kCoul != 0.0;
    if (p.coulomb) p.kCoul = kCoul;
    modifier_setup(p.model, layerRc);
    me->P(jtype, itype, layer) = p;
  }
}

extern "C" {

// EmDeeCode.f90:309-359
void EmDee_set_pair_model(tEmDee md, int itype, int jtype, void* model, double kCoul) {
  const char* task = "pair model setup";
  System* me = S(md);
  if (me->initialized) fatal(task, "cannot set model after coordinates have been defined");
  if (!ranged({ itype, jtype }, me->ntypes)) fatal(task, "provided type index is out of range");
  if (!model) fatal(task, "a valid pair model must be provided");
  const Model& m = *static_cast<Model*>(model);
  if (!is_pair(m)) fatal(task, "a valid pair model must be provided");
  for (int l = 1; l <= me->nlayers; l++) set_pair_type(me, itype, jtype, l, m, kCoul);
  me->T2(me->multilayer, itype, jtype) = 0;
  me->T2(me->multilayer, jtype, itype) = 0;
  if (itype == jtype) {
    for (int k = 1; k <= me->ntypes; k++)
      if (k != itype && me->T2(me->overridable, itype, k)) {
        me->T2(me->multilayer, itype, k) = me->T2(me->multilayer, k, k);
        me->T2(me->multilayer, k, itype) = me->T2(me->multilayer, k, k);
      }
  } else {
    me->T2(me->overridable, itype, jtype) = 0;
    me->T2(me->overridable, jtype, itype) = 0;
  }
}

// EmDeeCode.f90:363-412
void EmDee_set_pair_multimodel(tEmDee md, int itype, int jtype, void* model[], double kCoul[]) {
  const char* task = "pair multimodel setup";
  System* me = S(md);
  if (me->initialized) fatal(task, "cannot set model after coordinates have been defined");
  if (!ranged({ itype, jtype }, me->ntypes)) fatal(task, "provided type index is out of range");
  for (int l = 1; l <= me->nlayers; l++) {
    if (!model[l - 1]) fatal(task, std::to_string(me->nlayers) + " valid pair models must be provided");
    const Model& m = *static_cast<Model*>(model[l - 1]);
    if (!is_pair(m)) fatal(task, "a valid pair model must be provided");
    set_pair_type(me, itype, jtype, l, m, kCoul[l - 1]);
  }
  me->T2(me->multilayer, itype, jtype) = 1;
  me->T2(me->multilayer, jtype, itype) = 1;
  if (itype == jtype) {
    for (int k = 1; k <= me->ntypes; k++)
      if (k != itype && me->T2(me->overridable, itype, k)) { me->T2(me->multilayer, itype, k) = 1; me->T2(me->multilayer, k, itype) = 1; }
  } else {
    me->T2(me->overridable, itype, jtype) = 0;
    me->T2(me->overridable, jtype, itype) = 0;
  }
}

// EmDeeCode.f90:416-444
void EmDee_set_kspace_model(tEmDee md, void* model) {
  const char* task = "kspace model setup";
  System* me = S(md);
  if (me->initialized) fatal(task, "cannot set model after coordinates have been defined");
  if (!model) fatal(task, "a valid kspace model must be provided");
  const Model& m = *static_cast<Model*>(model);
  if (m.kind != KSPACE_EWALD) fatal(task, "a valid kspace model must be provided");
  me->kspace = m;
  me->kspace_active = true;
}

// EmDeeCode.f90:448-482
void EmDee_set_coul_model(tEmDee md, void* model) {
  const char* task = "coulomb model setup";
  System* me = S(md);
  if (me->initialized) fatal(task, "cannot set model after coordinates have been defined");
  if (!model) fatal(task, "a valid coulomb model must be provided");
  const Model& m = *static_cast<Model*>(model);
  if (!is_coul(m)) fatal(task, "a valid coulomb model must be provided");
  for (int l = 1; l <= me->nlayers; l++) {
    double layerRc = me->useInRc[l] ? me->InRc : me->Rc;
    me->coul[l - 1] = m;
    coul_cutoff_setup(me->coul[l - 1], layerRc);
    modifier_setup(me->coul[l - 1], layerRc);
  }
  me->multilayer_coulomb = false;
}

// EmDeeCode.f90:486-520
void EmDee_set_coul_multimodel(tEmDee md, void* model[]) {
  const char* task = "coulomb multimodel setup";
  System* me = S(md);
  if (me->initialized) fatal(task, "cannot set model after coordinates have been defined");
  for (int l = 1; l <= me->nlayers; l++) {
    double layerRc = me->useInRc[l] ? me->InRc : me->Rc;
    if (!model[l - 1]) fatal(task, std::to_string(me->nlayers) + " valid coulomb models must be provided");
    const Model& m = *static_cast<Model*>(model[l - 1]);
    if (!is_coul(m)) fatal(task, "a valid coulomb model must be provided");
    me->coul[l - 1] = m;
    coul_cutoff_setup(me->coul[l - 1], layerRc);
    modifier_setup(me->coul[l - 1], layerRc);
  }
  me->multilayer_coulomb = true;
}

// EmDeeCode.f90:524-570
void EmDee_ignore_pair(tEmDee md, int i, int j) {
  System* me = S(md);
  if (i != j && ranged({ i, j }, me->natoms)) { add_exclusion(me, i, j); add_exclusion(me, j, i); }
}

// EmDeeCode.f90:574-597
void EmDee_add_bond(tEmDee md, int i, int j, void* model) {
  System* me = S(md);
  if (!ranged({ i, j }, me->natoms)) fatal("add_bond", "atom index out of range");
  if (!model) fatal("add_bond", "a valid model must be provided");
  const Model& m = *static_cast<Model*>(model);
  if (m.kind != BOND_HARMONIC && m.kind != BOND_NONE) fatal("add_bond", "the provided model must be a bond model");
  me->nbonds++;
  me->bonds.push_back(DevStruct{ i - 1, j - 1, -1, m.kind, m.half_k, m.minus_k, m.x0, 0.0 });
  EmDee_ignore_pair(md, i, j);
}

// EmDeeCode.f90:600-626
void EmDee_add_angle(tEmDee md, int i, int j, int k, void* model) {
  System* me = S(md);
  if (!ranged({ i, j, k }, me->natoms)) fatal("add_angle", "atom index out of range");
  if (!model) fatal("add_angle", "a valid model must be provided");
  const Model& m = *static_cast<Model*>(model);
  if (m.kind != ANGLE_HARMONIC && m.kind != ANGLE_NONE) fatal("add_angle", "the provided model must be an angle model");
  me->nangles++;
  me->angles.push_back(DevStruct{ i - 1, j - 1, k - 1, m.kind, m.half_k, m.minus_k, m.x0, 0.0 });
  EmDee_ignore_pair(md, i, j);
  EmDee_ignore_pair(md, i, k);
  EmDee_ignore_pair(md, j, k);
}

// EmDeeCode.f90:628-655
void EmDee_add_dihedral(tEmDee md, int i, int j, int k, int l, void* model) {
  System* me = S(md);
  if (!ranged({ i, j, k, l }, me->natoms)) fatal("add_dihedral", "atom index out of range");
  if (!model) fatal("add_dihedral", "a valid model must be provided");
  const Model& m = *static_cast<Model*>(model);
  if (m.kind != DIHEDRAL_NONE && m.kind != DIHEDRAL_HARMONIC)
    fatal("add_dihedral", "the provided model must be a dihedral model");
  me->ndihedrals++;
  me->dihedrals.push_back(DevDihedral{ i - 1, j - 1, k - 1, l - 1, m.kind, 0, m.half_k, m.minus_k, m.x0, 0.0 });
  EmDee_ignore_pair(md, i, j); EmDee_ignore_pair(md, i, k); EmDee_ignore_pair(md, i, l);
  EmDee_ignore_pair(md, j, k); EmDee_ignore_pair(md, j, l); EmDee_ignore_pair(md, k, l);
}

// EmDeeCode.f90:659-801
void EmDee_download(tEmDee md, char* option, double* address) {
  System* me = S(md);
  PhaseScope ps(me, false);
  std::string item(option);
  if (!address) fatal("download", "provided address is invalid");
  int nb = me->nbodies;
  if (item == "box") *address = me->L;
  else if (item == "coordinates") {
    if (!me->hasR) fatal("download", "coordinates have not been allocated");
    dev_download_positions(me->dev, address);
  } else if (item == "momenta") dev_download_momenta(me->dev, address);
  else if (item == "forces") {
    if (!me->forcesUpToDate[me->layer]) compute_forces(&md);
    dev_download_forces(me->dev, me->layer - 1, address);
  } else if (item == "centersOfMass") {
    std::vector<double> R(3 * (size_t)me->natoms);
    dev_download_positions(me->dev, R.data());
    dev_bodies_get(me->dev, BF_RCM, address);
    for (int j = 0; j < me->nfree; j++) for (int x = 0; x < 3; x++) address[3 * (j + nb) + x] = R[3 * (me->freeList[j] - 1) + x];
  } else if (item == "quaternions") dev_bodies_get(me->dev, BF_Q, address);
  else if (item == "quatmom") dev_bodies_get(me->dev, BF_PI, address);
  else if (item == "quattau") {
    // C(q) (2 tau): EmDeeCode.f90:768
    std::vector<double> q(4 * (size_t)nb), tau(3 * (size_t)nb);
    dev_bodies_get(me->dev, BF_Q, q.data());
    dev_bodies_get(me->dev, BF_TAU, tau.data());
    for (int b = 0; b < nb; b++) {
      const double* Q = &q[4 * b];
      double t[3] = { 2.0 * tau[3 * b], 2.0 * tau[3 * b + 1], 2.0 * tau[3 * b + 2] };
      double C[12] = { -Q[1], -Q[2], -Q[3], Q[0], Q[3], -Q[2], -Q[3], Q[0], Q[1], Q[2], -Q[1], Q[0] };
      for (int r = 0; r < 4; r++) address[4 * b + r] = C[3 * r] * t[0] + C[3 * r + 1] * t[1] + C[3 * r + 2] * t[2];
    }
  } else if (item == "angmom") dev_bodies_get(me->dev, BF_OMEGA, address);
  else if (item == "bodycoord") dev_bodies_get(me->dev, BF_RCM, address);
  else if (item == "bodymom") dev_bodies_get(me->dev, BF_PCM, address);
  else if (item == "bodyforces") dev_bodies_get(me->dev, BF_F, address);
  else if (item == "torques") dev_bodies_get(me->dev, BF_TAU, address);
  else if (item == "inertia") dev_bodies_get(me->dev, BF_MOI, address);
  else fatal("download", "invalid option");
}

// EmDeeCode.f90:805-925
void EmDee_upload(tEmDee* md, char* option, double* address) {
  System* me = S(*md);
  PhaseScope ps(me, true);
  std::string item(option);
  if (!address) fatal("upload", "provided address is invalid");
  if (item == "box") {
    me->L = *address;
    me->hasL = true;
    if (me->initialized) mark_stale(me, md);
    else if (me->hasR) initialize_system(me, md);
  } else if (item == "coordinates") {
    dev_upload_positions(me->dev, address);
    me->hasR = true;
    if (me->initialized) {
      mark_stale(me, md);
      if (md->Options.AutoBodyUpdate && me->nbodies > 0) dev_bodies_update(me->dev, me->L);
    } else if (me->hasL) initialize_system(me, md);
  } else if (item == "momenta") {
    if (!me->initialized) fatal("upload", "box and coordinates have not been defined");
    dev_upload_momenta(me->dev, address);
    double kb[6];
    dev_bodies_assign_momenta(me->dev, kb);
    // assign_momenta, EmDeeData.f90:157-189: free atoms, then bodies
    double t[3] = { 0, 0, 0 }, r[3];
    for (int j : me->freeList)
      for (int x = 0; x < 3; x++) t[x] += me->invMass[j - 1] * (address[3 * (j - 1) + x] * address[3 * (j - 1) + x]);
    for (int x = 0; x < 3; x++) { t[x] += kb[x]; r[x] = kb[3 + x]; }
    set_kinetic(md, t, r);
  } else if (item == "forces") {
    if (!me->initialized) fatal("upload", "box and coordinates have not been defined");
    dev_upload_forces(me->dev, me->layer - 1, address);
  } else if (item == "charges") {
    if (me->initialized) fatal("upload", "cannot set charges after box and coordinates initialization");
    for (int i = 0; i < me->natoms; i++) { me->charge[i] = address[i]; me->charged[i] = std::fabs(address[i]) > EPS; }
    mark_stale(me, md);
  } else fatal("upload", "invalid option");
}

// EmDeeCode.f90:929-946
void EmDee_switch_model_layer(tEmDee* md, int layer) {
  PhaseScope ps(S(*md), false);
  switch_layer(S(*md), md, layer);
}

// pi = B(q) (2 I omega): tBody_assign_momenta, ArBee.f90:356-362
static void body_pi_from_omega(const double* q, const double* MoI, const double* w, double* pi) {
  double t[3] = { 2.0 * MoI[0] * w[0], 2.0 * MoI[1] * w[1], 2.0 * MoI[2] * w[2] };
  double B[12] = { -q[1], -q[2], -q[3], q[0], -q[3], q[2], q[3], q[0], -q[1], -q[2], q[1], q[0] };
  for (int r = 0; r < 4; r++) pi[r] = B[3 * r] * t[0] + B[3 * r + 1] * t[1] + B[3 * r + 2] * t[2];
}

// EmDeeCode.f90:950-1020.  The generator is serial (KISS + ziggurat), so the
// draws run on the host in the reference order -- bodies (3 normals for the
// centre-of-mass momentum, 3 for the angular velocity), then free atoms --
// and the results are uploaded.
void EmDee_random_momenta(tEmDee* md, double kT, bool adjust, int seed) {
  System* me = S(*md);
  PhaseScope ps(me, true);
  if (me->rng.seeding_required) me->rng.setup(seed);
  double twoKEt[3] = { 0, 0, 0 }, twoKEr[3] = { 0, 0, 0 };
  int N = me->natoms, nb = me->nbodies;
  std::vector<double> P(3 * (size_t)N, 0.0);
  std::vector<double> q(4 * (size_t)nb), MoI(3 * (size_t)nb), invMoI(3 * (size_t)nb), pcm(3 * (size_t)nb),
      omega(3 * (size_t)nb), pi(4 * (size_t)nb);
  if (nb != 0) {
    if (!me->initialized) fatal("random_momenta", "coordinates have not defined");
    dev_bodies_get(me->dev, BF_Q, q.data());
    dev_bodies_get(me->dev, BF_MOI, MoI.data());
    dev_bodies_get(me->dev, BF_INVMOI, invMoI.data());
    for (int b = 0; b < nb; b++) {
      double f = std::sqrt(me->bodyMass[b] * kT);
      for (int x = 0; x < 3; x++) pcm[3 * b + x] = f * me->rng.normal();
      for (int x = 0; x < 3; x++) omega[3 * b + x] = std::sqrt(invMoI[3 * b + x] * kT) * me->rng.normal();
      body_pi_from_omega(&q[4 * b], &MoI[3 * b], &omega[3 * b], &pi[4 * b]);
      double im = 1.0 / me->bodyMass[b];
      for (int x = 0; x < 3; x++) {
        twoKEt[x] += im * (pcm[3 * b + x] * pcm[3 * b + x]);
        twoKEr[x] += MoI[3 * b + x] * omega[3 * b + x] * omega[3 * b + x];
      }
    }
  }
  for (int i : me->freeList) {
    double f = std::sqrt(me->mass[i - 1] * kT);
    for (int x = 0; x < 3; x++) P[3 * (i - 1) + x] = f * me->rng.normal();
    for (int x = 0; x < 3; x++) twoKEt[x] += me->invMass[i - 1] * (P[3 * (i - 1) + x] * P[3 * (i - 1) + x]);
  }
  if (adjust) {   // adjust_momenta, EmDeeCode.f90:984-1018
    double vcm[3];
    for (int x = 0; x < 3; x++) {
      double s1 = 0, s2 = 0;
      for (int i : me->freeList) s1 += P[3 * (i - 1) + x];
      for (int b = 0; b < nb; b++) s2 += pcm[3 * b + x];
      vcm[x] = (s1 + s2) / me->totalMass;
    }
    for (int x = 0; x < 3; x++) twoKEt[x] = 0;
    for (int i : me->freeList)
      for (int x = 0; x < 3; x++) {
        P[3 * (i - 1) + x] = P[3 * (i - 1) + x] - me->mass[i - 1] * vcm[x];
        twoKEt[x] += me->invMass[i - 1] * (P[3 * (i - 1) + x] * P[3 * (i - 1) + x]);
      }
    for (int b = 0; b < nb; b++)
      for (int x = 0; x < 3; x++) {
        pcm[3 * b + x] = pcm[3 * b + x] - me->bodyMass[b] * vcm[x];
        twoKEt[x] += (1.0 / me->bodyMass[b]) * (pcm[3 * b + x] * pcm[3 * b + x]);
      }
    int dofsum = 6 * nb;
    double tot = (twoKEt[0] + twoKEr[0]) + (twoKEt[1] + twoKEr[1]) + (twoKEt[2] + twoKEr[2]);
    double factor = std::sqrt((double)(3 * me->nfree + dofsum - 3) * kT / tot);
    for (int i : me->freeList) for (int x = 0; x < 3; x++) P[3 * (i - 1) + x] *= factor;
    for (int b = 0; b < nb; b++) {
      double w[3];
      for (int x = 0; x < 3; x++) { pcm[3 * b + x] = factor * pcm[3 * b + x]; w[x] = factor * omega[3 * b + x]; }
      for (int x = 0; x < 3; x++) omega[3 * b + x] = w[x];
      body_pi_from_omega(&q[4 * b], &MoI[3 * b], w, &pi[4 * b]);
    }
    for (int x = 0; x < 3; x++) { twoKEt[x] = factor * factor * twoKEt[x]; twoKEr[x] = factor * factor * twoKEr[x]; }
  }
  dev_upload_momenta(me->dev, P.data());
  if (nb != 0) {
    dev_bodies_set(me->dev, BF_PCM, pcm.data());
    dev_bodies_set(me->dev, BF_OMEGA, omega.data());
    dev_bodies_set(me->dev, BF_PI, pi.data());
  }
  set_kinetic(md, twoKEt, twoKEr);
}

// EmDeeCode.f90:1024-1065
void EmDee_boost(tEmDee* md, double lambda, double alpha, double dt) {
  System* me = S(*md);
  PhaseScope ps(me, true);
  double CF = phi_fn(alpha * dt) * dt;
  double CP = 1.0 - alpha * CF;
  CF = lambda * CF;
  // a force evaluation that is due is queued before the kick, and its sums
  // are read after the kick's kinetic energy (one wait for both)
  ForceSums fs;
  bool pending = false;
  if (lambda != 0.0 && !me->forcesUpToDate[me->layer]) {
    pending = forces_begin(md, true, fs);
    if (!pending) forces_end(md, fs);
  }
  bool compute = md->Options.Compute, translate = md->Options.Translate, rotate = md->Options.Rotate;
  double twoKEt[3] = { 0, 0, 0 }, kb[6] = { 0, 0, 0, 0, 0, 0 };
  dev_bodies_boost(me->dev, me->layer - 1, CP, CF, translate, rotate, compute ? kb : nullptr);
  dev_boost(me->dev, me->layer - 1, CP, CF, translate, compute ? twoKEt : nullptr);
  if (pending) {
    dev_pairs_result(me->dev, &fs);
    forces_end(md, fs);
  }
  if (compute) {
    if (translate) for (int x = 0; x < 3; x++) md->Kinetic.TransPart[x] = 0.5 * (kb[x] + twoKEt[x]);
    if (rotate) {
      for (int x = 0; x < 3; x++) md->Kinetic.RotPart[x] = 0.5 * kb[3 + x];
      md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
    }
    md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
  }
  md->Kinetic.UpToDate = compute;
}

// EmDeeCode.f90:1069-1103
static void displace(tEmDee* md, double lambda, double alpha, double dt) {
  System* me = S(*md);
  PhaseScope ps(me, true);
  md->Time.Motion -= wtime();
  double CR, CP;
  if (alpha != 0.0) {
    CP = phi_fn(alpha * dt) * dt;
    CR = 1.0 - alpha * CP;
    me->L = CR * me->L;
  } else {
    CP = dt;
    CR = 1.0;
  }
  CP = lambda * CP;
  dev_bodies_move(me->dev, CR, CP, dt, md->Options.Translate, md->Options.Rotate, md->Options.RotationMode);
  if (md->Options.Translate) dev_displace(me->dev, CR, CP);
  mark_stale(me, md);
  md->Time.Motion += wtime();
}
void EmDee_displace(tEmDee* md, double lambda, double alpha, double dt) { displace(md, lambda, alpha, dt); }
void EmDee_move(tEmDee* md, double lambda, double alpha, double dt) { displace(md, lambda, alpha, dt); }

// EmDeeCode.f90:1107-1211
void EmDee_verlet_step(tEmDee* md, double dt) {
  System* me = S(*md);
  PhaseScope ps(me, true);
  double dt_2 = 0.5 * dt;
  bool compute = md->Options.Compute;
  int mode = md->Options.RotationMode, l = me->layer;
  // pre_force: shadow bookkeeping, boost by dt/2, move by dt (translation and
  // rotation regardless of the options, as the reference)
  if (compute) dev_shadow_pre(me->dev, l - 1, dt, mode);
  dev_bodies_boost(me->dev, l - 1, 1.0, dt_2, true, true, nullptr);
  dev_boost(me->dev, l - 1, 1.0, dt_2, true, nullptr);
  dev_bodies_move(me->dev, 1.0, dt, dt, true, true, mode);
  dev_displace(me->dev, 1.0, dt);
  compute_forces(md);
  // post_force
  double twoKEt[3] = { 0, 0, 0 }, kb[6] = { 0, 0, 0, 0, 0, 0 };
  dev_bodies_boost(me->dev, l - 1, 1.0, dt_2, true, true, compute ? kb : nullptr);
  dev_boost(me->dev, l - 1, 1.0, dt_2, true, compute ? twoKEt : nullptr);
  if (compute) {
    double sums[3];
    dev_shadow_post(me->dev, l - 1, dt, mode, sums);
    double Ks_t = sums[0], Ks_r = sums[1], Us = sums[2];
    for (int x = 0; x < 3; x++) { md->Kinetic.TransPart[x] = 0.5 * (kb[x] + twoKEt[x]); md->Kinetic.RotPart[x] = 0.5 * kb[3 + x]; }
    md->Kinetic.Rotational = md->Kinetic.RotPart[0] + md->Kinetic.RotPart[1] + md->Kinetic.RotPart[2];
    md->Kinetic.Kinetic = (md->Kinetic.TransPart[0] + md->Kinetic.TransPart[1] + md->Kinetic.TransPart[2]) + md->Kinetic.Rotational;
    md->Energy.ShadowPotential = md->Energy.ShadowPotential - dt * dt * Us / 24.0;
    md->Kinetic.ShadowRotational = Ks_r / (6.0 * dt);
    md->Kinetic.ShadowKinetic = (Ks_t + Ks_r) / (6.0 * dt);
  }
  md->Kinetic.UpToDate = compute;
}

// EmDeeCode.f90:1215-1277
void EmDee_compute_forces(tEmDee* md) {
  PhaseScope ps(S(*md), false);
  compute_forces(md);
}

// EmDeeCode.f90:1281-1395 (analysis observable, not part of the force engine)
void EmDee_rdf(tEmDee md, int bins, double Rc, int pairs, int itype[], int jtype[], double g[]) {
  const char* task = "radial distribution calculation";
  const double Pi4_3 = 4.188790204786391;
  System* me = S(md);
  PhaseScope ps(me, false);
  int nt = me->ntypes;
  for (int p = 0; p < pairs; p++)
    if (!ranged({ itype[p], jtype[p] }, nt)) fatal(task, "at least one provided type index is out of range");
  std::vector<uint8_t> pairOn((size_t)(nt + 1) * (nt + 1), 0);
  int maxtype = 0;
  for (int p = 0; p < pairs; p++) {
    pairOn[itype[p] * (nt + 1) + jtype[p]] = pairOn[jtype[p] * (nt + 1) + itype[p]] = 1;
    maxtype = std::max(maxtype, std::max(itype[p], jtype[p]));
  }
  auto symm1D = [](int i, int j) { int x = std::min(i, j) - 1, y = std::max(i, j) - 1; return x + (y + 1) * y / 2 + 1; };
  int npt = maxtype > 0 ? symm1D(maxtype, maxtype) : 0;
  std::vector<long long> count;
  if (npt > 0) dev_rdf(me->dev, bins, Rc, me->L, pairOn, npt, count);
  double invL = 1.0 / me->L;
  double norm = Pi4_3 * std::pow(Rc * invL / bins, 3);
  for (int p = 0; p < pairs; p++) {
    int i = itype[p], j = jtype[p];
    long Ni = 0, Nj = 0;
    for (int a = 0; a < me->natoms; a++) { Ni += me->atomType[a] == i; Nj += me->atomType[a] == j; }
    for (int b = 0; b < bins; b++) {
      double r = count[(size_t)b + (size_t)bins * (symm1D(i, j) - 1)] / norm / (3 * (b + 1) * b + 1);
      g[(size_t)b + (size_t)bins * p] = (i == j ? 2.0 * r : r) / ((double)Ni * (double)Nj);
    }
  }
}

// ---- extensions (include/emdee_gpu.h) --------------------------------------

// EmDee_rdf with no allocation (no counterpart in the reference): the same
// checks, counts and normalisation as EmDee_rdf above, with the device
// buffers in the library's static scratch array and no host containers.
void EmDeeGPU_rdf_noalloc(tEmDee md, int bins, double Rc, int pairs, int itype[], int jtype[], double g[]) {
  const char* task = "radial distribution calculation";
  const double Pi4_3 = 4.188790204786391;
  System* me = S(md);
  int nt = me->ntypes;
  for (int p = 0; p < pairs; p++)
    if (!ranged({ itype[p], jtype[p] }, nt)) fatal(task, "at least one provided type index is out of range");
  if (bins < 1) fatal(task, "the number of bins must be positive");
  if (pairs > RDF_NOALLOC_MAX_PAIRS) fatal(task, "more type pairs than the allocation-free form holds (64)");
  int maxtype = 0;
  for (int p = 0; p < pairs; p++) maxtype = std::max(maxtype, std::max(itype[p], jtype[p]));
  if (maxtype > RDF_NOALLOC_MAX_TYPE) fatal(task, "type index above the allocation-free form's limit (31)");
  if (pairs < 1) return;
  int npt = (maxtype - 1) + maxtype * (maxtype - 1) / 2 + 1;   // symm1D(maxtype, maxtype)
  if ((long)bins * npt > RDF_NOALLOC_MAX_COUNTS || (long)bins * pairs > RDF_NOALLOC_MAX_VALUES)
    fatal(task, "more bins than the allocation-free form holds (65536 per type pair set)");
  PhaseScope ps(me, false);
  double NiNj[RDF_NOALLOC_MAX_PAIRS];
  for (int p = 0; p < pairs; p++) {
    long Ni = 0, Nj = 0;
    for (int a = 0; a < me->natoms; a++) { Ni += me->atomType[a] == itype[p]; Nj += me->atomType[a] == jtype[p]; }
    NiNj[p] = (double)Ni * (double)Nj;
  }
  double invL = 1.0 / me->L;
  double norm = Pi4_3 * std::pow(Rc * invL / bins, 3);
  dev_rdf_noalloc(me->dev, bins, Rc, me->L, itype, jtype, pairs, NiNj, npt, norm, g);
}

void EmDeeGPU_set_precision(tEmDee md, int mode) {
  System* me = S(md);
  me->precision = mode ? 1 : 0;
  for (int l = 1; l <= me->nlayers; l++) me->forcesUpToDate[l] = 0;
}
int EmDeeGPU_get_precision(tEmDee md) { return S(md)->precision; }
int EmDeeGPU_neighbor_pairs(tEmDee md, int max, int* pairs) { return dev_neighbor_pairs(S(md)->dev, max, pairs); }
void EmDeeGPU_profile(tEmDee md, int on) { dev_set_profiling(S(md)->dev, on != 0); }
void EmDeeGPU_kernel_stats(tEmDee md, double* stats) { dev_kernel_stats(S(md)->dev, stats, 7); }
void EmDeeGPU_synchronize(tEmDee md) { dev_sync(S(md)->dev); }
void EmDeeGPU_atom_energies(tEmDee* md, double* e) {
  System* me = S(*md);
  if (!me->initialized) fatal("atom energies", "box and coordinates have not been defined");
  me->atomEnergies = true;
  compute_forces(md);
  me->atomEnergies = false;
  dev_get_atom_energies(me->dev, e);
  dev_set_atom_energies(me->dev, false);
}
void EmDeeGPU_attach_surrogate(tEmDee md, EmDeeSurrogate* s, int layer) {
  System* me = S(md);
  if (s && (layer < 1 || layer > me->nlayers)) fatal("surrogate attachment", "selected layer is out of range");
  if (s && surrogate_cutoff(s) > me->Rc) fatal("surrogate attachment", "the descriptor cutoff exceeds the system's cutoff");
  me->surrogate = s;
  me->surrogateLayer = s ? layer : 0;
  for (int l = 1; l <= me->nlayers; l++) me->forcesUpToDate[l] = 0;
}
void EmDeeGPU_surrogate_sample(tEmDee md, EmDeeSurrogate* s, const double* labels) {
  System* me = S(md);
  if (!me->initialized) fatal("surrogate sampling", "box and coordinates have not been defined");
  if (surrogate_cutoff(s) > me->Rc) fatal("surrogate sampling", "the descriptor cutoff exceeds the system's cutoff");
  PhaseScope ps(me, false);
  dev_surrogate_add_sample(me->dev, s, surrogate_context(me), labels);
}
void EmDeeGPU_destroy(tEmDee* md) {
  if (!md || !md->Data) return;
  System* me = S(*md);
  phase_leave(me);
  dev_destroy(me->dev);
  delete me;
  md->Data = nullptr;
}

}  // extern "C"
