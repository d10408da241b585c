// models.cpp -- model registry of the EmDee engine (host side).
//
// The reference generates one C constructor per model file from its "!<>"
// tagged parameters (src/make_models_module.sh:76-135) and wraps each model
// in a heap container whose address is the opaque handle (modelClass.f90:51-60).
// Here every handle points to one `Model` record; the constructors below keep
// the reference names and parameter order (include/emdee.h).  Set-up
// arithmetic (shifts, switching factors, mixing) follows the cited reference
// routines in the same order, in IEEE FP64 (this file is built without FMA
// contraction).
#include <cmath>
#include <cstring>
#include <string>

#include "engine.h"
#include "models.h"

namespace emdee {

static const double PI_ = 3.14159265358979324;   // global.f90:38

static Model* make(int kind) {
  Model* m = new Model();
  m->kind = kind;
  return m;
}

bool is_pair(const Model& m) { return m.kind >= PAIR_NONE && m.kind <= PAIR_SOFTCORE_CUT; }
bool is_coul(const Model& m) { return m.kind >= COUL_NONE && m.kind <= COUL_SHIFTED_SQUARE_SMOOTHED; }

// pair_lj_cut.f90:52-69
void setup_lj(Model& m, double eps, double sig) {
  m.epsilon = eps; m.sigma = sig;
  m.c0 = 4.0 * m.epsilon;   // eps4
  m.c1 = 24.0 * m.epsilon;  // eps24
  m.c2 = m.sigma * m.sigma; // sigsq
  m.c3 = 0;
}
// pair_softcore_cut.f90:58-82 (alpha = 1/2, n = p = 1)
void setup_softcore(Model& m, double eps, double sig, double lam) {
  m.epsilon = eps; m.sigma = sig; m.lambda = lam;
  if (m.lambda < 0.0 || m.lambda > 1.0) fatal("pair_softcore_cut setup", "out-of-range parameter lambda");
  m.c0 = 4.0 * m.epsilon * m.lambda;   // prefactor
  m.c1 = 6.0 * m.c0;                   // prefactor6
  m.c2 = 1.0 / (m.sigma * m.sigma);    // invSigSq
  m.c3 = 0.5 * (1.0 - m.lambda);       // shift
}

// Host evaluation of a non-bonded model at one distance (used by set-up):
// the "compute" bodies of the model files.
static void nb_compute(const Model& m, double invR, double invR2, double& E, double& W) {
  switch (m.kind) {
    case PAIR_LJ_CUT: {
      double sr2 = m.c2 * invR2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
      E = m.c0 * (sr12 - sr6); W = m.c1 * (sr12 + sr12 - sr6); return;
    }
    case PAIR_SOFTCORE_CUT: {
      double rsig2 = m.c2 / invR2, rsig6 = rsig2 * rsig2 * rsig2;
      double sinv = 1.0 / (rsig6 + m.c3), sinvSq = sinv * sinv, sinvCb = sinv * sinvSq;
      E = m.c0 * (sinvSq - sinv); W = m.c1 * rsig6 * (sinvCb + sinvCb - sinvSq); return;
    }
    case COUL_CUT: E = invR; W = invR; return;
    case COUL_SF: { double rFc = m.fshift / invR; E = invR + m.eshift + rFc; W = invR - rFc; return; }
    case COUL_DAMPED: case COUL_LONG: {
      double x = m.alpha / invR, e = std::exp(-x * x);
      double t = 1.0 / (1.0 + 0.327591100 * x);
      E = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429)))) * e * invR;
      W = E + m.beta * e; return;
    }
    case COUL_DAMPED_SMOOTHED: {
      double r = 1.0 / invR, x = m.alpha * r, e = std::exp(-x * x);
      double t = 1.0 / (1.0 + 0.327591100 * x);
      E = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429)))) * e * invR;
      W = E + m.beta * e;
      if (r > m.Rm) {
        double u = m.factor * (r - m.Rm), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0), WG = -30.0 * u2 * (2.0 * u - u2 - 1.0) * m.factor * r;
        W = W * G + E * WG; E = E * G;
      }
      return;
    }
    case COUL_DAMPED_SQUARE_SMOOTHED: {
      double x = m.alpha / invR, e = std::exp(-x * x);
      double t = 1.0 / (1.0 + 0.327591100 * x);
      E = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429)))) * e * invR;
      W = E + m.beta * e;
      if (invR < m.invRm) {
        double r2 = 1.0 / invR2, u = m.factor * (r2 - m.Rm2), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0), WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * m.factor * r2;
        W = W * G + E * WG; E = E * G;
      }
      return;
    }
    case COUL_SQUARE_SMOOTHED: case COUL_SHIFTED_SQUARE_SMOOTHED: {
      W = invR;
      E = m.kind == COUL_SQUARE_SMOOTHED ? W : W + m.eshift;
      if (invR < m.invRm) {
        double r2 = 1.0 / invR2, u = m.factor * (r2 - m.Rm2), u2 = u * u, u3 = u * u2;
        double G = 1.0 + u3 * (15.0 * u - 6.0 * u2 - 10.0), WG = -60.0 * u2 * (2.0 * u - u2 - 1.0) * m.factor * r2;
        W = W * G + E * WG; E = E * G;
      }
      return;
    }
    default: E = 0; W = 0; return;
  }
}

// modelClass_nonbonded.f90:245-296
void modifier_setup(Model& m, double cutoff) {
  double Ec = 0, Wc = 0, Es = 0, Ws = 0;
  m.fshift = 0.0;
  m.eshift = 0.0;
  m.Rm = cutoff - m.skin;
  m.RmSq = m.Rm * m.Rm;
  bool shifting = m.modifier == MOD_SHIFTED || m.modifier == MOD_SHIFTED_FORCE ||
                  m.modifier == MOD_SHIFTED_SMOOTHED || m.modifier == MOD_SHIFTED_SQUARE_SMOOTHED;
  if (shifting) nb_compute(m, 1.0 / cutoff, 1.0 / (cutoff * cutoff), Ec, Wc);
  switch (m.modifier) {
    case MOD_SHIFTED: m.eshift = -Ec; break;
    case MOD_SHIFTED_FORCE: m.eshift = -(Ec + Wc); m.fshift = Wc / cutoff; break;
    case MOD_SMOOTHED: case MOD_SHIFTED_SMOOTHED:
      if (shifting) { nb_compute(m, 1.0 / m.Rm, 1.0 / m.RmSq, Es, Ws); m.eshift = -0.5 * (Es + Ec); }
      m.factor = 1.0 / (cutoff - m.Rm);
      m.Rm2fac = m.factor * m.Rm;
      break;
    case MOD_SQUARE_SMOOTHED: case MOD_SHIFTED_SQUARE_SMOOTHED:
      if (shifting) { nb_compute(m, 1.0 / m.Rm, 1.0 / m.RmSq, Es, Ws); m.eshift = -0.5 * (Es + Ec); }
      m.factor = 1.0 / (cutoff * cutoff - m.RmSq);
      m.Rm2fac = m.factor * m.RmSq;
      break;
    default: break;
  }
}

// apply_cutoff of the smoothed Coulomb models (coul_damped_smoothed.f90:76-85,
// coul_damped_square_smoothed.f90:75-83, coul_square_smoothed.f90:68-76,
// coul_shifted_square_smoothed.f90:71-79).
static void coul_apply_cutoff(Model& m, double Rc) {
  switch (m.kind) {
    case COUL_DAMPED_SMOOTHED:
      m.Rm = Rc - m.skinWidth; m.Rm2 = m.Rm * m.Rm; m.invRm = 1.0 / m.Rm; m.factor = 1.0 / (Rc - m.Rm);
      break;
    case COUL_DAMPED_SQUARE_SMOOTHED: case COUL_SQUARE_SMOOTHED: case COUL_SHIFTED_SQUARE_SMOOTHED:
      m.Rm2 = (Rc - m.skinWidth) * (Rc - m.skinWidth);
      m.invRm = 1.0 / (Rc - m.skinWidth);
      m.factor = 1.0 / (Rc * Rc - m.Rm2);
      break;
    default: break;
  }
}

// modelClass_coul.f90:62-93
void coul_cutoff_setup(Model& m, double cutoff) {
  m.fshift = 0.0;
  m.eshift = 0.0;
  if (m.shifted || m.shifted_force) {
    double invR = 1.0 / cutoff, invR2 = invR * invR, E, W;
    nb_compute(m, invR, invR2, E, W);
    if (m.shifted_force) { m.fshift = W / cutoff; m.eshift = -(E + W); }
    else { m.fshift = 0.0; m.eshift = -E; }
  }
  coul_apply_cutoff(m, cutoff);
}

// coul_long.f90:66-73 / modelClass_coul.f90:106-111
void coul_kspace_setup(Model& m, double alpha) {
  if (m.kind == COUL_LONG) { m.alpha = alpha; m.beta = 2.0 * alpha / std::sqrt(PI_); }
  else if (m.requires_kspace) m.alpha = alpha;
}

static const char* model_name(const Model& m) {
  switch (m.kind) {
    case PAIR_LJ_CUT: return "lj_cut";
    case PAIR_SOFTCORE_CUT: return "softcore_cut";
    default: return "none";
  }
}

// "this%mix(other)": pair_lj_cut.f90:122-138, pair_softcore_cut.f90:139-163,
// modelClass_pair.f90:190-200.  The mixed model is freshly set up, so it
// carries no modifier.
static bool mix_rule(const Model& self, const Model& other, Model& out) {
  out = Model();
  switch (self.kind) {
    case PAIR_NONE: out.kind = PAIR_NONE; return true;
    case PAIR_LJ_CUT:
      if (other.kind == PAIR_LJ_CUT) {
        out.kind = PAIR_LJ_CUT;
        setup_lj(out, std::sqrt(self.epsilon * other.epsilon), 0.5 * (self.sigma + other.sigma));
        return true;
      }
      return false;
    case PAIR_SOFTCORE_CUT:
      if (other.kind == PAIR_SOFTCORE_CUT) {
        out.kind = PAIR_SOFTCORE_CUT;
        setup_softcore(out, std::sqrt(self.epsilon * other.epsilon), 0.5 * (self.sigma + other.sigma), self.lambda * other.lambda);
        return true;
      }
      if (other.kind == PAIR_LJ_CUT) {
        out.kind = PAIR_SOFTCORE_CUT;
        setup_softcore(out, std::sqrt(self.epsilon * other.epsilon), 0.5 * (self.sigma + other.sigma), self.lambda);
        return true;
      }
      return false;
    default: return false;
  }
}

// modelClass_pair.f90:120-140
PairEntry mix_pairs(const PairEntry& a, const PairEntry& b) {
  PairEntry c;
  if (!mix_rule(b.model, a.model, c.model) && !mix_rule(a.model, b.model, c.model)) {
    c.model = Model();
    c.model.kind = PAIR_NONE;
    warning(std::string("no mixing rule found for models ") + model_name(a.model) + " and " + model_name(b.model));
  }
  c.coulomb = a.coulomb && b.coulomb;
  c.kCoul = c.coulomb ? std::sqrt(a.kCoul * b.kCoul) : 0.0;
  return c;
}

DevPair pack_pair(const PairEntry& p) {
  DevPair d;
  std::memset(&d, 0, sizeof d);
  d.kind = p.model.kind;
  d.modifier = p.model.modifier;
  d.coulomb = p.coulomb ? 1 : 0;
  d.kCoul = p.kCoul;
  d.c0 = p.model.c0; d.c1 = p.model.c1; d.c2 = p.model.c2; d.c3 = p.model.c3;
  d.eshift = p.model.eshift; d.fshift = p.model.fshift; d.factor = p.model.factor; d.Rm2fac = p.model.Rm2fac;
  return d;
}

DevCoul pack_coul(const Model& m) {
  DevCoul d;
  std::memset(&d, 0, sizeof d);
  d.kind = m.kind;
  d.modifier = m.modifier;
  d.alpha = m.alpha; d.beta = m.beta; d.eshift = m.eshift; d.fshift = m.fshift;
  d.Rm = m.Rm; d.Rm2 = m.Rm2; d.invRm = m.invRm; d.factor = m.factor; d.Rm2fac = m.Rm2fac;
  return d;
}

}  // namespace emdee

using namespace emdee;

// ------------------------------------------------------------------ C ABI

// Modifier constructors: modelClass_nonbonded.f90:83-241
static void* modified(void* model, int modifier, double skin, bool setSkin, const char* task) {
  if (!model) fatal(task, "a valid pair model must be provided");
  const Model* src = static_cast<const Model*>(model);
  if (!is_pair(*src) && !is_coul(*src)) fatal(task, "a valid pair model must be provided");
  Model* m = new Model(*src);
  m->modifier = modifier;
  if (setSkin) m->skin = skin;
  return m;
}

extern "C" {

void* EmDee_shifted(void* model) { return modified(model, MOD_SHIFTED, 0, false, "shifted potential assignment"); }
void* EmDee_shifted_force(void* model) { return modified(model, MOD_SHIFTED_FORCE, 0, false, "shifted-force potential assignment"); }
void* EmDee_smoothed(void* model, double skin) { return modified(model, MOD_SMOOTHED, skin, true, "smoothed potential assignment"); }
void* EmDee_shifted_smoothed(void* model, double skin) { return modified(model, MOD_SHIFTED_SMOOTHED, skin, true, "shifted-smoothed potential assignment"); }
void* EmDee_square_smoothed(void* model, double skin) { return modified(model, MOD_SQUARE_SMOOTHED, skin, true, "square-smoothed potential assignment"); }
void* EmDee_shifted_square_smoothed(void* model, double skin) { return modified(model, MOD_SHIFTED_SQUARE_SMOOTHED, skin, true, "shifted-square-smoothed potential assignment"); }

void* EmDee_pair_none(void) { return make(PAIR_NONE); }
void* EmDee_coul_none(void) { return make(COUL_NONE); }
void* EmDee_bond_none(void) { return make(BOND_NONE); }
void* EmDee_angle_none(void) { return make(ANGLE_NONE); }
void* EmDee_dihedral_none(void) { return make(DIHEDRAL_NONE); }

void* EmDee_pair_lj_cut(double epsilon, double sigma) {
  Model* m = make(PAIR_LJ_CUT); setup_lj(*m, epsilon, sigma); return m;
}
void* EmDee_pair_softcore_cut(double epsilon, double sigma, double lambda) {
  Model* m = make(PAIR_SOFTCORE_CUT); setup_softcore(*m, epsilon, sigma, lambda); return m;
}
void* EmDee_coul_cut(void) { return make(COUL_CUT); }
void* EmDee_coul_sf(void) { Model* m = make(COUL_SF); m->shifted_force = true; return m; }
void* EmDee_coul_damped(double damp) {
  Model* m = make(COUL_DAMPED); m->damp = damp; m->alpha = damp; m->beta = 2.0 * m->alpha / std::sqrt(PI_); return m;
}
void* EmDee_coul_long(void) { Model* m = make(COUL_LONG); m->requires_kspace = true; return m; }
void* EmDee_coul_damped_smoothed(double damp, double skinWidth) {
  Model* m = make(COUL_DAMPED_SMOOTHED);
  m->damp = damp; m->skinWidth = skinWidth; m->alpha = m->damp; m->beta = 2.0 * m->alpha / std::sqrt(PI_); return m;
}
void* EmDee_coul_damped_square_smoothed(double damp, double skinWidth) {
  Model* m = make(COUL_DAMPED_SQUARE_SMOOTHED);
  m->damp = damp; m->skinWidth = skinWidth; m->alpha = m->damp; m->beta = 2.0 * m->alpha / std::sqrt(PI_); return m;
}
void* EmDee_coul_square_smoothed(double skinWidth) { Model* m = make(COUL_SQUARE_SMOOTHED); m->skinWidth = skinWidth; return m; }
void* EmDee_coul_shifted_square_smoothed(double skinWidth) {
  Model* m = make(COUL_SHIFTED_SQUARE_SMOOTHED); m->skinWidth = skinWidth; m->shifted = true; return m;
}
void* EmDee_bond_harmonic(double k, double r0) {
  Model* m = make(BOND_HARMONIC); m->k = k; m->x0 = r0; m->minus_k = -m->k; m->half_k = 0.5 * m->k; return m;
}
void* EmDee_angle_harmonic(double k, double theta0) {
  Model* m = make(ANGLE_HARMONIC); m->k = k; m->x0 = theta0; m->minus_k = -m->k; m->half_k = 0.5 * m->k; return m;
}
// A harmonic dihedral, E = k/2 (phi - phi0)^2 with the difference taken in
// (-pi, pi]: the reference defines the dihedral model class
// (modelClass_dihedral.f90:28-45: compute(Ed, Fd, phi), factor14) but ships
// no concrete model.  Laid out after bond_harmonic and angle_harmonic; no
// scaled 1-4 pair term (factor14 = 0).
void* EmDee_dihedral_harmonic(double k, double phi0) {
  Model* m = make(DIHEDRAL_HARMONIC); m->k = k; m->x0 = phi0; m->minus_k = -m->k; m->half_k = 0.5 * m->k; return m;
}
void* EmDee_kspace_ewald(double accuracy) { Model* m = make(KSPACE_EWALD); m->accuracy = accuracy; return m; }

}  // extern "C"
