// kspace.cpp -- Ewald set-up on the host: the reference's k-space
// initialisation (modelClass_kspace.f90:111-261) and k-vector generation
// (kspace_ewald.f90:82-184), producing the device layout of kspace.h.
// It runs once per system (the reference never re-runs update when the box
// changes later, and neither does this engine).
#include "kspace.h"

#include <algorithm>
#include <cmath>

namespace emdee {

namespace {
const double PI_ = 3.14159265358979324;     // global.f90:38
const double EPS = 2.220446049250313e-16;   // epsilon(1.0_rb)

// A&S 7.1.26 (math.f90:35-40, 685-691)
double uerfc(double x, double expmx2) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027, a5 = 1.061405429,
               p = 0.327591100;
  double t = 1.0 / (1.0 + p * x);
  return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * expmx2;
}

// modelClass_kspace.f90:321-334
void discount(double alpha, double beta, double rsq, double QiQj, double* E, double* W) {
  double r = std::sqrt(rsq), x = alpha * r, expmx2 = std::exp(-x * x);
  *E = -QiQj * (1.0 - uerfc(x, expmx2)) / r;
  *W = *E + QiQj * beta * expmx2;
}

// math.f90:695-702
int symm1D(int i, int j) {
  int x = std::min(i, j) - 1, y = std::max(i, j) - 1;
  return x + (y + 1) * y / 2 + 1;
}
}  // namespace

double inverse_of_x_plus_ln_x(double y) {
  const double tol = 1.0e-12;
  double x = y > 0.5671432904097839 ? y - std::log(y) : std::exp(y);
  double x0 = x + 1.0;
  while (std::fabs(x - x0) > tol * x0) {
    x0 = x;
    x = x * (y + 1.0 - std::log(x)) / (x + 1.0);
  }
  return x;
}

EwaldHost ewald_initialize(double accuracy, double Rc, double L, int nlayers, const std::vector<int>& atomType,
                           const std::vector<double>& charge, const std::vector<int>& bStart,
                           const std::vector<int>& bOrig, const double* R,
                           const std::function<double(int, int, int)>& kCoul) {
  const char* task = "kspace model initialization";
  EwaldHost h;
  int N = (int)atomType.size();
  std::vector<char> charged(N);
  int ncharged = 0;
  for (int i = 0; i < N; i++) { charged[i] = std::fabs(charge[i]) > EPS; ncharged += charged[i]; }
  if (ncharged == 0) fatal(task, "system has no charged atoms");
  // distinct charged types, ascending (global.f90:68-97)
  std::vector<int> types;
  for (int i = 0; i < N; i++)
    if (charged[i] && std::find(types.begin(), types.end(), atomType[i]) == types.end()) types.push_back(atomType[i]);
  std::sort(types.begin(), types.end());
  int nt = (int)types.size();
  int maxType = *std::max_element(atomType.begin(), atomType.end());
  std::vector<int> local(maxType + 1, 0);
  for (int t = 0; t < nt; t++) local[types[t]] = t + 1;

  // Ewald parameters: kspace_ewald.f90:82-99
  double s = std::sqrt(inverse_of_x_plus_ln_x(-std::log(accuracy)));
  h.alpha = s / Rc;
  h.kmax = 2.0 * h.alpha * s;
  h.beta = 2.0 * h.alpha / std::sqrt(PI_);
  h.dev.alpha = h.alpha;

  // slots: every atom of the charged types, as cfirst/clast/citem of
  // modelClass_kspace.f90:141-150 (all atoms of a charged type are listed)
  EwaldSetup& e = h.dev;
  e.nt = nt;
  e.nl = nlayers;
  e.typeGroups.assign(nt + 1, 0);
  std::vector<std::vector<double>> typeQ(nt);
  for (int t = 0; t < nt; t++) {
    for (int i = 0; i < N; i++)
      if (atomType[i] == types[t]) { e.slotItem.push_back(i); e.slotQ.push_back(charge[i]); typeQ[t].push_back(charge[i]); }
    while (e.slotItem.size() % EWALD_GROUP) { e.slotItem.push_back(-1); e.slotQ.push_back(0.0); }
    e.typeGroups[t + 1] = (int)(e.slotItem.size() / EWALD_GROUP);
    for (int g = e.typeGroups[t]; g < e.typeGroups[t + 1]; g++) e.groupType.push_back(t);
  }
  e.nslots = (int)e.slotItem.size();

  // intrabody pairs and self energy: modelClass_kspace.f90:155-229
  int nTypePairs = nt * (nt - 1) / 2 + nt;
  std::vector<double> Erigid(nTypePairs + 1, 0.0);
  double invL = 1.0 / L;
  int nb = bStart.empty() ? 0 : (int)bStart.size() - 1;
  for (int b = 0; b < nb; b++) {
    for (int ii = bStart[b]; ii < bStart[b + 1] - 1; ii++) {
      int i = bOrig[ii];
      if (!charged[i]) continue;
      for (int jj = ii + 1; jj < bStart[b + 1]; jj++) {
        int j = bOrig[jj];
        if (!charged[j]) continue;
        double Rij[3];
        for (int x = 0; x < 3; x++) {
          Rij[x] = R[3 * i + x] - R[3 * j + x];
          Rij[x] = Rij[x] - L * std::round(invL * Rij[x]);
        }
        double rsq = Rij[0] * Rij[0] + Rij[1] * Rij[1] + Rij[2] * Rij[2], E, W;
        discount(h.alpha, h.beta, rsq, charge[i] * charge[j], &E, &W);
        Erigid[symm1D(local[atomType[i]], local[atomType[j]])] += E;
      }
    }
  }
  for (int t = 1; t <= nt; t++) {
    double q2 = 0;
    for (double q : typeQ[t - 1]) q2 += q * q;
    int k = symm1D(t, t);
    Erigid[k] = Erigid[k] - h.alpha * q2 / std::sqrt(PI_);
  }
  e.lambda.assign((size_t)nt * nt * nlayers, 0.0);
  h.rigidEnergy.assign(nlayers, 0.0);
  for (int l = 1; l <= nlayers; l++) {
    std::vector<double> lam1D(nTypePairs + 1, 0.0);
    for (int i = 1; i <= nt; i++)
      for (int j = i; j <= nt; j++) {
        double lam = kCoul(types[i - 1], types[j - 1], l);
        e.lambda[(size_t)(l - 1) * nt * nt + (i - 1) + nt * (j - 1)] = lam;
        e.lambda[(size_t)(l - 1) * nt * nt + (j - 1) + nt * (i - 1)] = lam;
        lam1D[symm1D(i, j)] = lam;
      }
    double sum = 0;
    for (int k = 1; k <= nTypePairs; k++) sum += lam1D[k] * Erigid[k];
    h.rigidEnergy[l - 1] = sum;
  }

  // k-vectors: kspace_ewald.f90:103-184
  double unit = 2.0 * PI_ / L, unitSq = unit * unit;
  int nmax = (int)std::ceil(h.kmax / unit);
  int M = 2 * nmax + 1, maxnvecs = M * M * M / 2, M2M3 = M * M;
  double kmaxSq = unitSq * nmax * nmax;
  std::vector<int> n;
  for (int i = maxnvecs + 1; i <= 2 * maxnvecs; i++) {
    int n1 = i / M2M3, j = i - n1 * M2M3, n2 = j / M, n3 = j - n2 * M;
    int kv[3] = { n1 - nmax, n2 - nmax, n3 - nmax };
    if (unitSq * kv[0] * kv[0] + unitSq * kv[1] * kv[1] + unitSq * kv[2] * kv[2] <= kmaxSq) n.insert(n.end(), kv, kv + 3);
  }
  int nvecs = (int)n.size() / 3;
  e.nm = nmax;
  e.nvecs = nvecs;
  e.kn.resize(4 * (size_t)nvecs);
  e.kv.resize(4 * (size_t)nvecs);
  double B = -0.25 / (h.alpha * h.alpha);
  h.volume = L * L * L;
  double fourPiByV = 4.0 * PI_ / h.volume;
  for (int i = 0; i < nvecs; i++) {
    double k[3];
    for (int x = 0; x < 3; x++) k[x] = unit * n[3 * i + x];
    double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    double prefac = fourPiByV * std::exp(ksq * B) / ksq;
    for (int x = 0; x < 3; x++) e.kv[4 * i + x] = 2.0 * prefac * k[x];
    e.kv[4 * i + 3] = prefac;
    bool cont = i > 0 && n[3 * i] == n[3 * i - 3] && n[3 * i + 1] == n[3 * i - 2] && n[3 * i + 2] == n[3 * i - 1] + 1;
    e.kn[4 * i] = n[3 * i];
    e.kn[4 * i + 1] = n[3 * i + 1];
    e.kn[4 * i + 2] = n[3 * i + 2];
    e.kn[4 * i + 3] = cont ? 0 : 1;
  }
  return h;
}

// The k-vectors of the same half space with |k| <= kcut (not the reference's
// set: a partial sum over the longest wavelengths, used by the learned model
// as an exact long-range channel), in the reference's enumeration order (n3
// fastest, kspace_ewald.f90:115-138) and with its coefficients
// (kspace_ewald.f90:174-180); charges, slots and the kCoul weights of layer
// `layer` (1-based) are those of the full set-up.
EwaldSetup lowk_setup(const EwaldSetup& full, int layer, double alpha, double L, double kcut) {
  EwaldSetup e = full;
  e.nl = 1;
  e.alpha = alpha;
  int nt = full.nt;
  e.lambda.assign(full.lambda.begin() + (size_t)(layer - 1) * nt * nt, full.lambda.begin() + (size_t)layer * nt * nt);
  double unit = 2.0 * PI_ / L, unitSq = unit * unit, kcutSq = kcut * kcut;
  int nmax = std::max(1, (int)std::ceil(kcut / unit));
  std::vector<int> n;
  for (int n1 = 0; n1 <= nmax; n1++)
    for (int n2 = -nmax; n2 <= nmax; n2++)
      for (int n3 = -nmax; n3 <= nmax; n3++) {
        bool upper = n1 > 0 || (n1 == 0 && (n2 > 0 || (n2 == 0 && n3 > 0)));
        if (!upper) continue;
        if (unitSq * n1 * n1 + unitSq * n2 * n2 + unitSq * n3 * n3 <= kcutSq) {
          n.push_back(n1); n.push_back(n2); n.push_back(n3);
        }
      }
  int nvecs = (int)n.size() / 3;
  e.nm = nmax;
  e.nvecs = nvecs;
  e.kn.assign(4 * (size_t)nvecs, 0);
  e.kv.assign(4 * (size_t)nvecs, 0.0);
  double B = -0.25 / (alpha * alpha), fourPiByV = 4.0 * PI_ / (L * L * L);
  for (int i = 0; i < nvecs; i++) {
    double k[3];
    for (int x = 0; x < 3; x++) k[x] = unit * n[3 * i + x];
    double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    double prefac = fourPiByV * std::exp(ksq * B) / ksq;
    for (int x = 0; x < 3; x++) e.kv[4 * i + x] = 2.0 * prefac * k[x];
    e.kv[4 * i + 3] = prefac;
    bool cont = i > 0 && n[3 * i] == n[3 * i - 3] && n[3 * i + 1] == n[3 * i - 2] && n[3 * i + 2] == n[3 * i - 1] + 1;
    e.kn[4 * i] = n[3 * i];
    e.kn[4 * i + 1] = n[3 * i + 1];
    e.kn[4 * i + 2] = n[3 * i + 2];
    e.kn[4 * i + 3] = cont ? 0 : 1;
  }
  return e;
}

}  // namespace emdee
