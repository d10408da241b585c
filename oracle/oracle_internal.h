/*
 * oracle_internal.h -- data structures of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C, FP64 restatement of the
 * reference EmDee algorithm, used solely as the checker by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  No product
 * code may include, link or call anything under oracle/.
 *
 * Each structure below restates a reference Fortran type; the citation names
 * the type it follows.
 */
#ifndef ORACLE_INTERNAL_H
#define ORACLE_INTERNAL_H

#include <stdint.h>
#include "../include/emdee.h"

/* Model kinds (one per reference module; the file stem is the C name). */
enum {
  K_PAIR_NONE = 1, K_PAIR_LJ_CUT, K_PAIR_SOFTCORE_CUT,
  K_COUL_NONE, K_COUL_CUT, K_COUL_SF, K_COUL_DAMPED, K_COUL_LONG,
  K_COUL_DAMPED_SMOOTHED, K_COUL_DAMPED_SQUARE_SMOOTHED,
  K_COUL_SQUARE_SMOOTHED, K_COUL_SHIFTED_SQUARE_SMOOTHED,
  K_BOND_NONE, K_BOND_HARMONIC, K_ANGLE_NONE, K_ANGLE_HARMONIC,
  K_DIHEDRAL_NONE, K_DIHEDRAL_HARMONIC, K_KSPACE_EWALD
};

/* Modifier codes: modelClass_nonbonded.f90:27-33 */
enum { MOD_NONE = 0, MOD_SHIFTED, MOD_SHIFTED_FORCE, MOD_SMOOTHED, MOD_SHIFTED_SMOOTHED,
       MOD_SQUARE_SMOOTHED, MOD_SHIFTED_SQUARE_SMOOTHED };

/* Flat restatement of cModel + cNonBondedModel + concrete model fields
   (modelClass.f90, modelClass_nonbonded.f90:36-50, modelClass_coul.f90:27-38,
   pair_lj_cut.f90:35-40, pair_softcore_cut.f90:40-46, coul_*.f90). */
typedef struct {
  int kind;
  /* cNonBondedModel */
  int modifier;
  double eshift, fshift, skin, Rm, RmSq, factor, Rm2fac;
  /* pair_lj_cut */
  double epsilon, sigma, eps4, eps24, sigsq;
  /* pair_softcore_cut */
  double lambda, prefactor, prefactor6, invSigSq, shift;
  /* cCoulModel */
  int shifted, shifted_force, requires_kspace;
  double alpha, beta, damp, skinWidth, Rm2, invRm;
  /* bond/angle/dihedral harmonic (r0: equilibrium distance or angle) */
  double k, r0, minus_k, half_k;
  /* cDihedralModel (modelClass_dihedral.f90:28-31) */
  double factor14;
  /* kspace_ewald */
  double accuracy;
} omodel_t;

/* pairContainer: modelClass_pair.f90:66-74 */
typedef struct { omodel_t model; int coulomb; double kCoul; } opair_t;

/* tList: lists.f90:26-40 (CSR with first/middle/last) */
typedef struct {
  int nitems, count;
  int *first, *middle, *last;   /* 1-based positions into item/value */
  int *item;                    /* 1-based atom indices */
  double *value;
} olist_t;

/* structList item: structs.f90:28-31 */
typedef struct { int i, j, k, l; omodel_t model; } ostruct_t;

/* tBody: ArBee.f90:29-72 */
typedef struct {
  int NP, dof;
  double mass, MoI[3], rcm[3], pcm[3], q[4], pi[4], omega[3], F[3], tau[3];
  int *index;          /* 1-based atom indices */
  double *M;
  double *d;           /* d(3,NP) column-major: d[3*k + x] */
  double *delta;       /* delta(3,NP) */
  double invMass, invMoI[3];
  double I113, I313, I223, I212, m312;
} obody_t;

/* kiss + i32rng: math.f90:42-177 */
typedef struct {
  int seeding_required;
  int32_t kn[128];
  double wn[128], fn[128];
  int32_t x, y, z, w;
} okiss_t;

/* kspace_ewald + cKspaceModel: kspace_ewald.f90:37-55, modelClass_kspace.f90:38-70 */
typedef struct {
  omodel_t base;
  double alpha, beta, kmax, Volume;
  int ntypes;                 /* number of distinct charged types */
  int *type;                  /* [ntypes] global type of each local charged type */
  int ncharged;
  int *cfirst, *clast;        /* per local type, 1-based into citem */
  int *citem;                 /* charged atoms grouped by type, 1-based */
  double *cvalue;
  int nTypePairs;
  double *Erigid;
  int nRigidPairs;
  int *rp_i, *rp_j, *rp_it, *rp_jt;
  double *rp_FbyR;
  int nlayers;
  double *lambda;             /* lambda(ntypes,ntypes,nlayers) column-major */
  double *lambda1D;           /* lambda1D(nTypePairs,nlayers) */
  int nmax[3], nvecs;
  int *n;                     /* n(nvecs,3) */
  double *prefac, *vec;       /* vec(3,nvecs) */
  double *qeikr;              /* complex qeikr(nvecs,ncharged): re,im interleaved */
  double *S;                  /* complex S(nvecs,ntypes) */
  int nthreads;
} okspace_t;

/* tData: EmDeeData.f90:66-151 */
typedef struct {
  int natoms, mcells, ncells, maxcells, maxatoms, maxpairs, ntypes;
  int nbodies, nfree, nthreads, threadAtoms, threadFreeAtoms, threadBodies;
  int nlayers, layer;
  double *Lbox;
  double Rc, skin, RcSq, xRc, xRcSq, skinSq, InRc, InRcSq, xInRcSq, totalMass, startTime;
  int initialized;
  okiss_t random;
  /* cellAtom */
  int *cellFirst, *cellLast, *cellItem;
  int *threadCellFirst, *threadCellLast;
  /* excluded: per-atom sorted lists */
  int *exCount, *exCap, **exItem;
  ostruct_t *bonds; int nbonds, capbonds;
  ostruct_t *angles; int nangles, capangles;
  ostruct_t *dihedrals; int ndihedrals, capdihedrals;
  int *atomType, *atomCell, *atomBody, *free_, *atomsInCell;
  double *R, *P;             /* R(3,N) */
  obody_t *body;
  double *F;                 /* points into layerF */
  double *charge, *mass, *invMass, *R0;
  int *charged;
  int *cellNeighbor;         /* cell(icell)%neighbor(62) */
  olist_t *neighbor;         /* per thread */
  opair_t *pair;             /* pair(ntypes,ntypes,nlayers) */
  omodel_t *coul;            /* coul(nlayers) */
  okspace_t *kspace;
  int *multilayer, *overridable, *interact, *pairs_exist;
  int *bonded, *useInRc, *forcesUpToDate;
  double *layerF;            /* layerF(3,N,nlayers) */
  void *layerEnergy;         /* copies of tEmDee.Energy per layer */
  void *layerVirial;
  int multilayer_coulomb, kspace_active;
  int Rowned, Powned, bodyOwned, Lowned;   /* phase-space sharing bookkeeping */
} odata_t;

#endif
