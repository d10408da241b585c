/*
 * emdee.h -- C interface of the EmDee force engine (MI355X-native build).
 *
 * This header is the drop-in boundary: a host program written against the
 * reference library's generated header (src/emdee_header.h +
 * src/make_c_header.sh) compiles unchanged against it.  Every entry point is
 * a plain C-ABI function taking pointers and sizes; device memory never
 * crosses this boundary.
 *
 * Struct layout.  The reference library's bind(C) struct
 * (src/EmDeeCode.f90:36-61, src/EmDeeData.f90:40-64) disagrees with both of
 * its own headers (src/emdee_header.h:1-40 lacks Kinetic.UpToDate;
 * src/emdee_header.f03:28-80 adds Options.AutoForceCompute).  This build uses
 * ONE layout, the Fortran-header one (240 bytes), so that C hosts, Fortran
 * hosts (include/emdee.f03) and the library agree byte for byte:
 *   Builds@0 Time@8 Energy@40 Kinetic@104 Virial@192 BodyVirial@200
 *   DoF@208 RotDoF@212 Data@216 Options@224.
 * The C field names of the reference header (Kinetic.Kinetic, Virial,
 * BodyVirial) are kept so that src/testc.c compiles unchanged.
 *
 * Conventions (src/EmDeeCode.f90): atom, type and layer indices are 1-based;
 * coordinates, momenta and forces are double[N][3]; errors print
 * "Error in <task>: <msg>." on stderr and exit(1) (src/global.f90:51-56).
 */
#ifndef EMDEE_H
#define EMDEE_H

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#endif

typedef struct {
  int    Builds;             /* Number of neighbor-list builds */
  struct {
    double Pair;             /* Time taken in force calculations */
    double Motion;
    double Neighbor;
    double Total;            /* Total time since initialization */
  } Time;
  struct {
    double Potential;        /* Total potential energy of the system */
    double Dispersion;       /* Dispersion (vdW) part of the potential energy */
    double Coulomb;          /* Electrostatic part of the potential energy */
    double Bond;
    double Angle;
    double Dihedral;
    double ShadowPotential;
    bool   UpToDate;         /* Flag to attest whether energies have been computed */
  } Energy;
  struct {
    double Kinetic;          /* Total kinetic energy of the system */
    double TransPart[3];     /* Translational kinetic energy at each dimension */
    double Rotational;       /* Rotational kinetic energy of the system */
    double RotPart[3];       /* Rotational kinetic energy around each principal axis */
    double ShadowKinetic;
    double ShadowRotational;
    bool   UpToDate;         /* Flag to attest whether kinetic energies are current */
  } Kinetic;
  double Virial;             /* Total internal virial of the system */
  double BodyVirial;         /* Rigid body contribution to the internal virial */
  int    DoF;                /* Total number of degrees of freedom */
  int    RotDoF;             /* Number of rotational degrees of freedom */
  void*  Data;               /* Pointer to system data */
  struct {
    bool   Translate;        /* Flag to activate/deactivate translations */
    bool   Rotate;           /* Flag to activate/deactivate rotations */
    int    RotationMode;     /* 0 = exact free rotation, n > 0 = NO_SQUISH with n substeps */
    bool   AutoForceCompute; /* Present in the reference headers; not read by the engine */
    bool   AutoBodyUpdate;   /* Re-derive rigid-body state on coordinate upload */
    bool   Compute;          /* Flag to activate/deactivate energy computations */
  } Options;
} tEmDee;

/* ---- system and models: src/EmDeeCode.f90 -------------------------------- */

/* EmDeeCode.f90:69-208 */
tEmDee EmDee_system( int threads, int layers, double rc, double skin, int N, int* types,
                     double* masses, int* bodies );

/* EmDeeCode.f90:212-235 ("coordinates" | "momenta" | "forces" | "layerForces") */
void* EmDee_memory_address( tEmDee md, char *option );

/* EmDeeCode.f90:239-269 */
void EmDee_share_phase_space( tEmDee mdkeep, tEmDee* mdlose );

/* EmDeeCode.f90:273-305 */
void EmDee_layer_based_parameters( tEmDee md, double InternalRc, int* Apply, int* Bonded );

/* EmDeeCode.f90:309-359 */
void EmDee_set_pair_model( tEmDee md, int itype, int jtype, void* model, double kCoul );

/* EmDeeCode.f90:363-412 */
void EmDee_set_pair_multimodel( tEmDee md, int itype, int jtype, void* model[], double kCoul[] );

/* EmDeeCode.f90:416-444 */
void EmDee_set_kspace_model( tEmDee md, void* model );

/* EmDeeCode.f90:448-482 */
void EmDee_set_coul_model( tEmDee md, void* model );

/* EmDeeCode.f90:486-520 */
void EmDee_set_coul_multimodel( tEmDee md, void* model[] );

/* EmDeeCode.f90:524-570 */
void EmDee_ignore_pair( tEmDee md, int i, int j );

/* EmDeeCode.f90:574-597 */
void EmDee_add_bond( tEmDee md, int i, int j, void* model );

/* EmDeeCode.f90:600-626 */
void EmDee_add_angle( tEmDee md, int i, int j, int k, void* model );

/* EmDeeCode.f90:628-655 */
void EmDee_add_dihedral( tEmDee md, int i, int j, int k, int l, void* model );

/* EmDeeCode.f90:659-801 */
void EmDee_download( tEmDee md, char *option, double* address );

/* EmDeeCode.f90:805-925 */
void EmDee_upload( tEmDee* md, char *option, double* address );

/* EmDeeCode.f90:929-946 */
void EmDee_switch_model_layer( tEmDee* md, int layer );

/* EmDeeCode.f90:950-1020 */
void EmDee_random_momenta( tEmDee* md, double kT, bool adjust, int seed );

/* EmDeeCode.f90:1024-1065 */
void EmDee_boost( tEmDee* md, double lambda, double alpha, double dt );

/* EmDeeCode.f90:1069-1103 */
void EmDee_displace( tEmDee* md, double lambda, double alpha, double dt );

/* Alias of EmDee_displace under the name used for the drift step in some
   host code; no separate symbol exists in the reference snapshot. */
void EmDee_move( tEmDee* md, double lambda, double alpha, double dt );

/* EmDeeCode.f90:1107-1211 */
void EmDee_verlet_step( tEmDee* md, double dt );

/* EmDeeCode.f90:1215-1277 */
void EmDee_compute_forces( tEmDee* md );

/* EmDeeCode.f90:1281-1395 */
void EmDee_rdf( tEmDee md, int bins, double Rc, int pairs, int itype[], int jtype[], double g[] );

/* ---- potential modifiers: src/modelClass_nonbonded.f90:83-241 ------------ */

void* EmDee_shifted( void* model );
void* EmDee_shifted_force( void* model );
void* EmDee_smoothed( void* model, double skin );
void* EmDee_shifted_smoothed( void* model, double skin );
void* EmDee_square_smoothed( void* model, double skin );
void* EmDee_shifted_square_smoothed( void* model, double skin );

/* ---- "none" models: modelClass_pair.f90:146, modelClass_coul.f90:138,
        modelClass_bond.f90:57, modelClass_angle.f90:57, modelClass_dihedral.f90:58 ---- */

void* EmDee_pair_none( void );
void* EmDee_coul_none( void );
void* EmDee_bond_none( void );
void* EmDee_angle_none( void );
void* EmDee_dihedral_none( void );

/* ---- model constructors (src/make_models_module.sh: name = file stem,
        parameters = "!<>"-tagged fields in declaration order) ------------- */

void* EmDee_pair_lj_cut( double epsilon, double sigma );                      /* pair_lj_cut.f90:36-37 */
void* EmDee_pair_softcore_cut( double epsilon, double sigma, double lambda );  /* pair_softcore_cut.f90:41-43 */
void* EmDee_coul_cut( void );                                                  /* coul_cut.f90 */
void* EmDee_coul_sf( void );                                                   /* coul_sf.f90 */
void* EmDee_coul_damped( double damp );                                        /* coul_damped.f90:36 */
void* EmDee_coul_long( void );                                                 /* coul_long.f90 */
void* EmDee_coul_damped_smoothed( double damp, double skinWidth );             /* coul_damped_smoothed.f90:36-37 */
void* EmDee_coul_damped_square_smoothed( double damp, double skinWidth );      /* coul_damped_square_smoothed.f90:36-37 */
void* EmDee_coul_square_smoothed( double skinWidth );                          /* coul_square_smoothed.f90:35 */
void* EmDee_coul_shifted_square_smoothed( double skinWidth );                  /* coul_shifted_square_smoothed.f90:35 */
void* EmDee_bond_harmonic( double k, double r0 );                              /* bond_harmonic.f90:36-37 */
void* EmDee_angle_harmonic( double k, double theta0 );                         /* angle_harmonic.f90:36-37 */
/* E = k/2 (phi - phi0)^2: no reference model file (modelClass_dihedral.f90:28-45 defines the class) */
void* EmDee_dihedral_harmonic( double k, double phi0 );
void* EmDee_kspace_ewald( double accuracy );                                   /* kspace_ewald.f90:38 */

#ifdef __cplusplus
}
#endif

#endif /* EMDEE_H */
