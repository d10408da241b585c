/*
 * emdee_gpu.h -- extensions of the MI355X build beyond the reference C ABI.
 *
 * None of these entry points has a counterpart in the reference library
 * (src/EmDeeCode.f90); they exist for hosts that want to choose the
 * arithmetic form, inspect the neighbor list or time the device kernels.  A
 * host that uses only include/emdee.h never needs them.
 */
#ifndef EMDEE_GPU_H
#define EMDEE_GPU_H

#include "emdee.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic form of the pair loop: 0 = plain FP64 (default),
   1 = FP32 pair arithmetic on FP64-formed distances with FP64 accumulation
   of forces, energies and virials.  The environment variable
   EMDEE_PRECISION=mixed selects 1 for every new system. */
void EmDeeGPU_set_precision( tEmDee md, int mode );
int  EmDeeGPU_get_precision( tEmDee md );

/* Unordered pair set of the current Verlet list (pairs with r^2 below
   (rc+skin)^2 at the latest build): writes up to `max` pairs as 1-based
   (i, j) with i < j into pairs[2*max] and returns the total count. */
int EmDeeGPU_neighbor_pairs( tEmDee md, int max, int* pairs );

/* Kernel timing with HIP events on the system's stream.  After
   EmDeeGPU_profile(md, 1), EmDeeGPU_kernel_stats fills stats[0..6] with
   {pair-kernel ms (sum), pair launches, list-build ms (sum), list builds,
    neighbor-list capacity per atom, cells per box side,
    full-list entries at the latest build (twice the pair count)}. */
void EmDeeGPU_profile( tEmDee md, int on );
void EmDeeGPU_kernel_stats( tEmDee md, double* stats );

/* Batched hybrid Monte Carlo of independent Lennard-Jones replicas
   (pair_lj_cut, epsilon = sigma = mass = 1, cutoff rc) in one persistent
   launch per graph replay.  Each cycle of a replica draws momenta at its kT,
   runs nsteps velocity-Verlet steps of size dt (boost dt/2, displace dt,
   boost dt/2, as EmDee_boost/EmDee_displace) and accepts the trajectory with
   probability min(1, exp(-dH/kT)), restoring the coordinates otherwise.
   R holds 3 * sum(natoms) coordinates, replica after replica.  precision:
   0 = FP64 pair arithmetic, 1 = FP32 pair arithmetic with FP64 accumulation.
   No counterpart in the reference: a reference host loops over one tEmDee
   per replica. */
typedef struct EmDeeHMC EmDeeHMC;
EmDeeHMC* EmDeeHMC_create( int nrep, const int* natoms, const double* box, const double* kT, const double* R,
                           double rc, double dt, int nsteps, unsigned long long seed, int precision );
/* Times one cycle of every launch shape (workgroup size x cell division)
   from the same state, restores the state and keeps the fastest; returns
   its index.  EmDeeHMC_variant reports the choice and the six timings (ms,
   -1 for shapes that do not fit). */
int  EmDeeHMC_tune( EmDeeHMC* h );
void EmDeeHMC_variant( EmDeeHMC* h, int* blockSize, int* cellDivision, double* tuneMs );
/* Replays `launches` times a captured launch of `cycles` cycles per replica. */
void EmDeeHMC_run( EmDeeHMC* h, int launches, int cycles );
void EmDeeHMC_synchronize( EmDeeHMC* h );
void EmDeeHMC_get( EmDeeHMC* h, double* R, double* U, long long* accepted, double* lastDH );
/* Potential energies after each of the last `cycles` cycles (at most 1024
   and at most the cycles run), oldest first, U[c * nrep + replica], recorded
   on the device, so a host can run many cycles without reading back after
   each; returns the number of cycles copied. */
int  EmDeeHMC_history( EmDeeHMC* h, int cycles, double* U );
void EmDeeHMC_destroy( EmDeeHMC* h );
/* The run's random numbers: momenta (3n, unit mass) and Metropolis uniform
   of a (replica, cycle). */
void   EmDeeHMC_momenta( unsigned long long seed, int replica, int cycle, int n, double kT, double* P );
double EmDeeHMC_uniform( unsigned long long seed, int replica, int cycle );

/* Atom energies of the current layer in original atom order (evaluates the
   forces): half of each pair term of the atom, its k-space (reciprocal)
   energy and self energy, and its share of an attached learned model.  Bonded
   terms and the constant intrabody k-space correction are not included. */
void EmDeeGPU_atom_energies( tEmDee* md, double* e );

/* Learned per-atom energy model (radial-Gaussian descriptors, 32 per
   neighbour species within rcut, a 2 x 128 SiLU MLP per species on BF16 MFMA
   with FP32 accumulation), trained on per-atom energy labels.  Attached to a
   model layer it adds its energy, forces and virial to that layer, standing
   in for a costly term (e.g. the k-space electrostatics).  rcut must not
   exceed the system's cutoff.  No counterpart in the reference. */
typedef struct EmDeeSurrogate EmDeeSurrogate;
EmDeeSurrogate* EmDeeSurrogate_create( int ntypes, double rcut, unsigned long long seed );
/* Optional exact long-range channel: the model adds the reciprocal Ewald sum
   over the k-vectors with |k| <= kcut (1/distance; the system's kspace model
   gives the splitting parameter and charges), and the network learns the
   labels minus that sum's atom shares.  Set before the first sample; 0
   disables it.  _long_range_vectors reports the number of k-vectors in use. */
void   EmDeeSurrogate_set_long_range( EmDeeSurrogate* s, double kcut );
int    EmDeeSurrogate_long_range_vectors( EmDeeSurrogate* s );
/* Stores the descriptors of md's current configuration (its neighbor list
   must be current, e.g. after a force evaluation) with one label per atom. */
void   EmDeeGPU_surrogate_sample( tEmDee md, EmDeeSurrogate* s, const double* labels );
int    EmDeeSurrogate_rows( EmDeeSurrogate* s );
/* Adam on the per-atom mean squared error of the normalised labels; returns
   the last epoch's loss.  The first call fixes the normalisation. */
double EmDeeSurrogate_train( EmDeeSurrogate* s, int epochs, int batchRows, double lr );
int    EmDeeSurrogate_param_count( EmDeeSurrogate* s );
void   EmDeeSurrogate_get_params( EmDeeSurrogate* s, float* params, float* norm );
void   EmDeeSurrogate_set_params( EmDeeSurrogate* s, const float* params, const float* norm );
void   EmDeeSurrogate_destroy( EmDeeSurrogate* s );
/* layer 1..layers, or s = NULL to detach */
void   EmDeeGPU_attach_surrogate( tEmDee md, EmDeeSurrogate* s, int layer );

/* EmDee_rdf (EmDeeCode.f90:1281-1395) with no memory allocation.  No
   counterpart in the reference: it is for callers that cannot allocate
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments, checks and results as EmDee_rdf (bit for
   bit); its device buffers live in a static 1 MiB array that each GPU holds
   until the process ends.  Limits: at most 64 type pairs per call, type
   indices up to 31, bins * symm1D(maxtype, maxtype) <= 65536 and
   bins * pairs <= 65536; larger requests end with an error.  Calls that
   overlap on one GPU share that array and race. */
void EmDeeGPU_rdf_noalloc( tEmDee md, int bins, double Rc, int pairs, int itype[], int jtype[], double g[] );

/* Waits for all device work of the system. */
void EmDeeGPU_synchronize( tEmDee md );

/* Releases every host and device resource of a system (the reference has
   no destroy call; see EmDeeCode.f90:69-208). */
void EmDeeGPU_destroy( tEmDee* md );

#ifdef __cplusplus
}
#endif

#endif /* EMDEE_GPU_H */
