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

/* Waits for all device work of the system. */
void EmDeeGPU_synchronize( tEmDee md );

/* Releases every host and device resource of a system (the reference has
   no destroy call; see EmDeeCode.f90:69-208). */
void EmDeeGPU_destroy( tEmDee* md );

#ifdef __cplusplus
}
#endif

#endif /* EMDEE_GPU_H */
