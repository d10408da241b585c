"""Multiple-time-step (r-RESPA) integration over model layers, composed from
the reference API alone: EmDee_boost, EmDee_displace and
EmDee_switch_model_layer (EmDeeCode.f90:929-946, 1024-1103).

The reference's mechanism for a 1:n multiple time step is its model layers
(EmDeeCode.f90:273-305, 363-412, 486-520): layer `full` holds the complete
model and layer `fast` a cheap short-range one.  The full force splits as
F_full = (F_full - F_fast) + F_fast; the slow part is applied with the outer
step as two kicks, lambda = +1 on layer `full` and lambda = -1 on layer
`fast` (boost scales the layer's force by lambda, EmDeeCode.f90:1034-1036),
and the fast part drives n velocity-Verlet steps of dt/n.  Forces of a layer
are computed lazily by boost when its positions are stale, so the costly
layer is evaluated once per outer step.  With n = 1 the step is velocity
Verlet on layer `full` up to rounding.
"""


def respa_step(lib, md, dt, n, full=1, fast=2):
    """One outer step of length dt with n inner steps on layer `fast`."""
    h = 0.5 * dt
    lib.switch_model_layer(md, full)
    lib.boost(md, 1.0, 0.0, h)
    lib.switch_model_layer(md, fast)
    lib.boost(md, -1.0, 0.0, h)
    for _ in range(n):
        lib.boost(md, 1.0, 0.0, h / n)
        lib.displace(md, 1.0, 0.0, dt / n)
        lib.boost(md, 1.0, 0.0, h / n)
    lib.boost(md, -1.0, 0.0, h)
    lib.switch_model_layer(md, full)
    lib.boost(md, 1.0, 0.0, h)


def respa(lib, md, nsteps, dt, n, full=1, fast=2):
    """nsteps outer steps; returns the total energy (layer `full` potential
    plus kinetic) after each."""
    H = []
    for _ in range(nsteps):
        respa_step(lib, md, dt, n, full, fast)
        H.append(md.Energy.Potential + md.Kinetic.Total)
    return H
