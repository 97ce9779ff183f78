/*
 * bdem_adjoint.h - gradients of a drum run with respect to the contact
 * parameters, by a hand-written adjoint of the sphere step (libbdem.so).
 * No reference counterpart: the reference only steps forward
 * (Device/DeviceInterface.cu:1825-2513).
 *
 * A run is a configured sphere context (reference pair law,
 * Collision_Response.cuh:374-494; one particle type; cylinder world objects
 * with the rotating drum surface, Collision_Response.cuh:116-233, and flat
 * end plates, :504-526; no lifters, cones or discharge) stepped nsteps
 * times from the state it holds when the run is created.  Its parameters
 * theta are
 *   theta[0]  kinetic friction of the particle pair (PP.Fric_kinetic),
 *   theta[1]  particle restitution, giving PP.Cn as the deck reader derives
 *             it (Set_SpringConstantsParticleKN, Input/KSimulationData.cpp:
 *             955-1000: shifted restitution, reduced mass m/2),
 *   theta[2]  particle normal stiffness PP.Kn,
 *   theta[3]  rolling resistance (SimulationInfo::Roll_Res).
 * Its observables are, for each window of `window` consecutive steps, the
 * window's mean over its steps of: the bed's centre of mass x and y, and
 * the bed's inclination 0.5 atan2(2 Cxy, Cxx - Cyy) (radians), C the
 * mass-weighted covariance of the particle centres in the xy plane (the
 * drum's cross-section): the direction of the bed's long axis, which lies
 * along its free surface.  3 observables per window.
 *
 * bdem_adj_forward runs the steps (the library's own step, identical to
 * bdem_update_sim), saving the state every checkpoint_every steps.
 * bdem_adj_backward takes the cotangent of the observables (e.g. dL/dobs of
 * a loss) and returns dL/dtheta: it re-runs each checkpoint segment
 * recording every step's sorted input, contact lists and sort permutation,
 * then runs the adjoint of the steps in reverse.  The contact set is frozen
 * within each step.  Gradients are accumulated in double.
 */
#ifndef BDEM_ADJOINT_H_
#define BDEM_ADJOINT_H_

#include "bdem_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdem_adjoint bdem_adjoint;

/* nsteps a multiple of window and of checkpoint_every.  The context must
 * hold an uploaded state and must not have stepped since; it is used by the
 * run (its state after a run is unspecified) and must outlive it. */
int bdem_adj_create(bdem_adjoint** out, bdem_ctx* ctx, int nsteps, int window, int checkpoint_every);
int bdem_adj_destroy(bdem_adjoint* a);
const char* bdem_adj_last_error(const bdem_adjoint* a);
/* 3 * nsteps / window */
int bdem_adj_num_observables(const bdem_adjoint* a);
/* theta as above; required before the first forward run. */
int bdem_adj_set_params(bdem_adjoint* a, const double* theta);
/* Run from the initial state; obs receives the observables. */
int bdem_adj_forward(bdem_adjoint* a, double* obs);
/* After bdem_adj_forward with the same theta: dtheta = (dobs/dtheta)^T cot. */
int bdem_adj_backward(bdem_adjoint* a, const double* cot, double* dtheta);
/* Milliseconds of the last forward and backward runs (HIP events on the
 * context's stream). */
int bdem_adj_timing(const bdem_adjoint* a, double* forward_ms, double* backward_ms);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* BDEM_ADJOINT_H_ */
