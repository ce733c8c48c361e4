/*
 * The adjoint entry points (mgx_feautrier_adjoint_device,
 * mgx_feautrier_adjoint_width_device, mgx_image_adjoint_device,
 * mgx_image_adjoint_width_device, mgx_release_workspace) and the kernels they
 * launch: the solver's source compiled a second time with MGX_ADJOINT_TU, so
 * that the adjoint kernels get instruction-scheduling options of their own
 * (ADJOINT_FLAGS in the build).  Their reverse pass waits on the tape and on
 * gathers at 2 waves/SIMD; a scheduler that groups memory clauses and tracks
 * register pressure exactly takes the FP64 J adjoint of the 32^3 wind from
 * 77.5 to 72.8 ms, while the same options slow the forward solve, which stays
 * in mgx_solver.hip with the default scheduler.  The state the two units
 * share (kernel timing, per-call workspace, walk-capacity hints) is defined in
 * mgx_solver.hip.
 */
#define MGX_ADJOINT_TU 1
#include "mgx_solver.hip"
