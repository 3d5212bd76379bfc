/* fieldgrid discrete adjoint: d(objective)/d(eps_r) over a design box of cells.
 *
 * No counterpart in the reference, which has no gradient path. This entry point is for callers
 * that optimise materials: inverse design, and fitting eps to measurements.
 *
 * The contract is that of a torch.autograd.Function:
 *   forward(eps_box) -> series   E components at observation nodes after every step (fg_run's
 *                                FP32 engine; point probes, include/fieldgrid.h)
 *   backward(dseries) -> grad    dJ/d eps_r per box cell, given dJ/d series from the caller's
 *                                objective J(series)
 * The backward pass is the exact transpose of the discrete FDTD step: the curl updates, the CPML
 * psi recursions, the resistive ports (their (ca, cb)), the Debye ADE (cpu.rs:239-260; its pole
 * states are state), lumped inductors (cpu.rs:298-303; their currents are state) and thin-wire H
 * corrections (cpu.rs:270-273). The step is linear and time-invariant
 * in the fields, so the adjoint recursion runs backward in time on its own fields. The forward
 * fields enter only through E and the curl term on the design edges. The forward pass keeps a
 * snapshot of the fields every `every` steps. The backward pass recomputes one segment at a time
 * from its snapshot, taping the design-edge values, and runs the adjoint over that segment on a
 * second stream while the next (earlier) segment is recomputed. Memory is the snapshots plus two
 * segments' tapes. With fg_design.whole_tape the forward tapes every step instead and the
 * backward runs the adjoint alone: more memory where the design box is large against the grid,
 * no recompute.
 *
 * Supported problems: every Sim feature but resistive sheets, whose face weights follow the H^2
 * the run has accumulated (cpu.rs:169-193), so a step with sheets is not linear in the fields:
 * they give FG_UNSUPPORTED. The box must not touch a port, lumped or resistive edge, come within
 * two nodes of a thin wire (the wire rescales the edges around it), nor reach into the CPML
 * (FG_INVALID). Snapshots hold the pole states and inductor currents with the fields.
 *
 * fg_adjoint_forward64 runs the same problem in f64 with the box's edge coefficients evaluated
 * in f64 from a given eps_r: a smooth objective for central finite differences.
 */
#ifndef FIELDGRID_ADJOINT_H
#define FIELDGRID_ADJOINT_H

#include "fieldgrid.h"
#include "fieldgrid_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fg_design {
    uint32_t lo[3], hi[3];      /* design box [lo, hi) in cells per axis; cell (i,j,k) spans
                                   nodes i..i+1, j..j+1, k..k+1                               */
    uint32_t n_obs;
    uint32_t every;             /* steps between field snapshots (0 = 100)                     */
    const fg_point_probe *obs;  /* observed E components (field 0..2) at nodes               */
    uint64_t steps;             /* steps of every forward run                                   */
    uint32_t whole_tape;        /* 0: checkpoint-and-recompute (snapshots every `every` steps);
                                   1: the forward tapes the design edges of every step (8 B per
                                   design edge per step, in place of the snapshots) and the
                                   backward recomputes nothing; FG_TOO_LARGE if that tape does
                                   not fit in free device memory                               */
    uint32_t pad;
} fg_design;

typedef struct fg_adjoint fg_adjoint;

/* Build the problem `desc` (as fg_sim_new) with port `driven` excited by `pulse`. `desc` and
   the arrays it points at must stay valid until fg_adjoint_free; the eps array is copied. */
int fg_adjoint_new(fg_device *dev, const fg_sim_desc *desc, const fg_design *design, uint32_t driven,
                   const fg_pulse *pulse, fg_adjoint **out);
/* Forward run with snapshots (or, with whole_tape, the tape of every step). eps_box (k fastest,
   NULL = the values in desc or the last call) replaces the box's eps_r; series: steps * n_obs
   floats. */
int fg_adjoint_forward(fg_adjoint *adj, const float *eps_box, float *series);
/* Backward pass for the last forward run: dseries (steps * n_obs doubles, dJ/ds) -> grad
   (box cells, k fastest, dJ/d eps_r). */
int fg_adjoint_backward(fg_adjoint *adj, const double *dseries, double *grad);
/* f64 forward run of the same problem with the box's eps_r = eps_box (NULL = current). */
int fg_adjoint_forward64(fg_adjoint *adj, const double *eps_box, double *series);
/* Wall times (ms) of the last forward and backward passes, then the busy times (HIP events) of
   the backward's two streams: the recompute of the segments and the adjoint steps. The backward
   recomputes segment k-1 while the adjoint sweeps segment k, so the two overlap and their sum
   exceeds the backward's wall time (with whole_tape there is no recompute: 0). Also the number
   of snapshots kept (0 with whole_tape). */
int fg_adjoint_timing(const fg_adjoint *adj, double ms[4], uint64_t *snapshots);
void fg_adjoint_free(fg_adjoint *adj);

#ifdef __cplusplus
}
#endif

#endif /* FIELDGRID_ADJOINT_H */
