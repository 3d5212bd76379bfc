/*
 * fieldgrid_sim.h — host-side problem builder, the C++ mirror of sims-field's FDTD problem API.
 *
 * In the reference this code is Rust and runs on the host before the GPU driver:
 *   crates/sims-field/src/fdtd/sim.rs:352-554   `Sim::try_new`  (coefficients, ports, Debye, lumped)
 *   crates/sims-field/src/fdtd/sim.rs:556-635   `Sim::thin_wire`
 *   crates/sims-field/src/fdtd/sim.rs:29-143    `axis`, `time_step`, `psi_len`
 *   crates/sims-field/src/fdtd/run.rs:61-127    `CoefSets`, `port_loop`
 *   crates/sims-field/src/fdtd/run.rs:205-376   `pack`
 *   crates/sims-field/src/fdtd/ntff.rs:11-59    `patches`
 * A Rust caller keeps its own `Sim` and `pack()` and fills an fg_problem directly; this
 * builder exists for C, C++ and Python callers, and it lifts the reference's host limits
 * (no f64 per-edge temporaries, multi-threaded, integer indices).
 */
#ifndef FIELDGRID_SIM_H
#define FIELDGRID_SIM_H

#include "fieldgrid.h"

#ifdef __cplusplus
extern "C" {
#endif

/* `Edge` (sim.rs:145-149) */
typedef struct fg_edge {
    uint32_t comp;
    uint32_t at[3];
} fg_edge;

/* `Element` (sim.rs:151-156) */
#define FG_RESISTOR 0
#define FG_CAPACITOR 1
#define FG_INDUCTOR 2

/* `Lumped` (sim.rs:158-163) */
typedef struct fg_lumped {
    uint32_t kind;      /* FG_RESISTOR / FG_CAPACITOR / FG_INDUCTOR */
    uint32_t n_edges;
    double value;       /* ohm / farad / henry */
    const fg_edge *edges;
} fg_lumped;

/* `PortDef` (sim.rs:165-171): columns are concatenated in `edges`, column c has column_len[c] edges. */
typedef struct fg_port_def {
    double r;
    uint32_t reference_above;
    uint32_t n_columns;
    const uint32_t *column_len;
    const fg_edge *edges;
} fg_port_def;

/* `Resistive` (sim.rs:208-215) */
typedef struct fg_resistive {
    uint32_t comp;
    uint32_t adaptive;
    uint64_t id;
    double r;
    uint64_t faces[2];
} fg_resistive;

/* one entry of `Media.wires` (sim.rs:236-241): (edge, radius) */
typedef struct fg_wire {
    fg_edge edge;
    double radius;
} fg_wire;

/* Everything `Sim::try_new` takes (sim.rs:352-360), flattened. */
typedef struct fg_sim_desc {
    const double *line[3];      /* Grid.x, Grid.y, Grid.z node lines (grid.rs:1-7)             */
    uint32_t len[3];
    uint32_t pml;               /* Grid.pml                                                    */
    /* `Materials` (sim.rs:217-222), one value per cell, cell id (i*(n1-1)+j)*(n2-1)+k.
       A NULL array takes Materials::new's default (eps 1, sigma 0, inf 1, debye 0).          */
    const float *eps;
    const float *sigma;
    const float *inf;
    const float *debye;
    /* The `pec(c, p)` closure as a mask over flat edges c*nn+id (nonzero = PEC); NULL = none. */
    const uint8_t *pec;
    uint32_t n_lumped;
    uint32_t n_ports;
    const fg_lumped *lumped;
    const fg_port_def *ports;
    uint32_t n_resistive;
    uint32_t n_wires;
    const fg_resistive *resistive;
    const fg_wire *wires;
    /* `Debye.poles` (sim.rs:243-246): n_poles (x = 1/tau, a) pairs. */
    uint32_t n_poles;
    uint32_t n_pole_sets;       /* extension, 0 or 1 = the reference's single global table      */
    const double *poles;        /* pole set s occupies pairs [set_first[s], set_first[s+1])    */
    const uint32_t *set_first;  /* n_pole_sets+1 entries, or NULL for one set                  */
    const uint8_t *debye_set;   /* per-cell pole set, or NULL = set 0                          */
    /* `Sheets` (sim.rs:224-234) */
    uint32_t sheet_states;
    uint32_t sheet_table_len;
    const float *sheet_table;
} fg_sim_desc;

typedef struct fg_sim fg_sim;

/* Sim::try_new (sim.rs:352-554). On error returns FG_INVALID with the reference's message. */
int fg_sim_new(const fg_sim_desc *desc, fg_sim **out);
void fg_sim_free(fg_sim *sim);
/* Sim.dt, Grid::dims */
double fg_sim_dt(const fg_sim *sim);
void fg_sim_dims(const fg_sim *sim, uint32_t n[3]);
/* Sim.ca / Sim.cb, 3*nn each (sim.rs:177-178) */
const float *fg_sim_ca(const fg_sim *sim);
const float *fg_sim_cb(const fg_sim *sim);
/* counts of sim.debye_edges, sim.inductors, sim.sheets, sim.wires */
void fg_sim_counts(const fg_sim *sim, uint64_t counts[4]);

/* pack (run.rs:205-376) + CoefSets (run.rs:61-100): fills `out` with tables owned by `sim`,
   valid until the next fg_sim_pack or fg_sim_free. `narrow` is the CoefSets narrow limit
   (0 = the reference's 65536). */
int fg_sim_pack(fg_sim *sim, uint64_t narrow, fg_problem *out);

/* patches (ntff.rs:11-59): writes count*13 u32 to `gpu`, count*3 f64 to `pos` and count f64
   to `area` (any may be NULL) and returns the patch count. */
uint64_t fg_patches(const fg_sim *sim, uint32_t inset, uint32_t *gpu, double *pos, double *area);

#ifdef __cplusplus
}
#endif

#endif /* FIELDGRID_SIM_H */
