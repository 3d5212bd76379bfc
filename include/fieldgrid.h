/*
 * fieldgrid.h — C-ABI of the MI355X Yee-grid FDTD engine.
 *
 * The engine replaces the GPU driver of sims-field's FDTD path:
 *   crates/sims-field/src/fdtd/run.rs:390-758   `run(g, sim, driven, pulse, extras, progress)`
 *   crates/sims-field/src/fdtd/run.rs:378-388   `run_blocking`
 *   crates/sims-field/src/gpu.rs:36-110         `Gpu`, `Gpu::open_blocking`, `max_binding`
 *
 * Everything here is plain C: pointers, sizes, integers. Inputs are borrowed for the
 * duration of a call; outputs are caller-allocated. All device memory a call allocates
 * is freed before it returns, so a device handle holds no state between runs.
 *
 * Index conventions follow the reference (crates/sims-field/src/fdtd/sim.rs:8-10):
 *   dims n = [n0, n1, n2] are node counts; nn = n0*n1*n2;
 *   node id = (i*n1 + j)*n2 + k, so k (the third axis) is contiguous;
 *   field component c of node id lives at flat index f = c*nn + id.
 * Unlike the reference's packed tables, which carry indices as f32 (run.rs:243-340)
 * and so refuse ids >= 2^24 (run.rs:341-343), every index here is an integer.
 */
#ifndef FIELDGRID_H
#define FIELDGRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes; fg_last_error() holds the message ------------------------------ */
enum fg_status {
    FG_OK = 0,
    FG_CANCELLED = 1,   /* progress returned 0: "cancelled" (run.rs:742-744)                    */
    FG_UNSTABLE = 2,    /* non-finite energy: "the FDTD run went unstable" (run.rs:736-738)     */
    FG_TOO_LARGE = 3,   /* "{cells} cells is more than this GPU can hold" (run.rs:401-403)      */
    FG_UNSUPPORTED = 4, /* a feature an entry point does not run (caller falls back to CPU)    */
    FG_HIP = 5,         /* a HIP runtime error                                                 */
    FG_INVALID = 6      /* malformed input (bad index, size mismatch, NULL output buffer)       */
};

/* ---- packed problem: the device-side mirror of `Packed` (run.rs:178-203) ------------- */

/* One port (probe_def, run.rs:305-340, plus its port_src range, sim.rs:510-523). */
typedef struct fg_port {
    uint32_t comp;         /* axis of the port's edges (probe_def[0])                  */
    uint32_t first_edge;   /* range into port_edges (probe_def[1..3])                  */
    uint32_t n_edges;
    uint32_t first_loop;   /* range into loops (probe_def[3..5])                       */
    uint32_t n_loops;
    uint32_t first_source; /* range into sources: this port's sim.port_src entries     */
    uint32_t n_sources;
    uint32_t pad;
} fg_port;

typedef struct fg_port_edge { /* V = -sum E[comp*nn+id]*w  (cpu.rs:318-322) */
    uint64_t id;
    float w;
    uint32_t pad;
} fg_port_edge;

typedef struct fg_loop_term { /* I = sum H[comp*nn+id]*w  (cpu.rs:324-328) */
    uint64_t id;
    uint32_t comp; /* H component 0..2 (the reference stores 3+comp) */
    float w;
} fg_loop_term;

typedef struct fg_source { /* E[comp*nn+id] -= src * pulse((n+1/2)dt)  (cpu.rs:294-297) */
    uint64_t id;
    uint32_t comp;
    float src;
} fg_source;

typedef struct fg_inductor { /* x = E - k_e*I; E = x; I += k_i*x  (cpu.rs:298-303) */
    uint64_t id;
    uint32_t comp;
    float k_e;
    float k_i;
    float i0; /* initial current (0 in the reference) */
} fg_inductor;

typedef struct fg_debye_edge { /* (c*nn+id, dd) of sim.debye_edges (run.rs:291-294) */
    uint64_t f;
    float dd;
    uint32_t set; /* pole set; the reference has a single global set 0 (run.rs:296-301) */
} fg_debye_edge;

typedef struct fg_wire_h { /* H[h] += sum_t c[t]*E[e[t]]  (cpu.rs:270-273, sim.rs:556-635) */
    uint64_t h;
    uint64_t e[4];
    float c[4];
} fg_wire_h;

typedef struct fg_problem {
    uint32_t n[3];
    uint32_t pml;
    double dt;                  /* Sim.dt (kernels use it cast to f32, as run.rs:440)          */
    const float *axes;          /* (n0+n1+n2) rows of 10 f32, pack() order (run.rs:226-242):
                                   inv_d, inv_dd, raw_inv_d, raw_inv_dd, be, ce, bh, ch,
                                   slot_e, slot_h (slots as f32 integers, -1 = none)           */
    /* CoefSets (run.rs:61-100): sets[0] = (0,0); index over flat f = c*nn+id, either
       u16 pairs packed in u32 (wide = 0) or one u32 per edge (wide = 1).                     */
    uint32_t coef_wide;
    uint32_t n_sets;
    const float *coef_sets;     /* n_sets x (ca, cb)                                         */
    const uint32_t *coef_index; /* wide ? 3*nn : ceil(3*nn/2) words                          */
    /* ports and their V/I probes */
    uint32_t n_ports;
    uint32_t n_sources;
    uint32_t n_port_edges;
    uint32_t n_loops;
    const fg_port *ports;
    const fg_port_edge *port_edges;
    const fg_loop_term *loops;
    const fg_source *sources;
    /* lumped inductors */
    uint32_t n_inductors;
    uint32_t pad0;
    const fg_inductor *inductors;
    /* Debye ADE: n_pole_sets tables of n_poles (ka, kb) pairs each (run.rs:296-301);
       a set shorter than n_poles is padded with (0, 0), which leaves its state at 0.          */
    uint64_t n_debye;
    const fg_debye_edge *debye_edges;
    uint32_t n_poles;
    uint32_t n_pole_sets;
    const float *debye_table;   /* n_pole_sets * n_poles * 2                                  */
    /* resistive sheets (run.rs:259-290) and thin wires (run.rs:251-258)                     */
    uint64_t n_sheets;
    const float *sheets;        /* 16 f32 per sheet, pack() layout, ids as u32 bit patterns  */
    uint32_t sheet_states;
    uint32_t n_branches;
    const float *branches;      /* sim.sheet_table                                           */
    uint64_t n_wires;
    const fg_wire_h *wires;
} fg_problem;

/* `Pulse` (run.rs:47-50): Gaussian-modulated cosine, sim.rs:646-657. */
typedef struct fg_pulse {
    double f0;
    double fc;
} fg_pulse;

/* A point probe (an extension: the reference has none). field 0..2 = Ex,Ey,Ez; 3..5 = Hx,Hy,Hz.
   It records the field at node (i,j,k) after every step, like the port V/I series. */
typedef struct fg_point_probe {
    uint32_t field;
    uint32_t i, j, k;
} fg_point_probe;

#define FG_MAX_FREQS 4
#define FG_ENGINE_AUTO 0  /* fastest engine for the problem */
#define FG_ENGINE_SPLIT 1 /* separate H and E sweeps, in place (update_h/update_e, fdtd.wgsl:154-184) */
#define FG_ENGINE_FUSED 2 /* one fused H->E sweep with ping-pong buffers (fdtd.wgsl:186-330)        */

/* `Extras` (run.rs:34-45) plus point probes and an engine choice. */
typedef struct fg_extras {
    uint64_t min_steps;
    uint64_t max_steps;
    double decay_db;            /* +inf runs exactly max_steps (tests/fdtd.rs:117)            */
    uint32_t nf;                /* number of DFT frequencies, <= 4 (run.rs:347)              */
    uint32_t engine;            /* FG_ENGINE_*; the reference's fused_e = true maps to FUSED  */
    double freqs[FG_MAX_FREQS];
    int64_t plane_k;            /* -1 = no plane DFT                                          */
    uint64_t n_patches;
    const uint32_t *ntff;       /* 13 u32 per patch (ntff.rs:11-59), NULL when n_patches = 0 */
    uint64_t check_every;       /* 0 = 1000 (run.rs:681)                                      */
    uint32_t profile;           /* print mean kernel times to stderr (run.rs:133-176)         */
    uint32_t n_probes;
    const fg_point_probe *probes;
} fg_extras;

/* `Record` (run.rs:52-59). Every buffer is caller-allocated; capacities are in floats. */
typedef struct fg_record {
    uint64_t steps;             /* out: steps run (<= max_steps on an early stop)            */
    uint32_t ports;             /* out                                                       */
    uint32_t pad;
    double decay_db;            /* out                                                       */
    float *series;              /* [(n*ports + p)*2 + {V, I}]; cap >= max_steps*ports*2      */
    uint64_t series_cap;
    float *plane;               /* [((i*n1+j)*nf + f)*10 + ...]; cap >= n0*n1*nf*10          */
    uint64_t plane_cap;
    float *ntff;                /* [(g*nf + f)*12 + ...]; cap >= n_patches*nf*12             */
    uint64_t ntff_cap;
    float *probes;              /* [n*n_probes + q]; cap >= max_steps*n_probes               */
    uint64_t probes_cap;
    float *e_out;               /* optional final E, 3*nn (NULL = not wanted)                */
    float *h_out;               /* optional final H, 3*nn (NULL = not wanted)                */
} fg_record;

/* progress(user, steps, decay_db) once per chunk on the calling thread; 0 cancels
   (the reference's `progress: &mut dyn FnMut(usize, f64) -> bool`, run.rs:396, 742). */
typedef int (*fg_progress_fn)(void *user, uint64_t steps, double decay_db);

/* ---- device handle: replaces `Gpu` (gpu.rs:36-41) -------------------------------------- */
typedef struct fg_device fg_device;

/* Gpu::open / open_blocking (gpu.rs:86-108): open HIP device `ordinal`. */
int fg_open(int ordinal, fg_device **out);
/* drop(Gpu). Sessions, adjoint and far-field objects opened on `dev` are freed first: they
   hand their streams back to the device, which destroys them here. */
void fg_close(fg_device *dev);
/* Gpu.name (gpu.rs:36-41) */
const char *fg_name(const fg_device *dev);
/* Gpu::max_binding (gpu.rs:110) restated as the largest node count one run can hold. */
uint64_t fg_max_cells(const fg_device *dev);

/* run / run_blocking (run.rs:378-758): one FDTD run of `problem` driven at port `driven`.
   Blocking. Returns FG_OK and fills `out`, or a status with fg_last_error() set. */
int fg_run(fg_device *dev, const fg_problem *problem, uint32_t driven, const fg_pulse *pulse,
           const fg_extras *extras, fg_progress_fn progress, void *user, fg_record *out);

/* The checks fg_run makes on the host before it allocates anything, with no device needed:
   the reference refuses a problem it cannot run in pack() and run() (run.rs:341-347: port
   indices past f32's exact range; run.rs:401-403: more nodes than a binding holds). Here:
   FG_INVALID for out-of-grid ids, table ranges and NULL tables, and FG_TOO_LARGE for a grid
   whose n1*n2 planes exceed the 32-bit per-plane byte offsets of the march kernels
   (536 870 908 nodes per plane). Capacity against free device memory is checked by fg_run. */
int fg_problem_check(const fg_problem *problem, uint32_t driven, const fg_extras *extras);

/* Thread-local text of the last error (the reference returns Result<_, String>). */
const char *fg_last_error(void);

/* ---- sessions: run() split into its phases -------------------------------------------------
   The reference's run() allocates, uploads, steps in chunks and reads back in one call
   (run.rs:466-757); a session exposes those phases so a caller can keep the problem resident
   in HBM across calls (benchmarks, the adjoint's checkpoints). fg_run is built on them.    */
typedef struct fg_session fg_session;

/* Allocate and upload (run.rs:466-494); fields start at zero. */
int fg_session_new(fg_device *dev, const fg_problem *problem, uint32_t driven, const fg_pulse *pulse,
                   const fg_extras *extras, fg_session **out);
/* Enqueue `count` further steps on the session's stream (run.rs:689-722); asynchronous. */
int fg_session_step(fg_session *s, uint64_t count);
/* Per-step kernel timing with HIP events on the session's stream (run.rs:133-176 times
   kernels with timestamp queries). Returns the summed milliseconds since the last call of
   [step sweep(s), sparse fixes, observers] and the number of steps timed, then switches
   timing on (enable != 0) or off and restarts the totals. */
int fg_session_timing(fg_session *s, int enable, double ms[3], uint64_t *steps);
/* Steps the session's sweep advances per pass over HBM: 2 where only port sources and Debye
   media act between steps and no port or probe reads a Debye edge (temporal blocking), else 1.
   No reference counterpart (the reference runs one step per dispatch chain); for callers that
   account bytes or time per pass. */
int fg_session_steps_per_pass(const fg_session *s, uint32_t *out);
/* Wait for every step enqueued so far. */
int fg_session_sync(fg_session *s);
/* Sum of E^2 + (376.73 H)^2 over the grid, synchronously (run.rs:716-735). */
int fg_session_energy(fg_session *s, double *out);
/* Read series/plane/ntff/probes and optionally the fields into `out` (run.rs:749-757). */
int fg_session_read(fg_session *s, fg_record *out);
void fg_session_free(fg_session *s);
/* hipDeviceSynchronize on the handle's device. */
int fg_device_synchronize(fg_device *dev);

#ifdef __cplusplus
}
#endif

#endif /* FIELDGRID_H */
