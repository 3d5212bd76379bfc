/* fieldgrid batched sweep: many independent FDTD runs on one GPU in shared launches.
 *
 * The reference runs one problem per `run()` call (crates/sims-field/src/fdtd/run.rs:390-758);
 * a port sweep for S-parameters or a parameter sweep is a loop of such calls, each one a
 * sequence of small dispatches that leaves most of a large GPU idle. fg_run_batch takes the
 * whole list at once. Each run keeps its own problem, driven port, pulse, extras and record,
 * with the semantics of fg_run (include/fieldgrid.h): its steps, early stop on decay, and
 * outputs are those a single fg_run of it gives. It adds port V/I spectra computed on the
 * device, as the sims-antenna caller does on the host (crates/sims-antenna/src/fdtd/mod.rs:229-266).
 *
 * All runs advance together from step 0. Every pass is one launch per block class of the fused
 * march over the blocks of every run still going, plus one launch of the sparse fixes and one
 * of the observers for all of them; a pass advances one step, or two (temporal blocking) where
 * every run allows it. Chunks of passes are captured in a hipGraph and replayed. The march's
 * tile shape (rows per block, planes per block), the steps per pass and the graph chunk length
 * come from an autotuner, whose choices can be cached in a JSON file.
 */
#ifndef FIELDGRID_BATCH_H
#define FIELDGRID_BATCH_H

#include "fieldgrid.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tuning of one batch: in/out. Zero asks the autotuner; the choices used come back. */
typedef struct fg_tune {
    uint32_t tj;          /* rows per march block, 8 or 14; 0 = tune                           */
    uint32_t seg;         /* i planes per march block, 1..64; 0 = tune                         */
    uint32_t chunk;       /* steps per captured hipGraph (rounded up to an even number of
                             passes); 0 = tune; 1 = no graph (every pass launched directly)    */
    uint32_t depth;       /* steps per march pass (step-fusion depth): 1, or 2 = two steps per
                             pass over HBM where every run allows it (no inductors, plane DFT,
                             NTFF, and no observed Debye edge); 0 = tune                        */
    const char *cache;    /* JSON file of tuned choices per batch shape: read if present and
                             updated after tuning; NULL = no cache                              */
    /* out */
    uint32_t used_tj, used_seg, used_chunk, from_cache;
    uint32_t used_depth;
    uint32_t kept_steps;  /* steps the tuner's timed trials advanced every run by and the batch
                             went on from (0: the trials were rolled back, because a run would
                             have reached a decay check or its last step within them)          */
    double tune_ms;       /* time spent tuning (0 when the cache or the caller decided),
                             including the kept trial steps                                    */
    double step_ms;       /* mean time of one step of the whole batch, as the tuner measured it */
    uint32_t launch;      /* in: how two-step passes are launched: 1 = per pass (hipGraph chunks
                             or direct launches, as chunk says); 2 = persistent: one launch runs
                             many passes, each run going on to its next pass as soon as its own
                             previous one is done (switched off in this build: refused with
                             FG_UNSUPPORTED); 0 = tune                                          */
    uint32_t used_launch; /* out: the launch mode used (1 or 2)                                */
} fg_tune;

/* Port V/I spectra of one run (crates/sims-antenna/src/fdtd/mod.rs:231-246, per port):
   V(f) = sum_n V_n dt e^{-i w (n+1) dt},  I(f) = sum_n I_n dt e^{-i w (n+1/2) dt},  in f64. */
typedef struct fg_spectra {
    uint32_t nf;          /* number of frequencies                                              */
    uint32_t pad;
    const double *freqs;  /* Hz                                                                 */
    double *v;            /* out: [(port*nf + f)*2 + {re, im}], ports*nf*2 doubles              */
    double *i;            /* out: same layout                                                   */
} fg_spectra;

/* Run `count` problems together (a loop of run(), run.rs:390-758, in one batch).
   problems/driven/pulses/extras/records are arrays of `count`; `spectra` is NULL or an array of
   `count` (an entry with nf = 0 is skipped). progress(user, steps, decay) is called on the
   calling thread whenever any run reaches a decay check, with the batch's step count and the
   smallest decay among the runs still going; returning 0 cancels the whole batch. Runs with
   resistive sheets or thin wires are batched too; they keep the batch at one step per pass and,
   for wires, at 14-row tiles. A run that goes unstable fails the
   batch with FG_UNSTABLE and "run <r>: the FDTD run went unstable". The runs' device state is
   built on up to 8 host threads that end before the batch steps; an invalid problem fails the
   batch with its own status and message, as fg_run would. */
int fg_run_batch(fg_device *dev, uint32_t count, const fg_problem *const *problems, const uint32_t *driven,
                 const fg_pulse *pulses, const fg_extras *const *extras, fg_tune *tune, fg_progress_fn progress,
                 void *user, fg_record *records, fg_spectra *spectra);

/* finish()'s spectra of one record (crates/sims-antenna/src/fdtd/mod.rs:229-249) on the device,
   as fg_spectra computes them for a batch: series as fg_record.series (steps * ports * 2
   floats), v and i as in fg_spectra ([(port*nf + f)*2 + {re, im}]). The caller forms Z = V/I and
   p_in = Re(V I*)/2 as finish() does. */
int fg_port_spectra(fg_device *dev, const float *series, uint64_t steps, uint32_t ports, double dt, uint32_t nf,
                    const double *freqs, double *v, double *i);

/* fg_port_spectra with no device allocation. It has no counterpart in the reference: it is for
   callers that cannot allocate device memory (code under HIP graph capture, tight serving or
   sweep loops, harnesses that forbid hipMalloc). Same arguments and bit-identical results; the
   series, frequencies and outputs are staged through a static 32 MiB device scratch that each
   GPU holds until the process ends, on the null stream. It refuses with FG_INVALID a call
   that does not fit: 8*steps*ports + 8*nf + 32*ports*nf bytes (each rounded up to 256) above
   32 MiB. Calls on one device share the scratch: within a process they are serialised by a
   lock; calls from two processes on one GPU use separate copies. */
int fg_port_spectra_static(fg_device *dev, const float *series, uint64_t steps, uint32_t ports, double dt,
                           uint32_t nf, const double *freqs, double *v, double *i);

#ifdef __cplusplus
}
#endif

#endif /* FIELDGRID_BATCH_H */
