// Batched sweep behind include/fieldgrid_batch.h: many independent FDTD runs stepped together.
//
// The reference runs one problem per run() (crates/sims-field/src/fdtd/run.rs:390-758). A sweep
// of small problems that way leaves an MI355X mostly idle: a 100^3 grid is ~50 march blocks on
// 256 CUs, and every step pays five launch latencies. Here every run keeps its own session
// (fields, psi, coefficient tables, ports, records: engine.hip), and one table of per-run
// descriptors (BatchRun) lets a single launch serve blocks of any run:
//   * the march blocks of all runs still going are merged into lists per kernel variant
//     (CPML class x coefficient-index width x LDS-staged sets), ordered for the XCDs;
//   * the sparse fixes and observers of all runs are one launch each (blockIdx.y = run);
//   * all runs start at step 0 and advance in lockstep, so they share one step counter and
//     one ping-pong parity; a run that stops leaves the lists;
//   * chunks of steps are captured in a hipGraph and replayed;
//   * rows per block, planes per block and the chunk length are autotuned per batch shape and
//     cached in a JSON file.
// Each run's arithmetic is the single-run engine's, node for node, so a batched run's record
// is bit-identical to an fg_run of the same problem.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fieldgrid_batch.h"
#include "host.h"

namespace fg {

// Everything a kernel needs to step one run, reachable through one table index.
struct BatchRun {
    Geo g;
    float *E[2], *H[2], *psi[2];
    FixArgs fix;
    ObsArgs obs;
    // two-step passes: the first step's sources and Debye ADE (march2), and the sparse fixes
    // of the second step, which read the pole states and E the pass wrote
    Src2 sr;
    Deb2 db;
    FixArgs fix2;
    WireArgs wa;     // thin wires (wa.w NULL: none): the run's blocks take the wire march
    SheetArgs sheet;  // resistive sheets, corrected after the march (sheet.n 0: none)
    Cap2 cp;          // two-step runs: step n's port, loop and probe values, captured in the pass
};

struct BatchBlock {
    uint32_t bx, by, bz, run;
};

template <int TJ, int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ + 2)) void k_march_batch(const BatchRun *__restrict__ runs,
                                                               const BatchBlock *__restrict__ blocks, uint32_t par,
                                                               uint32_t seg) {
    const BatchBlock bb = blocks[blockIdx.x];
    if (bb.bx == ~0u) return;  // padding of the XCD-ordered list (uniform over the block)
    const BatchRun &r = runs[bb.run];
    march<TJ, PML, WIDE, LDS_SETS>(r.g, MarchBlock{bb.bx, bb.by, bb.bz}, r.E[par], r.E[par ^ 1u], r.H[par],
                                   r.H[par ^ 1u], r.psi[par], r.psi[par ^ 1u], seg);
}

// The same for runs with thin wires: the H correction of every copy of a wire's H node inside
// the march, always 14 rows per block (k_march_wires of the single-run engine)
template <int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * 16) void k_march_wires_batch(const BatchRun *__restrict__ runs,
                                                               const BatchBlock *__restrict__ blocks, uint32_t par,
                                                               uint32_t seg) {
    const BatchBlock bb = blocks[blockIdx.x];
    if (bb.bx == ~0u) return;
    const BatchRun &r = runs[bb.run];
    march<14, PML, WIDE, LDS_SETS, true>(r.g, MarchBlock{bb.bx, bb.by, bb.bz}, r.E[par], r.E[par ^ 1u], r.H[par],
                                         r.H[par ^ 1u], r.psi[par], r.psi[par ^ 1u], seg, r.wa);
}

// Resistive sheets of the runs in `list` after their march (k_sheet of the single-run engine:
// the face weights from the step's input E and new H, then the sheet current out of the new E),
// one thread per sheet, or one thread for a run whose sheets share an edge
__global__ __launch_bounds__(256) void k_sheet_batch(const BatchRun *__restrict__ runs, const uint32_t *__restrict__ list,
                                                     uint32_t par) {
    const BatchRun &r = runs[list[blockIdx.y]];
    const SheetArgs &a = r.sheet;
    const float *Eo = r.E[par], *Hn = r.H[par ^ 1u];
    float *En = r.E[par ^ 1u];
    const uint64_t t = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    auto one = [&](uint64_t q) {
        const float *sh = a.sh + q * 16;
        sheet_before(a, q, Eo[__float_as_uint(sh[0])], Hn[__float_as_uint(sh[1])], Hn[__float_as_uint(sh[2])]);
        sheet_after(a, q, En);
    };
    if (a.serial) {
        if (t == 0)
            for (uint64_t q = 0; q < a.n; q++) one(q);
    } else if (t < a.n) {
        one(t);
    }
}

constexpr int TJ2 = 12;  // rows per block of the two-step march

// PML -1: the block list mixes the CPML classes, each block's class in the two top bits of its
// bz; the class instances share one LDS copy, and one launch per pass replaces three.
template <int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ2 + 4)) void k_march2_batch(const BatchRun *__restrict__ runs,
                                                                 const BatchBlock *__restrict__ blocks, uint32_t par,
                                                                 uint32_t seg, uint32_t s_off) {
    const BatchBlock bb = blocks[blockIdx.x];
    if (bb.bx == ~0u) return;
    const BatchRun &r = runs[bb.run];
    const uint32_t n = r.fix.step[0] + s_off;
    const float pl = n < r.fix.n_pulse ? r.fix.pulse[n] : pulse_at(r.fix, (float(n) + 0.5f) * r.fix.dt);
    __shared__ March2LdsEither<TJ2, LDS_SETS> lds;
    const MarchBlock mb{bb.bx, bb.by, bb.bz & BZ_SEG};
    const int cls = PML >= 0 ? PML : int(bb.bz >> 30);
    // a block whose tile holds a Debye edge (march_classes2's SIDE_BLOCK) runs the in-pass ADE,
    // every other block the lean march, which computes the same values there
#define FG_M2(C, S, L)                                                                                                  \
    march2_ahead<TJ2, C, WIDE, LDS_SETS, S>(L, r.g, mb, r.E[par], r.E[par ^ 1u], r.H[par], r.H[par ^ 1u], r.psi[par],  \
                                            r.psi[par ^ 1u], seg, r.sr, pl, r.fix, r.db, Cap2{})
    if (bb.bz & SIDE_BLOCK) {
        if (cls == 2) FG_M2(2, true, lds.side);
        else if (cls == 1) FG_M2(1, true, lds.side);
        else FG_M2(0, true, lds.side);
    } else {
        if (cls == 2) FG_M2(2, false, lds.lean);
        else if (cls == 1) FG_M2(1, false, lds.lean);
        else FG_M2(0, false, lds.lean);
    }
#undef FG_M2
}

__global__ __launch_bounds__(256) void k_fix_batch(const BatchRun *__restrict__ runs, const uint32_t *__restrict__ list,
                                                   uint32_t par, uint32_t s_off, uint32_t second) {
    const BatchRun &r = runs[list[blockIdx.y]];
    fix_step(r.g, second ? r.fix2 : r.fix, r.E[par ^ 1u], r.E[par], s_off,
             uint64_t(blockIdx.x) * blockDim.x + threadIdx.x);
}

// both steps' observers of a two-step pass (observe_pair; the batch's pairs capture nothing)
__global__ __launch_bounds__(256) void k_observe_pair_batch(const BatchRun *__restrict__ runs,
                                                            const uint32_t *__restrict__ list, uint32_t par,
                                                            uint32_t s_off) {
    const BatchRun &r = runs[list[blockIdx.y]];
    observe_pair(r.g, r.obs, r.fix, r.E[par], r.H[par], r.psi[par], r.E[par ^ 1u], r.H[par ^ 1u], s_off, blockIdx.x,
                 Cap2{}, threadIdx.x);
}

__global__ __launch_bounds__(256) void k_observe_batch(const BatchRun *__restrict__ runs,
                                                       const uint32_t *__restrict__ list, uint32_t par,
                                                       uint32_t s_off) {
    const BatchRun &r = runs[list[blockIdx.y]];
    observe(r.g, r.obs, r.E[par ^ 1u], r.H[par ^ 1u], s_off, blockIdx.x, threadIdx.x);
}

// ---- persistent two-step passes ----------------------------------------------------------------
//
// One launch runs `passes` two-step passes of every run it lists, with no grid-wide boundary
// between passes: a run's next pass starts as soon as its previous one is done, while other runs
// are still in theirs. The work items are the runs' march blocks (k_march2_batch's); the same
// list serves every pass. A run's pass ends with its tail: the second step's sparse fixes
// (k_fix_batch), then both steps' observers (k_observe_pair_batch, reading step n's port and
// probe values from the pass's captures). Within one run the order is that of the per-pass
// launches: all march blocks of pass p, then its tail, then pass p+1's blocks (which overwrite
// the copy pass p read, and read what its tail wrote). A per-run counter of finished work
// carries it: each block adds 1 when done, and the block that completes the run's pass runs the
// tail and adds 1 more, so a block of pass p waits until the counter reaches p * (blocks + 1).
//
// The items sit in one queue per XCD (a workgroup takes from queue blockIdx.x % xcds, the XCD
// the dispatcher places it on), each run's blocks split into contiguous tile ranges over the
// queues as xcd_order does, so neighbouring tiles share an L2. Every queue is ordered by pass,
// and a block depends only on blocks (and tails) of earlier passes. With every workgroup
// resident (a cooperative launch) the launch cannot deadlock: the blocks of the earliest
// unfinished pass have all they need, and each queue reaches its own ones before any later
// pass's. A wait that exceeds `limit` ticks of the wall clock sets the abort flag, which every
// workgroup checks before it takes an item and while it waits, so the launch ends and the host
// reports the error instead of the GPU hanging.
//
// Results are those of the per-pass launches bit for bit: each item runs the same body on the
// same inputs. Workgroups (possibly on different XCDs, whose L2s are separate) hand data on
// with agent-scope release fences before a count and acquire fences after a wait.

// Thread 0 takes the workgroup's next item from queue q and waits until its run has finished
// the passes before the item's. Returns false when the queue is drained or the launch aborted;
// else the item's position in `items`, its pass and its run.
__device__ __forceinline__ bool persist_take(const BatchBlock *__restrict__ items, const uint32_t *__restrict__ per_pass,
                                             uint32_t *next, uint32_t *done, uint32_t *abort, uint32_t base,
                                             uint32_t qlen, uint32_t passes, uint64_t limit, uint32_t &pos, uint32_t &p,
                                             uint32_t &run) {
    if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    const uint32_t idx = __hip_atomic_fetch_add(next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (idx >= qlen * passes) return false;
    p = idx / qlen;
    pos = base + idx - p * qlen;
    run = items[pos].run;
    const uint32_t need = p * per_pass[run];
    // bounded twice: by the wall clock, and by a poll count that does not depend on the clock's
    // rate being what the runtime reports (2^24 polls of at least one s_sleep each: over a second)
    const uint64_t t0 = uint64_t(wall_clock64());
    for (uint32_t polls = 0; __hip_atomic_load(&done[run], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need; polls++) {
        if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (uint64_t(wall_clock64()) - t0 > limit || polls >= (1u << 24)) {
            __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(8);
    }
    return true;
}

// The launch's arguments other than the run table and the item list, in device memory.
struct PersistArgs {
    const uint32_t *queues;    // queue x: items[queues[x] .. queues[x] + queues[xcds + x])
    const uint32_t *per_pass;  // per run: march blocks per pass + 1 (the tail)
    uint32_t *next;            // per queue: items taken (zero at launch)
    uint32_t *done;            // per run: blocks and tails finished in this launch (zero at launch)
    uint32_t *abort;           // nonzero: a wait timed out
    uint64_t limit;            // wall-clock ticks a wait may last
    uint32_t passes, par0, seg, xcds, s_base;  // s_base: step offset of the launch's first pass
};

// A copy of p the compiler cannot see through: loads through it are not hoisted out of a loop.
template <class T>
__device__ __forceinline__ const T *fresh(const T *p) {
    asm volatile("" : "+s"(p));
    return p;
}

// Nothing but three pointers is held in registers across the march: the run table and the item
// list (restrict kernel arguments, read with scalar loads as k_march2_batch reads them) and the
// other arguments' address. Those arguments are re-read where each part of the loop needs them,
// and what crosses a barrier (the item's position, pass and run, whether the block completed its
// run's pass) goes through LDS, made uniform by readfirstlane. The march body then keeps
// k_march2_batch's registers: it sits at the 128-VGPR limit of a 1024-thread block, and values
// held across it spilled. The tail reads step n's observers from captures for the same reason:
// recomputing them (observe_mid) next to the march made the march spill.
template <bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ2 + 4)) void k_march2_persist(const BatchRun *__restrict__ runs,
                                                                   const BatchBlock *__restrict__ items,
                                                                   const PersistArgs *args) {
    __shared__ March2Lds<TJ2, LDS_SETS, true> lds;
    __shared__ uint32_t s_go, s_pos, s_p, s_run, s_last;
    const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
    for (;;) {
        if (lead) {
            const PersistArgs &A = *fresh(args);
            const uint32_t q = blockIdx.x % A.xcds;
            uint32_t pos = 0, p = 0, run = 0;
            s_go = persist_take(items, A.per_pass, A.next + q, A.done, A.abort, A.queues[q], A.queues[A.xcds + q],
                                A.passes, A.limit, pos, p, run);
            s_pos = pos, s_p = p, s_run = run;
        }
        __syncthreads();
        if (!__builtin_amdgcn_readfirstlane(s_go)) break;  // queue drained or aborted
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        {
            const PersistArgs &A = *fresh(args);
            const uint32_t p = __builtin_amdgcn_readfirstlane(s_p);
            const BatchBlock bb = items[__builtin_amdgcn_readfirstlane(s_pos)];
            const BatchRun &r = runs[bb.run];
            const uint32_t par = (__builtin_amdgcn_readfirstlane(A.par0) + p) & 1u;
            const uint32_t s_off = __builtin_amdgcn_readfirstlane(A.s_base) + 2u * p;
            const uint32_t seg = __builtin_amdgcn_readfirstlane(A.seg);
            const uint32_t n = r.fix.step[0] + s_off;
            const float pl = n < r.fix.n_pulse ? r.fix.pulse[n] : pulse_at(r.fix, (float(n) + 0.5f) * r.fix.dt);
            const MarchBlock mb{bb.bx, bb.by, bb.bz & BZ_SEG};
            const int cls = int(bb.bz >> 30);
#define FG_M2(C)                                                                                                    \
    march2_ahead<TJ2, C, WIDE, LDS_SETS, true>(lds, r.g, mb, r.E[par], r.E[par ^ 1u], r.H[par], r.H[par ^ 1u],    \
                                               r.psi[par], r.psi[par ^ 1u], seg, r.sr, pl, r.fix, r.db, r.cp)
            if (cls == 2) FG_M2(2);
            else if (cls == 1) FG_M2(1);
            else FG_M2(0);
#undef FG_M2
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (lead) {
            const PersistArgs &A = *fresh(args);
            const uint32_t run = s_run;
            const uint32_t before = __hip_atomic_fetch_add(&A.done[run], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            s_last = before + 2u == (s_p + 1u) * A.per_pass[run];  // this block completed the run's pass
        }
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(s_last)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const PersistArgs &A = *fresh(args);
            const uint32_t p = __builtin_amdgcn_readfirstlane(s_p), run = __builtin_amdgcn_readfirstlane(s_run);
            const BatchRun &r = runs[run];
            const uint32_t par = (__builtin_amdgcn_readfirstlane(A.par0) + p) & 1u;
            const uint32_t s_off = __builtin_amdgcn_readfirstlane(A.s_base) + 2u * p;
            const uint32_t flat = threadIdx.y * 64 + threadIdx.x;
            const FixArgs &f2 = r.fix2;
            const uint64_t ns = f2.n_src > f2.n_ind ? f2.n_src : f2.n_ind, nf = f2.n_debye > ns ? f2.n_debye : ns;
            for (uint64_t t = flat; t < nf; t += 64 * (TJ2 + 4))
                fix_step(r.g, f2, r.E[par ^ 1u], r.E[par], s_off + 1, t);
            __syncthreads();
            const uint32_t nob = r.obs.blk_ports + r.obs.blk_probes;  // two-step batches: no plane DFT, NTFF
            for (uint32_t vb = flat / 256; vb < nob; vb += (64 * (TJ2 + 4)) / 256)
                observe_pair_ports_cap(r.g, r.obs, r.E[par ^ 1u], r.H[par ^ 1u], s_off, vb, r.cp, flat % 256);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (lead) __hip_atomic_fetch_add(&fresh(args)->done[run], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void k_set_persist_args(PersistArgs *dst, PersistArgs v) { *dst = v; }

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) p[t] = 0u;
}

__global__ void k_batch_set_step(uint32_t *step, uint32_t value) { step[0] = value; }
__global__ void k_batch_add_step(uint32_t *step, uint32_t value) { step[0] += value; }

// One (run, port, frequency) of the port spectra: the f64 phasor recursion of
// crates/sims-antenna/src/fdtd/mod.rs:231-246 over that port's (V, I) pairs.
struct SpecJob {
    const float *series;
    uint64_t steps;
    uint32_t ports, port;
    double dt, w;
};

// The sums of one (port, frequency): V and I phasors into out[0..4) = {V re, V im, I re, I im}.
__device__ __forceinline__ void spectrum(const float *__restrict__ series, uint64_t steps, uint32_t ports,
                                         uint32_t port, double dt, double w, double *__restrict__ out) {
    const double re = cos(-w * dt), ri = sin(-w * dt);   // rot_e = e^{-i w dt}
    double pr = dt * re, pi = dt * ri;                    // pe = dt e^{-i w dt}
    double qr = dt * cos(-w * 0.5 * dt), qi = dt * sin(-w * 0.5 * dt);  // ph
    double vr = 0, vi = 0, ir = 0, ii = 0;
    for (uint64_t s = 0; s < steps; s++) {
        const double v = series[(s * ports + port) * 2], c = series[(s * ports + port) * 2 + 1];
        vr += pr * v, vi += pi * v;
        ir += qr * c, ii += qi * c;
        const double pr2 = pr * re - pi * ri, pi2 = pr * ri + pi * re;
        const double qr2 = qr * re - qi * ri, qi2 = qr * ri + qi * re;
        pr = pr2, pi = pi2, qr = qr2, qi = qi2;
    }
    out[0] = vr, out[1] = vi, out[2] = ir, out[3] = ii;
}

__global__ __launch_bounds__(256) void k_spectra(const SpecJob *__restrict__ jobs, uint32_t n, double *__restrict__ out) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const SpecJob j = jobs[t];
    spectrum(j.series, j.steps, j.ports, j.port, j.dt, j.w, out + 4 * size_t(t));
}

// The spectra of one series with no job table: thread t = port * nf + f, written straight into
// the caller's layout (v and i at [t * 2 + {re, im}]). w is formed as the host forms it for
// k_spectra's jobs, so both kernels add the same terms.
__global__ __launch_bounds__(256) void k_spectra_flat(const float *__restrict__ series, uint64_t steps, uint32_t ports,
                                                      double dt, const double *__restrict__ freqs, uint32_t nf,
                                                      double *__restrict__ v, double *__restrict__ i) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ports * nf) return;
    double o[4];
    spectrum(series, steps, ports, t / nf, dt, 2.0 * 3.14159265358979323846 * freqs[t % nf], o);
    v[2 * size_t(t)] = o[0], v[2 * size_t(t) + 1] = o[1];
    i[2 * size_t(t)] = o[2], i[2 * size_t(t) + 1] = o[3];
}

// ---- host -----------------------------------------------------------------------------------

constexpr int N_LISTS = 24;  // thin wires (2) x CPML class (3) x wide index (2) x LDS sets (2)

// The persistent launches (k_march2_persist) are switched off: their first GPU run (a 5-run
// batch of small grids, tuned with persistent launches) did not finish, and its cause is not
// known. The per-pass launches serve every batch meanwhile, and the sessions build no observer
// captures (which only the persistent launches read).
constexpr bool PERSIST_ENABLED = false;

struct Batch {
    fg_device *dev = nullptr;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Session>> runs;
    DevBuf<BatchRun> table;
    DevBuf<uint32_t> step;
    DevBuf<BatchBlock> lists[N_LISTS], lists2[N_LISTS];  // one-step and two-step march blocks
    uint32_t list_n[N_LISTS] = {}, list2_n[N_LISTS] = {};
    // steps per march pass (1, or 2 when every run allows it) and the current ping-pong copy
    uint32_t depth = 1, par = 0;
    bool pairs_ok = false;
    uint32_t seg2 = 32, seg2_base = 32;
    uint32_t seg2_cap = SEG_MAX;  // the two-step march's longest segment for every run's planes
    DevBuf<uint32_t> fix_list, obs_list, sheet_list;
    uint32_t n_fix = 0, n_obs = 0, fix_gx = 0, obs_gx = 0, n_sheet = 0, sheet_gx = 0;
    bool wires = false;  // some run has thin wires: 14-row tiles, one step per pass
    uint32_t tj = 14, seg = 32, chunk = 32, xcds = 8;
    uint32_t seg_base = 32;  // tuned planes per block; plan() halves it while too few blocks remain
    uint32_t cus = 256;
    // persistent two-step launches (k_march2_persist): per key (index width x LDS sets) the XCD
    // queues of one pass, their offsets and lengths (2 x xcds), the runs' items per pass, and the
    // launch's counters
    bool persist = false;      // two-step passes go through persistent launches
    bool persist_ok = false;   // the device takes cooperative launches of the kernel
    uint32_t persist_grid = 0;
    uint64_t persist_limit = 0;  // wall-clock ticks a dependency wait may last
    DevBuf<BatchBlock> pitems[4];
    DevBuf<uint32_t> pq[4];
    uint32_t p_qmax[4] = {};   // longest queue of each key (0: the key has no runs)
    DevBuf<uint32_t> per_pass, pnext, pdone, pabort;
    DevBuf<PersistArgs> pargs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~Batch() {
        const bool idle = stream && hipStreamSynchronize(stream) == hipSuccess;  // an error may leave work in flight
        drop_graph();
        for (auto e : ev) (void)hipEventDestroy(e);
        if (idle) give_stream(dev, stream);
        else if (stream) (void)hipStreamDestroy(stream);
    }
};

// The persistent launches' queues for the runs still going (k_march2_persist): per key, XCD
// queue x holds, for each run in turn, the x-th contiguous range of each of its block classes
// (x/y-slab blocks, then z-slab, then interior: the longest first). The same queues serve every
// pass of a launch; per run, its blocks per pass plus one (the tail).
static void plan_persist(Batch &b, const std::vector<uint8_t> &active) {
    const uint32_t X = b.xcds;
    std::vector<uint32_t> pp(b.runs.size(), 0);
    std::vector<MarchBlock> cls[3];
    for (int key = 0; key < 4; key++) {
        std::vector<std::vector<BatchBlock>> mq(X);
        bool any = false;
        for (uint32_t r = 0; r < b.runs.size(); r++) {
            if (!active[r]) continue;
            const Session &s = *b.runs[r];
            if ((s.geo.wide ? 2 : 0) + (s.geo.n_sets <= uint32_t(SETS_LDS) ? 1 : 0) != key) continue;
            march_classes2(s, TJ2, b.seg2, cls);
            uint32_t m = 0;
            for (int c = 2; c >= 0; c--) {
                const size_t N = cls[c].size(), per = (N + X - 1) / X;
                for (size_t L = 0; L < N; L++) {
                    const MarchBlock &mb = cls[c][L];
                    mq[L / per].push_back(BatchBlock{mb.bx, mb.by, mb.bz | (uint32_t(c) << 30), r});
                }
                m += uint32_t(N);
            }
            pp[r] = m + 1;
            any = true;
        }
        std::vector<BatchBlock> items;
        std::vector<uint32_t> q(2 * X);
        uint32_t qmax = 0;
        for (uint32_t x = 0; x < X; x++) {
            q[x] = uint32_t(items.size());
            items.insert(items.end(), mq[x].begin(), mq[x].end());
            q[X + x] = uint32_t(items.size()) - q[x];
            qmax = std::max(qmax, q[X + x]);
        }
        b.p_qmax[key] = any ? std::max(qmax, 1u) : 0u;
        if (!any) continue;
        b.pitems[key].refill(items.data(), items.size(), b.stream);
        b.pq[key].refill(q.data(), q.size(), b.stream);
        FG_CHECK(hipStreamSynchronize(b.stream));  // items and q go out of scope
    }
    b.per_pass.refill(pp.data(), pp.size(), b.stream);
    FG_CHECK(hipStreamSynchronize(b.stream));  // the host vectors go out of scope
}

// Rebuild the merged block lists and the fix/observe run lists for the runs still going.
static void plan(Batch &b, const std::vector<uint8_t> &active) {
    b.drop_graph();
    std::vector<BatchBlock> L[N_LISTS], L2[N_LISTS];
    std::vector<MarchBlock> cls[3];
    std::vector<uint32_t> fl, ol, sl;
    b.fix_gx = b.obs_gx = b.sheet_gx = 0;
    // as runs finish, fewer blocks remain than the GPU holds: shorter segments (each block
    // marches fewer planes) keep ~6 blocks per CU in flight (3 for the two-step march, whose
    // blocks are twice as long-lived). The results do not depend on it.
    b.seg = b.seg_base;
    b.seg2 = std::min(b.seg2_base, b.seg2_cap);
    if (b.depth == 2)
        for (;;) {
            uint64_t blocks = 0;
            for (uint32_t r = 0; r < b.runs.size(); r++)
                if (active[r]) {
                    const Session &s = *b.runs[r];
                    blocks += uint64_t((s.n[2] + OWN_K2 - 1) / OWN_K2) * ((s.n[1] + TJ2 - 1) / TJ2) *
                              ((s.n[0] + b.seg2 - 1) / b.seg2);
                }
            if (blocks >= 3ull * b.cus || b.seg2 <= 4) break;
            b.seg2 /= 2;
        }
    for (;;) {
        uint64_t blocks = 0;
        for (uint32_t r = 0; r < b.runs.size(); r++)
            if (active[r]) {
                const Session &s = *b.runs[r];
                blocks += uint64_t((s.n[2] + OWN_K - 1) / OWN_K) * ((s.n[1] + b.tj - 1) / b.tj) * ((s.n[0] + b.seg - 1) / b.seg);
            }
        if (blocks >= 6ull * b.cus || b.seg <= 4) break;
        b.seg /= 2;
    }
    for (uint32_t r = 0; r < b.runs.size(); r++) {
        if (!active[r]) continue;
        const Session &s = *b.runs[r];
        const int key = (s.geo.wide ? 2 : 0) + (s.geo.n_sets <= uint32_t(SETS_LDS) ? 1 : 0);
        const int wl = s.wires.n ? 12 : 0;  // a run with thin wires takes the wire march's lists
        // one-step lists always (a two-step batch still takes an odd last step alone)
        march_classes(s, b.tj, b.seg, cls);
        for (int c = 0; c < 3; c++)
            for (const MarchBlock &m : cls[c]) L[wl + c * 4 + key].push_back(BatchBlock{m.bx, m.by, m.bz, r});
        if (s.sheet.n)
            sl.push_back(r), b.sheet_gx = std::max(b.sheet_gx, s.sheet.serial ? 1u : uint32_t((s.sheet.n + 255) / 256));
        if (b.depth == 2) {
            march_classes2(s, TJ2, b.seg2, cls, 0, true);
            for (int c = 0; c < 3; c++)
                for (const MarchBlock &m : cls[c]) L2[c * 4 + key].push_back(BatchBlock{m.bx, m.by, m.bz, r});
        }
        if (s.fix_blocks) fl.push_back(r), b.fix_gx = std::max(b.fix_gx, s.fix_blocks);
        if (s.obs_blocks) ol.push_back(r), b.obs_gx = std::max(b.obs_gx, s.obs_blocks);
    }
    if (b.depth == 2 && b.persist) plan_persist(b, active);
    for (int q = 0; q < N_LISTS; q++) {
        xcd_order(L[q], b.xcds, BatchBlock{~0u, ~0u, ~0u, ~0u});
        first_per_xcd(L2[q], b.xcds, [](const BatchBlock &m) { return (m.bz & SIDE_BLOCK) != 0; });
        xcd_order(L2[q], b.xcds, BatchBlock{~0u, ~0u, ~0u, ~0u});
    }
    // the two-step march runs each key's three classes in one mixed list (k_march2_batch<-1>),
    // class-tagged, the longest blocks (x/y, then z) first; each class list is padded to a
    // multiple of the XCD count, so the concatenation keeps its XCD order
    for (int key = 0; key < 4; key++) {
        std::vector<BatchBlock> mix;
        for (int c = 2; c >= 0; c--)
            for (BatchBlock m : L2[c * 4 + key]) {
                if (m.bx != ~0u) m.bz |= uint32_t(c) << 30;
                mix.push_back(m);
            }
        L2[key].swap(mix);
        L2[4 + key].clear(), L2[8 + key].clear();
    }
    for (int q = 0; q < N_LISTS; q++) {
        b.lists[q].refill(L[q].data(), L[q].size(), b.stream);
        b.list_n[q] = uint32_t(L[q].size());
        b.lists2[q].refill(L2[q].data(), L2[q].size(), b.stream);
        b.list2_n[q] = uint32_t(L2[q].size());
    }
    // the lists are rewritten in place (stream-ordered after the launches that read them)
    // as runs finish, so the batch allocates them once
    b.fix_list.refill(fl.data(), fl.size(), b.stream);
    b.obs_list.refill(ol.data(), ol.size(), b.stream);
    b.sheet_list.refill(sl.data(), sl.size(), b.stream);
    b.n_fix = uint32_t(fl.size());
    b.n_obs = uint32_t(ol.size());
    b.n_sheet = uint32_t(sl.size());
    FG_CHECK(hipStreamSynchronize(b.stream));
}

template <int TJ, int PML, bool W, bool L>
static void launch_list(Batch &b, int q, uint32_t par) {
    hipLaunchKernelGGL((k_march_batch<TJ, PML, W, L>), dim3(b.list_n[q]), dim3(64, TJ + 2), 0, b.stream, b.table.p,
                       b.lists[q].p, par, b.seg);
}

template <int PML, bool W, bool L>
static void launch_list2(Batch &b, int q, uint32_t par, uint32_t s_off) {
    hipLaunchKernelGGL((k_march2_batch<PML, W, L>), dim3(b.list2_n[q]), dim3(64, TJ2 + 4), 0, b.stream, b.table.p,
                       b.lists2[q].p, par, b.seg2, s_off);
}

static void launch_march2(Batch &b, uint32_t par, uint32_t s_off) {
    for (int q = 0; q < 4; q++) {  // the mixed lists, one per key (plan)
        if (!b.list2_n[q]) continue;
        switch (q) {
#define FG_CASE(Q, W, L) \
    case Q: launch_list2<-1, W, L>(b, q, par, s_off); break;
            FG_CASE(0, false, false) FG_CASE(1, false, true) FG_CASE(2, true, false) FG_CASE(3, true, true)
#undef FG_CASE
        }
    }
}

template <int PML, bool W, bool L>
static void launch_wires(Batch &b, int q, uint32_t par) {
    hipLaunchKernelGGL((k_march_wires_batch<PML, W, L>), dim3(b.list_n[q]), dim3(64, 16), 0, b.stream, b.table.p,
                       b.lists[q].p, par, b.seg);
}

template <int TJ>
static void launch_march(Batch &b, uint32_t par) {
    for (int q = 12; q < N_LISTS; q++) {  // runs with thin wires (b.tj is 14 whenever one exists)
        if (!b.list_n[q]) continue;
        switch (q - 12) {
#define FG_CASE(Q, PML, W, L) \
    case Q: launch_wires<PML, W, L>(b, q, par); break;
            FG_CASE(0, 0, false, false) FG_CASE(1, 0, false, true) FG_CASE(2, 0, true, false) FG_CASE(3, 0, true, true)
            FG_CASE(4, 1, false, false) FG_CASE(5, 1, false, true) FG_CASE(6, 1, true, false) FG_CASE(7, 1, true, true)
            FG_CASE(8, 2, false, false) FG_CASE(9, 2, false, true) FG_CASE(10, 2, true, false) FG_CASE(11, 2, true, true)
#undef FG_CASE
        }
    }
    for (int q = 0; q < 12; q++) {
        if (!b.list_n[q]) continue;
        switch (q) {
#define FG_CASE(Q, PML, W, L) \
    case Q: launch_list<TJ, PML, W, L>(b, q, par); break;
            FG_CASE(0, 0, false, false) FG_CASE(1, 0, false, true) FG_CASE(2, 0, true, false) FG_CASE(3, 0, true, true)
            FG_CASE(4, 1, false, false) FG_CASE(5, 1, false, true) FG_CASE(6, 1, true, false) FG_CASE(7, 1, true, true)
            FG_CASE(8, 2, false, false) FG_CASE(9, 2, false, true) FG_CASE(10, 2, true, false) FG_CASE(11, 2, true, true)
#undef FG_CASE
        }
    }
}

// One step of every run still going, at offset s_off from the device step counter.
static void launch_step(Batch &b, uint32_t s_off, uint32_t par) {
    if (b.tj == 8) launch_march<8>(b, par);
    else launch_march<14>(b, par);
    if (b.n_sheet)
        hipLaunchKernelGGL(k_sheet_batch, dim3(b.sheet_gx, b.n_sheet), dim3(256), 0, b.stream, b.table.p,
                           b.sheet_list.p, par);
    if (b.n_fix)
        hipLaunchKernelGGL(k_fix_batch, dim3(b.fix_gx, b.n_fix), dim3(256), 0, b.stream, b.table.p, b.fix_list.p, par,
                           s_off, 0u);
    if (b.n_obs)
        hipLaunchKernelGGL(k_observe_batch, dim3(b.obs_gx, b.n_obs), dim3(256), 0, b.stream, b.table.p, b.obs_list.p,
                           par, s_off);
}

template <bool W, bool L>
static void launch_persist_key(Batch &b, int key, uint32_t passes, uint32_t par, uint32_t s_base) {
    hipLaunchKernelGGL(k_zero_u32, dim3((uint32_t(b.runs.size()) + 255) / 256), dim3(256), 0, b.stream, b.pdone.p,
                       uint32_t(b.runs.size()));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, b.stream, b.pnext.p, b.xcds);
    // the arguments go to device memory in stream order (k_march2_persist reads them there)
    const PersistArgs pa{b.pq[key].p, b.per_pass.p, b.pnext.p, b.pdone.p, b.pabort.p, b.persist_limit,
                         passes, par, b.seg2, b.xcds, s_base};
    hipLaunchKernelGGL(k_set_persist_args, dim3(1), dim3(1), 0, b.stream, b.pargs.p, pa);
    const BatchRun *runs = b.table.p;
    const BatchBlock *items = b.pitems[key].p;
    const PersistArgs *args = b.pargs.p;
    // every entry is the address of one argument of k_march2_persist, in its order
    void *kargs[] = {&runs, &items, &args};
    FG_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_march2_persist<W, L>), dim3(b.persist_grid),
                                        dim3(64, TJ2 + 4), kargs, 0, b.stream));
}

// `passes` two-step passes of every run still going from step offset 0, as persistent launches:
// one per key, each of at most as many passes as keep its item count in 32 bits.
static void launch_persist(Batch &b, uint64_t passes, uint32_t par) {
    for (int key = 0; key < 4; key++) {
        if (!b.p_qmax[key]) continue;
        const uint64_t most = std::max<uint64_t>(1, std::min<uint64_t>(0xFFFFFFFFull / b.p_qmax[key], 1u << 16));
        for (uint64_t done = 0; done < passes;) {
            const uint32_t n = uint32_t(std::min<uint64_t>(most, passes - done));
            const uint32_t pr = par ^ uint32_t(done & 1u), s_base = uint32_t(2 * done);
            switch (key) {
                case 0: launch_persist_key<false, false>(b, key, n, pr, s_base); break;
                case 1: launch_persist_key<false, true>(b, key, n, pr, s_base); break;
                case 2: launch_persist_key<true, false>(b, key, n, pr, s_base); break;
                default: launch_persist_key<true, true>(b, key, n, pr, s_base); break;
            }
            done += n;
        }
    }
}

// Two steps (s_off, s_off + 1) of every run still going in one pass of the two-step march.
static void launch_pair(Batch &b, uint32_t s_off, uint32_t par) {
    launch_march2(b, par, s_off);
    if (b.n_fix)
        hipLaunchKernelGGL(k_fix_batch, dim3(b.fix_gx, b.n_fix), dim3(256), 0, b.stream, b.table.p, b.fix_list.p, par,
                           s_off + 1, 1u);
    if (b.n_obs)
        hipLaunchKernelGGL(k_observe_pair_batch, dim3(b.obs_gx, b.n_obs), dim3(256), 0, b.stream, b.table.p,
                           b.obs_list.p, par, s_off);
}

// Capture `chunk` steps starting from copy 0, ending on copy 0 with the counter advanced by
// chunk (an even number of passes: chunk is even for one step per pass, a multiple of 4 for two).
static void capture(Batch &b) {
    b.drop_graph();
    FG_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeRelaxed));
    if (b.depth == 2)
        for (uint32_t q = 0; q < b.chunk; q += 2) launch_pair(b, q, (q / 2) & 1u);
    else
        for (uint32_t q = 0; q < b.chunk; q++) launch_step(b, q, q & 1u);
    hipLaunchKernelGGL(k_batch_add_step, dim3(1), dim3(1), 0, b.stream, b.step.p, b.chunk);
    FG_CHECK(hipStreamEndCapture(b.stream, &b.graph));
    FG_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

// Enqueue `count` steps from step `from`: graph replays while on copy 0 and a whole chunk fits,
// else direct passes (two steps each at depth 2, a single step for an odd remainder).
static void advance(Batch &b, uint64_t from, uint64_t count, bool graphs) {
    if (b.persist && b.depth == 2) {
        // persistent launches for the pairs, then a single step for an odd remainder
        hipLaunchKernelGGL(k_batch_set_step, dim3(1), dim3(1), 0, b.stream, b.step.p, uint32_t(from));
        const uint64_t passes = count / 2;
        if (passes) launch_persist(b, passes, b.par);
        b.par ^= uint32_t(passes & 1u);
        if (count & 1u) {
            launch_step(b, uint32_t(2 * passes), b.par);
            b.par ^= 1u;
        }
        FG_CHECK(hipGetLastError());
        return;
    }
    uint64_t done = 0;
    const bool replay = graphs && b.chunk >= 2;
    while (done < count) {
        const uint64_t at = from + done;
        hipLaunchKernelGGL(k_batch_set_step, dim3(1), dim3(1), 0, b.stream, b.step.p, uint32_t(at));
        if (replay && b.par == 0 && count - done >= b.chunk) {
            if (!b.exec) capture(b);
            while (count - done >= b.chunk) {
                FG_CHECK(hipGraphLaunch(b.exec, b.stream));
                done += b.chunk;
            }
            continue;
        }
        // direct passes: one to get back to copy 0 where replays can start, or the remainder
        // shorter than a chunk
        const uint64_t rem = count - done;
        uint64_t q = 0;
        const bool one = replay && b.par != 0 && rem >= b.chunk;
        while (q < rem) {
            if (b.depth == 2 && rem - q >= 2) {
                launch_pair(b, uint32_t(q), b.par);
                q += 2;
            } else {
                launch_step(b, uint32_t(q), b.par);
                q += 1;
            }
            b.par ^= 1u;
            if (one) break;
        }
        done += q;
    }
    FG_CHECK(hipGetLastError());
}

static hipEvent_t event(Batch &b, size_t i) {
    while (b.ev.size() <= i) {
        hipEvent_t e;
        FG_CHECK(hipEventCreate(&e));
        b.ev.push_back(e);
    }
    return b.ev[i];
}

// Mean ms per step over `steps` steps after `warm` untimed ones (graphs as in a real run),
// from step `at`, which moves on by warm + steps.
static double time_steps(Batch &b, uint64_t &at, uint32_t warm, uint32_t steps, bool graphs) {
    advance(b, at, warm, graphs);
    at += warm;
    FG_CHECK(hipEventRecord(event(b, 0), b.stream));
    advance(b, at, steps, graphs);
    at += steps;
    FG_CHECK(hipEventRecord(event(b, 1), b.stream));
    FG_CHECK(hipEventSynchronize(event(b, 1)));
    float ms = 0.f;
    FG_CHECK(hipEventElapsedTime(&ms, event(b, 0), event(b, 1)));
    return double(ms) / steps;
}

// Whether the device takes the persistent launches, and their grid: every workgroup resident at
// once (a cooperative launch), one per CU slot that all four instances fit, a multiple of the
// XCD count so each queue has as many workgroups.
static void persist_setup(Batch &b, uint32_t count) {
    if (!PERSIST_ENABLED) return;
    int coop = 0, khz = 0;
    FG_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, b.dev->ordinal));
    FG_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b.dev->ordinal));
    if (!coop || khz <= 0 || b.xcds == 0) return;
    int least = 1 << 30;
    const void *fns[4] = {reinterpret_cast<const void *>(&k_march2_persist<false, false>),
                          reinterpret_cast<const void *>(&k_march2_persist<false, true>),
                          reinterpret_cast<const void *>(&k_march2_persist<true, false>),
                          reinterpret_cast<const void *>(&k_march2_persist<true, true>)};
    for (const void *f : fns) {
        int nb = 0;
        FG_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * (TJ2 + 4), 0));
        least = std::min(least, nb);
    }
    const uint32_t grid = uint32_t(std::max(least, 0)) * b.cus / b.xcds * b.xcds;
    if (!grid) return;
    b.persist_grid = grid;
    b.persist_limit = uint64_t(khz) * 2000ull;  // 2 s: a legitimate wait is one block's lifetime
    b.pnext.alloc(b.xcds);
    b.pdone.alloc(count);
    b.pabort.alloc(1);
    b.pargs.alloc(1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, b.stream, b.pabort.p, 1u);
    FG_CHECK(hipGetLastError());
    b.persist_ok = true;
}

// Fail the batch when a persistent launch gave up on a dependency wait (k_march2_persist).
static void check_abort(Batch &b) {
    if (!b.persist_ok) return;
    uint32_t flag = 0;
    FG_CHECK(hipMemcpyAsync(&flag, b.pabort.p, 4, hipMemcpyDeviceToHost, b.stream));
    FG_CHECK(hipStreamSynchronize(b.stream));
    if (flag)
        throw HipError{"fg_run_batch: a persistent pass waited too long for its inputs; the launch was abandoned",
                       FG_HIP};
}

// ---- tuning cache: {"<shape key>": [tj, seg, chunk, step_ms, depth, persistent], ...} ----------

static std::map<std::string, std::vector<double>> read_cache(const char *path) {
    std::map<std::string, std::vector<double>> m;
    if (!path) return m;
    std::ifstream f(path);
    if (!f) return m;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    size_t p = 0;
    while ((p = t.find('"', p)) != std::string::npos) {
        const size_t e = t.find('"', p + 1);
        if (e == std::string::npos) break;
        const std::string key = t.substr(p + 1, e - p - 1);
        const size_t lb = t.find('[', e), rb = t.find(']', e);
        if (lb == std::string::npos || rb == std::string::npos || rb < lb) break;
        std::vector<double> v;
        std::stringstream nums(t.substr(lb + 1, rb - lb - 1));
        std::string tok;
        while (std::getline(nums, tok, ',')) v.push_back(std::atof(tok.c_str()));
        if (v.size() >= 3) m[key] = v;
        p = rb + 1;
    }
    return m;
}

static void write_cache(const char *path, const std::map<std::string, std::vector<double>> &m) {
    if (!path) return;
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp);
        if (!f) return;  // an unwritable cache only loses the memo
        f << "{\n";
        size_t q = 0;
        for (const auto &kv : m) {
            f << "  \"" << kv.first << "\": [";
            for (size_t i = 0; i < kv.second.size(); i++) f << (i ? ", " : "") << kv.second[i];
            f << "]" << (++q < m.size() ? "," : "") << "\n";
        }
        f << "}\n";
    }
    std::rename(tmp.c_str(), path);
}

static std::string shape_key(const Batch &b) {
    uint64_t nodes = 0, psi = 0;
    uint32_t nmax = 0;
    for (const auto &s : b.runs) {
        nodes += s->nn;
        psi += s->psi[0].n;
        nmax = std::max({nmax, s->n[0], s->n[1], s->n[2]});
    }
    std::ostringstream k;
    k << b.dev->name << "|runs=" << b.runs.size() << "|nodes=" << nodes << "|psi=" << psi << "|nmax=" << nmax;
    return k.str();
}

// Pick the tile, depth and graph chunk (from the cache, or by timing candidates: at most 174
// steps, a 4-step probe, up to 9 candidates of 2 + 8 steps and the graph-chunk trials). The
// trial steps are real steps of every run, since every tiling gives bit-identical fields: when
// they end before first_check (the earliest decay check or last step over the runs), the runs
// keep them and the returned count is the step the batch goes on from; otherwise every run is
// reset and 0 is returned.
static uint64_t autotune(Batch &b, fg_tune *t, std::vector<uint8_t> &active, uint64_t first_check) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::string key = shape_key(b);
    auto cache = read_cache(t ? t->cache : nullptr);
    const bool ask_tj = !t || !t->tj, ask_seg = !t || !t->seg, ask_chunk = !t || !t->chunk;
    // step-fusion depth: two steps per pass only where every run allows it
    const bool ask_depth = (!t || !t->depth) && b.pairs_ok;
    b.depth = t && t->depth == 2 && b.pairs_ok ? 2 : 1;
    b.tj = t && t->tj && !b.wires ? (t->tj == 8 ? 8 : 14) : 14;
    const uint32_t seg_in = t && t->seg ? std::min<uint32_t>(std::max<uint32_t>(t->seg, 1), SEG_MAX) : 32;
    b.seg = b.seg_base = b.seg2 = b.seg2_base = seg_in;
    b.chunk = t && t->chunk ? t->chunk : 32;
    bool from_cache = false, tuned = false;
    double step_ms = 0.0;
    uint64_t at = 0;  // steps the trial passes have taken
    auto it = cache.find(key);
    // persistent launches (two steps per pass only): tuned unless the caller chose
    if (t && t->launch == 2 && !b.persist_ok)
        throw HipError{"fg_run_batch: persistent launches are not available (switched off, or the device takes no "
                       "cooperative launches)",
                       FG_UNSUPPORTED};
    const bool ask_launch = (!t || !t->launch) && b.persist_ok;
    b.persist = t && t->launch == 2 && b.persist_ok;
    const bool ask = ask_tj || ask_seg || ask_chunk || ask_depth || ask_launch;
    if (ask && it != cache.end()) {
        const std::vector<double> &v = it->second;
        const uint32_t depth = v.size() > 4 && uint32_t(v[4]) == 2 && b.pairs_ok ? 2 : 1;
        if (ask_depth) b.depth = depth;
        if (ask_tj) b.tj = uint32_t(v[0]) == 8 && !b.wires ? 8 : 14;
        if (ask_seg) {
            const uint32_t sg = std::min<uint32_t>(std::max<uint32_t>(uint32_t(v[1]), 1), SEG_MAX);
            b.seg = b.seg_base = b.seg2 = b.seg2_base = sg;
        }
        if (ask_chunk) b.chunk = std::max<uint32_t>(uint32_t(v[2]), 1);
        if (ask_launch) b.persist = v.size() > 5 && v[5] != 0.0;
        step_ms = v.size() > 3 ? v[3] : 0.0;
        from_cache = true;
    } else if (ask) {
        tuned = true;
        // tile shape and depth: direct launches, so only the kernels' own time differs. The
        // first candidate sets how many steps each one gets: about 40 ms of stepping per
        // candidate, an even number so two-step passes cover them exactly.
        double best = 1e30;
        uint32_t btj = b.tj, bseg = seg_in, bdepth = b.depth, steps = 0;
        for (uint32_t depth : {1u, 2u}) {
            if (!ask_depth && depth != b.depth) continue;
            for (uint32_t tj : {14u, 8u}) {
                if ((depth == 2 || b.wires) && tj != 14) continue;  // one tile height for these marches
                if (depth == 1 && !ask_tj && tj != b.tj) continue;
                // where two steps per pass are possible, one step per pass is only tried with
                // 14-row tiles (8-row tiles re-read more halo and have not won there)
                if (depth == 1 && tj == 8 && ask_depth && ask_tj) continue;
                // the caller's segment length is timed as given, also one outside the list
                const std::vector<uint32_t> segs = ask_seg ? std::vector<uint32_t>{32u, 16u, 8u} : std::vector<uint32_t>{seg_in};
                for (uint32_t seg : segs) {
                    b.depth = depth, b.tj = tj, b.seg = b.seg_base = b.seg2 = b.seg2_base = seg;
                    plan(b, active);
                    if (!steps) {
                        const double one = time_steps(b, at, 2, 2, false);
                        steps = 2 * std::min<uint32_t>(4, std::max<uint32_t>(1, uint32_t(20.0 / one)));
                    }
                    const double ms = time_steps(b, at, 2, steps, false);
                    if (ms < best) best = ms, btj = tj, bseg = seg, bdepth = depth;
                }
            }
        }
        b.depth = bdepth, b.tj = btj, b.seg = b.seg_base = b.seg2 = b.seg2_base = bseg;
        plan(b, active);
        step_ms = best;
        if (ask_chunk) {
            // graph chunk: replays only pay off where launches are a visible share of a step
            // (short steps); there pick the shorter chunk unless the longer one is 2 % faster,
            // since the batch shrinks as runs end and every shrink recaptures the graph
            b.chunk = 32;
            if (best < 1.0) {
                b.chunk = 8;
                b.drop_graph();
                const double c8 = time_steps(b, at, 8, 8, true);
                b.chunk = 32;
                b.drop_graph();
                const double c32 = time_steps(b, at, 32, 32, true);
                b.chunk = c32 < 0.98 * c8 ? 32 : 8;
                step_ms = std::min(step_ms, std::min(c8, c32));
            }
            b.drop_graph();
        }
        if (ask_launch && b.depth == 2) {
            b.chunk = b.chunk > 1 ? (b.chunk + 3) / 4 * 4 : b.chunk;  // whole pairs of passes per replay
            // persistent launches against the per-pass launches just chosen, over the same
            // number of steps (a whole graph chunk, at least 16)
            const uint32_t n = std::max<uint32_t>(16, (b.chunk + 3) / 4 * 4);
            b.persist = false;
            plan(b, active);
            const double per = time_steps(b, at, 4, n, b.chunk >= 2);
            b.persist = true;
            plan(b, active);
            const double pst = time_steps(b, at, 4, n, false);
            check_abort(b);
            b.persist = pst < per;
            step_ms = std::min(per, pst);
            b.drop_graph();
        }
        FG_CHECK(hipStreamSynchronize(b.stream));
        if (at >= first_check) {
            for (auto &s : b.runs) reset(*s);
            b.par = 0;
            at = 0;
        }
        cache[key] = {double(b.tj), double(b.depth == 2 ? b.seg2_base : b.seg_base), double(b.chunk), step_ms,
                      double(b.depth), b.persist ? 1.0 : 0.0};
        write_cache(t ? t->cache : nullptr, cache);
    }
    // replays start and end on copy 0: an even number of passes per chunk
    if (b.chunk > 1) {
        const uint32_t m = b.depth == 2 ? 4 : 2;
        b.chunk = (b.chunk + m - 1) / m * m;
    }
    plan(b, active);
    if (t) {
        t->used_tj = b.depth == 2 ? TJ2 : b.tj, t->used_seg = b.depth == 2 ? b.seg2_base : b.seg_base;
        t->used_chunk = b.chunk, t->from_cache = from_cache ? 1 : 0, t->used_depth = b.depth;
        t->step_ms = step_ms;
        t->used_launch = b.persist && b.depth == 2 ? 2u : 1u;
        t->kept_steps = uint32_t(at);
        t->tune_ms = tuned ? std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() : 0.0;
    }
    return at;
}

static void spectra(Batch &b, uint32_t count, fg_spectra *sp) {
    std::vector<SpecJob> jobs;
    struct Dest {
        double *v, *i;
        size_t at;
    };
    std::vector<Dest> dest;
    for (uint32_t r = 0; r < count; r++) {
        if (!sp || !sp[r].nf) continue;
        const Session &s = *b.runs[r];
        if (!sp[r].freqs || !sp[r].v || !sp[r].i) throw HipError{"fg_run_batch: spectra buffers are NULL", FG_INVALID};
        const uint64_t steps = std::min(s.steps, s.max_steps);
        for (uint32_t p = 0; p < s.n_ports; p++)
            for (uint32_t f = 0; f < sp[r].nf; f++) {
                jobs.push_back(SpecJob{s.series.p, steps, s.n_ports, p, s.dt, 2.0 * 3.14159265358979323846 * sp[r].freqs[f]});
                dest.push_back(Dest{sp[r].v, sp[r].i, (size_t(p) * sp[r].nf + f) * 2});
            }
    }
    if (jobs.empty()) return;
    DevBuf<SpecJob> dj;
    DevBuf<double> out;
    dj.upload(jobs.data(), jobs.size(), b.stream);
    out.alloc(4 * jobs.size());
    const uint32_t n = uint32_t(jobs.size());
    hipLaunchKernelGGL(k_spectra, dim3((n + 255) / 256), dim3(256), 0, b.stream, dj.p, n, out.p);
    FG_CHECK(hipGetLastError());
    std::vector<double> h(4 * jobs.size());
    FG_CHECK(hipMemcpyAsync(h.data(), out.p, h.size() * 8, hipMemcpyDeviceToHost, b.stream));
    FG_CHECK(hipStreamSynchronize(b.stream));
    for (size_t q = 0; q < jobs.size(); q++) {
        dest[q].v[dest[q].at] = h[4 * q], dest[q].v[dest[q].at + 1] = h[4 * q + 1];
        dest[q].i[dest[q].at] = h[4 * q + 2], dest[q].i[dest[q].at + 1] = h[4 * q + 3];
    }
}

// The runs' sessions, built on up to 8 host threads. A session's set-up is mostly host work
// (tables built, copied into pinned memory) and waits on its own stream, so the set-ups of
// different runs overlap. The first error is rethrown once every thread has stopped.
static void make_sessions(Batch &b, uint32_t count, const fg_problem *const *problems, const uint32_t *driven,
                          const fg_pulse *pulses, const fg_extras *const *extras) {
    std::vector<fg_extras> xs(count);
    for (uint32_t r = 0; r < count; r++) xs[r] = *extras[r], xs[r].engine = FG_ENGINE_FUSED;
    b.runs.resize(count);
    std::atomic<uint32_t> next{0};
    std::mutex m;
    std::exception_ptr first;
    auto work = [&] {
        try {
            FG_CHECK(hipSetDevice(b.dev->ordinal));
            for (uint32_t r; (r = next.fetch_add(1)) < count;)
                b.runs[r].reset(session_new(b.dev, problems[r], driven[r], &pulses[r], &xs[r], PERSIST_ENABLED));
        } catch (...) {
            std::lock_guard<std::mutex> lock(m);
            if (!first) first = std::current_exception();
            next.store(count);
        }
    };
    const uint32_t threads = std::min({count, std::max(1u, std::thread::hardware_concurrency()), 8u});
    std::vector<std::thread> pool;
    for (uint32_t q = 1; q < threads; q++) {
        try {
            pool.emplace_back(work);
        } catch (const std::system_error &) {
            break;  // fewer threads
        }
    }
    work();
    for (auto &t : pool) t.join();
    if (first) std::rethrow_exception(first);
}

}  // namespace fg

extern "C" int fg_run_batch(fg_device *dev, uint32_t count, const fg_problem *const *problems, const uint32_t *driven,
                            const fg_pulse *pulses, const fg_extras *const *extras, fg_tune *tune,
                            fg_progress_fn progress, void *user, fg_record *records, fg_spectra *sp) {
    using namespace fg;
    return guarded([&] {
        if (!dev || (count && (!problems || !driven || !pulses || !extras || !records)))
            throw HipError{"fg_run_batch: NULL argument", FG_INVALID};
        Batch b;
        b.dev = dev;
        FG_CHECK(hipSetDevice(dev->ordinal));
        b.stream = take_stream(dev);
        b.xcds = dev->cus >= 64 ? uint32_t(dev->cus / 32) : 1u;  // gfx950: 32 CUs per XCD
        b.cus = uint32_t(std::max(dev->cus, 1));
        std::vector<BatchRun> table(count);
        b.step.alloc(1);
        bool all_pairs = true;
        for (uint32_t r = 0; r < count; r++) {
            if (!problems[r] || !extras[r]) throw HipError{"fg_run_batch: NULL problem or extras", FG_INVALID};
        }
        {
            // the runs' sessions are built concurrently, so each one's own capacity check cannot
            // see the others': check their sum first, extras included
            double need = 0.0;
            for (uint32_t r = 0; r < count; r++)
                need += session_bytes(problems[r], true) + session_extra_bytes(problems[r], extras[r]);
            size_t free_b = 0, total_b = 0;
            FG_CHECK(hipMemGetInfo(&free_b, &total_b));
            if (need > double(free_b)) {
                char msg[160];
                std::snprintf(msg, sizeof msg, "the batch's %u runs need %.1f GB of device memory; %.1f GB are free",
                              count, need / 1e9, double(free_b) / 1e9);
                throw HipError{msg, FG_TOO_LARGE};
            }
        }
        make_sessions(b, count, problems, driven, pulses, extras);
        for (uint32_t r = 0; r < count; r++) {
            Session &s = *b.runs[r];
            BatchRun &t = table[r];
            t.g = s.geo;
            for (int c = 0; c < 2; c++) t.E[c] = s.E[c].p, t.H[c] = s.H[c].p, t.psi[c] = s.psi[c].p;
            t.fix = s.fix;
            t.obs = s.obs;
            t.fix.step = t.obs.step = b.step.p;  // lockstep: one counter for the whole batch
            t.sr = Src2{s.src2_start.p, s.src2_ent.p};
            t.db = Deb2{s.deb2_row.p, s.deb2_map.p, s.deb2_jn.p, s.deb2_eo.p};
            t.fix2 = t.fix;
            if (t.fix.n_debye) t.fix2.dsrc = s.deb2_jn.p, t.fix2.dold = s.deb2_eo.p;
            // the per-pass launches hand observers no captures: no plane DFT or NTFF in a pair
            if (!s.fuse2 || s.obs.plane_n || s.obs.n_patches) all_pairs = false;
            if (s.fuse2 && s.cap_obs.p)
                t.cp = Cap2{s.cap_start.p, s.cap_seg.p, s.cap_buf.p, s.cap_ntff.p, s.cap_plane.p, s.cap_obs.p,
                            s.cap_loop0, s.cap_probe0};
            b.seg2_cap = std::min(b.seg2_cap, march2_seg_cap(march2_plane_bytes(s.n, s.pml)));
            t.wa = WireArgs{s.wire_start.p, s.wire_ent.p, s.wires.n ? s.wires.p : nullptr};
            t.sheet = s.sheet;
            b.wires |= s.wires.n > 0;
        }
        b.pairs_ok = all_pairs && count > 0;
        b.table.upload(table.data(), table.size(), b.stream);
        if (b.pairs_ok) persist_setup(b, count);
        std::vector<uint8_t> active(count, 0);
        for (uint32_t r = 0; r < count; r++) active[r] = extras[r]->max_steps > 0;
        // per-run decay checks at multiples of check_every and at max_steps (run.rs:681-748)
        std::vector<uint64_t> every(count), pulse_end(count);
        for (uint32_t r = 0; r < count; r++) {
            every[r] = extras[r]->check_every ? extras[r]->check_every : 1000;
            const double t0 = 4.0 * std::sqrt(std::log(10.0)) / (3.14159265358979323846 * pulses[r].fc);
            pulse_end[r] = uint64_t(2.0 * t0 / problems[r]->dt);
        }
        auto next_check = [&](uint32_t r, uint64_t at) {
            return std::min((at / every[r] + 1) * every[r], extras[r]->max_steps);
        };
        uint64_t first_check = UINT64_MAX;
        for (uint32_t r = 0; r < count; r++)
            if (active[r]) first_check = std::min(first_check, next_check(r, 0));
        // the batch goes on from the tuner's trial steps when it kept them
        uint64_t global = count ? autotune(b, tune, active, first_check) : 0;
        const bool graphs = b.chunk >= 2;
        bool dirty = false;
        while (std::find(active.begin(), active.end(), 1) != active.end()) {
            uint64_t next = UINT64_MAX;
            for (uint32_t r = 0; r < count; r++)
                if (active[r]) next = std::min(next, next_check(r, global));
            if (next <= global) throw HipError{"fg_run_batch: no step left before the next check", FG_INVALID};
            if (dirty) plan(b, active), dirty = false;
            advance(b, global, next - global, graphs);
            global = next;
            FG_CHECK(hipStreamSynchronize(b.stream));
            check_abort(b);
            double least = INFINITY;
            for (uint32_t r = 0; r < count; r++) {
                if (!active[r]) continue;
                Session &s = *b.runs[r];
                s.steps = global;
                s.cur = int(b.par);
                if (global % every[r] && global != extras[r]->max_steps) continue;
                const double e = energy(s);
                if (!std::isfinite(e))  // run.rs:736-738
                    throw HipError{"run " + std::to_string(r) + ": the FDTD run went unstable", FG_UNSTABLE};
                s.peak = std::max(s.peak, e);
                s.decay = e > 0.0 && s.peak > 0.0 ? 10.0 * std::log10(s.peak / e) : 0.0;
                least = std::min(least, s.decay);
                const bool done = global >= extras[r]->max_steps ||
                                  (global > std::max(pulse_end[r], extras[r]->min_steps) && s.decay >= extras[r]->decay_db);
                if (done) {
                    read_record(s, &records[r]);
                    active[r] = 0;
                    dirty = true;
                }
            }
            if (progress && std::isfinite(least) && !progress(user, global, least))  // run.rs:742-744
                throw HipError{"cancelled", FG_CANCELLED};
        }
        for (uint32_t r = 0; r < count; r++)
            if (extras[r]->max_steps == 0) read_record(*b.runs[r], &records[r]);
        spectra(b, count, sp);
        return int(FG_OK);
    });
}

// finish()'s spectra (crates/sims-antenna/src/fdtd/mod.rs:229-249) of one record's series, on
// the device: k_spectra over every (port, frequency).
extern "C" int fg_port_spectra(fg_device *dev, const float *series, uint64_t steps, uint32_t ports, double dt,
                               uint32_t nf, const double *freqs, double *v, double *i) {
    using namespace fg;
    return guarded([&] {
        if (!dev || (steps && ports && !series) || (nf && (!freqs || !v || !i)))
            throw HipError{"fg_port_spectra: NULL argument", FG_INVALID};
        const uint64_t n = uint64_t(ports) * nf;
        if (!n) return int(FG_OK);
        if (n > 0x7fffff00ull) throw HipError{"fg_port_spectra: too many ports x frequencies", FG_INVALID};
        FG_CHECK(hipSetDevice(dev->ordinal));
        hipStream_t st = nullptr;
        FG_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s); }
        } guard{st};
        DevBuf<float> ds;
        if (steps) ds.upload(series, steps * ports * 2, st);
        else ds.alloc(1);
        std::vector<SpecJob> jobs;
        for (uint32_t p = 0; p < ports; p++)
            for (uint32_t f = 0; f < nf; f++)
                jobs.push_back(SpecJob{ds.p, steps, ports, p, dt, 2.0 * 3.14159265358979323846 * freqs[f]});
        DevBuf<SpecJob> dj;
        DevBuf<double> out;
        dj.upload(jobs.data(), jobs.size(), st);
        out.alloc(4 * jobs.size());
        hipLaunchKernelGGL(k_spectra, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, dj.p, uint32_t(n), out.p);
        FG_CHECK(hipGetLastError());
        std::vector<double> h(4 * n);
        FG_CHECK(hipMemcpyAsync(h.data(), out.p, h.size() * 8, hipMemcpyDeviceToHost, st));
        FG_CHECK(hipStreamSynchronize(st));
        for (uint64_t q = 0; q < n; q++) {
            v[2 * q] = h[4 * q], v[2 * q + 1] = h[4 * q + 1];
            i[2 * q] = h[4 * q + 2], i[2 * q + 1] = h[4 * q + 3];
        }
        return int(FG_OK);
    });
}

// ---- port spectra with no allocation ---------------------------------------------------------

namespace fg {

// The device scratch of fg_port_spectra_static: the series staged from the host, the
// frequencies and the two output arrays, each at a 256-byte boundary. Every GPU the library
// runs on holds its own copy until the process ends. Host code never passes `scratch` itself
// (a host-side stand-in): it uses the address hipGetSymbolAddress gives for the device.
constexpr uint64_t SPECTRA_SCRATCH_BYTES = 32ull << 20;
static __device__ __align__(256) unsigned char scratch[SPECTRA_SCRATCH_BYTES];

void spectra_scratch_init(fg_device *dev) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    dev->scratch = e == hipSuccess ? p : nullptr;
    dev->scratch_error = e;
}

}  // namespace fg

// fg_port_spectra without allocating: the static scratch above instead of hipMalloc'd buffers,
// the null stream instead of a created one, and one kernel that forms each (port, frequency)
// from the frequency table. Bit-identical to fg_port_spectra (the same sums, through
// spectrum()).
extern "C" int fg_port_spectra_static(fg_device *dev, const float *series, uint64_t steps, uint32_t ports, double dt,
                                      uint32_t nf, const double *freqs, double *v, double *i) {
    using namespace fg;
    return guarded([&] {
        if (!dev || (steps && ports && !series) || (nf && (!freqs || !v || !i)))
            throw HipError{"fg_port_spectra_static: NULL argument", FG_INVALID};
        const uint64_t n = uint64_t(ports) * nf;
        if (!n) return int(FG_OK);
        auto up = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
        const uint64_t cap = SPECTRA_SCRATCH_BYTES;
        // every term is bounded before it is multiplied, so nothing below overflows
        const bool fits = steps <= cap / 8 && n <= cap / 32 && nf <= cap / 8 &&
                          up(steps * ports * 8) + up(uint64_t(nf) * 8) + 2 * up(n * 16) <= cap;
        if (!fits)
            throw HipError{"fg_port_spectra_static: " + std::to_string(steps) + " steps x " + std::to_string(ports) +
                               " ports x " + std::to_string(nf) + " frequencies do not fit the " +
                               std::to_string(cap >> 20) + " MiB static scratch (use fg_port_spectra)",
                           FG_INVALID};
        if (!dev->scratch)
            throw HipError{std::string("fg_port_spectra_static: no static scratch on this device: ") +
                               hipGetErrorString(dev->scratch_error),
                           FG_HIP};
        const uint64_t b_series = up(steps * ports * 8), b_freqs = up(uint64_t(nf) * 8), b_out = up(n * 16);
        char *base = static_cast<char *>(dev->scratch);
        float *ds = reinterpret_cast<float *>(base);
        double *df = reinterpret_cast<double *>(base + b_series);
        double *dv = reinterpret_cast<double *>(base + b_series + b_freqs);
        double *di = reinterpret_cast<double *>(base + b_series + b_freqs + b_out);
        std::lock_guard<std::mutex> lock(dev->scratch_mutex);
        FG_CHECK(hipSetDevice(dev->ordinal));
        hipStream_t st = nullptr;  // the null stream: nothing is created
        if (steps) FG_CHECK(hipMemcpyAsync(ds, series, steps * ports * 8, hipMemcpyHostToDevice, st));
        FG_CHECK(hipMemcpyAsync(df, freqs, uint64_t(nf) * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_spectra_flat, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, ds, steps, ports, dt, df,
                           nf, dv, di);
        FG_CHECK(hipGetLastError());
        FG_CHECK(hipMemcpyAsync(v, dv, n * 16, hipMemcpyDeviceToHost, st));
        FG_CHECK(hipMemcpyAsync(i, di, n * 16, hipMemcpyDeviceToHost, st));
        FG_CHECK(hipStreamSynchronize(st));
        return int(FG_OK);
    });
}
