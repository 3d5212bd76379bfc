// Host-side state shared by the single-run engine (engine.hip) and the batched sweep
// (batch.hip): errors, device buffers, the device handle and one run's session.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fieldgrid.h"
#include "common.h"
#include "device.h"

namespace fg {

extern thread_local std::string g_error;

struct HipError {
    std::string msg;
    int status;
};

#define FG_CHECK(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::fg::HipError{std::string(#expr) + ": " + hipGetErrorString(e_),                \
                                 e_ == hipErrorOutOfMemory ? FG_TOO_LARGE : FG_HIP};               \
    } while (0)

constexpr int ENERGY_BLOCKS = 1024;

// Host-to-device copies through pinned host memory. From pageable memory the runtime stages
// each copy itself and returns only once it is done (~0.1 ms per small copy, ~20 per session
// set-up; ~3 GB/s for a large one); from pinned memory a copy is an asynchronous DMA. While a
// Stager is installed on this thread, DevBuf::upload on the Stager's stream copies through it:
//   - up to MAX_BYTES: the source goes into a pinned chunk of CHUNK bytes, shared with the next
//     small copies, and the DMA reads it from there;
//   - larger: in pieces of SLOT bytes through a ring of RING pinned slots, the CPU filling one
//     slot while the DMA engine reads the others (each slot's event says when its last DMA is
//     done).
// Chunks and slots come from process-wide pools (hipHostMalloc) that keep them until the
// process ends. The destructor synchronizes the stream before they go back to the pools.
class Stager {
  public:
    explicit Stager(hipStream_t s);
    ~Stager();
    Stager(const Stager &) = delete;
    Stager &operator=(const Stager &) = delete;
    // enqueue dst <- src on stream s through pinned memory; false when s is another stream (the
    // caller copies directly)
    bool copy(void *dst, const void *src, size_t bytes, hipStream_t s);
    static constexpr size_t CHUNK = 1u << 20, MAX_BYTES = 256u << 10, SLOT = 4u << 20, RING = 4;
    struct Slot {
        char *host;
        hipEvent_t done;
        bool busy;  // a DMA from this slot was enqueued since the slot was last waited for
    };

  private:
    bool copy_large(void *dst, const void *src, size_t bytes);
    hipStream_t stream_;
    Stager *prev_;
    std::vector<char *> chunks_;
    size_t used_ = CHUNK;  // bytes used in chunks_.back()
    std::vector<Slot> ring_;
    size_t next_ = 0;  // the ring slot the next piece goes through (mod RING)
};
extern thread_local Stager *t_stager;

// Device memory for one session's small tables, carved at 256-byte alignment from chunks of
// CHUNK bytes: a session makes one hipMalloc and one hipFree per chunk instead of one per table
// (each call costs tens of microseconds, and a session has about 40 tables). Requests above
// MAX_BYTES are refused (take returns nullptr) and get their own hipMalloc. The chunks are freed
// with the arena, so it must outlive every buffer carved from it.
class DevArena {
  public:
    DevArena() = default;
    ~DevArena();
    DevArena(const DevArena &) = delete;
    DevArena &operator=(const DevArena &) = delete;
    void *take(size_t bytes);
    static constexpr size_t CHUNK = 2u << 20, MAX_BYTES = 256u << 10;

  private:
    std::vector<void *> chunks_;
    size_t used_ = CHUNK;  // bytes used in chunks_.back()
};
// While an ArenaScope is alive on this thread, DevBuf::alloc carves small buffers from its arena.
extern thread_local DevArena *t_arena;
class ArenaScope {
  public:
    explicit ArenaScope(DevArena &a) : prev_(t_arena) { t_arena = &a; }
    ~ArenaScope() { t_arena = prev_; }
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;

  private:
    DevArena *prev_;
};

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0, cap = 0;  // elements in use, elements allocated
    bool pooled = false;    // p was carved from a DevArena or another buffer, which frees it
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && !pooled) (void)hipFree(p);
        p = nullptr;
        pooled = false;
        cap = 0;
    }
    void alloc(size_t count) {
        release();
        n = cap = count;
        if (!count) return;
        if (t_arena && (p = static_cast<T *>(t_arena->take(count * sizeof(T))))) {
            pooled = true;
            return;
        }
        FG_CHECK(hipMalloc(&p, count * sizeof(T)));
    }
    // view count elements at `at`, memory that another buffer owns and frees
    void carve(T *at, size_t count) {
        release();
        p = at, n = cap = count, pooled = true;
    }
    void zero(hipStream_t s) {
        if (n) FG_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    void upload(const T *src, size_t count, hipStream_t s) {
        alloc(count);
        copy_in(src, count, s);
    }
    // upload into the present allocation when it holds count elements (the copy is ordered on
    // stream s after the work already there), else into a new one
    void refill(const T *src, size_t count, hipStream_t s) {
        if (p && count <= cap) n = count;
        else alloc(count);
        copy_in(src, count, s);
    }

  private:
    void copy_in(const T *src, size_t count, hipStream_t s) {
        if (!count || (t_stager && t_stager->copy(p, src, count * sizeof(T), s))) return;
        FG_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

}  // namespace fg

struct fg_device {
    int ordinal;
    std::string name;
    uint64_t total_mem;
    int cus;
    // idle streams that sessions of this device have finished with (creating a stream costs
    // ~0.5 ms and destroying one ~0.2 ms); destroyed by fg_close
    std::vector<hipStream_t> idle_streams;
    std::mutex stream_mutex;
    // this device's copy of the static spectra scratch (hipGetSymbolAddress at fg_open), and
    // the lock that keeps two calls of one process from using it at once
    void *scratch = nullptr;
    hipError_t scratch_error = hipSuccess;
    std::mutex scratch_mutex;
};

namespace fg {

// A non-blocking stream from the device's idle streams, or a new one; give_stream hands an idle
// (synchronized) stream back.
hipStream_t take_stream(fg_device *dev);
void give_stream(fg_device *dev, hipStream_t s);

// Looks up this device's address of the static spectra scratch (batch.hip), with the device
// current; fg_open calls it once per device handle.
void spectra_scratch_init(fg_device *dev);

struct Session {
    fg_device *dev;
    DevArena arena;  // the small tables below; declared before them, so destroyed after them
    hipStream_t stream = nullptr;
    Geo geo{};
    FixArgs fix{};
    ObsArgs obs{};
    uint64_t nn = 0;
    uint32_t n[3];
    double dt = 0.0;  // Sim.dt
    bool fused = true;
    uint32_t tj = 14, seg = 32, xcds = 8;
    uint64_t max_steps = 0, steps = 0;
    uint32_t n_ports = 0, n_probes = 0, nf = 0;
    uint64_t plane_len = 0, ntff_len = 0;
    double peak = 0.0, decay = 0.0;
    int cur = 0;  // which ping-pong copy holds the latest fields (fused)
    uint32_t fix_blocks = 0, obs_blocks = 0;
    bool timing = false;
    bool observe = true;  // false: steps skip the observers (a recompute that already has them)
    double kernel_ms[3] = {0, 0, 0};  // step kernel(s), fix, observe
    uint64_t kernel_count = 0;
    // buffers
    DevBuf<char> field_mem;  // holds E, H and psi; declared before them, so freed after them
    DevBuf<float> E[2], H[2], psi[2];
    DevBuf<uint32_t> coef_index, step;
    DevBuf<float> coef_sets, axes, dtab, dstate, dold, ind_i, series, plane, ntff, probes, pulse, phase;
    DevBuf<int> slots;
    DevBuf<fg_debye_edge> debye;
    DevBuf<fg_source> src;
    DevBuf<fg_inductor> ind;
    DevBuf<fg_port> ports;
    DevBuf<fg_port_edge> edges;
    DevBuf<fg_loop_term> loops;
    DevBuf<uint64_t> probe_f;
    DevBuf<uint32_t> patch;
    DevBuf<double> partial;
    DevBuf<MarchBlock> blk_int, blk_z, blk_pml;  // march blocks clear of CPML / touching z slabs only / any
    // two steps per launch (march2): on when nothing but sources sits between two steps
    bool fuse2 = false;
    uint32_t tj2 = 12, seg2 = 32;
    // segment lengths for the two-step march to time at the start of the run (enqueue_steps
    // picks the fastest and drops the list); empty: seg2 as set
    std::vector<uint32_t> seg2_try;
    uint32_t seg2_base = 32;  // seg2 as set before any tuning
    DevBuf<MarchBlock> blk2_int, blk2_z, blk2_pml;
    DevBuf<MarchBlock> blk2_mix;  // the three, class-tagged, for k_march2_mix
    bool mix2 = true;             // one mixed launch per pass where the classes would run serially
    DevBuf<uint32_t> src2_start;
    DevBuf<uint4> src2_ent;
    DevBuf<int32_t> deb2_row, deb2_map;  // Debye k maps of the two-step march (Deb2)
    DevBuf<float> deb2_jn, deb2_eo;
    DevBuf<uint32_t> cap_start, cap_ntff, cap_plane;  // observed points of the two-step march (Cap2)
    DevBuf<uint32_t> cap_obs;  // observer captures (Cap2::obs): port edges, loop terms, probes
    uint32_t cap_loop0 = 0, cap_probe0 = 0;
    DevBuf<uint4> cap_seg;
    DevBuf<float> cap_buf;
    // the two-step march's side work per (i, j) row, as masks over its 60-lane k tiles (bit t:
    // tile t): side_deb, the tiles whose lanes (k0-2 .. k0+61) hold a Debye node of the row;
    // side_cap, the tiles that own a captured node of the row. Empty: no such work. With more
    // than 64 k tiles every bit of a row with any such node is set. (march_classes2's flags)
    std::vector<uint64_t> side_deb, side_cap;
    // resistive sheets (records restored from sheets0 on reset) and thin wires
    SheetArgs sheet{};
    DevBuf<float> sheets, sheet_cur, branches;
    std::vector<float> sheets0;
    DevBuf<fg_wire_h> wires;
    DevBuf<uint32_t> wire_start;
    DevBuf<uint4> wire_ent;
    bool wire_serial = false;
    std::vector<std::pair<size_t, uint32_t>> units;  // timed launch units: (event slot, steps)
    std::vector<hipEvent_t> events;
    // host copies the batched sweep and reset() need
    uint32_t pml = 0, nmax = 0;
    std::vector<int> slots_host;  // per axis: slot_e row, slot_h row, nmax entries each
    std::vector<float> ind_i0;    // initial inductor currents

    // the block classes of one pass run concurrently on stream, side[0] and side[1] (forked and
    // joined with events), so each class's last round of blocks overlaps the others
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};

    ~Session() {
        for (auto e : events) (void)hipEventDestroy(e);
        for (int q = 0; q < 2; q++) {
            if (side[q] && hipStreamSynchronize(side[q]) == hipSuccess) give_stream(dev, side[q]);
            else if (side[q]) (void)hipStreamDestroy(side[q]);
            if (join[q]) (void)hipEventDestroy(join[q]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (stream && hipStreamSynchronize(stream) == hipSuccess) give_stream(dev, stream);
        else if (stream) (void)hipStreamDestroy(stream);
    }
};

// Every C-ABI entry point runs its body here: a thrown HipError becomes its status code with
// the message in fg_last_error().
template <class F>
int guarded(F &&body) {
    try {
        return body();
    } catch (const HipError &e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc &) {
        return set_error(FG_TOO_LARGE, "out of host memory");
    }
}

// The device bytes of a session's fields, psi (both ping-pong copies when fused) and
// coefficient index, in double so that no extent overflows it.
double session_bytes(const fg_problem *P, bool fused);
// the session's other device memory (Debye, records, observers, tables), bounded from above
double session_extra_bytes(const fg_problem *P, const fg_extras *X);
// One run's device state: validated, allocated and uploaded, fields at zero (engine.hip).
// obs_capture: where the run takes two steps per pass, its port edges, loop terms and probes are
// captured too (Cap2::obs), for launches that read step n's observers from captures.
Session *session_new(fg_device *dev, const fg_problem *P, uint32_t driven, const fg_pulse *pulse, const fg_extras *X,
                     bool obs_capture = false);
// Zero every field, psi, Debye, inductor-current and accumulator buffer and the step count on
// the session's stream, so the run starts over exactly as from session_new.
void reset(Session &s);
// The fused march's blocks of TJ rows x `seg` planes, split by CPML reach: out[0] clear of
// every slab, out[1] touching only z (k) slabs, out[2] touching an x or y slab.
void march_classes(const Session &s, uint32_t tj, uint32_t seg, std::vector<MarchBlock> out[3]);
// The same for the two-step march (60 lanes x tj rows owned, two-cell halo, two-plane lag).
// The two-step march's block classes; `reach` more planes past a segment's end are read (the
// two-step adjoint reads one: its first step's lamH at the last plane needs lamE one plane on).
// side_flags: a block whose tile holds none of the session's side work (a Debye edge anywhere in
// the planes, rows and lanes it computes, a captured node among those it owns) gets no flag; the
// others get SIDE_BLOCK in bz, and the side-work kernels run the lean march for the former.
void march_classes2(const Session &s, uint32_t tj, uint32_t seg, std::vector<MarchBlock> out[3], uint32_t reach = 0,
                    bool side_flags = false);
// The largest per-plane byte count among the arrays the two-step march addresses plane by plane
// (a field plane, one plane's rows of a z or y CPML psi slab): the argument of march2_seg_cap.
inline uint64_t march2_plane_bytes(const uint32_t n[3], uint32_t pml) {
    const uint64_t p2 = 2ull * pml, a = uint64_t(n[1]) * n[2], b = uint64_t(n[1]) * p2, c = p2 * n[2];
    return 4ull * (a > b ? (a > c ? a : c) : (b > c ? b : c));
}

// Within the contiguous run of a block list that xcd_order hands each XCD, the blocks for which
// `first` holds go first (stable otherwise): the side-work blocks of the two-step march, which
// take longer, start in the first round instead of trailing the pass.
template <class T, class F>
void first_per_xcd(std::vector<T> &v, size_t xcds, F first) {
    const size_t N = v.size(), X = xcds > 1 ? xcds : 1, per = (N + X - 1) / X;
    for (size_t x = 0; x * per < N; x++)
        std::stable_partition(v.begin() + x * per, v.begin() + std::min(N, (x + 1) * per), first);
}

// Workgroups go to the XCDs round-robin (launch index b runs on XCD b % xcds). Permute a block
// list so that XCD x receives a contiguous run of the logical order (k tiles fastest, then j
// tiles): neighbouring tiles then share one L2, and the cache lines their halos and unaligned
// edges have in common are fetched from HBM once. Padding entries carry `pad`.
template <class T>
void xcd_order(std::vector<T> &v, size_t xcds, const T &pad) {
    const size_t N = v.size(), X = xcds, per = (N + X - 1) / X;
    if (X <= 1 || N == 0) return;
    std::vector<T> out(per * X, pad);
    for (size_t b = 0; b < per * X; b++) {
        const size_t L = (b % X) * per + b / X;
        if (L < N) out[b] = v[L];
    }
    v.swap(out);
}

// Enqueue `count` further steps on the session's stream (asynchronous).
void enqueue_steps(Session &s, uint64_t count);
// The pieces of enqueue_steps: load the device step counter with s.steps, then launch the
// step at offset s_off from it (the caller advances s.steps afterwards).
void set_step_counter(Session &s);
void launch_step(Session &s, uint32_t s_off, uint64_t global_step);
// Streams for the three block classes of a pass (interior, z-slab, x/y-slab): forked from
// the session stream when the session has side streams, else the session stream three times;
// join_classes makes the session stream wait for them.
void fork_classes(Session &s, hipStream_t out[3]);
void join_classes(Session &s);
// The two-step march's block lists for segments of s.seg2 planes.
void plan_march2(Session &s);
// The segment tuner of the two-step march: times two-step passes of each length in s.seg2_try
// (emptied) through chunk(k), which runs the session's next k steps (an even number), keeps the
// fastest and re-plans the march with it. Every tiling computes the same fields bit for bit, so
// the timed passes are steps of the run. Returns the steps taken, seg2_tune_steps(s) before.
uint64_t tune_seg2(Session &s, const std::function<void(uint64_t)> &chunk);
uint64_t seg2_tune_steps(const Session &s);
// Two steps (s_off, s_off + 1) in one march2 launch; only when s.fuse2.
void launch_pair(Session &s, uint32_t s_off);

// Sum of E^2 + (376.73 H)^2 over the latest fields (cpu.rs:417-421); synchronizes the stream.
double energy(Session &s);
// Copy the run's outputs into a caller-allocated record (run.rs:746-757).
void read_record(Session &s, fg_record *out);

}  // namespace fg
