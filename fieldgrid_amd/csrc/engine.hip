// MI355X (gfx950) FDTD engine: device kernels and the run driver behind include/fieldgrid.h.
//
// Replaces the reference's GPU driver and WGSL kernels:
//   crates/sims-field/src/fdtd/run.rs:390-758   run()
//   crates/sims-field/src/fdtd/fdtd.wgsl:1-630   update_h, update_e, step_even/odd, debye_*,
//                                               lumped, probe, energy, field_dft, ntff
// with the step order of crates/sims-field/src/fdtd/cpu.rs:267-307.
//
// Layout in HBM: E and H are three nn-float planes each (component-major, k contiguous),
// exactly the reference's E[c*nn+id]. The fused engine keeps two copies (ping-pong) of E, H
// and the CPML psi store, so each step reads one copy and writes the other: one pass over
// the grid per step, 54 B/node of compulsory traffic (24 B read + 24 B written fields, 6 B
// of u16 coefficient-set indices). The split engine updates in place with two passes.


#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/fieldgrid_diag.h"
#include "host.h"

namespace fg {

thread_local std::string g_error;
thread_local Stager *t_stager = nullptr;
thread_local DevArena *t_arena = nullptr;

namespace {
std::mutex g_pin_mutex;
std::vector<char *> g_pin_free;         // pinned chunks of Stager::CHUNK bytes not in use
std::vector<Stager::Slot> g_slot_free;  // pinned ring slots of Stager::SLOT bytes not in use
}  // namespace

hipStream_t take_stream(fg_device *dev) {
    {
        std::lock_guard<std::mutex> lock(dev->stream_mutex);
        if (!dev->idle_streams.empty()) {
            hipStream_t s = dev->idle_streams.back();
            dev->idle_streams.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    FG_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void give_stream(fg_device *dev, hipStream_t s) {
    std::lock_guard<std::mutex> lock(dev->stream_mutex);
    dev->idle_streams.push_back(s);
}

Stager::Stager(hipStream_t s) : stream_(s), prev_(t_stager) { t_stager = this; }

Stager::~Stager() {
    t_stager = prev_;
    if (chunks_.empty() && ring_.empty()) return;
    // the DMAs have read the pinned memory; if that cannot be known, it stays out of the pools
    if (hipStreamSynchronize(stream_) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_pin_mutex);
    for (char *c : chunks_) g_pin_free.push_back(c);
    for (Slot sl : ring_) sl.busy = false, g_slot_free.push_back(sl);
}

bool Stager::copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (s != stream_) return false;
    if (bytes > MAX_BYTES) return copy_large(dst, src, bytes);
    const size_t at = (used_ + 255) & ~size_t(255);
    if (at + bytes > CHUNK) {
        char *c = nullptr;
        {
            std::lock_guard<std::mutex> lock(g_pin_mutex);
            if (!g_pin_free.empty()) c = g_pin_free.back(), g_pin_free.pop_back();
        }
        if (!c) FG_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c), CHUNK, hipHostMallocDefault));
        chunks_.push_back(c);
        used_ = 0;
        return copy(dst, src, bytes, s);
    }
    std::memcpy(chunks_.back() + at, src, bytes);
    FG_CHECK(hipMemcpyAsync(dst, chunks_.back() + at, bytes, hipMemcpyHostToDevice, s));
    used_ = at + bytes;
    return true;
}

bool Stager::copy_large(void *dst, const void *src, size_t bytes) {
    if (ring_.empty()) {
        std::lock_guard<std::mutex> lock(g_pin_mutex);
        while (ring_.size() < RING && !g_slot_free.empty()) ring_.push_back(g_slot_free.back()), g_slot_free.pop_back();
    }
    while (ring_.size() < RING) {
        Slot sl{nullptr, nullptr, false};
        FG_CHECK(hipHostMalloc(reinterpret_cast<void **>(&sl.host), SLOT, hipHostMallocDefault));
        const hipError_t e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e != hipSuccess) (void)hipHostFree(sl.host);
        FG_CHECK(e);
        ring_.push_back(sl);
    }
    for (size_t at = 0; at < bytes; at += SLOT) {
        Slot &sl = ring_[next_++ % RING];
        if (sl.busy) FG_CHECK(hipEventSynchronize(sl.done));  // its previous DMA has read it
        const size_t len = std::min(SLOT, bytes - at);
        std::memcpy(sl.host, static_cast<const char *>(src) + at, len);
        FG_CHECK(hipMemcpyAsync(static_cast<char *>(dst) + at, sl.host, len, hipMemcpyHostToDevice, stream_));
        FG_CHECK(hipEventRecord(sl.done, stream_));
        sl.busy = true;
    }
    return true;
}

DevArena::~DevArena() {
    for (void *c : chunks_) (void)hipFree(c);
}

void *DevArena::take(size_t bytes) {
    if (bytes > MAX_BYTES) return nullptr;
    const size_t at = (used_ + 255) & ~size_t(255);
    if (at + bytes > CHUNK) {
        void *c = nullptr;
        FG_CHECK(hipMalloc(&c, CHUNK));
        chunks_.push_back(c);
        used_ = 0;
        return take(bytes);
    }
    used_ = at + bytes;
    return static_cast<char *>(chunks_.back()) + at;
}

int set_error(int status, const std::string &msg) {
    g_error = msg;
    return status;
}

// ---- split engine: in-place H sweep then E sweep (fdtd.wgsl:154-184) ---------------------

__global__ __launch_bounds__(256) void k_update_h(Geo g, const float *__restrict__ E, float *__restrict__ H, float *psi) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= g.n2 || j >= g.n1) return;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const float hx = h_next<0>(g, E, H, psi, psi, i, j, k);
    const float hy = h_next<1>(g, E, H, psi, psi, i, j, k);
    const float hz = h_next<2>(g, E, H, psi, psi, i, j, k);
    H[id] = hx;
    H[g.nn + id] = hy;
    H[2 * g.nn + id] = hz;
}

__global__ __launch_bounds__(256) void k_update_e(Geo g, float *__restrict__ E, const float *__restrict__ H, float *psi) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= g.n2 || j >= g.n1) return;
    const uint64_t nn = g.nn;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    // -1 neighbours, clamped in-bounds at the low faces (those edges have cb == 0)
    const uint64_t im = i ? id - stride<0>(g) : id, jm = j ? id - stride<1>(g) : id, km = k ? id - 1 : id;
    const float hx = H[id], hy = H[nn + id], hz = H[2 * nn + id];
    const float ex = e_next<0>(g, E[id], hz - H[2 * nn + jm], hy - H[nn + km], i, j, k, id, psi, psi);
    const float ey = e_next<1>(g, E[nn + id], hx - H[km], hz - H[2 * nn + im], i, j, k, id, psi, psi);
    const float ez = e_next<2>(g, E[2 * nn + id], hy - H[nn + im], hx - H[jm], i, j, k, id, psi, psi);
    E[id] = ex;
    E[nn + id] = ey;
    E[2 * nn + id] = ez;
}

// One step of the fused march over one run's block list (see march() for the scheme).
template <int TJ, int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ + 2)) void k_march(Geo g, const MarchBlock *__restrict__ blocks,
                                                         const float *__restrict__ Eo, float *__restrict__ En,
                                                         const float *__restrict__ Ho, float *__restrict__ Hn,
                                                         const float *__restrict__ psio, float *__restrict__ psin,
                                                         uint32_t seg) {
    const MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;  // padding of the XCD-ordered list (uniform over the block)
    march<TJ, PML, WIDE, LDS_SETS>(g, mb, Eo, En, Ho, Hn, psio, psin, seg);
}

// The same sweep for a problem with thin wires (always 14 rows per block)
template <int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * 16) void k_march_wires(Geo g, const MarchBlock *__restrict__ blocks,
                                                         const float *__restrict__ Eo, float *__restrict__ En,
                                                         const float *__restrict__ Ho, float *__restrict__ Hn,
                                                         const float *__restrict__ psio, float *__restrict__ psin,
                                                         uint32_t seg, WireArgs wa) {
    const MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;
    march<14, PML, WIDE, LDS_SETS, true>(g, mb, Eo, En, Ho, Hn, psio, psin, seg, wa);
}

// Thin wires, split engine: H[h] += sum_t c_t E[e_t] after the H sweep; one thread per entry, or
// one thread for all in order where entries share an H node (`serial`)
__global__ __launch_bounds__(256) void k_wires(const fg_wire_h *__restrict__ w, uint64_t n, uint32_t serial,
                                               const float *__restrict__ E, float *__restrict__ H) {
    const uint64_t t = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    auto one = [&](uint64_t q) {
        const fg_wire_h &wh = w[q];
        float sw = 0.f;
        for (int u = 0; u < 4; u++) sw += wh.c[u] * E[wh.e[u]];
        H[wh.h] += sw;
    };
    if (serial) {
        if (t == 0)
            for (uint64_t q = 0; q < n; q++) one(q);
    } else if (t < n) {
        one(t);
    }
}

// Resistive sheets after a fused sweep (fdtd.wgsl's sheet_fix): the face weights from the step's
// input E and new H, then the sheet current out of the new E.
__global__ __launch_bounds__(256) void k_sheet(SheetArgs a, float *__restrict__ En, const float *__restrict__ Eo,
                                               const float *__restrict__ Hn) {
    const uint64_t t = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    auto one = [&](uint64_t s) {
        const float *sh = a.sh + s * 16;
        sheet_before(a, s, Eo[__float_as_uint(sh[0])], Hn[__float_as_uint(sh[1])], Hn[__float_as_uint(sh[2])]);
        sheet_after(a, s, En);
    };
    if (a.serial) {
        if (t == 0)
            for (uint64_t s = 0; s < a.n; s++) one(s);
    } else if (t < a.n) {
        one(t);
    }
}

// Split engine: the face weights between the H and E sweeps (E still the step's input) ...
__global__ __launch_bounds__(256) void k_sheet_pre(SheetArgs a, const float *__restrict__ E, const float *__restrict__ H) {
    const uint64_t s = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (s >= a.n) return;
    const float *sh = a.sh + s * 16;
    sheet_before(a, s, E[__float_as_uint(sh[0])], H[__float_as_uint(sh[1])], H[__float_as_uint(sh[2])]);
}

// ... and the sheet currents after the E sweep
__global__ __launch_bounds__(256) void k_sheet_post(SheetArgs a, float *__restrict__ E) {
    const uint64_t t = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (a.serial) {
        if (t == 0)
            for (uint64_t s = 0; s < a.n; s++) sheet_after(a, s, E);
    } else if (t < a.n) {
        sheet_after(a, t, E);
    }
}

__global__ __launch_bounds__(256) void k_debye_save(uint64_t n, const fg_debye_edge *__restrict__ d,
                                                     const float *__restrict__ E, float *__restrict__ old) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) old[t] = E[d[t].f];
}

__global__ __launch_bounds__(256) void k_fix(Geo g, FixArgs a, float *__restrict__ E, const float *__restrict__ Eo,
                                             uint32_t s_off) {
    fix_step(g, a, E, Eo, s_off, uint64_t(blockIdx.x) * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_observe(Geo g, ObsArgs a, const float *__restrict__ E,
                                                 const float *__restrict__ H, uint32_t s_off) {
    observe(g, a, E, H, s_off, blockIdx.x, threadIdx.x);
}

// The LDS of a two-step march kernel: with side work (DEB) either march's, else the lean one's.
template <int TJ, bool LDS_SETS, bool DEB>
using March2Shared = std::conditional_t<DEB, March2LdsEither<TJ, LDS_SETS>, March2Lds<TJ, LDS_SETS, false>>;

template <int TJ, bool LDS_SETS, bool DEB>
__device__ __forceinline__ March2Lds<TJ, LDS_SETS, false> &lean_lds(March2Shared<TJ, LDS_SETS, DEB> &l) {
    if constexpr (DEB) return l.lean;
    else return l;
}

// Two steps of the march per launch (see march2): the first step's pulse value comes from the
// device step counter and the host table, as in k_fix. With side work (DEB), a block flagged
// SIDE_BLOCK runs the side-work march and every other block the lean one (march2_ahead for the
// interior class), which computes the same values where no side work falls.
template <int TJ, int PML, bool WIDE, bool LDS_SETS, bool DEB>
__global__ __launch_bounds__(64 * (TJ + 4)) void k_march2(Geo g, const MarchBlock *__restrict__ blocks,
                                                          const float *__restrict__ Eo, float *__restrict__ En,
                                                          const float *__restrict__ Ho, float *__restrict__ Hn,
                                                          const float *__restrict__ psio, float *__restrict__ psin,
                                                          uint32_t seg, Src2 sr, FixArgs f, Deb2 db, Cap2 cp,
                                                          uint32_t s_off) {
    MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;
    const bool side = DEB && (mb.bz & SIDE_BLOCK);
    mb.bz &= BZ_SEG;
    const uint32_t n = f.step[0] + s_off;
    const float pl = n < f.n_pulse ? f.pulse[n] : pulse_at(f, (float(n) + 0.5f) * f.dt);
    __shared__ March2Shared<TJ, LDS_SETS, DEB> lds;
    if constexpr (DEB) {
        if (side) {
            march2<TJ, PML, WIDE, LDS_SETS, true>(lds.side, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
            return;
        }
    }
    auto &ll = lean_lds<TJ, LDS_SETS, DEB>(lds);
    if constexpr (PML == 0) march2_ahead<TJ, 0, WIDE, LDS_SETS, false>(ll, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
    else march2<TJ, PML, WIDE, LDS_SETS, false>(ll, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
}

// The same for a block list that mixes the CPML classes: the class is in the two top bits of
// the block's bz (march_classes2's out[c] tagged c), each block runs its class's march2, and the
// instances share one LDS copy. One launch replaces the three class launches of a pass where
// those would run one after another (no side streams), so a small grid's pass lasts about one
// block instead of three.
template <int TJ, bool WIDE, bool LDS_SETS, bool DEB>
__global__ __launch_bounds__(64 * (TJ + 4)) void k_march2_mix(Geo g, const MarchBlock *__restrict__ blocks,
                                                              const float *__restrict__ Eo, float *__restrict__ En,
                                                              const float *__restrict__ Ho, float *__restrict__ Hn,
                                                              const float *__restrict__ psio, float *__restrict__ psin,
                                                              uint32_t seg, Src2 sr, FixArgs f, Deb2 db, Cap2 cp,
                                                              uint32_t s_off) {
    MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;
    const uint32_t cls = mb.bz >> 30;
    const bool side = DEB && (mb.bz & SIDE_BLOCK);
    mb.bz &= BZ_SEG;
    const uint32_t n = f.step[0] + s_off;
    const float pl = n < f.n_pulse ? f.pulse[n] : pulse_at(f, (float(n) + 0.5f) * f.dt);
    __shared__ March2Shared<TJ, LDS_SETS, DEB> lds;
    if constexpr (DEB) {
        if (side) {
            if (cls == 2) march2<TJ, 2, WIDE, LDS_SETS, true>(lds.side, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
            else if (cls == 1) march2<TJ, 1, WIDE, LDS_SETS, true>(lds.side, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
            else march2<TJ, 0, WIDE, LDS_SETS, true>(lds.side, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
            return;
        }
    }
    auto &ll = lean_lds<TJ, LDS_SETS, DEB>(lds);
    if (cls == 2) march2<TJ, 2, WIDE, LDS_SETS, false>(ll, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
    else if (cls == 1) march2<TJ, 1, WIDE, LDS_SETS, false>(ll, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
    else march2_ahead<TJ, 0, WIDE, LDS_SETS, false>(ll, g, mb, Eo, En, Ho, Hn, psio, psin, seg, sr, pl, f, db, cp);
}

// Both steps' observers of a two-step pass (observe_pair): the first step exists only inside the
// pass, so its port and probe observers recompute what they read from the untouched previous
// copy (observe_mid) and its DFT observers read what the pass captured.
__global__ __launch_bounds__(256) void k_observe_pair(Geo g, ObsArgs a, FixArgs f, const float *__restrict__ Eo,
                                                      const float *__restrict__ Ho, const float *__restrict__ psio,
                                                      const float *__restrict__ En, const float *__restrict__ Hn,
                                                      uint32_t s_off, Cap2 cp) {
    observe_pair(g, a, f, Eo, Ho, psio, En, Hn, s_off, blockIdx.x, cp, threadIdx.x);
}

// The largest coefficient-set index of the per-edge index table (u32 words: one index when
// wide, two u16 indices when narrow), into out[0] (zeroed before).
__global__ __launch_bounds__(256) void k_index_max(const uint32_t *__restrict__ idx, uint64_t words, bool wide,
                                                   uint32_t *__restrict__ out) {
    uint32_t m = 0;
    for (uint64_t q = uint64_t(blockIdx.x) * 256 + threadIdx.x; q < words; q += uint64_t(gridDim.x) * 256) {
        const uint32_t w = idx[q];
        m = max(m, wide ? w : max(w & 0xFFFFu, w >> 16));
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), o, 64)));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ---- energy: sum E^2 + (376.73 H)^2 (cpu.rs:417-421, fdtd.wgsl:547-568) -----------------

__global__ __launch_bounds__(256) void k_energy(const float *__restrict__ E, const float *__restrict__ H, uint64_t total,
                                                double *__restrict__ partial) {
    __shared__ double red[4];
    double s = 0.0;
    for (uint64_t q = uint64_t(blockIdx.x) * 256 + threadIdx.x; q < total; q += uint64_t(gridDim.x) * 256) {
        const double x = E[q];
        const double y = double(H[q]) * 376.73;
        s += x * x + y * y;
    }
    s = wave_sum(s);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_set_step(uint32_t *step, uint32_t value) { step[0] = value; }

// Plain streaming copy, 16 B per lane per access: the bandwidth probe that calibrates the
// roofline and the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     uint64_t count) {
    for (uint64_t q = uint64_t(blockIdx.x) * 256 + threadIdx.x; q < count; q += uint64_t(gridDim.x) * 256)
        dst[q] = src[q];
}

static void validate(const fg_problem *P, uint32_t driven, const fg_extras *X) {
    auto bad = [](const std::string &m) { throw HipError{m, FG_INVALID}; };
    if (!P || !X) bad("NULL problem or extras");
    for (int a = 0; a < 3; a++)
        if (P->n[a] < 2) bad("a grid axis needs at least two nodes");
    const uint64_t nn = uint64_t(P->n[0]) * P->n[1] * P->n[2];
    // The march kernels address each plane of a field, coefficient-index or psi array with a
    // 32-bit byte offset from a plane-uniform base, and send lanes that must not touch memory to
    // OOB (device.h), which must lie past every such plane: n1*n2 nodes of 4 bytes stay at or
    // below it, or an access would wrap into another node instead of being dropped.
    if (uint64_t(P->n[1]) * P->n[2] * 4 > OOB)
        throw HipError{"a grid plane of " + std::to_string(P->n[1]) + " x " + std::to_string(P->n[2]) +
                           " nodes is more than this engine addresses (n1*n2 must stay at or below " +
                           std::to_string(OOB / 4) + ")",
                       FG_TOO_LARGE};
    if (!P->axes || !P->coef_sets || !P->coef_index || P->n_sets == 0) bad("missing axes or coefficient tables");
    if (P->n_ports && driven >= P->n_ports) bad("driven port " + std::to_string(driven) + " does not exist");
    if (X->nf > FG_MAX_FREQS) bad("at most 4 DFT frequencies");
    // every table with entries needs its pointer
    if ((P->n_ports && !P->ports) || (P->n_port_edges && !P->port_edges) || (P->n_loops && !P->loops) ||
        (P->n_sources && !P->sources) || (P->n_inductors && !P->inductors) ||
        (P->n_debye && (!P->debye_edges || !P->debye_table)) || (X->n_probes && !X->probes))
        bad("a table with entries is NULL");
    if (X->plane_k >= int64_t(P->n[2])) bad("plane_k lies outside the grid");
    if (X->n_patches && !X->ntff) bad("NTFF patch table is NULL");
    for (uint64_t q = 0; q < X->n_patches; q++)
        for (int t = 1; t < 13; t++)
            if (X->ntff[q * 13 + t] >= nn) bad("an NTFF patch index lies outside the grid");
    for (uint32_t q = 0; q < X->n_probes; q++) {
        const fg_point_probe &p = X->probes[q];
        if (p.field > 5 || p.i >= P->n[0] || p.j >= P->n[1] || p.k >= P->n[2]) bad("a point probe lies outside the grid");
    }
    for (uint32_t q = 0; q < P->n_ports; q++) {
        const fg_port &p = P->ports[q];
        if (p.comp > 2 || uint64_t(p.first_edge) + p.n_edges > P->n_port_edges ||
            uint64_t(p.first_loop) + p.n_loops > P->n_loops || uint64_t(p.first_source) + p.n_sources > P->n_sources)
            bad("a port's table range is out of bounds");
    }
    for (uint32_t q = 0; q < P->n_port_edges; q++)
        if (P->port_edges[q].id >= nn) bad("a port edge lies outside the grid");
    for (uint32_t q = 0; q < P->n_loops; q++)
        if (P->loops[q].id >= nn || P->loops[q].comp > 2) bad("a port loop term lies outside the grid");
    for (uint32_t q = 0; q < P->n_sources; q++)
        if (P->sources[q].id >= nn || P->sources[q].comp > 2) bad("a port source lies outside the grid");
    for (uint32_t q = 0; q < P->n_inductors; q++)
        if (P->inductors[q].id >= nn || P->inductors[q].comp > 2) bad("an inductor lies outside the grid");
    for (uint64_t q = 0; q < P->n_debye; q++)
        if (P->debye_edges[q].f >= 3 * nn || P->debye_edges[q].set >= std::max(1u, P->n_pole_sets))
            bad("a Debye edge lies outside the grid");
    if (P->n_wires && !P->wires) bad("thin-wire table is NULL");
    for (uint64_t q = 0; q < P->n_wires; q++) {
        const fg_wire_h &w = P->wires[q];
        if (w.h >= 3 * nn || w.e[0] >= 3 * nn || w.e[1] >= 3 * nn || w.e[2] >= 3 * nn || w.e[3] >= 3 * nn)
            bad("a thin-wire term lies outside the grid");
    }
    if (P->n_sheets) {
        // the sheet table: [nr, nc, ..] then per face a row of 2 + nr + 2 nc weights (run.rs:259-290)
        if (!P->sheets || !P->branches || P->n_branches < 4) bad("resistive sheet tables are NULL or short");
        const float fnr = P->branches[0], fnc = P->branches[1];
        if (!(fnr >= 0.f && fnc >= 0.f && fnr < 65536.f && fnc < 65536.f) || fnr != float(uint32_t(fnr)) ||
            fnc != float(uint32_t(fnc)))
            bad("the sheet table's branch counts are not small whole numbers");
        const uint64_t nr = uint32_t(fnr), nc = uint32_t(fnc), stride = 2 + nr + 2 * nc, faces = 4 + 2 * nr + 4 * nc;
        if (P->sheet_states < nr + 2 * nc) bad("sheet_states is smaller than the sheet table's branches");
        for (uint64_t q = 0; q < P->n_sheets; q++) {
            const float *sh = P->sheets + 16 * q;
            uint32_t id[3], fc[2];
            std::memcpy(id, sh, 12);
            std::memcpy(fc, sh + 14, 8);
            if (id[0] >= 3 * nn || id[1] >= 3 * nn || id[2] >= 3 * nn) bad("a resistive sheet edge lies outside the grid");
            for (uint32_t f : fc)
                if (faces + (uint64_t(f) + 1) * stride > P->n_branches) bad("a resistive sheet's face row is past the sheet table");
        }
    }
}

static uint64_t psi_len(const uint32_t n[3], int a, uint32_t pml) {
    if (!pml) return 0;
    uint64_t m[3] = {n[0], n[1], n[2]};
    m[a] = 2ull * pml;
    return m[0] * m[1] * m[2];
}

double session_bytes(const fg_problem *P, bool fused) {
    const double nn = double(P->n[0]) * P->n[1] * P->n[2], pml = P->pml;
    double psi = 0.0;  // 12 psi arrays, two per (component, other axis); a slab of 2 pml along it
    for (int a = 0; a < 3; a++) psi += pml ? 4.0 * (nn / P->n[a]) * 2.0 * pml : 0.0;
    return (fused ? 2.0 : 1.0) * (24.0 * nn + 4.0 * psi) + (P->coef_wide ? 12.0 : 6.0) * nn;
}

// Device bytes a session holds beyond session_bytes, bounded from above: Debye pole states and
// the two-step march's Debye tables, port series and probe records, plane-DFT and NTFF
// accumulators with their capture buffers, sheet and wire tables, and the small tables' arena
// chunks with their rounding.
double session_extra_bytes(const fg_problem *P, const fg_extras *X) {
    const double n0 = P->n[0], n1 = P->n[1], n2 = P->n[2], nd = double(P->n_debye), np = P->n_poles;
    double b = 0.0;
    if (P->n_debye) {
        b += nd * (sizeof(fg_debye_edge) + 4.0 * 3.0 * np + 4.0);  // edges, states (+ two-step J, E)
        b += 4.0 * n0 * n1 + 12.0 * n2 * std::min(nd, n0 * n1);      // row slots, k maps
    }
    if (X) {
        const double steps = double(X->max_steps), nf = X->nf;
        b += 8.0 * steps * P->n_ports + 4.0 * steps * X->n_probes;    // V/I series, probes
        if (X->plane_k >= 0) b += n0 * n1 * nf * 10.0 * 8.0 + n0 * n1 * 5.0 * 4.0 * 2.0;
        b += double(X->n_patches) * (nf * 12.0 * 8.0 + 13.0 * 4.0 + 12.0 * 4.0 * 2.0);
    }
    b += 64.0 * double(P->n_sheets) + 64.0 * double(P->n_wires);
    return b + double(8u << 20);  // arena chunks (2 MiB each) and allocation rounding
}

// Split the march blocks into those whose reach (their own nodes plus the one-node halo on
// every side) touches a CPML slab position and those that do not.
void march_classes(const Session &s, uint32_t tj, uint32_t seg, std::vector<MarchBlock> out[3]) {
    const uint32_t nmax = s.nmax;
    std::vector<uint32_t> pre[3];
    for (int a = 0; a < 3; a++) {
        pre[a].assign(s.n[a] + 1, 0);
        for (uint32_t q = 0; q < s.n[a]; q++) {
            const bool slab = s.pml && (s.slots_host[(size_t(a) * 2) * nmax + q] >= 0 ||
                                        s.slots_host[(size_t(a) * 2 + 1) * nmax + q] >= 0);
            pre[a][q + 1] = pre[a][q] + (slab ? 1 : 0);
        }
    }
    auto touches = [&](int a, int64_t lo, int64_t hi) {  // any slab position in [lo, hi]
        lo = std::max<int64_t>(lo, 0);
        hi = std::min<int64_t>(hi, int64_t(s.n[a]) - 1);
        return hi >= lo && pre[a][size_t(hi) + 1] > pre[a][size_t(lo)];
    };
    const uint32_t gx = (s.n[2] + OWN_K - 1) / OWN_K, gy = (s.n[1] + tj - 1) / tj, gz = (s.n[0] + seg - 1) / seg;
    for (int c = 0; c < 3; c++) out[c].clear();
    for (uint32_t bz = 0; bz < gz; bz++)
        for (uint32_t by = 0; by < gy; by++)
            for (uint32_t bx = 0; bx < gx; bx++) {
                const int64_t k0 = int64_t(bx) * OWN_K, j0 = int64_t(by) * tj, i0 = int64_t(bz) * seg;
                const bool z = touches(2, k0 - 1, k0 + OWN_K);
                const bool xy = touches(1, j0 - 1, j0 + tj) || touches(0, i0 - 1, i0 + seg);
                out[xy ? 2 : z ? 1 : 0].push_back(MarchBlock{bx, by, bz});
            }
}

// The 60-lane k tiles of the two-step march (mask bit t: tile t) that compute node k (halo: its
// lanes k0-2 .. k0+61) or own it (k0 .. k0+59); with more than 64 tiles, all.
static uint64_t side_tiles(uint32_t k, uint32_t n2, bool halo) {
    if ((n2 + OWN_K2 - 1) / OWN_K2 > 64) return ~0ull;
    uint64_t m = 1ull << (k / OWN_K2);
    if (halo) {
        if (k % OWN_K2 < 2 && k >= OWN_K2) m |= 1ull << (k / OWN_K2 - 1);  // the tile before reads it as k0+60, k0+61
        if (k % OWN_K2 >= OWN_K2 - 2 && (k / OWN_K2 + 1) * OWN_K2 < n2) m |= 1ull << (k / OWN_K2 + 1);  // the next as k0-2, k0-1
    }
    return m;
}

void march_classes2(const Session &s, uint32_t tj, uint32_t seg, std::vector<MarchBlock> out[3], uint32_t reach,
                    bool side_flags) {
    const uint32_t nmax = s.nmax;
    std::vector<uint32_t> pre[3];
    for (int a = 0; a < 3; a++) {
        pre[a].assign(s.n[a] + 1, 0);
        for (uint32_t q = 0; q < s.n[a]; q++) {
            const bool slab = s.pml && (s.slots_host[(size_t(a) * 2) * nmax + q] >= 0 ||
                                        s.slots_host[(size_t(a) * 2 + 1) * nmax + q] >= 0);
            pre[a][q + 1] = pre[a][q] + (slab ? 1 : 0);
        }
    }
    auto touches = [&](int a, int64_t lo, int64_t hi) {
        lo = std::max<int64_t>(lo, 0);
        hi = std::min<int64_t>(hi, int64_t(s.n[a]) - 1);
        return hi >= lo && pre[a][size_t(hi) + 1] > pre[a][size_t(lo)];
    };
    const uint32_t gx = (s.n[2] + OWN_K2 - 1) / OWN_K2, gy = (s.n[1] + tj - 1) / tj, gz = (s.n[0] + seg - 1) / seg;
    // the blocks with side work: a Debye node in the planes (i0-2 .. i1), rows (j0-2 .. j0+tj+1)
    // and lanes (k0-2 .. k0+61) march2 computes (its in-pass ADE runs on every copy of an edge),
    // or a captured node it owns (i0 .. i1-1, j0 .. j0+tj-1, k0 .. k0+59): per (bz, by), the
    // rows' k-tile masks ORed over those windows
    std::vector<uint64_t> side;  // per (bz, by): the k tiles with side work
    if (side_flags && (!s.side_deb.empty() || !s.side_cap.empty())) {
        side.assign(size_t(gz) * gy, 0);
        const int64_t n0 = s.n[0], n1 = s.n[1];
        for (uint32_t bz = 0; bz < gz; bz++) {
            const int64_t i0 = int64_t(bz) * seg, i1 = std::min<int64_t>(i0 + seg, n0);
            for (uint32_t by = 0; by < gy; by++) {
                const int64_t j0 = int64_t(by) * tj;
                uint64_t m = 0;
                if (!s.side_deb.empty())
                    for (int64_t t = std::max<int64_t>(i0 - 2, 0); t <= std::min<int64_t>(i1, n0 - 1); t++)
                        for (int64_t j = std::max<int64_t>(j0 - 2, 0); j < std::min<int64_t>(j0 + tj + 2, n1); j++)
                            m |= s.side_deb[size_t(t * n1 + j)];
                if (!s.side_cap.empty())
                    for (int64_t t = i0; t < i1; t++)
                        for (int64_t j = j0; j < std::min<int64_t>(j0 + tj, n1); j++) m |= s.side_cap[size_t(t * n1 + j)];
                side[size_t(bz) * gy + by] = m;
            }
        }
    }
    for (int c = 0; c < 3; c++) out[c].clear();
    for (uint32_t bz = 0; bz < gz; bz++)
        for (uint32_t by = 0; by < gy; by++)
            for (uint32_t bx = 0; bx < gx; bx++) {
                const int64_t k0 = int64_t(bx) * OWN_K2, j0 = int64_t(by) * tj, i0 = int64_t(bz) * seg;
                const int64_t i1 = std::min<int64_t>(i0 + seg, s.n[0]);
                const bool z = touches(2, k0 - 2, k0 + OWN_K2 + 1);
                const bool xy = touches(1, j0 - 2, j0 + tj + 1) || touches(0, i0 - 2, i1 + int64_t(reach));
                const bool sw = !side.empty() && ((side[size_t(bz) * gy + by] >> std::min<uint32_t>(bx, 63)) & 1u);
                out[xy ? 2 : z ? 1 : 0].push_back(MarchBlock{bx, by, bz | (sw ? SIDE_BLOCK : 0u)});
            }
}

// the segment tuner of the two-step march: runs of at least SEG2_TUNE_MIN_STEPS steps time
// SEG2_TUNE_PAIRS passes of each candidate
constexpr uint64_t SEG2_TUNE_MIN_STEPS = 800;
constexpr uint32_t SEG2_TUNE_PAIRS = 3;

void plan_march2(Session &s) {
    std::vector<MarchBlock> cls[3];
    march_classes2(s, s.tj2, s.seg2, cls, 0, true);
    for (int c = 0; c < 3; c++) {
        first_per_xcd(cls[c], s.xcds, [](const MarchBlock &m) { return (m.bz & SIDE_BLOCK) != 0; });
        xcd_order(cls[c], s.xcds, MarchBlock{~0u, ~0u, ~0u});
    }
    s.blk2_int.upload(cls[0].data(), cls[0].size(), s.stream);
    s.blk2_z.upload(cls[1].data(), cls[1].size(), s.stream);
    s.blk2_pml.upload(cls[2].data(), cls[2].size(), s.stream);
    // the three lists in one for k_march2_mix, each block tagged with its class, the longest
    // (x/y, then z) first; every list is padded to a multiple of the XCD count, so the
    // concatenation keeps each list's XCD order
    std::vector<MarchBlock> mix;
    for (int c = 2; c >= 0; c--)
        for (MarchBlock m : cls[c]) {
            if (m.bx != ~0u) m.bz |= uint32_t(c) << 30;
            mix.push_back(m);
        }
    s.blk2_mix.upload(mix.data(), mix.size(), s.stream);
}

static void plan_march(Session &s) {
    std::vector<MarchBlock> cls[3];
    march_classes(s, s.tj, s.seg, cls);
    for (int c = 0; c < 3; c++) xcd_order(cls[c], s.xcds, MarchBlock{~0u, ~0u, ~0u});
    s.blk_int.upload(cls[0].data(), cls[0].size(), s.stream);
    s.blk_z.upload(cls[1].data(), cls[1].size(), s.stream);
    s.blk_pml.upload(cls[2].data(), cls[2].size(), s.stream);
}

// The points the first step of a two-step pass must hand to its plane-DFT and NTFF observers
// (Cap2): every (field, node) they read, grouped per (i, j) row into segments of consecutive k
// with the same fields, and the per-observer index tables into the capture buffer. With `obs`,
// also every port edge, loop term and probe (Cap2::obs).
static void capture_points(Session &s, const fg_problem *P, const fg_extras *X, bool obs) {
    hipStream_t st = s.stream;
    const uint64_t nn = s.nn, n1 = P->n[1], n2 = P->n[2];
    const ObsArgs &o = s.obs;
    std::vector<uint64_t> pts;  // field * nn + id, field 0..5 = Ex Ey Ez Hx Hy Hz
    std::vector<uint64_t> pkeys, nkeys;
    if (o.plane_n)
        for (uint32_t q = 0; q < o.plane_n; q++) {
            const uint64_t id = (uint64_t(q / n1) * n1 + q % n1) * n2 + uint64_t(X->plane_k);
            for (uint64_t c : {0, 1, 2, 3, 4}) pkeys.push_back(c * nn + id);
        }
    for (uint64_t g = 0; g < o.n_patches; g++) {
        const uint32_t *ix = X->ntff + g * 13;
        const uint64_t ax = ix[0] & 3u, u = (ax + 1) % 3, v = (ax + 2) % 3;
        const uint64_t f[12] = {u, u, v, v, 3 + u, 3 + u, 3 + u, 3 + u, 3 + v, 3 + v, 3 + v, 3 + v};
        for (int t = 0; t < 12; t++) nkeys.push_back(f[t] * nn + ix[1 + t]);
    }
    // port edges, loop terms, probes, at their index in Cap2::obs (P->port_edges, then P->loops,
    // then X->probes); an entry no port refers to stays unset (~0) and is never read
    const uint64_t unset = ~0ull;
    std::vector<uint64_t> okeys;
    if (obs) {
        okeys.assign(uint64_t(P->n_port_edges) + P->n_loops + X->n_probes, unset);
        s.cap_loop0 = P->n_port_edges;
        s.cap_probe0 = P->n_port_edges + P->n_loops;
        for (uint32_t p = 0; p < P->n_ports; p++) {
            const fg_port &d = P->ports[p];
            for (uint32_t q = 0; q < d.n_edges; q++)
                okeys[d.first_edge + q] = uint64_t(d.comp) * nn + P->port_edges[d.first_edge + q].id;
            for (uint32_t q = 0; q < d.n_loops; q++) {
                const fg_loop_term &lt = P->loops[d.first_loop + q];
                okeys[s.cap_loop0 + d.first_loop + q] = (3ull + lt.comp) * nn + lt.id;
            }
        }
        for (uint32_t q = 0; q < X->n_probes; q++) {
            const fg_point_probe &pr = X->probes[q];
            okeys[s.cap_probe0 + q] = uint64_t(pr.field) * nn + (uint64_t(pr.i) * n1 + pr.j) * n2 + pr.k;
        }
    }
    pts.insert(pts.end(), pkeys.begin(), pkeys.end());
    pts.insert(pts.end(), nkeys.begin(), nkeys.end());
    for (uint64_t key : okeys)
        if (key != unset) pts.push_back(key);
    // per node id: the mask of captured fields
    std::vector<std::pair<uint64_t, uint32_t>> nodes;  // (id, mask)
    nodes.reserve(pts.size());
    for (uint64_t key : pts) nodes.push_back({key % nn, 1u << uint32_t(key / nn)});
    std::sort(nodes.begin(), nodes.end());
    std::vector<std::pair<uint64_t, uint32_t>> merged;
    for (const auto &p : nodes) {
        if (!merged.empty() && merged.back().first == p.first) merged.back().second |= p.second;
        else merged.push_back(p);
    }
    s.side_cap.assign(uint64_t(P->n[0]) * n1, 0);
    for (const auto &p : merged) s.side_cap[p.first / n2] |= side_tiles(uint32_t(p.first % n2), uint32_t(n2), false);
    // segments: consecutive k of one row with one mask; node -> first index of its fields
    std::vector<uint32_t> start(uint64_t(P->n[0]) * n1 + 1, 0);
    std::vector<uint4> seg;
    std::vector<uint64_t> first(merged.size());
    uint64_t count = 0;
    for (size_t q = 0; q < merged.size(); q++) {
        const uint64_t id = merged[q].first, row = id / n2, k = id % n2;
        const uint32_t mask = merged[q].second;
        const bool extend = q > 0 && !seg.empty() && merged[q - 1].first + 1 == id && merged[q - 1].second == mask &&
                            merged[q - 1].first / n2 == row;
        if (extend) {
            seg.back().y = uint32_t(k + 1);
        } else {
            seg.push_back(make_uint4(uint32_t(k), uint32_t(k + 1), mask, uint32_t(count)));
            start[row + 1]++;
        }
        first[q] = count;
        count += uint32_t(__builtin_popcount(mask));
    }
    if (count >= (1ull << 32)) throw HipError{"too many observed points for the two-step march", FG_TOO_LARGE};
    for (size_t r = 1; r < start.size(); r++) start[r] += start[r - 1];
    auto index = [&](uint64_t key) {
        const uint64_t id = key % nn, c = key / nn;
        const size_t q = size_t(std::lower_bound(merged.begin(), merged.end(), std::make_pair(id, 0u)) - merged.begin());
        const uint32_t mask = merged[q].second;
        return uint32_t(first[q] + uint32_t(__builtin_popcount(mask & ((1u << c) - 1u))));
    };
    std::vector<uint32_t> pidx(pkeys.size()), nidx(nkeys.size()), oidx(okeys.size());
    for (size_t q = 0; q < pkeys.size(); q++) pidx[q] = index(pkeys[q]);
    for (size_t q = 0; q < nkeys.size(); q++) nidx[q] = index(nkeys[q]);
    for (size_t q = 0; q < okeys.size(); q++) oidx[q] = okeys[q] == unset ? 0u : index(okeys[q]);
    if (obs) {
        if (oidx.empty()) s.cap_obs.alloc(1);
        else s.cap_obs.upload(oidx.data(), oidx.size(), st);
    }
    s.cap_start.upload(start.data(), start.size(), st);
    s.cap_seg.upload(seg.data(), seg.size(), st);
    s.cap_buf.alloc(std::max<uint64_t>(count, 1));
    if (pidx.empty()) s.cap_plane.alloc(1);
    else s.cap_plane.upload(pidx.data(), pidx.size(), st);
    if (nidx.empty()) s.cap_ntff.alloc(1);
    else s.cap_ntff.upload(nidx.data(), nidx.size(), st);
    FG_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

Session *session_new(fg_device *dev, const fg_problem *P, uint32_t driven, const fg_pulse *pulse,
                            const fg_extras *X, bool obs_capture) {
    validate(P, driven, X);
    auto S = std::make_unique<Session>();
    Session &s = *S;
    s.dev = dev;
    FG_CHECK(hipSetDevice(dev->ordinal));
    for (int a = 0; a < 3; a++) s.n[a] = P->n[a];
    const uint64_t nn = s.nn = uint64_t(P->n[0]) * P->n[1] * P->n[2];
    s.fused = X->engine != FG_ENGINE_SPLIT;
    s.dt = P->dt;
    // capacity check before touching HBM: "{cells} cells is more than this GPU can hold"
    uint64_t psi_total = 0, off[18];
    for (int e = 0; e < 2; e++)
        for (int c = 0; c < 3; c++)
            for (int a = 0; a < 3; a++) {
                off[e * 9 + c * 3 + a] = psi_total;
                if (a != c) psi_total += psi_len(P->n, a, P->pml);
            }
    const int copies = s.fused ? 2 : 1;
    size_t free_b = 0, total_b = 0;
    FG_CHECK(hipMemGetInfo(&free_b, &total_b));
    if (session_bytes(P, s.fused) > double(free_b)) {
        const uint64_t cells = uint64_t(P->n[0] - 1) * (P->n[1] - 1) * (P->n[2] - 1);
        throw HipError{std::to_string(cells) + " cells is more than this GPU can hold", FG_TOO_LARGE};
    }
    s.stream = take_stream(dev);
    hipStream_t st = s.stream;
    Stager stage(st);               // the tables below go up through pinned memory
    ArenaScope arena_scope(s.arena);  // and the small ones into the session's arena
    // concurrent block classes: on a 16.7 M-node grid the classes' last rounds overlap for ~10 %;
    // below ~8 M nodes a pass lasts ~100 us and the fork/join latency costs more than it saves,
    // and from ~16 blocks per CU (32-plane two-step blocks) the one mixed launch, longest blocks
    // first, wins (768x768x384, 42 blocks per CU: 2.247 against 2.318 ms per step)
    const char *cc = getenv("FIELDGRID_STREAMS");
    const uint64_t blocks32 = uint64_t((P->n[2] + OWN_K2 - 1) / OWN_K2) * ((P->n[1] + 11) / 12) * ((P->n[0] + 31) / 32);
    const bool side = cc ? cc[0] != '0' : nn >= (8ull << 20) && blocks32 < 16ull * uint64_t(std::max(dev->cus, 1));
    if (const char *m = getenv("FIELDGRID_MIX2")) s.mix2 = m[0] != '0';
    if (s.fused && side) {
        for (int q = 0; q < 2; q++) {
            s.side[q] = take_stream(dev);
            FG_CHECK(hipEventCreateWithFlags(&s.join[q], hipEventDisableTiming));
        }
        FG_CHECK(hipEventCreateWithFlags(&s.fork, hipEventDisableTiming));
    }
    {
        // E, H and psi of every copy in one allocation (one hipMalloc/hipFree instead of six),
        // each array at a 256-byte boundary, all of it zeroed
        auto round = [](uint64_t bytes) { return (bytes + 255) & ~uint64_t(255); };
        const uint64_t eb = round(12 * nn), pb = round(4 * std::max<uint64_t>(psi_total, 1));
        s.field_mem.alloc(copies * (2 * eb + pb));
        s.field_mem.zero(st);
        char *at = s.field_mem.p;
        for (int c = 0; c < copies; c++) {
            s.E[c].carve(reinterpret_cast<float *>(at), 3 * nn), at += eb;
            s.H[c].carve(reinterpret_cast<float *>(at), 3 * nn), at += eb;
            s.psi[c].carve(reinterpret_cast<float *>(at), std::max<uint64_t>(psi_total, 1)), at += pb;
        }
    }
    s.coef_index.upload(P->coef_index, P->coef_wide ? 3 * nn : (3 * nn + 1) / 2, st);
    s.coef_sets.upload(P->coef_sets, 2 * size_t(P->n_sets), st);
    {
        // every index must name a set: the kernels read sets[index] unchecked
        DevBuf<uint32_t> top;
        top.alloc(1);
        top.zero(st);
        const uint64_t words = s.coef_index.n;
        const uint32_t blocks = uint32_t(std::min<uint64_t>((words + 255) / 256, 4096));
        hipLaunchKernelGGL(k_index_max, dim3(blocks), dim3(256), 0, st, s.coef_index.p, words, bool(P->coef_wide), top.p);
        uint32_t m = 0;
        FG_CHECK(hipMemcpyAsync(&m, top.p, 4, hipMemcpyDeviceToHost, st));
        FG_CHECK(hipStreamSynchronize(st));
        if (m >= P->n_sets)
            throw HipError{"coefficient-set index " + std::to_string(m) + " is past the " + std::to_string(P->n_sets) +
                               "-entry set table", FG_INVALID};
    }
    // axes: AoS rows of 10 -> SoA per axis (8 float fields + 2 int slots)
    const uint32_t nmax = std::max({P->n[0], P->n[1], P->n[2]});
    std::vector<float> axf(3 * 8 * size_t(nmax), 0.f);
    std::vector<int> axs(3 * 2 * size_t(nmax), -1);
    uint32_t row = 0;
    for (int a = 0; a < 3; a++)
        for (uint32_t i = 0; i < P->n[a]; i++, row++) {
            const float *r = P->axes + 10 * size_t(row);
            for (int f = 0; f < 8; f++) axf[(size_t(a) * 8 + f) * nmax + i] = r[f];
            // slots address the CPML psi slabs: -1 or a whole number below 2 pml
            for (int f = 8; f < 10; f++)
                if (!(r[f] >= -1.f && r[f] < 2.f * float(P->pml)) || r[f] != float(int(r[f])))
                    throw HipError{"an axes-table CPML slot is not -1 or below 2 pml", FG_INVALID};
            axs[(size_t(a) * 2 + 0) * nmax + i] = int(r[8]);
            axs[(size_t(a) * 2 + 1) * nmax + i] = int(r[9]);
        }
    s.axes.upload(axf.data(), axf.size(), st);
    s.slots.upload(axs.data(), axs.size(), st);
    Geo &g = s.geo;
    g.n0 = P->n[0], g.n1 = P->n[1], g.n2 = P->n[2], g.pml = P->pml;
    g.nn = nn;
    g.k_mu = float(P->dt / 1.25663706212e-6);
    g.wide = P->coef_wide;
    for (int a = 0; a < 3; a++) {
        const float *b = s.axes.p + size_t(a) * 8 * nmax;
        g.ax[a] = AxDev{b, b + nmax, b + 2 * nmax, b + 3 * nmax, b + 4 * nmax, b + 5 * nmax, b + 6 * nmax, b + 7 * nmax,
                        s.slots.p + (size_t(a) * 2) * nmax, s.slots.p + (size_t(a) * 2 + 1) * nmax};
    }
    for (int q = 0; q < 18; q++) g.psi_off[q] = off[q];
    g.coef_index = s.coef_index.p;
    g.coef_sets = reinterpret_cast<const float2 *>(s.coef_sets.p);
    g.n_sets = P->n_sets;
    // sparse fixes
    s.max_steps = X->max_steps;
    s.step.alloc(1);
    s.step.zero(st);
    FixArgs &f = s.fix;
    f.n_debye = P->n_debye;
    f.np = P->n_poles;
    if (P->n_debye) {
        s.debye.upload(P->debye_edges, P->n_debye, st);
        s.dtab.upload(P->debye_table, size_t(std::max(1u, P->n_pole_sets)) * P->n_poles * 2, st);
        s.dstate.alloc(std::max<uint64_t>(uint64_t(P->n_poles) * P->n_debye, 1));
        s.dstate.zero(st);
        if (!s.fused) s.dold.alloc(P->n_debye);
    }
    f.debye = s.debye.p;
    f.dtab = s.dtab.p;
    f.dstate = s.dstate.p;
    f.dold = s.fused ? nullptr : s.dold.p;
    if (P->n_ports) {
        const fg_port &dp = P->ports[driven];
        f.n_src = dp.n_sources;
        s.src.upload(P->sources + dp.first_source, dp.n_sources, st);
    }
    f.src = s.src.p;
    f.n_ind = P->n_inductors;
    if (P->n_inductors) {
        s.ind.upload(P->inductors, P->n_inductors, st);
        std::vector<float> i0(P->n_inductors);
        for (uint32_t q = 0; q < P->n_inductors; q++) i0[q] = P->inductors[q].i0;
        s.ind_i.upload(i0.data(), i0.size(), st);
        s.ind_i0 = i0;
    }
    f.ind = s.ind.p;
    // resistive sheets: records (state updated in place, restored by reset), branch currents
    if (P->n_sheets) {
        s.sheets0.assign(P->sheets, P->sheets + 16 * P->n_sheets);
        s.sheets.upload(P->sheets, 16 * P->n_sheets, st);
        s.branches.upload(P->branches, P->n_branches, st);
        const uint64_t nb = uint64_t(P->branches[0]) + 2 * uint64_t(P->branches[1]);
        s.sheet_cur.alloc(std::max<uint64_t>(P->n_sheets * nb, 1));
        s.sheet_cur.zero(st);
        std::vector<uint32_t> fl(P->n_sheets);
        for (uint64_t q = 0; q < P->n_sheets; q++) std::memcpy(&fl[q], P->sheets + 16 * q, 4);
        std::sort(fl.begin(), fl.end());
        const bool shared = std::adjacent_find(fl.begin(), fl.end()) != fl.end();
        s.sheet = SheetArgs{P->n_sheets, s.sheets.p, s.sheet_cur.p, s.branches.p, shared ? 1u : 0u};
    }
    // thin wires: the table (split engine) and its per-plane list of H nodes (fused march)
    if (P->n_wires) {
        s.wires.upload(P->wires, P->n_wires, st);
        std::vector<uint64_t> hs(P->n_wires);
        for (uint64_t q = 0; q < P->n_wires; q++) hs[q] = P->wires[q].h;
        std::sort(hs.begin(), hs.end());
        s.wire_serial = std::adjacent_find(hs.begin(), hs.end()) != hs.end();
        std::vector<std::vector<uint4>> by_plane(P->n[0]);
        for (uint64_t q = 0; q < P->n_wires; q++) {
            const uint64_t h = P->wires[q].h, c = h / nn, id = h % nn;
            const uint32_t i = uint32_t(id / (uint64_t(P->n[1]) * P->n[2])), j = uint32_t((id / P->n[2]) % P->n[1]),
                           k = uint32_t(id % P->n[2]);
            by_plane[i].push_back(make_uint4(j, k, uint32_t(c), uint32_t(q)));
        }
        std::vector<uint32_t> start(P->n[0] + 1, 0);
        std::vector<uint4> ent;
        for (uint32_t i = 0; i < P->n[0]; i++) {
            start[i] = uint32_t(ent.size());
            ent.insert(ent.end(), by_plane[i].begin(), by_plane[i].end());
        }
        start[P->n[0]] = uint32_t(ent.size());
        s.wire_start.upload(start.data(), start.size(), st);
        s.wire_ent.upload(ent.data(), ent.size(), st);
    }
    f.ind_i = s.ind_i.p;
    double t0, tau;
    tau = std::sqrt(std::log(10.0)) / (3.14159265358979323846 * pulse->fc);  // pulse_shape, sim.rs:646-649
    t0 = 4.0 * tau;
    f.dt = float(P->dt);
    f.t0 = float(t0);
    f.tau = float(tau);
    f.w0 = float(2.0 * 3.14159265358979323846 * pulse->f0);
    f.step = s.step.p;
    // pulse((n+1/2)dt) in f32, the expression of cpu.rs:262-265 evaluated on the host
    if (f.n_src && X->max_steps) {
        std::vector<float> pl(X->max_steps);
        const float dtf = float(P->dt), t0f = float(t0), tauf = float(tau), w0f = f.w0;
        for (uint64_t n = 0; n < X->max_steps; n++) {
            const float t = (float(double(n)) + 0.5f) * dtf;
            const float x = (t - t0f) / tauf;
            pl[n] = std::exp(-x * x) * std::cos(w0f * (t - t0f));
        }
        s.pulse.upload(pl.data(), pl.size(), st);
        f.n_pulse = uint32_t(std::min<uint64_t>(X->max_steps, UINT32_MAX));
    }
    f.pulse = s.pulse.p;
    // sources and inductors that share an edge must be applied in order, not in parallel
    {
        std::vector<uint64_t> fx;
        for (uint32_t q = 0; q < f.n_src; q++) fx.push_back(uint64_t(P->sources[P->ports[driven].first_source + q].comp) * nn +
                                                           P->sources[P->ports[driven].first_source + q].id);
        for (uint32_t q = 0; q < f.n_ind; q++) fx.push_back(uint64_t(P->inductors[q].comp) * nn + P->inductors[q].id);
        std::sort(fx.begin(), fx.end());
        f.serial = std::adjacent_find(fx.begin(), fx.end()) != fx.end() ? 1u : 0u;
    }
    s.fix_blocks = uint32_t((std::max<uint64_t>({f.n_debye, f.n_src, f.n_ind}) + 255) / 256);
    // observers
    ObsArgs &o = s.obs;
    s.n_ports = o.n_ports = P->n_ports;
    s.n_probes = o.n_probes = X->n_probes;
    s.nf = o.nf = X->nf;
    o.cap = X->max_steps;
    if (P->n_ports) {
        s.ports.upload(P->ports, P->n_ports, st);
        s.edges.upload(P->port_edges, P->n_port_edges, st);
        s.loops.upload(P->loops, P->n_loops, st);
        s.series.alloc(X->max_steps * P->n_ports * 2);
        s.series.zero(st);
    }
    o.ports = s.ports.p;
    o.edges = s.edges.p;
    o.loops = s.loops.p;
    o.series = s.series.p;
    if (X->n_probes) {
        std::vector<uint64_t> pf(X->n_probes);
        for (uint32_t q = 0; q < X->n_probes; q++) {
            const fg_point_probe &p = X->probes[q];
            const uint64_t id = (uint64_t(p.i) * P->n[1] + p.j) * P->n[2] + p.k;
            pf[q] = p.field < 3 ? p.field * nn + id : 3 * nn + (p.field - 3) * nn + id;
        }
        s.probe_f.upload(pf.data(), pf.size(), st);
        s.probes.alloc(X->max_steps * X->n_probes);
        s.probes.zero(st);
    }
    o.probe_f = s.probe_f.p;
    o.probes = s.probes.p;
    o.plane_n = X->plane_k >= 0 && X->nf ? P->n[0] * P->n[1] : 0;
    o.plane_k = X->plane_k >= 0 ? uint32_t(X->plane_k) : 0;
    s.plane_len = uint64_t(o.plane_n) * X->nf * 10;
    if (s.plane_len) {
        s.plane.alloc(s.plane_len);
        s.plane.zero(st);
    }
    o.plane = s.plane.p;
    o.n_patches = X->nf ? X->n_patches : 0;
    s.ntff_len = o.n_patches * X->nf * 12;
    if (s.ntff_len) {
        s.patch.upload(X->ntff, X->n_patches * 13, st);
        s.ntff.alloc(s.ntff_len);
        s.ntff.zero(st);
    }
    o.patch = s.patch.p;
    o.ntff = s.ntff.p;
    for (uint32_t q = 0; q < FG_MAX_FREQS; q++) o.omega[q] = q < X->nf ? float(2.0 * 3.14159265358979323846 * X->freqs[q]) : 0.f;
    o.dt = float(P->dt);
    // DFT phases per step in f32, the expressions of cpu.rs:352-353 evaluated on the host
    if (X->nf && (o.plane_n || o.n_patches) && X->max_steps) {
        std::vector<float> ph(X->max_steps * X->nf * 4);
        const float dtf = o.dt;
        for (uint64_t n = 0; n < X->max_steps; n++) {
            const float te = (float(double(n)) + 1.0f) * dtf, th = (float(double(n)) + 0.5f) * dtf;
            for (uint32_t k = 0; k < X->nf; k++) {
                const float w = o.omega[k];
                float *q = &ph[(n * X->nf + k) * 4];
                q[0] = std::cos(w * te) * dtf, q[1] = -std::sin(w * te) * dtf;
                q[2] = std::cos(w * th) * dtf, q[3] = -std::sin(w * th) * dtf;
            }
        }
        s.phase.upload(ph.data(), ph.size(), st);
        o.n_phase = uint32_t(std::min<uint64_t>(X->max_steps, UINT32_MAX));
    }
    o.phase = s.phase.p;
    o.blk_ports = (P->n_ports + 3) / 4;
    o.blk_probes = (X->n_probes + 255) / 256;
    o.blk_plane = (o.plane_n + 255) / 256;
    o.blk_ntff = uint32_t((o.n_patches + 255) / 256);
    o.step = s.step.p;
    s.obs_blocks = o.blk_ports + o.blk_probes + o.blk_plane + o.blk_ntff;
    s.partial.alloc(ENERGY_BLOCKS);
    // fused march: TJ owned rows (+2 halo rows) x 62 owned lanes, `seg` planes of i per block
    s.tj = 14;
    s.seg = 32;
    s.xcds = dev->cus >= 64 ? uint32_t(dev->cus / 32) : 1u;  // gfx950: 32 CUs per XCD
    if (const char *v = getenv("FIELDGRID_TJ")) s.tj = uint32_t(atoi(v));
    if (const char *v = getenv("FIELDGRID_SEG")) s.seg = std::max(1u, uint32_t(atoi(v)));
    if (const char *v = getenv("FIELDGRID_XCDS")) s.xcds = std::max(1u, uint32_t(atoi(v)));
    if (s.tj != 8 && s.tj != 14) s.tj = 14;
    if (P->n_wires) s.tj = 14;  // the thin-wire march comes in one tile height
    s.seg = std::min<uint32_t>(s.seg, SEG_MAX);
    // a grid too small to give ~6 blocks per CU at `seg` planes per block marches shorter
    // segments instead (the results do not depend on the tiling)
    if (!getenv("FIELDGRID_SEG"))
        while (s.seg > 4 && uint64_t((P->n[2] + OWN_K - 1) / OWN_K) * ((P->n[1] + s.tj - 1) / s.tj) *
                                    ((P->n[0] + s.seg - 1) / s.seg) < 6ull * uint64_t(std::max(dev->cus, 1)))
            s.seg /= 2;
    s.pml = P->pml;
    s.nmax = nmax;
    s.slots_host = axs;
    if (s.fused) plan_march(s);
    // two steps per launch when only port sources and Debye media act between steps, and no
    // observer of the first step reads a Debye edge (the first step's observers recompute their
    // values from the previous copy without the ADE)
    bool observed_debye = false;
    if (P->n_debye) {
        std::vector<uint64_t> dfl(P->n_debye);
        for (uint64_t t = 0; t < P->n_debye; t++) dfl[t] = P->debye_edges[t].f;
        std::sort(dfl.begin(), dfl.end());
        auto is_debye = [&](uint64_t flat) { return std::binary_search(dfl.begin(), dfl.end(), flat); };
        for (uint32_t p = 0; p < P->n_ports && !observed_debye; p++)
            for (uint32_t q = 0; q < P->ports[p].n_edges; q++)
                if (is_debye(P->ports[p].comp * nn + P->port_edges[P->ports[p].first_edge + q].id)) observed_debye = true;
        for (uint32_t q = 0; q < X->n_probes; q++) {
            const fg_point_probe &pp = X->probes[q];
            if (pp.field < 3 && is_debye(pp.field * nn + (uint64_t(pp.i) * P->n[1] + pp.j) * P->n[2] + pp.k))
                observed_debye = true;
        }
    }
    const char *f2 = getenv("FIELDGRID_FUSE2");
    const uint32_t seg2_cap = march2_seg_cap(march2_plane_bytes(P->n, P->pml));
    s.fuse2 = s.fused && !(f2 && f2[0] == '0') && !observed_debye && f.n_ind == 0 && !f.serial && P->n_sheets == 0 &&
              P->n_wires == 0 && seg2_cap > 0;
    if (s.fuse2) {
        s.seg2 = 32;
        if (const char *v = getenv("FIELDGRID_SEG2")) s.seg2 = std::max(1u, uint32_t(atoi(v)));
        s.seg2 = std::min<uint32_t>(s.seg2, seg2_cap);
        // Blocks are 1024 threads, one or two per CU, and a segment of S planes costs S + 3 plane
        // iterations: halve the segment only while the pass has fewer than 1.5 blocks per CU.
        // (Measured best: 32 planes at 768x768x384, 16 at 192x192x112 and 256^3 (32 within 3 %),
        // 4 at 64^3.)
        if (!getenv("FIELDGRID_SEG2")) {
            const uint64_t cus = uint64_t(std::max(dev->cus, 1));
            std::vector<MarchBlock> cls[3];
            while (s.seg2 > 4) {
                march_classes2(s, s.tj2, s.seg2, cls);
                if (2 * (cls[0].size() + cls[1].size() + cls[2].size()) >= 3 * cus) break;
                s.seg2 /= 2;
            }
            // The best length also depends on how evenly the segments cut the i extent and on
            // how the side work spreads over the blocks (measured at 256^3 with Debye media and
            // NTFF: 16 planes 4 % faster than 32; at 768x768x384: 24 planes 0.5 % faster): a
            // long run times a few passes of seg2, 3/4, 1/2, 3/8 and 1/4 of it first.
            const char *tu = getenv("FIELDGRID_SEG2_TUNE");
            if (!(tu && tu[0] == '0') && X->max_steps >= SEG2_TUNE_MIN_STEPS)
                for (uint32_t num : {8u, 6u, 4u, 3u, 2u}) {
                    const uint32_t c = s.seg2 * num / 8;
                    if (c >= 4 && (s.seg2_try.empty() || s.seg2_try.back() != c)) s.seg2_try.push_back(c);
                }
            if (s.seg2_try.size() < 2) s.seg2_try.clear();
        }
        s.seg2_base = s.seg2;
        // the driven port's sources by plane (i), for the first step inside the launch
        std::vector<std::vector<uint4>> by_plane(P->n[0]);
        if (P->n_ports)
            for (uint32_t q = 0; q < f.n_src; q++) {
                const fg_source &e = P->sources[P->ports[driven].first_source + q];
                const uint64_t id = e.id;
                const uint32_t i = uint32_t(id / (uint64_t(P->n[1]) * P->n[2])), j = uint32_t((id / P->n[2]) % P->n[1]),
                               k = uint32_t(id % P->n[2]);
                uint32_t bits;
                std::memcpy(&bits, &e.src, 4);
                by_plane[i].push_back(make_uint4(j, k, e.comp, bits));
            }
        std::vector<uint32_t> start(P->n[0] + 1, 0);
        std::vector<uint4> ent;
        for (uint32_t i = 0; i < P->n[0]; i++) {
            start[i] = uint32_t(ent.size());
            ent.insert(ent.end(), by_plane[i].begin(), by_plane[i].end());
        }
        start[P->n[0]] = uint32_t(ent.size());
        s.src2_start.upload(start.data(), start.size(), st);
        if (ent.empty()) s.src2_ent.alloc(1);
        else s.src2_ent.upload(ent.data(), ent.size(), st);
        // the Debye edges' k maps, one per (i, j) row that holds any
        if (P->n_debye) {
            if (P->n_debye >= (1ull << 31)) throw HipError{"too many Debye edges for the two-step march", FG_TOO_LARGE};
            std::vector<int32_t> row(uint64_t(P->n[0]) * P->n[1], -1), map;
            for (uint64_t t = 0; t < P->n_debye; t++) {
                const uint64_t fl = P->debye_edges[t].f, c = fl / nn, id = fl % nn;
                const uint64_t r = id / P->n[2], k = id % P->n[2];
                if (row[r] < 0) {
                    row[r] = int32_t(map.size() / (3ull * P->n[2]));
                    map.resize(map.size() + 3ull * P->n[2], -1);
                }
                map[(uint64_t(row[r]) * P->n[2] + k) * 3 + c] = int32_t(t);
            }
            s.deb2_row.upload(row.data(), row.size(), st);
            s.deb2_map.upload(map.data(), map.size(), st);
            s.side_deb.assign(uint64_t(P->n[0]) * P->n[1], 0);
            for (uint64_t t = 0; t < P->n_debye; t++) {
                const uint64_t id = P->debye_edges[t].f % nn;
                s.side_deb[id / P->n[2]] |= side_tiles(uint32_t(id % P->n[2]), P->n[2], true);
            }
            s.deb2_jn.alloc(std::max<uint64_t>(uint64_t(P->n_poles) * P->n_debye, 1));
            s.deb2_eo.alloc(P->n_debye);
        }
        if (o.plane_n || o.n_patches || obs_capture) capture_points(s, P, X, obs_capture);
        plan_march2(s);
    }
    FG_CHECK(hipStreamSynchronize(st));
    return S.release();
}

void reset(Session &s) {
    hipStream_t st = s.stream;
    if (s.sheet.n) {
        FG_CHECK(hipMemcpyAsync(s.sheets.p, s.sheets0.data(), s.sheets0.size() * 4, hipMemcpyHostToDevice, st));
        s.sheet_cur.zero(st);
    }
    for (int c = 0; c < 2; c++) {
        s.E[c].zero(st);
        s.H[c].zero(st);
        s.psi[c].zero(st);
    }
    s.dstate.zero(st);
    s.series.zero(st);
    s.plane.zero(st);
    s.ntff.zero(st);
    s.probes.zero(st);
    if (!s.ind_i0.empty())
        FG_CHECK(hipMemcpyAsync(s.ind_i.p, s.ind_i0.data(), s.ind_i0.size() * 4, hipMemcpyHostToDevice, st));
    FG_CHECK(hipStreamSynchronize(st));
    s.steps = 0;
    s.cur = 0;
    s.peak = s.decay = 0.0;
}

static hipEvent_t timing_event(Session &s, size_t slot) {
    while (s.events.size() <= slot) {
        hipEvent_t e;
        FG_CHECK(hipEventCreate(&e));
        s.events.push_back(e);
    }
    return s.events[slot];
}

// Fold the recorded event pairs of the last enqueue into the per-kernel-group totals.
static void collect_timing(Session &s) {
    if (!s.timing || s.units.empty()) return;
    FG_CHECK(hipStreamSynchronize(s.stream));
    for (const auto &u : s.units) {
        for (int g = 0; g < 3; g++) {
            float ms = 0.f;
            FG_CHECK(hipEventElapsedTime(&ms, s.events[u.first + g], s.events[u.first + g + 1]));
            s.kernel_ms[g] += ms;
        }
        s.kernel_count += u.second;
    }
    s.units.clear();
}

// Extras.profile (run.rs:133-176): the mean time per step of each kernel group, in the
// reference's wording, on stderr; then the totals restart.
static void report_timing(Session &s) {
    collect_timing(s);
    if (!s.kernel_count) return;
    const char *march = !s.fused ? "k_update_h + k_update_e" : s.fuse2 ? "k_march2 (two steps per pass)" : "k_march";
    const char *names[3] = {march, "k_fix", "k_observe"};
    double step = 0.0;
    for (int g = 0; g < 3; g++) {
        const double us = s.kernel_ms[g] * 1e3 / double(s.kernel_count);
        step += us;
        std::fprintf(stderr, "fdtd kernel %s (every step): %.1f us over %llu steps\n", names[g], us,
                     static_cast<unsigned long long>(s.kernel_count));
    }
    std::fprintf(stderr, "fdtd kernel time per step: %.1f us\n", step);
    s.kernel_ms[0] = s.kernel_ms[1] = s.kernel_ms[2] = 0.0;
    s.kernel_count = 0;
}

void fork_classes(Session &s, hipStream_t out[3]) {
    out[0] = out[1] = out[2] = s.stream;
    if (!s.side[0]) return;
    FG_CHECK(hipEventRecord(s.fork, s.stream));
    for (int q = 0; q < 2; q++) {
        FG_CHECK(hipStreamWaitEvent(s.side[q], s.fork, 0));
        out[1 + q] = s.side[q];
    }
}

void join_classes(Session &s) {
    if (!s.side[0]) return;
    for (int q = 0; q < 2; q++) {
        FG_CHECK(hipEventRecord(s.join[q], s.side[q]));
        FG_CHECK(hipStreamWaitEvent(s.stream, s.join[q], 0));
    }
}

// Enqueue one step at offset s_off from the device step counter.
void launch_step(Session &s, uint32_t s_off, uint64_t global_step) {
    hipStream_t st = s.stream;
    const bool timed = s.timing;
    const size_t slot = 4 * size_t(s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot), st));
    if (timed) s.units.push_back({slot, 1u});
    const Geo &g = s.geo;
    const int a = s.cur, b = 1 - s.cur;
    float *Ecur;
    const float *Eold = nullptr;
    if (s.fused) {
        dim3 block(64, s.tj + 2);
        const float *Eo = s.E[a].p, *Ho = s.H[a].p, *Po = s.psi[a].p;
        float *En = s.E[b].p, *Hn = s.H[b].p, *Pn = s.psi[b].p;
        const uint32_t ni = uint32_t(s.blk_int.n), nz = uint32_t(s.blk_z.n), np_ = uint32_t(s.blk_pml.n);
        const MarchBlock *bi = s.blk_int.p, *bz = s.blk_z.p, *bp = s.blk_pml.p;
        hipStream_t cs[3];
        fork_classes(s, cs);
#define FG_MARCH2(TJ, W, L)                                                                                       \
    do {                                                                                                          \
        if (ni) hipLaunchKernelGGL((k_march<TJ, 0, W, L>), dim3(ni), block, 0, cs[0], g, bi, Eo, En, Ho, Hn, Po, Pn, s.seg); \
        if (nz) hipLaunchKernelGGL((k_march<TJ, 1, W, L>), dim3(nz), block, 0, cs[1], g, bz, Eo, En, Ho, Hn, Po, Pn, s.seg); \
        if (np_) hipLaunchKernelGGL((k_march<TJ, 2, W, L>), dim3(np_), block, 0, cs[2], g, bp, Eo, En, Ho, Hn, Po, Pn, s.seg); \
    } while (0)
#define FG_MARCH(TJ)                                                   \
    do {                                                               \
        const bool lds = g.n_sets <= uint32_t(SETS_LDS);               \
        if (g.wide) {                                                  \
            if (lds) FG_MARCH2(TJ, true, true); else FG_MARCH2(TJ, true, false);   \
        } else {                                                       \
            if (lds) FG_MARCH2(TJ, false, true); else FG_MARCH2(TJ, false, false); \
        }                                                              \
    } while (0)
        if (s.wires.n) {
            const WireArgs wa{s.wire_start.p, s.wire_ent.p, s.wires.p};
#define FG_MARCHW(W, L)                                                                                            \
    do {                                                                                                           \
        if (ni) hipLaunchKernelGGL((k_march_wires<0, W, L>), dim3(ni), block, 0, cs[0], g, bi, Eo, En, Ho, Hn, Po, Pn, s.seg, wa); \
        if (nz) hipLaunchKernelGGL((k_march_wires<1, W, L>), dim3(nz), block, 0, cs[1], g, bz, Eo, En, Ho, Hn, Po, Pn, s.seg, wa); \
        if (np_) hipLaunchKernelGGL((k_march_wires<2, W, L>), dim3(np_), block, 0, cs[2], g, bp, Eo, En, Ho, Hn, Po, Pn, s.seg, wa); \
    } while (0)
            const bool lds = g.n_sets <= uint32_t(SETS_LDS);
            if (g.wide) {
                if (lds) FG_MARCHW(true, true); else FG_MARCHW(true, false);
            } else {
                if (lds) FG_MARCHW(false, true); else FG_MARCHW(false, false);
            }
#undef FG_MARCHW
        } else {
            switch (s.tj) {
                case 8: FG_MARCH(8); break;
                default: FG_MARCH(14); break;
            }
        }
#undef FG_MARCH2
#undef FG_MARCH
        join_classes(s);
        if (s.sheet.n)
            hipLaunchKernelGGL(k_sheet, dim3(s.sheet.serial ? 1u : uint32_t((s.sheet.n + 255) / 256)), dim3(256), 0, st,
                               s.sheet, En, Eo, Hn);
        Ecur = s.E[b].p;
        Eold = s.E[a].p;
        s.cur = b;
    } else {
        dim3 block(64, 4);
        dim3 grid((g.n2 + 63) / 64, (g.n1 + 3) / 4, g.n0);
        hipLaunchKernelGGL(k_update_h, grid, block, 0, st, g, s.E[0].p, s.H[0].p, s.psi[0].p);
        if (s.wires.n)
            hipLaunchKernelGGL(k_wires, dim3(s.wire_serial ? 1u : uint32_t((s.wires.n + 255) / 256)), dim3(256), 0, st,
                               s.wires.p, s.wires.n, s.wire_serial ? 1u : 0u, s.E[0].p, s.H[0].p);
        const uint32_t sb = s.sheet.serial ? 1u : uint32_t((s.sheet.n + 255) / 256);
        if (s.sheet.n)
            hipLaunchKernelGGL(k_sheet_pre, dim3(uint32_t((s.sheet.n + 255) / 256)), dim3(256), 0, st, s.sheet, s.E[0].p,
                               s.H[0].p);
        if (s.fix.n_debye)
            hipLaunchKernelGGL(k_debye_save, dim3(uint32_t((s.fix.n_debye + 255) / 256)), dim3(256), 0, st, s.fix.n_debye,
                               s.fix.debye, s.E[0].p, s.dold.p);
        hipLaunchKernelGGL(k_update_e, grid, block, 0, st, g, s.E[0].p, s.H[0].p, s.psi[0].p);
        if (s.sheet.n) hipLaunchKernelGGL(k_sheet_post, dim3(sb), dim3(256), 0, st, s.sheet, s.E[0].p);
        Ecur = s.E[0].p;
    }
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 1), st));
    if (s.fix_blocks) hipLaunchKernelGGL(k_fix, dim3(s.fix_blocks), dim3(256), 0, st, g, s.fix, Ecur, Eold, s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 2), st));
    if (s.obs_blocks && s.observe)
        hipLaunchKernelGGL(k_observe, dim3(s.obs_blocks), dim3(256), 0, st, g, s.obs, Ecur, s.H[s.fused ? s.cur : 0].p, s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 3), st));
    (void)global_step;
    FG_CHECK(hipGetLastError());
}

// Enqueue two steps (s_off and s_off + 1) with one launch per block class of march2: the first
// step's observers recompute from the copy the launch reads, the second's run as usual after
// its sparse fixes.
void launch_pair(Session &s, uint32_t s_off) {
    hipStream_t st = s.stream;
    const bool timed = s.timing;
    const size_t slot = 4 * size_t(s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot), st));
    if (timed) s.units.push_back({slot, 2u});
    const Geo &g = s.geo;
    const int a = s.cur, b = 1 - s.cur;
    const dim3 block(64, s.tj2 + 4);
    const float *Eo = s.E[a].p, *Ho = s.H[a].p, *Po = s.psi[a].p;
    float *En = s.E[b].p, *Hn = s.H[b].p, *Pn = s.psi[b].p;
    const uint32_t ni = uint32_t(s.blk2_int.n), nz = uint32_t(s.blk2_z.n), np_ = uint32_t(s.blk2_pml.n);
    const MarchBlock *bi = s.blk2_int.p, *bz = s.blk2_z.p, *bp = s.blk2_pml.p;
    const Src2 sr{s.src2_start.p, s.src2_ent.p};
    const Deb2 db{s.deb2_row.p, s.deb2_map.p, s.deb2_jn.p, s.deb2_eo.p};
    const Cap2 cp{s.cap_start.p, s.cap_seg.p, s.cap_buf.p, s.cap_ntff.p, s.cap_plane.p};
    // without side streams the three class launches would run one after another: one launch of
    // the mixed list instead
    const bool mix = !s.side[0] && s.mix2;
    hipStream_t cs[3];
    if (!mix) fork_classes(s, cs);
#define FG_MARCH2M(TJ, W, L, D)                                                                                      \
    hipLaunchKernelGGL((k_march2_mix<TJ, W, L, D>), dim3(uint32_t(s.blk2_mix.n)), block, 0, st, g, s.blk2_mix.p, Eo, En, \
                       Ho, Hn, Po, Pn, s.seg2, sr, s.fix, db, cp, s_off)
#define FG_MARCH2X(TJ, W, L, D)                                                                                      \
    do {                                                                                                             \
        if (ni) hipLaunchKernelGGL((k_march2<TJ, 0, W, L, D>), dim3(ni), block, 0, cs[0], g, bi, Eo, En, Ho, Hn, Po, Pn, s.seg2, sr, s.fix, db, cp, s_off); \
        if (nz) hipLaunchKernelGGL((k_march2<TJ, 1, W, L, D>), dim3(nz), block, 0, cs[1], g, bz, Eo, En, Ho, Hn, Po, Pn, s.seg2, sr, s.fix, db, cp, s_off); \
        if (np_) hipLaunchKernelGGL((k_march2<TJ, 2, W, L, D>), dim3(np_), block, 0, cs[2], g, bp, Eo, En, Ho, Hn, Po, Pn, s.seg2, sr, s.fix, db, cp, s_off); \
    } while (0)
#define FG_MARCH2D(W, L)                                     \
    do {                                                    \
        if (mix) {                                          \
            if (db.row || cp.start) FG_MARCH2M(12, W, L, true); \
            else FG_MARCH2M(12, W, L, false);               \
        } else if (db.row || cp.start) FG_MARCH2X(12, W, L, true); \
        else FG_MARCH2X(12, W, L, false);                   \
    } while (0)
    {
        // runs without Debye media or captured observers take the leaner kernel
        const bool lds = g.n_sets <= uint32_t(SETS_LDS);
        if (g.wide) {
            if (lds) FG_MARCH2D(true, true); else FG_MARCH2D(true, false);
        } else {
            if (lds) FG_MARCH2D(false, true); else FG_MARCH2D(false, false);
        }
    }
#undef FG_MARCH2D
#undef FG_MARCH2X
#undef FG_MARCH2M
    if (!mix) join_classes(s);
    s.cur = b;
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 1), st));
    // the second step's sparse fixes: Debye from the pole states and step-n E the pass wrote
    FixArgs f2 = s.fix;
    if (s.fix.n_debye) f2.dsrc = db.jn, f2.dold = db.eo;
    if (s.fix_blocks) hipLaunchKernelGGL(k_fix, dim3(s.fix_blocks), dim3(256), 0, st, g, f2, En, Eo, s_off + 1);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 2), st));
    if (s.obs_blocks && s.observe) {
        hipLaunchKernelGGL(k_observe_pair, dim3(s.obs_blocks), dim3(256), 0, st, g, s.obs, s.fix, Eo, Ho, Po, En, Hn,
                           s_off, cp);
    }
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 3), st));
    FG_CHECK(hipGetLastError());
}

// Steps [s.steps, s.steps + count) from offset 0 of the device counter: pairs through march2
// where it is on, single steps otherwise (and for an odd last one).
static void launch_steps(Session &s, uint64_t count) {
    uint64_t q = 0;
    if (s.fuse2)
        for (; q + 2 <= count; q += 2) launch_pair(s, uint32_t(q));
    for (; q < count; q++) launch_step(s, uint32_t(q), s.steps + q);
}

void set_step_counter(Session &s) {
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s.stream, s.step.p, uint32_t(s.steps));
}

uint64_t seg2_tune_steps(const Session &s) {
    return s.seg2_try.empty() ? 0 : 2 + 2ull * SEG2_TUNE_PAIRS * s.seg2_try.size();
}

uint64_t tune_seg2(Session &s, const std::function<void(uint64_t)> &chunk) {
    std::vector<uint32_t> cand;
    cand.swap(s.seg2_try);
    const uint32_t pairs = SEG2_TUNE_PAIRS;
    std::vector<hipEvent_t> ev(2 * cand.size());
    for (auto &e : ev) FG_CHECK(hipEventCreate(&e));
    chunk(2);
    for (size_t c = 0; c < cand.size(); c++) {
        s.seg2 = cand[c];
        plan_march2(s);
        FG_CHECK(hipEventRecord(ev[2 * c], s.stream));
        chunk(2 * pairs);
        FG_CHECK(hipEventRecord(ev[2 * c + 1], s.stream));
    }
    FG_CHECK(hipStreamSynchronize(s.stream));
    size_t best = 0;
    float best_ms = 0.f;
    for (size_t c = 0; c < cand.size(); c++) {
        float ms = 0.f;
        FG_CHECK(hipEventElapsedTime(&ms, ev[2 * c], ev[2 * c + 1]));
        if (c == 0 || ms < best_ms) best = c, best_ms = ms;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (best + 1 != cand.size()) {
        s.seg2 = cand[best];
        plan_march2(s);
        FG_CHECK(hipStreamSynchronize(s.stream));
    }
    return 2 + 2ull * pairs * cand.size();
}

void enqueue_steps(Session &s, uint64_t count) {
    if (!s.timing && !s.seg2_try.empty() && count >= seg2_tune_steps(s))
        count -= tune_seg2(s, [&](uint64_t k) {
            hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s.stream, s.step.p, uint32_t(s.steps));
            launch_steps(s, k);
            s.steps += k;
        });
    if (s.timing) {
        collect_timing(s);
        // timed steps go in batches so the event pool stays bounded
        for (uint64_t done = 0; done < count;) {
            const uint64_t k = std::min<uint64_t>(count - done, 256);
            hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s.stream, s.step.p, uint32_t(s.steps));
            launch_steps(s, k);
            s.steps += k;
            collect_timing(s);
            done += k;
        }
        return;
    }
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s.stream, s.step.p, uint32_t(s.steps));
    launch_steps(s, count);
    s.steps += count;
}

double energy(Session &s) {
    const float *E = s.E[s.fused ? s.cur : 0].p, *H = s.H[s.fused ? s.cur : 0].p;
    hipLaunchKernelGGL(k_energy, dim3(ENERGY_BLOCKS), dim3(256), 0, s.stream, E, H, 3 * s.nn, s.partial.p);
    FG_CHECK(hipGetLastError());
    std::vector<double> part(ENERGY_BLOCKS);
    FG_CHECK(hipMemcpyAsync(part.data(), s.partial.p, ENERGY_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    FG_CHECK(hipStreamSynchronize(s.stream));
    double e = 0.0;
    for (double v : part) e += v;
    return e;
}

void read_record(Session &s, fg_record *out) {
    auto need = [](const float *p, uint64_t cap, uint64_t want, const char *what) {
        if (want && (!p || cap < want))
            throw HipError{std::string("record buffer too small: ") + what, FG_INVALID};
    };
    const uint64_t steps = std::min(s.steps, s.max_steps);
    need(out->series, out->series_cap, steps * s.n_ports * 2, "series");
    need(out->plane, out->plane_cap, s.plane_len, "plane");
    need(out->ntff, out->ntff_cap, s.ntff_len, "ntff");
    need(out->probes, out->probes_cap, steps * s.n_probes, "probes");
    hipStream_t st = s.stream;
    if (steps * s.n_ports)
        FG_CHECK(hipMemcpyAsync(out->series, s.series.p, steps * s.n_ports * 2 * 4, hipMemcpyDeviceToHost, st));
    if (s.plane_len) FG_CHECK(hipMemcpyAsync(out->plane, s.plane.p, s.plane_len * 4, hipMemcpyDeviceToHost, st));
    if (s.ntff_len) FG_CHECK(hipMemcpyAsync(out->ntff, s.ntff.p, s.ntff_len * 4, hipMemcpyDeviceToHost, st));
    if (steps * s.n_probes)
        FG_CHECK(hipMemcpyAsync(out->probes, s.probes.p, steps * s.n_probes * 4, hipMemcpyDeviceToHost, st));
    const int c = s.fused ? s.cur : 0;
    if (out->e_out) FG_CHECK(hipMemcpyAsync(out->e_out, s.E[c].p, 3 * s.nn * 4, hipMemcpyDeviceToHost, st));
    if (out->h_out) FG_CHECK(hipMemcpyAsync(out->h_out, s.H[c].p, 3 * s.nn * 4, hipMemcpyDeviceToHost, st));
    FG_CHECK(hipStreamSynchronize(st));
    out->steps = s.steps;
    out->ports = s.n_ports;
    out->decay_db = s.decay;
}


}  // namespace fg

// ---- C-ABI -----------------------------------------------------------------------------------

extern "C" {

const char *fg_last_error(void) { return fg::g_error.c_str(); }

int fg_open(int ordinal, fg_device **out) {
    return fg::guarded([&] {
        if (!out) throw fg::HipError{"fg_open: NULL output", FG_INVALID};
        int count = 0;
        FG_CHECK(hipGetDeviceCount(&count));
        if (ordinal < 0 || ordinal >= count)
            throw fg::HipError{"no HIP device " + std::to_string(ordinal) + " (" + std::to_string(count) + " present)",
                               FG_HIP};
        hipDeviceProp_t prop;
        FG_CHECK(hipGetDeviceProperties(&prop, ordinal));
        FG_CHECK(hipSetDevice(ordinal));
        // the marketing name can be empty on some ROCm releases: name the part by its ISA then
        std::string name = prop.name;
        if (name.empty()) {
            const std::string arch = prop.gcnArchName;
            name = "AMD GPU " + arch.substr(0, arch.find(':'));
        }
        char info[160];
        std::snprintf(info, sizeof info, " (%s, %d CUs, %.0f GiB)", prop.gcnArchName, prop.multiProcessorCount,
                      double(prop.totalGlobalMem) / double(1ull << 30));
        std::unique_ptr<fg_device> dev(new fg_device{ordinal, name + info, uint64_t(prop.totalGlobalMem),
                                                     prop.multiProcessorCount});
        fg::spectra_scratch_init(dev.get());
        *out = dev.release();
        return int(FG_OK);
    });
}

void fg_close(fg_device *dev) {
    if (!dev) return;
    for (hipStream_t s : dev->idle_streams) (void)hipStreamDestroy(s);
    delete dev;
}

const char *fg_name(const fg_device *dev) { return dev->name.c_str(); }

uint64_t fg_max_cells(const fg_device *dev) {
    // fused engine: two copies of E and H (48 B) plus u16 coefficient indices (6 B) per node,
    // with headroom for the CPML store
    return dev->total_mem / 64;
}

int fg_session_new(fg_device *dev, const fg_problem *problem, uint32_t driven, const fg_pulse *pulse,
                   const fg_extras *extras, fg_session **out) {
    return fg::guarded([&] {
        if (!dev || !out || !pulse) throw fg::HipError{"fg_session_new: NULL argument", FG_INVALID};
        *out = reinterpret_cast<fg_session *>(fg::session_new(dev, problem, driven, pulse, extras));
        return int(FG_OK);
    });
}

int fg_session_step(fg_session *ss, uint64_t count) {
    return fg::guarded([&] {
        fg::Session &s = *reinterpret_cast<fg::Session *>(ss);
        fg::enqueue_steps(s, count);
        return int(FG_OK);
    });
}

int fg_session_energy(fg_session *ss, double *out) {
    return fg::guarded([&] {
        *out = fg::energy(*reinterpret_cast<fg::Session *>(ss));
        return int(FG_OK);
    });
}

int fg_session_read(fg_session *ss, fg_record *out) {
    return fg::guarded([&] {
        fg::read_record(*reinterpret_cast<fg::Session *>(ss), out);
        return int(FG_OK);
    });
}

void fg_session_free(fg_session *ss) { delete reinterpret_cast<fg::Session *>(ss); }

int fg_session_timing(fg_session *ss, int enable, double ms[3], uint64_t *steps) {
    return fg::guarded([&] {
        fg::Session &s = *reinterpret_cast<fg::Session *>(ss);
        fg::collect_timing(s);
        if (ms)
            for (int g = 0; g < 3; g++) ms[g] = s.kernel_ms[g];
        if (steps) *steps = s.kernel_count;
        s.timing = enable != 0;
        s.kernel_ms[0] = s.kernel_ms[1] = s.kernel_ms[2] = 0.0;
        s.kernel_count = 0;
        return int(FG_OK);
    });
}

int fg_session_steps_per_pass(const fg_session *ss, uint32_t *out) {
    return fg::guarded([&] {
        if (!ss || !out) throw fg::HipError{"fg_session_steps_per_pass: NULL argument", FG_INVALID};
        *out = reinterpret_cast<const fg::Session *>(ss)->fuse2 ? 2u : 1u;
        return int(FG_OK);
    });
}

int fg_session_sync(fg_session *ss) {
    return fg::guarded([&] {
        FG_CHECK(hipStreamSynchronize(reinterpret_cast<fg::Session *>(ss)->stream));
        return int(FG_OK);
    });
}

int fg_stream_copy(fg_device *dev, uint64_t bytes, uint32_t iterations, double *gbps) {
    return fg::guarded([&] {
        FG_CHECK(hipSetDevice(dev->ordinal));
        const uint64_t count = bytes / 16;
        fg::DevBuf<float4> a, b;
        a.alloc(count);
        b.alloc(count);
        hipStream_t st;
        FG_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        a.zero(st);
        const uint32_t blocks = uint32_t(std::min<uint64_t>((count + 255) / 256, uint64_t(dev->cus) * 64));
        hipEvent_t e0, e1;
        FG_CHECK(hipEventCreate(&e0));
        FG_CHECK(hipEventCreate(&e1));
        hipLaunchKernelGGL(fg::k_stream_copy, dim3(blocks), dim3(256), 0, st, a.p, b.p, count);  // warm-up
        FG_CHECK(hipEventRecord(e0, st));
        for (uint32_t q = 0; q < iterations; q++)
            hipLaunchKernelGGL(fg::k_stream_copy, dim3(blocks), dim3(256), 0, st, a.p, b.p, count);
        FG_CHECK(hipEventRecord(e1, st));
        FG_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        FG_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
        if (gbps) *gbps = 2.0 * double(count) * 16.0 * iterations / (double(ms) * 1e-3) / 1e9;
        return int(FG_OK);
    });
}

int fg_problem_check(const fg_problem *problem, uint32_t driven, const fg_extras *extras) {
    return fg::guarded([&] {
        fg::validate(problem, driven, extras);
        return int(FG_OK);
    });
}

int fg_device_synchronize(fg_device *dev) {
    return fg::guarded([&] {
        FG_CHECK(hipSetDevice(dev->ordinal));
        FG_CHECK(hipDeviceSynchronize());
        return int(FG_OK);
    });
}

int fg_run(fg_device *dev, const fg_problem *problem, uint32_t driven, const fg_pulse *pulse, const fg_extras *extras,
           fg_progress_fn progress, void *user, fg_record *out) {
    return fg::guarded([&] {
        if (!dev || !pulse || !out) throw fg::HipError{"fg_run: NULL argument", FG_INVALID};
        std::unique_ptr<fg::Session> s(fg::session_new(dev, problem, driven, pulse, extras));
        const uint64_t chunk = extras->check_every ? extras->check_every : 1000;  // run.rs:681
        const double t0 = 4.0 * std::sqrt(std::log(10.0)) / (3.14159265358979323846 * pulse->fc);
        const uint64_t pulse_end = uint64_t(2.0 * t0 / problem->dt);               // run.rs:685
        int status = FG_OK;
        while (s->steps < extras->max_steps) {
            const uint64_t k = std::min(chunk, extras->max_steps - s->steps);
            if (extras->profile && s->steps == 0 && k > 200) {
                // the reference times the first chunk's dispatches from its 200th step on
                // (run.rs:686-732) and prints their means (run.rs:152-176)
                fg::enqueue_steps(*s, 200);
                s->timing = true;
                fg::enqueue_steps(*s, k - 200);
                fg::report_timing(*s);
                s->timing = false;
            } else {
                fg::enqueue_steps(*s, k);
            }
            const double e = fg::energy(*s);
            if (!std::isfinite(e)) {  // run.rs:736-738
                status = fg::set_error(FG_UNSTABLE, "the FDTD run went unstable");
                break;
            }
            s->peak = std::max(s->peak, e);
            s->decay = e > 0.0 && s->peak > 0.0 ? 10.0 * std::log10(s->peak / e) : 0.0;
            if (progress && !progress(user, s->steps, s->decay)) {  // run.rs:742-744
                status = fg::set_error(FG_CANCELLED, "cancelled");
                break;
            }
            if (s->steps > std::max(pulse_end, extras->min_steps) && s->decay >= extras->decay_db) break;
        }
        if (status != FG_OK) return status;
        fg::read_record(*s, out);
        return int(FG_OK);
    });
}

}  // extern "C"
