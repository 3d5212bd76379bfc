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
#include <memory>
#include <string>
#include <vector>

#include "../../include/fieldgrid.h"
#include "common.h"

namespace fg {

thread_local std::string g_error;

int set_error(int status, const std::string &msg) {
    g_error = msg;
    return status;
}

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

// ---- device-side views --------------------------------------------------------------------

// One axis of the CPML/spacing tables (sim.rs:12-27), structure of arrays.
struct AxDev {
    const float *inv_d, *inv_dd, *raw_d, *raw_dd, *be, *ce, *bh, *ch;
    const int *slot_e, *slot_h;
};

struct Geo {
    uint32_t n0, n1, n2, pml;
    uint64_t nn;
    float k_mu;
    uint32_t wide;
    AxDev ax[3];
    uint64_t psi_off[18];  // e*9 + c*3 + a  (run.rs:214-225)
    const uint32_t *coef_index;
    const float2 *coef_sets;
    uint32_t n_sets;
};

template <int A>
__device__ __forceinline__ uint32_t coord(uint32_t i, uint32_t j, uint32_t k) {
    return A == 0 ? i : (A == 1 ? j : k);
}

template <int A>
__device__ __forceinline__ uint32_t dim(const Geo &g) {
    return A == 0 ? g.n0 : (A == 1 ? g.n1 : g.n2);
}

template <int A>
__device__ __forceinline__ uint64_t stride(const Geo &g) {
    return A == 0 ? uint64_t(g.n1) * g.n2 : (A == 1 ? uint64_t(g.n2) : uint64_t(1));
}

// psi_idx (sim.rs:659-665) plus the slab's base offset
template <int A>
__device__ __forceinline__ uint64_t psi_at(const Geo &g, int e, int c, uint32_t i, uint32_t j, uint32_t k, uint32_t s) {
    const uint64_t b = g.psi_off[e * 9 + c * 3 + A];
    const uint32_t w = 2u * g.pml;
    if (A == 0) return b + (uint64_t(s) * g.n1 + j) * g.n2 + k;
    if (A == 1) return b + (uint64_t(i) * w + s) * g.n2 + k;
    return b + (uint64_t(i) * g.n1 + j) * w + s;
}

__device__ __forceinline__ float2 coef_at(const Geo &g, uint64_t f) {
    const uint32_t s = g.wide ? g.coef_index[f] : uint32_t(reinterpret_cast<const uint16_t *>(g.coef_index)[f]);
    return g.coef_sets[s];
}

// New H component C at node (i,j,k): H -= k_mu*(dEv/du - dEu/dv + psi terms) (cpu.rs:48-107).
// Reads E/H/psi from the "old" buffers; writes psi only when psi_new is non-null (the
// owner of the node), so tiles may recompute a neighbour's H for their halo.
template <int C>
__device__ __forceinline__ float h_next(const Geo &g, const float *__restrict__ E, const float *__restrict__ H,
                                        const float *psi_old, float *psi_new, uint32_t i, uint32_t j, uint32_t k) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const float hold = H[C * g.nn + id];
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    if (pu + 1 >= dim<U>(g) || pv + 1 >= dim<V>(g)) return hold;
    const float ev = E[V * g.nn + id + stride<U>(g)] - E[V * g.nn + id];
    const float eu = E[U * g.nn + id + stride<V>(g)] - E[U * g.nn + id];
    float t1 = ev * g.ax[U].inv_d[pu];
    float t2 = eu * g.ax[V].inv_d[pv];
    if (g.pml) {
        const int su = g.ax[U].slot_h[pu];
        if (su >= 0) {
            const uint64_t q = psi_at<U>(g, 1, C, i, j, k, uint32_t(su));
            const float r = g.ax[U].bh[pu] * psi_old[q] + g.ax[U].ch[pu] * ev * g.ax[U].raw_d[pu];
            if (psi_new) psi_new[q] = r;
            t1 += r;
        }
        const int sv = g.ax[V].slot_h[pv];
        if (sv >= 0) {
            const uint64_t q = psi_at<V>(g, 1, C, i, j, k, uint32_t(sv));
            const float r = g.ax[V].bh[pv] * psi_old[q] + g.ax[V].ch[pv] * eu * g.ax[V].raw_d[pv];
            if (psi_new) psi_new[q] = r;
            t2 += r;
        }
    }
    return hold - g.k_mu * (t1 - t2);
}

// New E component C from the H differences hv = H_v(p)-H_v(p-u), hu = H_u(p)-H_u(p-v)
// (cpu.rs:109-167). Edges with cb == 0 (PEC, outer faces) keep their value.
template <int C>
__device__ __forceinline__ float e_next(const Geo &g, float eold, float hv, float hu, uint32_t i, uint32_t j, uint32_t k,
                                        uint64_t id, const float *psi_old, float *psi_new) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const float2 cf = coef_at(g, C * g.nn + id);
    if (cf.y == 0.f) return eold;
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    float t1 = hv * g.ax[U].inv_dd[pu];
    float t2 = hu * g.ax[V].inv_dd[pv];
    if (g.pml) {
        const int su = g.ax[U].slot_e[pu];
        if (su >= 0) {
            const uint64_t q = psi_at<U>(g, 0, C, i, j, k, uint32_t(su));
            const float r = g.ax[U].be[pu] * psi_old[q] + g.ax[U].ce[pu] * hv * g.ax[U].raw_dd[pu];
            psi_new[q] = r;
            t1 += r;
        }
        const int sv = g.ax[V].slot_e[pv];
        if (sv >= 0) {
            const uint64_t q = psi_at<V>(g, 0, C, i, j, k, uint32_t(sv));
            const float r = g.ax[V].be[pv] * psi_old[q] + g.ax[V].ce[pv] * hu * g.ax[V].raw_dd[pv];
            psi_new[q] = r;
            t2 += r;
        }
    }
    return cf.x * eold + cf.y * (t1 - t2);
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

// ---- fused engine: one H->E pass per step, ping-pong buffers ------------------------------
//
// The reference fuses H and E per step with 64x4 tiles and recomputed halos
// (fdtd.wgsl:186-330). Here a block is a column of ROWS = TJ+2 waves marching along i (the
// slowest axis) over `seg` planes:
//   * wave w holds row j = j0-1+w; rows 1..TJ are owned, row 0 recomputes new Hx/Hz at
//     j0-1 (the E update needs them), row TJ+1 only supplies E at j0+TJ (the H update needs it);
//   * lane L holds k = k0-1+L with k0 = 62*blockIdx.x; lanes 1..62 are owned, lanes 0 and 63
//     are k-halos whose values reach their neighbours through cross-lane shuffles;
//   * E(i) and new Hy/Hz(i-1) ride in registers from one plane to the next, so each plane
//     loads E(i+1) and H(i) once, and rows trade E and new H through LDS behind one barrier.
// Halo work reads only the old buffers and writes nothing, so no block reads what another
// block writes in the same step. Every node's own E, H, coefficient index and CPML psi are
// read once and written once per step.
constexpr int OWN_K = 62;
constexpr int SETS_LDS = 2048;  // coefficient sets staged in LDS per block (16 KB)
constexpr int SEG_MAX = 64;     // longest i segment one march block takes

// Block coordinates of one march block (k tile, j tile, i segment). Blocks that touch a CPML
// slab and blocks that do not are launched as two lists, so the interior kernel carries no
// psi code at all.
struct MarchBlock {
    uint32_t bx, by, bz;
};

// Raw buffer access: a plane-uniform base pointer in a buffer resource (scalar registers)
// and a 32-bit per-lane byte offset, so no lane carries a 64-bit address. `bytes` bounds the
// resource: an access past it reads 0 / is dropped by the hardware instead of faulting.
// 0x00020000 is the gfx9-family (incl. gfx950) resource word 3 for raw dword access.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, int(bytes), 0x00020000);
}
__device__ __forceinline__ float ld(const float *p, uint32_t byte_off, uint32_t bytes) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc(p, bytes), int(byte_off), 0, 0));
}
__device__ __forceinline__ uint32_t ld(const uint32_t *p, uint32_t byte_off, uint32_t bytes) {
    return __builtin_amdgcn_raw_buffer_load_b32(rsrc(p, bytes), int(byte_off), 0, 0);
}
__device__ __forceinline__ uint16_t ld(const uint16_t *p, uint32_t byte_off, uint32_t bytes) {
    return __builtin_amdgcn_raw_buffer_load_b16(rsrc(p, bytes), int(byte_off), 0, 0);
}
__device__ __forceinline__ void st(float *p, uint32_t byte_off, uint32_t bytes, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rsrc(p, bytes), int(byte_off), 0, 0);
}

constexpr uint32_t OOB = 0x7FFFFFF0u;  // a byte offset past every resource: load 0, store dropped

// psi = b*old + c*diff*raw (cpu.rs:78-99, 137-157), written to the new store at `off`.
__device__ __forceinline__ float cpml_next(float *pn, uint32_t bytes, uint32_t off, float b, float c, float raw,
                                           float old, float diff) {
    const float r = b * old + c * diff * raw;
    st(pn, off, bytes, r);
    return r;
}

// One CPML psi recursion, psi = b*psi + c*diff*raw (cpu.rs:78-99, 137-157), from the old store
// to the new one; po/pn are plane-uniform bases of the psi array, `off` the lane's byte offset.
__device__ __forceinline__ float cpml(const float *po, float *pn, uint32_t bytes, uint32_t off, float b, float c,
                                      float raw, float diff, bool store) {
    const float r = b * ld(po, off, bytes) + c * diff * raw;
    if (store) st(pn, off, bytes, r);
    return r;
}

// PML: 0 = no CPML slab in reach, 1 = only z (k) slabs, 2 = any slab (x, y and z psi code)
template <int TJ, int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ + 2)) void k_march(Geo g, const MarchBlock *__restrict__ blocks,
                                                         const float *__restrict__ Eo, float *__restrict__ En,
                                                         const float *__restrict__ Ho, float *__restrict__ Hn,
                                                         const float *__restrict__ psio, float *__restrict__ psin,
                                                         uint32_t seg) {
    constexpr int ROWS = TJ + 2;
    // row exchanges; rows are stored at index w+1 so that rows -1 and ROWS are zero pads
    __shared__ float xe[2][ROWS + 2][2][64];  // E(i+1) Ex, Ez of each row, for the row above
    __shared__ float xh[2][ROWS + 2][2][64];  // new H(i) Hx, Hz of each row, for the row below
    __shared__ float2 sets[LDS_SETS ? SETS_LDS : 1];
    constexpr bool PZ = PML >= 1, PXY = PML == 2;
    __shared__ float xa[SEG_MAX + 1][PXY ? 10 : 2];  // x-axis values of the segment's planes
    const MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;  // padding of the XCD-ordered list (uniform over the block)
    // blockDim.x is one wave, so the row index is wave-uniform: keep it (and every per-row
    // value) in scalar registers
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k = int(mb.bx) * OWN_K + lane - 1;
    const int j = int(mb.by) * TJ + w - 1;
    const uint32_t i0 = mb.bz * seg, i1 = min(i0 + seg, g.n0);
    const bool kin = k >= 0 && k < int(g.n2), jin = j >= 0 && j < int(g.n1);
    const bool act = kin && jin;
    const bool own = act && lane >= 1 && lane <= OWN_K && w >= 1 && w <= TJ;
    const bool bottom = w == TJ + 1;
    const uint32_t uk = kin ? uint32_t(k) : 0u, uj = jin ? uint32_t(j) : 0u;
    const uint64_t nn = g.nn, plane = uint64_t(g.n1) * g.n2;
    const uint32_t off = uj * g.n2 + uk;
    const bool kl = uk + 1 < g.n2, jl = uj + 1 < g.n1;  // k+1 / j+1 inside the grid
    const float idz = g.ax[2].inv_d[uk], iddz = g.ax[2].inv_dd[uk];
    const float idy = g.ax[1].inv_d[uj], iddy = g.ax[1].inv_dd[uj];
    // byte offsets of this lane; a lane that must not touch a plane gets OOB instead, so
    // every access below is unconditional (no exec-masked branches around memory ops,
    // which would make the compiler drain all outstanding loads)
    const uint32_t ob4 = off * 4u;
    const uint32_t ldE = act ? ob4 : OOB, ldH = act && !bottom ? ob4 : OOB, stO = own ? ob4 : OOB;
    const uint32_t ldC = own ? off * (WIDE ? 4u : 2u) : OOB;
    const uint32_t pl4 = uint32_t(plane) * 4u, plc = uint32_t(plane) * (WIDE ? 4u : 2u);
    // CPML slots and recursion coefficients of this lane (k) and row (j), and the lane's
    // loop-invariant byte offsets into the z- and y-slab psi planes
    float bhz = 0, chz = 0, rdz = 0, bez = 0, cez = 0, rddz = 0, bhy = 0, chy = 0, rdy = 0, bey = 0, cey = 0, rddy = 0;
    uint32_t ozh = OOB, oze = OOB, oyh = OOB, oye = OOB;
    const uint32_t p2 = 2u * g.pml;
    const uint32_t zbytes = g.n1 * p2 * 4u, ybytes = p2 * g.n2 * 4u;
    if constexpr (PML) {
        const int shz = g.ax[2].slot_h[uk], sez = g.ax[2].slot_e[uk];
        const int shy = g.ax[1].slot_h[uj], sey = g.ax[1].slot_e[uj];
        bhz = g.ax[2].bh[uk], chz = g.ax[2].ch[uk], rdz = g.ax[2].raw_d[uk];
        bez = g.ax[2].be[uk], cez = g.ax[2].ce[uk], rddz = g.ax[2].raw_dd[uk];
        bhy = g.ax[1].bh[uj], chy = g.ax[1].ch[uj], rdy = g.ax[1].raw_d[uj];
        bey = g.ax[1].be[uj], cey = g.ax[1].ce[uj], rddy = g.ax[1].raw_dd[uj];
        if (act) {
            ozh = shz >= 0 ? (uj * p2 + uint32_t(shz)) * 4u : OOB, oze = sez >= 0 ? (uj * p2 + uint32_t(sez)) * 4u : OOB;
            oyh = shy >= 0 ? (uint32_t(shy) * g.n2 + uk) * 4u : OOB, oye = sey >= 0 ? (uint32_t(sey) * g.n2 + uk) * 4u : OOB;
        }
    }
    const uint32_t *cix = g.coef_index;
    const uint16_t *cix16 = reinterpret_cast<const uint16_t *>(g.coef_index);
    auto cidx = [&](uint64_t ub, uint32_t bytes) {
        if constexpr (WIDE) return ld(cix + ub, ldC, bytes);
        else return uint32_t(ld(cix16 + ub, ldC, bytes));
    };
    auto coef = [&](uint32_t c) {
        if constexpr (LDS_SETS) return sets[c];
        else return g.coef_sets[c];
    };
    if constexpr (LDS_SETS)
        for (uint32_t t = w * 64 + lane; t < g.n_sets; t += 64 * ROWS) sets[t] = g.coef_sets[t];
    // per-plane (x) axis values staged in LDS: plain loads of them inside the loop would queue
    // behind the prefetches and force the wave to wait for all of them
    const uint32_t istart = i0 > 0 ? i0 - 1 : 0;
    for (uint32_t t = w * 64 + lane; t < i1 - istart; t += 64 * ROWS) {
        const uint32_t q = istart + t;
        xa[t][0] = g.ax[0].inv_d[q], xa[t][1] = g.ax[0].inv_dd[q];
        if constexpr (PXY) {
            xa[t][2] = g.ax[0].bh[q], xa[t][3] = g.ax[0].ch[q], xa[t][4] = g.ax[0].raw_d[q];
            xa[t][5] = g.ax[0].be[q], xa[t][6] = g.ax[0].ce[q], xa[t][7] = g.ax[0].raw_dd[q];
            xa[t][8] = __int_as_float(g.ax[0].slot_h[q]), xa[t][9] = __int_as_float(g.ax[0].slot_e[q]);
        }
    }
    // zero pads (rows -1 and ROWS of both buffers)
    if (w < 2) {
        for (int q = 0; q < 2; q++) {
            xe[q][w * (ROWS + 1)][0][lane] = 0.f, xe[q][w * (ROWS + 1)][1][lane] = 0.f;
            xh[q][w * (ROWS + 1)][0][lane] = 0.f, xh[q][w * (ROWS + 1)][1][lane] = 0.f;
        }
    }
    // software pipeline: E(i+1), H(i) and the coefficient indices of plane i are loaded one
    // plane ahead into the other of two register sets; the loop is unrolled by two so the sets
    // swap roles without copies (a copy would have to wait for the loads in flight)
    struct Pipe {
        float ex1, ey1, ez1, hx, hy, hz;
        uint32_t c0, c1, c2;
    };
    Pipe A, B;
    {
        const float *p = Eo + uint64_t(istart) * plane;
        B.ex1 = ld(p, ldE, pl4), B.ey1 = ld(p + nn, ldE, pl4), B.ez1 = ld(p + 2 * nn, ldE, pl4);
    }
    // E(istart) Ex/Ez of the row below, through the buffer the first plane does not write
    const int pb0 = (istart + 1) & 1;
    xe[pb0][w + 1][0][lane] = B.ex1;
    xe[pb0][w + 1][1][lane] = B.ez1;
    __syncthreads();
    float exn = xe[pb0][w + 2][0][lane], ezn = xe[pb0][w + 2][1][lane];
    {
        const uint64_t pb = uint64_t(istart) * plane;
        const uint32_t eb = istart + 1 < g.n0 ? pl4 : 0u;  // a zero-size resource reads 0
        const float *p = Eo + pb + (istart + 1 < g.n0 ? plane : 0);
        A.ex1 = ld(p, ldE, eb), A.ey1 = ld(p + nn, ldE, eb), A.ez1 = ld(p + 2 * nn, ldE, eb);
        const float *q = Ho + pb;
        A.hx = ld(q, ldH, pl4), A.hy = ld(q + nn, ldH, pl4), A.hz = ld(q + 2 * nn, ldH, pl4);
        A.c0 = cidx(pb, plc), A.c1 = cidx(nn + pb, plc), A.c2 = cidx(2 * nn + pb, plc);
    }
    float hy_p = 0.f, hz_p = 0.f;
#define PSI(e, c, a) g.psi_off[(e) * 9 + (c) * 3 + (a)]
    auto plane_step = [&](uint32_t i, Pipe &cur, Pipe &nxt) {
        const int b = i & 1;
        const bool wplane = i >= i0;
        const bool il = i + 1 < g.n0;
        const uint64_t pb = uint64_t(i) * plane;  // uniform
        // E(i) arrived two planes ago in the set this plane's prefetch now refills
        const float ex = nxt.ex1, ey = nxt.ey1, ez = nxt.ez1;
        // prefetch of plane i+1 (E of plane i+2); absent planes read through zero-size resources
        const bool more = i + 1 < i1, more2 = more && i + 2 < g.n0;
        const float *pe = Eo + pb + (more2 ? 2 * plane : 0);
        nxt.ex1 = ld(pe, ldE, more2 ? pl4 : 0u), nxt.ey1 = ld(pe + nn, ldE, more2 ? pl4 : 0u),
        nxt.ez1 = ld(pe + 2 * nn, ldE, more2 ? pl4 : 0u);
        const float *ph = Ho + pb + (more ? plane : 0);
        nxt.hx = ld(ph, ldH, more ? pl4 : 0u), nxt.hy = ld(ph + nn, ldH, more ? pl4 : 0u),
        nxt.hz = ld(ph + 2 * nn, ldH, more ? pl4 : 0u);
        const uint64_t cb_ = pb + (more ? plane : 0);
        nxt.c0 = cidx(cb_, more ? plc : 0u), nxt.c1 = cidx(nn + cb_, more ? plc : 0u),
        nxt.c2 = cidx(2 * nn + cb_, more ? plc : 0u);
        const float *xv = xa[i - istart];
        const float idx_ = xv[0], iddx = xv[1];
        // CPML: every psi value this lane needs on plane i is loaded here, up front, so the
        // loads overlap. A slot that does not exist maps to an offset past its resource (the
        // load returns 0 and the store is dropped) and has b = c = 0, so its term adds exactly 0.
        float bhx = 0, chx = 0, rdx = 0, bex = 0, cex = 0, rddx = 0;
        float hxy = 0, hxz = 0, hyz = 0, hyx = 0, hzx = 0, hzy = 0;  // old H psi (component, axis)
        float exy = 0, exz = 0, eyz = 0, eyx = 0, ezx = 0, ezy = 0;  // old E psi
        uint64_t zrow = 0, yrow = 0, xhb = 0, xeb = 0;
        uint32_t oxh = OOB, oxe = OOB;
        if constexpr (PZ) {
            zrow = uint64_t(i) * g.n1 * p2;  // a = 2: (i*n1 + j)*2p + s
            const uint32_t zh = ldH == OOB ? OOB : ozh, ze = stO == OOB ? OOB : oze;
            hxz = ld(psio + PSI(1, 0, 2) + zrow, zh, zbytes), hyz = ld(psio + PSI(1, 1, 2) + zrow, zh, zbytes);
            exz = ld(psio + PSI(0, 0, 2) + zrow, ze, zbytes), eyz = ld(psio + PSI(0, 1, 2) + zrow, ze, zbytes);
        }
        if constexpr (PXY) {
            const int shx = __float_as_int(xv[8]), sex = __float_as_int(xv[9]);
            bhx = xv[2], chx = xv[3], rdx = xv[4];
            bex = xv[5], cex = xv[6], rddx = xv[7];
            yrow = uint64_t(i) * p2 * g.n2;  // a = 1: (i*2p + s)*n2 + k
            xhb = uint64_t(max(shx, 0)) * plane, xeb = uint64_t(max(sex, 0)) * plane;  // a = 0
            oxh = shx >= 0 ? ldE : OOB, oxe = sex >= 0 ? stO : OOB;
            const uint32_t yh = ldH == OOB ? OOB : oyh, xh_ = ldH == OOB ? OOB : oxh;
            hxy = ld(psio + PSI(1, 0, 1) + yrow, yh, ybytes);
            hyx = ld(psio + PSI(1, 1, 0) + xhb, xh_, pl4);
            hzx = ld(psio + PSI(1, 2, 0) + xhb, xh_, pl4), hzy = ld(psio + PSI(1, 2, 1) + yrow, yh, ybytes);
            const uint32_t ye = stO == OOB ? OOB : oye;
            exy = ld(psio + PSI(0, 0, 1) + yrow, ye, ybytes);
            eyx = ld(psio + PSI(0, 1, 0) + xeb, oxe, pl4);
            ezx = ld(psio + PSI(0, 2, 0) + xeb, oxe, pl4), ezy = ld(psio + PSI(0, 2, 1) + yrow, ye, ybytes);
        }
        // stores of this plane: only owners, only on planes of their own segment
        const uint32_t sto = wplane ? stO : OOB;
        // new psi from its old value; stored (where the slot exists) only by the node's owner
#define CP(base, bytes, o, B, CC, R, old, diff) cpml_next(psin + (base), bytes, sto == OOB ? OOB : (o), B, CC, R, old, diff)
        // k+1 neighbours of E(i) from the next lane
        const float ex_k = __shfl_down(ex, 1, 64), ey_k = __shfl_down(ey, 1, 64);
        float hx, hy, hz;
        {
            // Hx: u = y, v = z
            const float ev = ezn - ez, eu = ey_k - ey;
            float t1 = ev * idy, t2 = eu * idz;
            const bool v = jl && kl;  // psi advances only where the H component does
            if constexpr (PXY) t1 += CP(PSI(1, 0, 1) + yrow, ybytes, v ? oyh : OOB, bhy, chy, rdy, hxy, ev);
            if constexpr (PZ) t2 += CP(PSI(1, 0, 2) + zrow, zbytes, v ? ozh : OOB, bhz, chz, rdz, hxz, eu);
            hx = jl && kl ? cur.hx - g.k_mu * (t1 - t2) : cur.hx;
        }
        {
            // Hy: u = z, v = x
            const float ev = ex_k - ex, eu = cur.ez1 - ez;
            float t1 = ev * idz, t2 = eu * idx_;
            const bool v = kl && il;
            if constexpr (PZ) t1 += CP(PSI(1, 1, 2) + zrow, zbytes, v ? ozh : OOB, bhz, chz, rdz, hyz, ev);
            if constexpr (PXY) t2 += CP(PSI(1, 1, 0) + xhb, pl4, v ? oxh : OOB, bhx, chx, rdx, hyx, eu);
            hy = kl && il ? cur.hy - g.k_mu * (t1 - t2) : cur.hy;
        }
        {
            // Hz: u = x, v = y
            const float ev = cur.ey1 - ey, eu = exn - ex;
            float t1 = ev * idx_, t2 = eu * idy;
            const bool v = il && jl;
            if constexpr (PXY) t1 += CP(PSI(1, 2, 0) + xhb, pl4, v ? oxh : OOB, bhx, chx, rdx, hzx, ev);
            if constexpr (PXY) t2 += CP(PSI(1, 2, 1) + yrow, ybytes, v ? oyh : OOB, bhy, chy, rdy, hzy, eu);
            hz = il && jl ? cur.hz - g.k_mu * (t1 - t2) : cur.hz;
        }
        {
            float *q = Hn + pb;
            st(q, sto, pl4, hx), st(q + nn, sto, pl4, hy), st(q + 2 * nn, sto, pl4, hz);
        }
        xh[b][w + 1][0][lane] = hx;
        xh[b][w + 1][1][lane] = hz;
        xe[b][w + 1][0][lane] = cur.ex1;
        xe[b][w + 1][1][lane] = cur.ez1;
        __syncthreads();
        const float hx_j = xh[b][w][0][lane], hz_j = xh[b][w][1][lane];
        const float exn1 = xe[b][w + 2][0][lane], ezn1 = xe[b][w + 2][1][lane];
        // k-1 neighbours of new H(i) from the previous lane
        const float hx_k = __shfl_up(hx, 1, 64), hy_k = __shfl_up(hy, 1, 64);
        {
            float *q = En + pb;
            // Ex: u = y, v = z
            {
                const float2 cf = coef(cur.c0);
                const float hv = hz - hz_j, hu = hy - hy_k;
                float t1 = hv * iddy, t2 = hu * iddz;
                const bool on = cf.y != 0.f;
                if constexpr (PXY) t1 += CP(PSI(0, 0, 1) + yrow, ybytes, on ? oye : OOB, bey, cey, rddy, exy, hv);
                if constexpr (PZ) t2 += CP(PSI(0, 0, 2) + zrow, zbytes, on ? oze : OOB, bez, cez, rddz, exz, hu);
                st(q, sto, pl4, cf.y != 0.f ? cf.x * ex + cf.y * (t1 - t2) : ex);
            }
            // Ey: u = z, v = x
            {
                const float2 cf = coef(cur.c1);
                const float hv = hx - hx_k, hu = hz - hz_p;
                float t1 = hv * iddz, t2 = hu * iddx;
                const bool on = cf.y != 0.f;
                if constexpr (PZ) t1 += CP(PSI(0, 1, 2) + zrow, zbytes, on ? oze : OOB, bez, cez, rddz, eyz, hv);
                if constexpr (PXY) t2 += CP(PSI(0, 1, 0) + xeb, pl4, on ? oxe : OOB, bex, cex, rddx, eyx, hu);
                st(q + nn, sto, pl4, cf.y != 0.f ? cf.x * ey + cf.y * (t1 - t2) : ey);
            }
            // Ez: u = x, v = y
            {
                const float2 cf = coef(cur.c2);
                const float hv = hy - hy_p, hu = hx - hx_j;
                float t1 = hv * iddx, t2 = hu * iddy;
                const bool on = cf.y != 0.f;
                if constexpr (PXY) t1 += CP(PSI(0, 2, 0) + xeb, pl4, on ? oxe : OOB, bex, cex, rddx, ezx, hv);
                if constexpr (PXY) t2 += CP(PSI(0, 2, 1) + yrow, ybytes, on ? oye : OOB, bey, cey, rddy, ezy, hu);
                st(q + 2 * nn, sto, pl4, cf.y != 0.f ? cf.x * ez + cf.y * (t1 - t2) : ez);
            }
        }
#undef CP
        exn = exn1, ezn = ezn1;
        hy_p = hy, hz_p = hz;
    };
    // both unrolled copies are unconditional (an odd plane count is peeled first) so the
    // compiler keeps them apart instead of merging them back into one body with copies
    if ((i1 - istart) & 1u) {
        plane_step(istart, A, B);
        for (uint32_t i = istart + 1; i < i1; i += 2) {
            plane_step(i, B, A);
            plane_step(i + 1, A, B);
        }
    } else {
        for (uint32_t i = istart; i < i1; i += 2) {
            plane_step(i, A, B);
            plane_step(i + 1, B, A);
        }
    }
#undef PSI
}

// ---- sparse per-step fixes: Debye ADE, port sources, inductors (cpu.rs:239-303) ----------

struct FixArgs {
    uint64_t n_debye;
    const fg_debye_edge *debye;
    const float *dtab;  // per set: np x (ka, kb)
    uint32_t np;
    float *dstate;      // np x n_debye, pole-major (run.rs:492-493)
    const float *dold;  // split engine: E saved before the E sweep; fused: NULL (read Eo)
    uint32_t n_src;
    const fg_source *src;
    uint32_t n_ind;
    const fg_inductor *ind;
    float *ind_i;
    float dt, t0, tau, w0;
    const uint32_t *step;
};

__device__ __forceinline__ float pulse_at(const FixArgs &a, float t) {
    const float x = (t - a.t0) / a.tau;
    return expf(-x * x) * cosf(a.w0 * (t - a.t0));
}

__global__ __launch_bounds__(256) void k_debye_save(uint64_t n, const fg_debye_edge *__restrict__ d,
                                                     const float *__restrict__ E, float *__restrict__ old) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) old[t] = E[d[t].f];
}

__global__ __launch_bounds__(256) void k_fix(Geo g, FixArgs a, float *__restrict__ E, const float *__restrict__ Eo,
                                             uint32_t s_off) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < a.n_debye) {
        const fg_debye_edge d = a.debye[t];
        const float old = a.dold ? a.dold[t] : Eo[d.f];
        const float cb = coef_at(g, d.f).y;
        const float *tab = a.dtab + size_t(d.set) * a.np * 2;
        float s0 = 0.f;
        for (uint32_t q = 0; q < a.np; q++) s0 += 0.5f * (1.0f + tab[2 * q]) * a.dstate[q * a.n_debye + t];
        const float next = E[d.f] - cb * s0;
        E[d.f] = next;
        const float de = next - old;
        for (uint32_t q = 0; q < a.np; q++) {
            float &st = a.dstate[q * a.n_debye + t];
            st = tab[2 * q] * st + d.dd * tab[2 * q + 1] * de;
        }
    }
    if (t < a.n_src) {
        const uint32_t n = a.step[0] + s_off;
        const float tt = (float(n) + 0.5f) * a.dt;
        const fg_source s = a.src[t];
        E[s.comp * g.nn + s.id] -= s.src * pulse_at(a, tt);
    }
    if (t < a.n_ind) {
        const fg_inductor q = a.ind[t];
        const uint64_t f = q.comp * g.nn + q.id;
        const float x = E[f] - q.k_e * a.ind_i[t];
        E[f] = x;
        a.ind_i[t] += q.k_i * x;
    }
}

// ---- per-step observers: port V/I, point probes, plane DFT, NTFF (cpu.rs:309-415) --------

struct ObsArgs {
    uint32_t n_ports;
    const fg_port *ports;
    const fg_port_edge *edges;
    const fg_loop_term *loops;
    float *series;
    uint64_t cap;
    uint32_t n_probes;
    const uint64_t *probe_f;  // flat index into E (field < 3) or H (+ 3*nn marker)
    float *probes;
    uint32_t plane_n, plane_k;
    float *plane;
    uint64_t n_patches;
    const uint32_t *patch;
    float *ntff;
    uint32_t nf;
    float omega[FG_MAX_FREQS];
    float dt;
    uint32_t blk_ports, blk_probes, blk_plane, blk_ntff;
    const uint32_t *step;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_observe(Geo g, ObsArgs a, const float *__restrict__ E,
                                                 const float *__restrict__ H, uint32_t s_off) {
    const uint32_t n = a.step[0] + s_off;
    const uint64_t nn = g.nn;
    uint32_t blk = blockIdx.x;
    if (blk < a.blk_ports) {
        // one wave per port; f32 products summed in f64 like cpu.rs:317-328
        const uint32_t p = blk * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
        if (p >= a.n_ports) return;
        const fg_port d = a.ports[p];
        double v = 0.0, cur = 0.0;
        for (uint32_t q = lane; q < d.n_edges; q += 64) {
            const fg_port_edge pe = a.edges[d.first_edge + q];
            v -= double(E[d.comp * nn + pe.id] * pe.w);
        }
        for (uint32_t q = lane; q < d.n_loops; q += 64) {
            const fg_loop_term lt = a.loops[d.first_loop + q];
            cur += double(H[lt.comp * nn + lt.id] * lt.w);
        }
        v = wave_sum(v);
        cur = wave_sum(cur);
        if (lane == 0 && n < a.cap) {
            a.series[(uint64_t(n) * a.n_ports + p) * 2] = float(v);
            a.series[(uint64_t(n) * a.n_ports + p) * 2 + 1] = float(cur);
        }
        return;
    }
    blk -= a.blk_ports;
    if (blk < a.blk_probes) {
        const uint32_t q = blk * 256 + threadIdx.x;
        if (q < a.n_probes && n < a.cap) {
            const uint64_t f = a.probe_f[q];
            a.probes[uint64_t(n) * a.n_probes + q] = f < 3 * nn ? E[f] : H[f - 3 * nn];
        }
        return;
    }
    blk -= a.blk_probes;
    const float dt = a.dt;
    const float te = (float(n) + 1.0f) * dt, th = (float(n) + 0.5f) * dt;
    if (blk < a.blk_plane) {
        const uint32_t q = blk * 256 + threadIdx.x;
        if (q >= a.plane_n) return;
        const uint32_t i = q / g.n1, j = q % g.n1;
        const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + a.plane_k;
        const float f[5] = {E[id], E[nn + id], E[2 * nn + id], H[id], H[nn + id]};
        float *out = a.plane + uint64_t(q) * a.nf * 10;
        for (uint32_t kf = 0; kf < a.nf; kf++) {
            const float w = a.omega[kf];
            const float ce0 = cosf(w * te) * dt, ce1 = -sinf(w * te) * dt;
            const float ch0 = cosf(w * th) * dt, ch1 = -sinf(w * th) * dt;
            for (int c = 0; c < 5; c++) {
                const float p0 = c >= 3 ? ch0 : ce0, p1 = c >= 3 ? ch1 : ce1;
                out[kf * 10 + 2 * c] += f[c] * p0;
                out[kf * 10 + 2 * c + 1] += f[c] * p1;
            }
        }
        return;
    }
    blk -= a.blk_plane;
    if (blk < a.blk_ntff) {
        const uint64_t gi = uint64_t(blk) * 256 + threadIdx.x;
        if (gi >= a.n_patches) return;
        const uint32_t *ix = a.patch + gi * 13;
        const uint32_t meta = ix[0];
        const uint32_t ax = meta & 3u;
        const float side = (meta & 4u) ? 1.f : -1.f;
        const uint32_t u = (ax + 1) % 3, v = (ax + 2) % 3;
        float ef[3] = {0.f, 0.f, 0.f}, hf[3] = {0.f, 0.f, 0.f};
        ef[u] = 0.5f * (E[u * nn + ix[1]] + E[u * nn + ix[2]]);
        ef[v] = 0.5f * (E[v * nn + ix[3]] + E[v * nn + ix[4]]);
        hf[u] = 0.25f * (H[u * nn + ix[5]] + H[u * nn + ix[6]] + H[u * nn + ix[7]] + H[u * nn + ix[8]]);
        hf[v] = 0.25f * (H[v * nn + ix[9]] + H[v * nn + ix[10]] + H[v * nn + ix[11]] + H[v * nn + ix[12]]);
        float nrm[3] = {0.f, 0.f, 0.f};
        nrm[ax] = side;
        const float jv[3] = {nrm[1] * hf[2] - nrm[2] * hf[1], nrm[2] * hf[0] - nrm[0] * hf[2],
                             nrm[0] * hf[1] - nrm[1] * hf[0]};
        const float mv[3] = {ef[1] * nrm[2] - ef[2] * nrm[1], ef[2] * nrm[0] - ef[0] * nrm[2],
                             ef[0] * nrm[1] - ef[1] * nrm[0]};
        float *out = a.ntff + gi * a.nf * 12;
        for (uint32_t kf = 0; kf < a.nf; kf++) {
            const float wf = a.omega[kf];
            const float pe0 = cosf(-wf * te) * dt, pe1 = sinf(-wf * te) * dt;
            const float ph0 = cosf(-wf * th) * dt, ph1 = sinf(-wf * th) * dt;
            float *o = out + kf * 12;
            for (int c = 0; c < 3; c++) {
                o[2 * c] += ph0 * jv[c];
                o[2 * c + 1] += ph1 * jv[c];
                o[6 + 2 * c] += pe0 * mv[c];
                o[6 + 2 * c + 1] += pe1 * mv[c];
            }
        }
    }
}

// ---- energy: sum E^2 + (376.73 H)^2 (cpu.rs:417-421, fdtd.wgsl:547-568) -----------------

constexpr int ENERGY_BLOCKS = 1024;

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

// ---- host side ------------------------------------------------------------------------------

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count) FG_CHECK(hipMalloc(&p, count * sizeof(T)));
    }
    void zero(hipStream_t s) {
        if (n) FG_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    void upload(const T *src, size_t count, hipStream_t s) {
        alloc(count);
        if (count) FG_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

}  // namespace fg

struct fg_device {
    int ordinal;
    std::string name;
    uint64_t total_mem;
    int cus;
};

namespace fg {

struct Session {
    fg_device *dev;
    hipStream_t stream = nullptr;
    Geo geo{};
    FixArgs fix{};
    ObsArgs obs{};
    uint64_t nn = 0;
    uint32_t n[3];
    bool fused = true;
    uint32_t tj = 14, seg = 32, xcds = 8;
    uint64_t max_steps = 0, steps = 0;
    uint32_t n_ports = 0, n_probes = 0, nf = 0;
    uint64_t plane_len = 0, ntff_len = 0;
    double peak = 0.0, decay = 0.0;
    int cur = 0;  // which ping-pong copy holds the latest fields (fused)
    uint32_t fix_blocks = 0, obs_blocks = 0;
    bool timing = false;
    double kernel_ms[3] = {0, 0, 0};  // step kernel(s), fix, observe
    uint64_t kernel_count = 0, pending_steps = 0;
    // buffers
    DevBuf<float> E[2], H[2], psi[2];
    DevBuf<uint32_t> coef_index, step;
    DevBuf<float> coef_sets, axes, dtab, dstate, dold, ind_i, series, plane, ntff, probes;
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
    std::vector<hipEvent_t> events;

    ~Session() {
        for (auto e : events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static void validate(const fg_problem *P, uint32_t driven, const fg_extras *X) {
    auto bad = [](const std::string &m) { throw HipError{m, FG_INVALID}; };
    if (!P || !X) bad("NULL problem or extras");
    for (int a = 0; a < 3; a++)
        if (P->n[a] < 2) bad("a grid axis needs at least two nodes");
    const uint64_t nn = uint64_t(P->n[0]) * P->n[1] * P->n[2];
    if (!P->axes || !P->coef_sets || !P->coef_index || P->n_sets == 0) bad("missing axes or coefficient tables");
    if (P->n_ports && driven >= P->n_ports) bad("driven port " + std::to_string(driven) + " does not exist");
    if (X->nf > FG_MAX_FREQS) bad("at most 4 DFT frequencies");
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
}

static uint64_t psi_len(const uint32_t n[3], int a, uint32_t pml) {
    if (!pml) return 0;
    uint64_t m[3] = {n[0], n[1], n[2]};
    m[a] = 2ull * pml;
    return m[0] * m[1] * m[2];
}

// Split the march blocks into those whose reach (their own nodes plus the one-node halo on
// every side) touches a CPML slab position and those that do not.
static void plan_march(Session &s, const fg_problem *P, const std::vector<int> &slots, uint32_t nmax) {
    std::vector<uint32_t> pre[3];
    for (int a = 0; a < 3; a++) {
        pre[a].assign(P->n[a] + 1, 0);
        for (uint32_t q = 0; q < P->n[a]; q++) {
            const bool slab = P->pml && (slots[(size_t(a) * 2) * nmax + q] >= 0 || slots[(size_t(a) * 2 + 1) * nmax + q] >= 0);
            pre[a][q + 1] = pre[a][q] + (slab ? 1 : 0);
        }
    }
    auto touches = [&](int a, int64_t lo, int64_t hi) {  // any slab position in [lo, hi]
        lo = std::max<int64_t>(lo, 0);
        hi = std::min<int64_t>(hi, int64_t(P->n[a]) - 1);
        return hi >= lo && pre[a][size_t(hi) + 1] > pre[a][size_t(lo)];
    };
    const uint32_t gx = (P->n[2] + OWN_K - 1) / OWN_K, gy = (P->n[1] + s.tj - 1) / s.tj, gz = (P->n[0] + s.seg - 1) / s.seg;
    std::vector<MarchBlock> in, zs, pm;
    for (uint32_t bz = 0; bz < gz; bz++)
        for (uint32_t by = 0; by < gy; by++)
            for (uint32_t bx = 0; bx < gx; bx++) {
                const int64_t k0 = int64_t(bx) * OWN_K, j0 = int64_t(by) * s.tj, i0 = int64_t(bz) * s.seg;
                const bool z = touches(2, k0 - 1, k0 + OWN_K);
                const bool xy = touches(1, j0 - 1, j0 + s.tj) || touches(0, i0 - 1, i0 + s.seg);
                (xy ? pm : z ? zs : in).push_back(MarchBlock{bx, by, bz});
            }
    // Workgroups go to the XCDs round-robin (launch index b runs on XCD b % xcds). Permute each
    // list so that XCD x receives a contiguous run of the logical order (k tiles fastest, then
    // j tiles): neighbouring tiles then share one L2, and the cache lines their halos and
    // unaligned edges have in common are fetched from HBM once.
    auto xcd_order = [&](std::vector<MarchBlock> &v) {
        const size_t N = v.size(), X = s.xcds, per = (N + X - 1) / X;
        if (X <= 1 || N == 0) return;
        std::vector<MarchBlock> out(per * X, MarchBlock{~0u, ~0u, ~0u});
        for (size_t b = 0; b < per * X; b++) {
            const size_t L = (b % X) * per + b / X;
            if (L < N) out[b] = v[L];
        }
        v.swap(out);
    };
    xcd_order(in);
    xcd_order(zs);
    xcd_order(pm);
    s.blk_int.upload(in.data(), in.size(), s.stream);
    s.blk_z.upload(zs.data(), zs.size(), s.stream);
    s.blk_pml.upload(pm.data(), pm.size(), s.stream);
}

static Session *session_new(fg_device *dev, const fg_problem *P, uint32_t driven, const fg_pulse *pulse,
                            const fg_extras *X) {
    validate(P, driven, X);
    if (P->n_sheets || P->n_wires)
        throw HipError{"resistive sheets and thin wires run on the CPU path (run_cpu)", FG_UNSUPPORTED};
    auto S = std::make_unique<Session>();
    Session &s = *S;
    s.dev = dev;
    FG_CHECK(hipSetDevice(dev->ordinal));
    for (int a = 0; a < 3; a++) s.n[a] = P->n[a];
    const uint64_t nn = s.nn = uint64_t(P->n[0]) * P->n[1] * P->n[2];
    s.fused = X->engine != FG_ENGINE_SPLIT;
    // capacity check before touching HBM: "{cells} cells is more than this GPU can hold"
    uint64_t psi_total = 0, off[18];
    for (int e = 0; e < 2; e++)
        for (int c = 0; c < 3; c++)
            for (int a = 0; a < 3; a++) {
                off[e * 9 + c * 3 + a] = psi_total;
                if (a != c) psi_total += psi_len(P->n, a, P->pml);
            }
    const int copies = s.fused ? 2 : 1;
    const uint64_t need = uint64_t(copies) * (24 * nn + 4 * psi_total) + (P->coef_wide ? 12 : 6) * nn;
    size_t free_b = 0, total_b = 0;
    FG_CHECK(hipMemGetInfo(&free_b, &total_b));
    if (need > free_b) {
        const uint64_t cells = uint64_t(P->n[0] - 1) * (P->n[1] - 1) * (P->n[2] - 1);
        throw HipError{std::to_string(cells) + " cells is more than this GPU can hold", FG_TOO_LARGE};
    }
    FG_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    hipStream_t st = s.stream;
    for (int c = 0; c < copies; c++) {
        s.E[c].alloc(3 * nn);
        s.H[c].alloc(3 * nn);
        s.psi[c].alloc(std::max<uint64_t>(psi_total, 1));
        s.E[c].zero(st);
        s.H[c].zero(st);
        s.psi[c].zero(st);
    }
    s.coef_index.upload(P->coef_index, P->coef_wide ? 3 * nn : (3 * nn + 1) / 2, st);
    s.coef_sets.upload(P->coef_sets, 2 * size_t(P->n_sets), st);
    // axes: AoS rows of 10 -> SoA per axis (8 float fields + 2 int slots)
    const uint32_t nmax = std::max({P->n[0], P->n[1], P->n[2]});
    std::vector<float> axf(3 * 8 * size_t(nmax), 0.f);
    std::vector<int> axs(3 * 2 * size_t(nmax), -1);
    uint32_t row = 0;
    for (int a = 0; a < 3; a++)
        for (uint32_t i = 0; i < P->n[a]; i++, row++) {
            const float *r = P->axes + 10 * size_t(row);
            for (int f = 0; f < 8; f++) axf[(size_t(a) * 8 + f) * nmax + i] = r[f];
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
    }
    f.ind = s.ind.p;
    f.ind_i = s.ind_i.p;
    double t0, tau;
    tau = std::sqrt(std::log(10.0)) / (3.14159265358979323846 * pulse->fc);  // pulse_shape, sim.rs:646-649
    t0 = 4.0 * tau;
    f.dt = float(P->dt);
    f.t0 = float(t0);
    f.tau = float(tau);
    f.w0 = float(2.0 * 3.14159265358979323846 * pulse->f0);
    f.step = s.step.p;
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
    s.seg = std::min<uint32_t>(s.seg, SEG_MAX);
    if (s.fused) plan_march(s, P, axs, nmax);
    FG_CHECK(hipStreamSynchronize(st));
    return S.release();
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
    if (!s.timing || s.pending_steps == 0) return;
    FG_CHECK(hipStreamSynchronize(s.stream));
    for (uint64_t q = 0; q < s.pending_steps; q++)
        for (int g = 0; g < 3; g++) {
            float ms = 0.f;
            FG_CHECK(hipEventElapsedTime(&ms, s.events[4 * q + g], s.events[4 * q + g + 1]));
            s.kernel_ms[g] += ms;
        }
    s.kernel_count += s.pending_steps;
    s.pending_steps = 0;
}

// Enqueue one step at offset s_off from the device step counter.
static void launch_step(Session &s, uint32_t s_off, uint64_t global_step) {
    hipStream_t st = s.stream;
    const bool timed = s.timing;
    const size_t slot = 4 * size_t(s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot), st));
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
#define FG_MARCH2(TJ, W, L)                                                                                       \
    do {                                                                                                          \
        if (ni) hipLaunchKernelGGL((k_march<TJ, 0, W, L>), dim3(ni), block, 0, st, g, bi, Eo, En, Ho, Hn, Po, Pn, s.seg); \
        if (nz) hipLaunchKernelGGL((k_march<TJ, 1, W, L>), dim3(nz), block, 0, st, g, bz, Eo, En, Ho, Hn, Po, Pn, s.seg); \
        if (np_) hipLaunchKernelGGL((k_march<TJ, 2, W, L>), dim3(np_), block, 0, st, g, bp, Eo, En, Ho, Hn, Po, Pn, s.seg); \
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
        switch (s.tj) {
            case 8: FG_MARCH(8); break;
            default: FG_MARCH(14); break;
        }
#undef FG_MARCH2
#undef FG_MARCH
        Ecur = s.E[b].p;
        Eold = s.E[a].p;
        s.cur = b;
    } else {
        dim3 block(64, 4);
        dim3 grid((g.n2 + 63) / 64, (g.n1 + 3) / 4, g.n0);
        hipLaunchKernelGGL(k_update_h, grid, block, 0, st, g, s.E[0].p, s.H[0].p, s.psi[0].p);
        if (s.fix.n_debye)
            hipLaunchKernelGGL(k_debye_save, dim3(uint32_t((s.fix.n_debye + 255) / 256)), dim3(256), 0, st, s.fix.n_debye,
                               s.fix.debye, s.E[0].p, s.dold.p);
        hipLaunchKernelGGL(k_update_e, grid, block, 0, st, g, s.E[0].p, s.H[0].p, s.psi[0].p);
        Ecur = s.E[0].p;
    }
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 1), st));
    if (s.fix_blocks) hipLaunchKernelGGL(k_fix, dim3(s.fix_blocks), dim3(256), 0, st, g, s.fix, Ecur, Eold, s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 2), st));
    if (s.obs_blocks)
        hipLaunchKernelGGL(k_observe, dim3(s.obs_blocks), dim3(256), 0, st, g, s.obs, Ecur, s.H[s.fused ? s.cur : 0].p, s_off);
    if (timed) FG_CHECK(hipEventRecord(timing_event(s, slot + 3), st));
    (void)global_step;
    FG_CHECK(hipGetLastError());
}

static void enqueue_steps(Session &s, uint64_t count) {
    if (s.timing) {
        collect_timing(s);
        // timed steps go in batches so the event pool stays bounded
        for (uint64_t done = 0; done < count;) {
            const uint64_t k = std::min<uint64_t>(count - done, 256);
            hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s.stream, s.step.p, uint32_t(s.steps));
            for (uint64_t q = 0; q < k; q++) launch_step(s, uint32_t(q), s.steps + q);
            s.steps += k;
            s.pending_steps = k;
            collect_timing(s);
            done += k;
        }
        return;
    }
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s.stream, s.step.p, uint32_t(s.steps));
    for (uint64_t q = 0; q < count; q++) launch_step(s, uint32_t(q), s.steps + q);
    s.steps += count;
}

static double energy(Session &s) {
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

static void read_record(Session &s, fg_record *out) {
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

template <class F>
static int guarded(F &&body) {
    try {
        return body();
    } catch (const HipError &e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc &) {
        return set_error(FG_TOO_LARGE, "out of host memory");
    }
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
        *out = new fg_device{ordinal, std::string(prop.name) + " (" + prop.gcnArchName + ")", uint64_t(prop.totalGlobalMem),
                             prop.multiProcessorCount};
        return int(FG_OK);
    });
}

void fg_close(fg_device *dev) { delete dev; }

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
            fg::enqueue_steps(*s, k);
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
