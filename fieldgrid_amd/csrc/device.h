// Device-side views and kernel bodies shared by the single-run engine (engine.hip) and the
// batched sweep (batch.hip). Every body here is a __device__ function; the __global__
// wrappers that bind it to one run or to a table of runs live in the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fieldgrid.h"

namespace fg {

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

// The +k / -k neighbour's value across the wave's 64 lanes as one DPP move (wave_shl:1 /
// wave_shr:1, a VALU op instead of an LDS-routed ds_bpermute): lane l gets lane l+1's (l-1's)
// value; lane 63 (lane 0), which has no such neighbour, keeps its own, as __shfl_down(v, 1) and
// __shfl_up(v, 1) return. Called with all 64 lanes active.
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xF, 0xF, false));
}

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

// A register move the compiler keeps where it stands (it neither drops nor moves an asm
// statement): march2 moves each plane's loaded values out at the plane that uses them.
__device__ __forceinline__ float vcopy(float x) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ uint32_t vcopy(uint32_t x) {
    uint32_t r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Loads and stores through a resource built once (see march2's block resources).
__device__ __forceinline__ float ldr(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(byte_off), 0, 0));
}
__device__ __forceinline__ uint32_t ldr32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, int(byte_off), 0, 0);
}
__device__ __forceinline__ uint32_t ldr16(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, int(byte_off), 0, 0);
}
__device__ __forceinline__ void str(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, int(byte_off), 0, 0);
}

// march2 builds one resource per component array before its plane loop, based at the block's
// first plane, and reaches plane p by adding (p - first) plane strides to the lane's offset.
// Lanes outside the grid start at OOB2 instead of an in-plane offset. The host keeps a block's
// planes within OOB2 bytes (march2_seg_cap), so every offset sum stays below 2^31, an in-grid
// sum below the range and every other one past it.
constexpr uint32_t OOB2 = 0x3FFFFFF0u;
// Longest i segment march2 may take when the largest per-plane byte count of the arrays it
// marches (a field plane, a z or y CPML psi row block) is `plane_bytes`: a block's resources
// span at most seg + 4 planes, and its offsets reach one plane past them. 0 when not even one
// plane fits; the host then takes one step per pass.
__host__ __device__ constexpr uint32_t march2_seg_cap(uint64_t plane_bytes) {
    return plane_bytes == 0 ? uint32_t(SEG_MAX)
           : OOB2 / plane_bytes < 6 ? 0u
           : OOB2 / plane_bytes - 5 < uint64_t(SEG_MAX) ? uint32_t(OOB2 / plane_bytes - 5)
                                                        : uint32_t(SEG_MAX);
}

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


// Thin wires (cpu.rs:270-273, sim.rs:556-635): after the H update, H[h] += sum_t c_t E[e_t], E
// being the step's input. In the fused march every copy of such an H node (owner or halo) adds
// it, so the E updates of every block see the same H. Entries of plane i are
// ent[start[i] .. start[i+1]) as (j, k, component, index into w).
struct WireArgs {
    const uint32_t *start;
    const uint4 *ent;
    const fg_wire_h *w;
};

// PML: 0 = no CPML slab in reach, 1 = only z (k) slabs, 2 = any slab (x, y and z psi code);
// WIRES: the step has thin-wire terms (wa)
template <int TJ, int PML, bool WIDE, bool LDS_SETS, bool WIRES = false>
__device__ __forceinline__ void march(const Geo &g, const MarchBlock mb, const float *__restrict__ Eo,
                                      float *__restrict__ En, const float *__restrict__ Ho, float *__restrict__ Hn,
                                      const float *__restrict__ psio, float *__restrict__ psin, uint32_t seg,
                                      const WireArgs &wa = WireArgs{}) {
    constexpr int ROWS = TJ + 2;
    // row exchanges; rows are stored at index w+1 so that rows -1 and ROWS are zero pads
    __shared__ float xe[2][ROWS + 2][2][64];  // E(i+1) Ex, Ez of each row, for the row above
    __shared__ float xh[2][ROWS + 2][2][64];  // new H(i) Hx, Hz of each row, for the row below
    __shared__ float2 sets[LDS_SETS ? SETS_LDS : 1];
    constexpr bool PZ = PML >= 1, PXY = PML == 2;
    __shared__ float xa[SEG_MAX + 1][PXY ? 10 : 2];  // x-axis values of the segment's planes
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
        const float ex_k = lane_next(ex), ey_k = lane_next(ey);
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
        if constexpr (WIRES) {
            const uint32_t q0 = wa.start[i], q1 = wa.start[i + 1];  // uniform
            for (uint32_t q = q0; q < q1; q++) {
                const uint4 e = wa.ent[q];
                if (act && e.x == uj && e.y == uk) {
                    const fg_wire_h &wh = wa.w[e.w];
                    float sw = 0.f;
                    for (int t = 0; t < 4; t++) sw += wh.c[t] * Eo[wh.e[t]];
                    if (e.z == 0) hx += sw;
                    else if (e.z == 1) hy += sw;
                    else hz += sw;
                }
            }
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
        const float hx_k = lane_prev(hx), hy_k = lane_prev(hy);
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
    const float *dsrc;  // where the step reads the pole states when not from dstate (NULL =
                        // in place): the second step of a two-step pass reads those the pass wrote
    const float *dold;  // E before the step's E update: split engine, saved before the E sweep;
                        // second step of a two-step pass, saved by the pass; else NULL (read Eo)
    uint32_t n_src;
    const fg_source *src;
    uint32_t n_ind;
    const fg_inductor *ind;
    float *ind_i;
    float dt, t0, tau, w0;
    const uint32_t *step;
    // pulse((n+1/2)dt) per step, evaluated on the host in f32 with the reference's expression
    // (cpu.rs:262-265), so the device adds exactly the source values of the CPU path
    const float *pulse;
    uint32_t n_pulse;
    // sources and inductors share an edge (or a source edge repeats): apply them one after
    // another in one thread, in the reference's order, instead of one thread per entry
    uint32_t serial;
};

// The same pulse on the device, for steps past the host table (a session stepped beyond
// max_steps): f32 like cpu.rs:262-265, but with the device's expf/cosf.
__device__ __forceinline__ float pulse_at(const FixArgs &a, float t) {
    const float x = (t - a.t0) / a.tau;
    return expf(-x * x) * cosf(a.w0 * (t - a.t0));
}

// Thread t of one run's sparse fixes, in the order of cpu.rs:281-303: Debye ADE, then the
// driven port's sources, then the inductors. E is the new field, Eo the previous step's.
__device__ __forceinline__ void fix_step(const Geo &g, const FixArgs &a, float *__restrict__ E,
                                         const float *__restrict__ Eo, uint32_t s_off, uint64_t t) {
    if (t < a.n_debye) {
        const fg_debye_edge d = a.debye[t];
        const float old = a.dold ? a.dold[t] : Eo[d.f];
        const float cb = coef_at(g, d.f).y;
        const float *tab = a.dtab + size_t(d.set) * a.np * 2;
        const float *js = a.dsrc ? a.dsrc : a.dstate;
        float s0 = 0.f;
        for (uint32_t q = 0; q < a.np; q++) s0 += 0.5f * (1.0f + tab[2 * q]) * js[q * a.n_debye + t];
        const float next = E[d.f] - cb * s0;
        E[d.f] = next;
        const float de = next - old;
        for (uint32_t q = 0; q < a.np; q++)
            a.dstate[q * a.n_debye + t] = tab[2 * q] * js[q * a.n_debye + t] + d.dd * tab[2 * q + 1] * de;
    }
    auto source = [&](uint64_t q) {
        const uint32_t n = a.step[0] + s_off;
        const float pl = n < a.n_pulse ? a.pulse[n] : pulse_at(a, (float(n) + 0.5f) * a.dt);
        const fg_source s = a.src[q];
        E[s.comp * g.nn + s.id] -= s.src * pl;
    };
    auto inductor = [&](uint64_t q) {
        const fg_inductor d = a.ind[q];
        const uint64_t f = d.comp * g.nn + d.id;
        const float x = E[f] - d.k_e * a.ind_i[q];
        E[f] = x;
        a.ind_i[q] += d.k_i * x;
    };
    if (a.serial) {
        if (t == 0) {
            for (uint32_t q = 0; q < a.n_src; q++) source(q);
            for (uint32_t q = 0; q < a.n_ind; q++) inductor(q);
        }
        return;
    }
    if (t < a.n_src) source(t);
    if (t < a.n_ind) inductor(t);
}

// ---- resistive sheets (cpu.rs:169-237, fdtd.wgsl:332-427) -------------------------------

// One sheet edge's state lives in its 16-f32 record (the pack() layout, run.rs:259-290): [0] E
// index, [1], [2] the H indices on either face, [3] length, [4] c, [5] r, [6] adaptive, [7..9]
// running aa, bb, ab, [10], [11] face weights, [12] E before the E update, [13] the previous
// current, [14], [15] the faces' rows in the sheet table. The table holds nr real and nc
// complex branches: per branch its recursion coefficients, per face row its weights.
struct SheetArgs {
    uint64_t n;
    float *sh;
    float *cur;         // n x (nr + 2 nc) branch currents
    const float *br;
    uint32_t serial;    // sheets share an edge: one thread applies them in order
};

// cpu.rs:169-193: face weights from the running H energy on either face
__device__ __forceinline__ void sheet_before(const SheetArgs &a, uint64_t s, float old, float ha, float hb) {
    float *sh = a.sh + s * 16;
    sh[12] = old;
    const float aa = sh[7] + ha * ha, bb = sh[8] + hb * hb, ab = sh[9] + ha * hb;
    sh[7] = aa, sh[8] = bb, sh[9] = ab;
    const float sum = aa + bb;
    if (sum <= 1e-30f) return;
    float scale = 1.0f;
    if (sh[6] != 0.0f) {
        scale = sh[10] + sh[11];
        const float tot = aa + bb - 2.0f * ab;
        if (tot > 1e-30f) {
            float x = sum / tot;  // clamp(x, 0.5, 4.0), NaN passing through as in Rust
            x = x < 0.5f ? 0.5f : x;
            scale = 4.0f < x ? 4.0f : x;
        }
    }
    sh[10] = scale * aa / sum;
    sh[11] = scale * bb / sum;
}

// cpu.rs:195-237: the sheet's current from the trapezoidal branch model, taken out of E
__device__ __forceinline__ void sheet_after(const SheetArgs &a, uint64_t s, float *__restrict__ E) {
    float *sh = a.sh + s * 16;
    const float *br = a.br;
    const uint64_t f = __float_as_uint(sh[0]);
    const float len = sh[3], c = sh[4], r = sh[5];
    const float ma = sh[10], mb = sh[11], old = sh[12], i_old = sh[13];
    const uint32_t nr = uint32_t(br[0]), nc = uint32_t(br[1]);
    const uint32_t stride = 2 + nr + 2 * nc, faces = 4 + 2 * nr + 4 * nc;
    const uint64_t fa = faces + uint64_t(__float_as_uint(sh[14])) * stride,
                   fb = faces + uint64_t(__float_as_uint(sh[15])) * stride;
    float *cu = a.cur + s * (nr + 2 * nc);
    float hist = (ma * br[fa + 1] + mb * br[fb + 1]) * i_old;
    for (uint32_t k = 0; k < nr; k++) hist -= (ma * br[fa + 2 + k] + mb * br[fb + 2 + k]) * cu[k];
    for (uint32_t k = 0; k < nc; k++) {
        const uint32_t o = 2 + nr + 2 * k;
        const float cr = ma * br[fa + o] + mb * br[fb + o];
        const float ci = ma * br[fa + o + 1] + mb * br[fb + o + 1];
        hist -= cr * cu[nr + 2 * k] - ci * cu[nr + 2 * k + 1];
    }
    const float ra = r * (ma * br[fa] + mb * br[fb]);
    const float ibar = (0.5f * len * (E[f] + old) + r * hist) / (ra + 0.5f * len * c);
    E[f] = E[f] - c * ibar;
    sh[13] = 2.0f * ibar - i_old;
    for (uint32_t k = 0; k < nr; k++) {
        const uint32_t q = 4 + 2 * k;
        cu[k] = br[q] * cu[k] + br[q + 1] * ibar;
    }
    for (uint32_t k = 0; k < nc; k++) {
        const uint32_t q = 4 + 2 * nr + 4 * k;
        const float ur = cu[nr + 2 * k], ui = cu[nr + 2 * k + 1];
        cu[nr + 2 * k] = br[q] * ur - br[q + 1] * ui + br[q + 2] * ibar;
        cu[nr + 2 * k + 1] = br[q] * ui + br[q + 1] * ur + br[q + 3] * ibar;
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
    // per step and frequency: cos(w te) dt, -sin(w te) dt, cos(w th) dt, -sin(w th) dt with
    // te = (n+1)dt, th = (n+1/2)dt, evaluated on the host in f32 as cpu.rs:352-353 and 404-405 do
    const float *phase;
    uint32_t n_phase;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One plane-DFT point q at step n (cpu.rs:336-361): f = {Ex, Ey, Ez, Hx, Hy} of the point.
__device__ __forceinline__ void plane_acc(const ObsArgs &a, uint32_t q, uint32_t n, const float f[5]) {
    const float dt = a.dt;
    const float te = (float(n) + 1.0f) * dt, th = (float(n) + 0.5f) * dt;
    float *out = a.plane + uint64_t(q) * a.nf * 10;
    for (uint32_t kf = 0; kf < a.nf; kf++) {
        const float w = a.omega[kf];
        const float *ph = a.phase + (uint64_t(n) * a.nf + kf) * 4;
        const bool tab = n < a.n_phase;
        const float ce0 = tab ? ph[0] : cosf(w * te) * dt, ce1 = tab ? ph[1] : -sinf(w * te) * dt;
        const float ch0 = tab ? ph[2] : cosf(w * th) * dt, ch1 = tab ? ph[3] : -sinf(w * th) * dt;
        for (int c = 0; c < 5; c++) {
            const float p0 = c >= 3 ? ch0 : ce0, p1 = c >= 3 ? ch1 : ce1;
            out[kf * 10 + 2 * c] += f[c] * p0;
            out[kf * 10 + 2 * c + 1] += f[c] * p1;
        }
    }
}

// One NTFF patch gi at step n (cpu.rs:363-415): e = {Eu(p), Eu(p+v), Ev(p), Ev(p+u)} and
// h = {Hu at its four nodes, Hv at its four nodes}, in the patch table's order.
__device__ __forceinline__ void ntff_acc(const ObsArgs &a, uint64_t gi, uint32_t n, uint32_t meta, const float e[4],
                                         const float h[8]) {
    const float dt = a.dt;
    const float te = (float(n) + 1.0f) * dt, th = (float(n) + 0.5f) * dt;
    const uint32_t ax = meta & 3u;
    const float side = (meta & 4u) ? 1.f : -1.f;
    const uint32_t u = (ax + 1) % 3, v = (ax + 2) % 3;
    float ef[3] = {0.f, 0.f, 0.f}, hf[3] = {0.f, 0.f, 0.f};
    ef[u] = 0.5f * (e[0] + e[1]);
    ef[v] = 0.5f * (e[2] + e[3]);
    hf[u] = 0.25f * (h[0] + h[1] + h[2] + h[3]);
    hf[v] = 0.25f * (h[4] + h[5] + h[6] + h[7]);
    float nrm[3] = {0.f, 0.f, 0.f};
    nrm[ax] = side;
    const float jv[3] = {nrm[1] * hf[2] - nrm[2] * hf[1], nrm[2] * hf[0] - nrm[0] * hf[2],
                         nrm[0] * hf[1] - nrm[1] * hf[0]};
    const float mv[3] = {ef[1] * nrm[2] - ef[2] * nrm[1], ef[2] * nrm[0] - ef[0] * nrm[2],
                         ef[0] * nrm[1] - ef[1] * nrm[0]};
    float *out = a.ntff + gi * a.nf * 12;
    for (uint32_t kf = 0; kf < a.nf; kf++) {
        const float wf = a.omega[kf];
        // cos and sin are exactly even and odd, so the table's values are these
        const float *ph = a.phase + (uint64_t(n) * a.nf + kf) * 4;
        const bool tab = n < a.n_phase;
        const float pe0 = tab ? ph[0] : cosf(-wf * te) * dt, pe1 = tab ? ph[1] : sinf(-wf * te) * dt;
        const float ph0 = tab ? ph[2] : cosf(-wf * th) * dt, ph1 = tab ? ph[3] : sinf(-wf * th) * dt;
        float *o = out + kf * 12;
        for (int c = 0; c < 3; c++) {
            o[2 * c] += ph0 * jv[c];
            o[2 * c + 1] += ph1 * jv[c];
            o[6 + 2 * c] += pe0 * mv[c];
            o[6 + 2 * c + 1] += pe1 * mv[c];
        }
    }
}

// The port V/I and probe blocks of observe() (blk < a.blk_ports + a.blk_probes).
__device__ __forceinline__ void observe_ports(const Geo &g, const ObsArgs &a, const float *__restrict__ E,
                                              const float *__restrict__ H, uint32_t s_off, uint32_t blk,
                                              uint32_t tid) {
    const uint32_t n = a.step[0] + s_off;
    const uint64_t nn = g.nn;
    if (blk < a.blk_ports) {
        // one wave per port; f32 products summed in f64 like cpu.rs:317-328
        const uint32_t p = blk * 4 + tid / 64, lane = tid % 64;
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
        const uint32_t q = blk * 256 + tid;
        if (q < a.n_probes && n < a.cap) {
            const uint64_t f = a.probe_f[q];
            a.probes[uint64_t(n) * a.n_probes + q] = f < 3 * nn ? E[f] : H[f - 3 * nn];
        }
    }
}

// Block `blk` of one run's observers (cpu.rs:309-415): port V/I waves, then point probes,
// then plane DFT, then NTFF patches, each section a range of 256-thread blocks; `tid` is the
// thread's index in its 256-thread block (its four waves whole).
__device__ __forceinline__ void observe(const Geo &g, const ObsArgs &a, const float *__restrict__ E,
                                        const float *__restrict__ H, uint32_t s_off, uint32_t blk, uint32_t tid) {
    if (blk < a.blk_ports + a.blk_probes) {
        observe_ports(g, a, E, H, s_off, blk, tid);
        return;
    }
    const uint32_t n = a.step[0] + s_off;
    const uint64_t nn = g.nn;
    blk -= a.blk_ports + a.blk_probes;
    if (blk < a.blk_plane) {
        const uint32_t q = blk * 256 + tid;
        if (q >= a.plane_n) return;
        const uint32_t i = q / g.n1, j = q % g.n1;
        const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + a.plane_k;
        const float f[5] = {E[id], E[nn + id], E[2 * nn + id], H[id], H[nn + id]};
        plane_acc(a, q, n, f);
        return;
    }
    blk -= a.blk_plane;
    if (blk < a.blk_ntff) {
        const uint64_t gi = uint64_t(blk) * 256 + tid;
        if (gi >= a.n_patches) return;
        const uint32_t *ix = a.patch + gi * 13;
        const uint32_t meta = ix[0], ax = meta & 3u, u = (ax + 1) % 3, v = (ax + 2) % 3;
        const float e[4] = {E[u * nn + ix[1]], E[u * nn + ix[2]], E[v * nn + ix[3]], E[v * nn + ix[4]]};
        const float h[8] = {H[u * nn + ix[5]], H[u * nn + ix[6]], H[u * nn + ix[7]], H[u * nn + ix[8]],
                            H[v * nn + ix[9]], H[v * nn + ix[10]], H[v * nn + ix[11]], H[v * nn + ix[12]]};
        ntff_acc(a, gi, n, meta, e, h);
    }
}

// ---- two steps per pass (temporal blocking) --------------------------------------------------
//
// march2 advances the fields by two steps in one pass over HBM: it reads step n-1's E, H, psi
// from one ping-pong copy and writes step n+1's into the other; step n lives only in registers
// and LDS. A block is TJ+4 rows x 64 lanes (j0-2..j0+TJ+1, k0-2..k0+61) and owns the inner
// TJ x 60. The first sub-step (step n) runs on the whole tile, the second (step n+1) on the
// tile less one ring, so the second sees exactly the step-n values the owners of its
// neighbours computed (the same loads and the same arithmetic). Along i the block marches
// with a two-plane lag: iteration t computes step n at plane t and step n+1 at plane t-1.
// Step n's port sources are applied in the kernel, to every copy of the edge (owner or halo),
// from a per-plane list, and so is its Debye ADE (DEB: Deb2 below); nothing else may sit between
// the two steps (no inductors, plane DFT or NTFF: those runs take one step per pass).
constexpr int OWN_K2 = 60;

// Step n's sources by plane: entries of plane i are ent[start[i] .. start[i+1]) as
// (j, k, component, bits of src); the value subtracted is src * pulse(n).
struct Src2 {
    const uint32_t *start;
    const uint4 *ent;
};

// Step n's Debye ADE (cpu.rs:239-260) inside a two-step pass. Every copy of a Debye edge (owner
// or halo) applies it from the pole states J^n, which nothing writes during the pass; the
// edge's owner writes J^{n+1} and the edge's final step-n E to side buffers, from which the
// sparse fixes of step n+1 read them. Edges are found per node through a k map of each (i, j)
// row that holds any: row[i*n1 + j] is the row's slot or -1, map[(slot*n2 + k)*3 + c] the
// Debye index of edge (c; i, j, k) or -1.
struct Deb2 {
    const int32_t *row;
    const int32_t *map;
    float *jn;  // J^{n+1}, pole-major like FixArgs::dstate
    float *eo;  // step n's final E per Debye edge
};

// Step n's values at the points its plane DFT and NTFF observers read, captured inside the pass
// by each point's owner into buf. Per (i, j) row, seg[start[r] .. start[r+1]) are segments of
// consecutive k with the same fields: (k_lo, k_hi, mask of fields Ex Ey Ez Hx Hy Hz, first
// index); field c at k sits at first + (k - k_lo) * popcount(mask) + (rank of c in mask). The
// observers find their values through per-point index tables (ntff: 12 per patch, plane: 5 per
// point), in the order the one-step observers read them.
struct Cap2 {
    const uint32_t *start;
    const uint4 *seg;
    float *buf;
    const uint32_t *ntff;
    const uint32_t *plane;
    // with observer captures (NULL: none): the buffer index of every port edge's E (in the
    // order of ObsArgs::edges), then from obs_loop0 of every loop term's H (ObsArgs::loops),
    // then from obs_probe0 of every probe's field (ObsArgs::probe_f)
    const uint32_t *obs;
    uint32_t obs_loop0, obs_probe0;
};

// march2's LDS, declared once by the kernel that calls it, so that a kernel running blocks of
// several CPML classes (several march2 instances) holds one copy. Row exchanges are
// double-buffered by plane parity; row w is stored at index w+1 so that rows -1 and ROWS read
// as zero pads. xr[0] E0 (Ex, Ez) for the row above's step-n H update, xr[1] step-n H (Hx, Hz)
// for the row below's step-n E update; xr[2], xr[3] the same for step n+1.
template <int TJ, bool LDS_SETS, bool SIDE>
struct March2Lds {
    static constexpr int ROWS = TJ + 4;
    static constexpr uint32_t CAP_SEGS = SIDE ? 1536 : 1;
    float xr[4][2][ROWS + 2][2][64];
    float2 sets[LDS_SETS ? SETS_LDS : 1];
    float xa[SEG_MAX + 3][10];  // x-axis values of the segment's planes (the CPML ones from [2])
    uint32_t xsrc[SEG_MAX + 4];
    int32_t xdeb[SIDE ? (SEG_MAX + 3) * ROWS : 1];
    uint32_t xcap[SIDE ? SEG_MAX * (TJ + 1) : 1];
    uint32_t xcoff[SIDE ? SEG_MAX + 1 : 1];
    uint4 xseg[CAP_SEGS];
};

// One LDS block for a kernel that runs either march of a block: with side work or lean.
template <int TJ, bool LDS_SETS>
union March2LdsEither {
    March2Lds<TJ, LDS_SETS, true> side;
    March2Lds<TJ, LDS_SETS, false> lean;
};

// Bit 29 of a two-step march block's bz (bits 30-31 carry the class in mixed lists): the block's
// tile holds side work (march_classes2's side flags); kernels that run side work run the lean
// march for a block without it, which computes the same values there.
constexpr uint32_t SIDE_BLOCK = 1u << 29, BZ_SEG = SIDE_BLOCK - 1u;

// SIDE: the first step has side work (Debye ADE through db, captures through cp; either may be
// absent, NULL tables)
template <int TJ, int PML, bool WIDE, bool LDS_SETS, bool SIDE>
__device__ __forceinline__ void march2(March2Lds<TJ, LDS_SETS, SIDE> &lds, const Geo &g, const MarchBlock mb,
                                       const float *__restrict__ Eo,
                                       float *__restrict__ En, const float *__restrict__ Ho, float *__restrict__ Hn,
                                       const float *__restrict__ psio, float *__restrict__ psin, uint32_t seg,
                                       const Src2 &sr, float pl, const FixArgs &fx, const Deb2 &db,
                                       const Cap2 &cp) {
    constexpr int ROWS = TJ + 4;
    constexpr bool PZ = PML >= 1, PXY = PML == 2;
    auto &xr = lds.xr;
    auto &sets = lds.sets;
    auto &xa = lds.xa;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k = int(mb.bx) * OWN_K2 + lane - 2;
    const int j = int(mb.by) * TJ + w - 2;
    const uint32_t i0 = mb.bz * seg, i1 = min(i0 + seg, g.n0);
    const bool kin = k >= 0 && k < int(g.n2), jin = j >= 0 && j < int(g.n1);
    const bool act = kin && jin;
    const bool own = act && lane >= 2 && lane < 2 + OWN_K2 && w >= 2 && w < 2 + TJ;
    const uint32_t uk = kin ? uint32_t(k) : 0u, uj = jin ? uint32_t(j) : 0u;
    const uint64_t nn = g.nn, plane = uint64_t(g.n1) * g.n2;
    const uint32_t off = uj * g.n2 + uk;
    const bool kl = uk + 1 < g.n2, jl = uj + 1 < g.n1;
    const float idz = g.ax[2].inv_d[uk], iddz = g.ax[2].inv_dd[uk];
    const float idy = g.ax[1].inv_d[uj], iddy = g.ax[1].inv_dd[uj];
    const uint32_t ob4 = off * 4u;
    const uint32_t ldA = act ? ob4 : OOB, stO = own ? ob4 : OOB;
    const uint32_t ldC = act ? off * (WIDE ? 4u : 2u) : OOB;
    const uint32_t pl4 = uint32_t(plane) * 4u, plc = uint32_t(plane) * (WIDE ? 4u : 2u);
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
    // step n runs on planes tstart..i1, step n+1 on planes max(i0,1)-1 .. i1-1 (the plane
    // before i0 feeds the -x neighbour of i0's E update)
    const uint32_t tstart = i0 >= 2 ? i0 - 2 : 0;
    for (uint32_t t = w * 64 + lane; t + tstart <= i1; t += 64 * ROWS) {
        const uint32_t q = min(tstart + t, g.n0 - 1);
        xa[t][0] = g.ax[0].inv_d[q], xa[t][1] = g.ax[0].inv_dd[q];
        if constexpr (PXY) {
            xa[t][2] = g.ax[0].bh[q], xa[t][3] = g.ax[0].ch[q], xa[t][4] = g.ax[0].raw_d[q];
            xa[t][5] = g.ax[0].be[q], xa[t][6] = g.ax[0].ce[q], xa[t][7] = g.ax[0].raw_dd[q];
            xa[t][8] = __int_as_float(g.ax[0].slot_h[q]), xa[t][9] = __int_as_float(g.ax[0].slot_e[q]);
        }
    }
    if (w < 2)
        for (int x = 0; x < 4; x++)
            for (int q = 0; q < 2; q++) xr[x][q][w * (ROWS + 1)][0][lane] = 0.f, xr[x][q][w * (ROWS + 1)][1][lane] = 0.f;
    // The per-plane entries of the source, Debye-row and capture tables, staged in LDS before the
    // march: the plane loop then reads them from LDS. Read from global memory inside the loop, a
    // table entry's wait (vmcnt counts every vector memory access in issue order) would also
    // wait for the plane loads issued one iteration ahead, once per plane.
    //   xsrc[t - tstart]: sr.start[t] for the planes tstart .. i1 + 1 (clamped to n0);
    //   xdeb[(t - tstart) * ROWS + w]: the Debye row slot of (t, row w), -1 outside the grid;
    //   xcap[(t - i0) * (TJ + 1) + r]: cp.start of owned row r of owned plane t (r = TJ: the end),
    //   xseg: the owned planes' capture segments, plane p's at xcoff[p] on, when they all fit.
    constexpr uint32_t CAP_SEGS = March2Lds<TJ, LDS_SETS, SIDE>::CAP_SEGS;
    auto &xsrc = lds.xsrc;
    auto &xdeb = lds.xdeb;
    auto &xcap = lds.xcap;
    auto &xcoff = lds.xcoff;
    auto &xseg = lds.xseg;
    for (uint32_t t = w * 64 + lane; t + tstart <= i1 + 1; t += 64 * ROWS) xsrc[t] = sr.start[min(tstart + t, g.n0)];
    bool seg_lds = false;
    if constexpr (SIDE) {
        const uint32_t np = i1 - tstart + 1, no = i1 - i0, j0 = mb.by * TJ;
        for (uint32_t x = w * 64 + lane; x < np * ROWS; x += 64 * ROWS) {
            const uint32_t t = tstart + x / ROWS;
            const int jj = int(j0) + int(x % ROWS) - 2;
            xdeb[x] = db.row && t < g.n0 && jj >= 0 && jj < int(g.n1) ? db.row[uint64_t(t) * g.n1 + uint32_t(jj)] : -1;
        }
        if (cp.start)
            for (uint32_t x = w * 64 + lane; x < no * (TJ + 1); x += 64 * ROWS)
                xcap[x] = cp.start[uint64_t(i0 + x / (TJ + 1)) * g.n1 + min(j0 + x % (TJ + 1), g.n1)];
        __syncthreads();
        if (cp.start && w == 0 && lane == 0) {
            uint32_t o = 0;
            for (uint32_t q = 0; q < no; q++) xcoff[q] = o, o += xcap[q * (TJ + 1) + TJ] - xcap[q * (TJ + 1)];
            xcoff[no] = o;
        }
        __syncthreads();
        seg_lds = cp.start && xcoff[no] <= CAP_SEGS;  // block-uniform
        if (seg_lds)
            for (uint32_t x = w * 64 + lane; x < xcoff[no]; x += 64 * ROWS) {
                uint32_t lo = 0, hi = no;  // the owned plane of entry x: xcoff[lo] <= x < xcoff[hi]
                while (hi - lo > 1) {
                    const uint32_t m = (lo + hi) / 2;
                    if (xcoff[m] <= x) lo = m;
                    else hi = m;
                }
                xseg[x] = cp.seg[xcap[lo * (TJ + 1)] + (x - xcoff[lo])];
            }
    }
    // plane p of a field array through a resource that is empty when p is outside the grid
    auto pbase = [&](const float *a, int64_t p) { return a + (p >= 0 && p < int64_t(g.n0) ? uint64_t(p) * plane : 0); };
    auto pbytes = [&](int64_t p, uint32_t bytes) { return p >= 0 && p < int64_t(g.n0) ? bytes : 0u; };
#define PSI(e, c, a) g.psi_off[(e) * 9 + (c) * 3 + (a)]
    // carried state: E0 of the current plane, step-n values of the previous plane, step-n+1 H
    // of the plane before, coefficient pairs and step-n psi of the previous plane
    float ex0, ey0, ez0;
    {
        const float *p = pbase(Eo, tstart);
        const uint32_t b = pbytes(tstart, pl4);
        ex0 = ld(p, ldA, b), ey0 = ld(p + nn, ldA, b), ez0 = ld(p + 2 * nn, ldA, b);
    }
    {
        const int pb = (tstart + 1) & 1;  // the buffer the first iteration reads
        xr[0][pb][w + 1][0][lane] = ex0, xr[0][pb][w + 1][1][lane] = ez0;
    }
    __syncthreads();
    float h1x = 0, h1y = 0, h1z = 0;           // step-n H at the previous plane
    float e1x = 0, e1y = 0, e1z = 0;           // step-n E at the previous plane
    float h2y = 0, h2z = 0;                    // step-n+1 Hy, Hz at the plane before that
    float2 cf0 = {0, 0}, cf1 = {0, 0}, cf2 = {0, 0};  // (ca, cb) of the previous plane
    // step-n psi of the previous plane (component, axis)
    float q_hxy = 0, q_hxz = 0, q_hyz = 0, q_hyx = 0, q_hzx = 0, q_hzy = 0;
    float q_exy = 0, q_exz = 0, q_eyz = 0, q_eyx = 0, q_ezx = 0, q_ezy = 0;
    // the loads of plane t: E0(t+1), H0(t), coefficient indices and step n-1 psi. They are issued
    // one iteration ahead (plane t+1's while plane t computes), so their latency overlaps a
    // whole iteration instead of stalling between the barriers.
    struct Pl {
        float exn, eyn, ezn, hx0, hy0, hz0;
        uint32_t c0, c1, c2;
        float hxy, hxz, hyz, hyx, hzx, hzy, exy, exz, eyz, eyx, ezx, ezy;
    };
    auto load_plane = [&](uint32_t t, Pl &P) {
        const uint64_t pb = uint64_t(t) * plane;
        const float *pe = pbase(Eo, int64_t(t) + 1);
        const uint32_t eb = pbytes(int64_t(t) + 1, pl4);
        P.exn = ld(pe, ldA, eb), P.eyn = ld(pe + nn, ldA, eb), P.ezn = ld(pe + 2 * nn, ldA, eb);
        const float *ph = pbase(Ho, t);
        const uint32_t hb = pbytes(t, pl4);
        P.hx0 = ld(ph, ldA, hb), P.hy0 = ld(ph + nn, ldA, hb), P.hz0 = ld(ph + 2 * nn, ldA, hb);
        const uint32_t cbytes = pbytes(t, plc);
        const uint64_t cb_ = t < g.n0 ? pb : 0;
        P.c0 = cidx(cb_, cbytes), P.c1 = cidx(nn + cb_, cbytes), P.c2 = cidx(2 * nn + cb_, cbytes);
        P.hxy = P.hxz = P.hyz = P.hyx = P.hzx = P.hzy = 0.f;
        P.exy = P.exz = P.eyz = P.eyx = P.ezx = P.ezy = 0.f;
        const uint32_t zb = t < g.n0 ? zbytes : 0u, yb = t < g.n0 ? ybytes : 0u, xb = t < g.n0 ? pl4 : 0u;
        if constexpr (PZ) {
            const uint64_t zrow = t < g.n0 ? uint64_t(t) * g.n1 * p2 : 0;
            P.hxz = ld(psio + PSI(1, 0, 2) + zrow, ozh, zb), P.hyz = ld(psio + PSI(1, 1, 2) + zrow, ozh, zb);
            P.exz = ld(psio + PSI(0, 0, 2) + zrow, oze, zb), P.eyz = ld(psio + PSI(0, 1, 2) + zrow, oze, zb);
        }
        if constexpr (PXY) {
            const float *xv = xa[t - tstart];
            const int shx = __float_as_int(xv[8]), sex = __float_as_int(xv[9]);
            const uint64_t yrow = t < g.n0 ? uint64_t(t) * p2 * g.n2 : 0;
            const uint64_t xhb = uint64_t(max(shx, 0)) * plane, xeb = uint64_t(max(sex, 0)) * plane;
            const uint32_t oxh = shx >= 0 ? ldA : OOB, oxe = sex >= 0 ? ldA : OOB;
            P.hxy = ld(psio + PSI(1, 0, 1) + yrow, oyh, yb);
            P.hyx = ld(psio + PSI(1, 1, 0) + xhb, oxh, xb);
            P.hzx = ld(psio + PSI(1, 2, 0) + xhb, oxh, xb), P.hzy = ld(psio + PSI(1, 2, 1) + yrow, oyh, yb);
            P.exy = ld(psio + PSI(0, 0, 1) + yrow, oye, yb);
            P.eyx = ld(psio + PSI(0, 1, 0) + xeb, oxe, xb);
            P.ezx = ld(psio + PSI(0, 2, 0) + xeb, oxe, xb), P.ezy = ld(psio + PSI(0, 2, 1) + yrow, oye, yb);
        }
    };
    Pl nxt;
    load_plane(tstart, nxt);
    for (uint32_t t = tstart; t <= i1; t++) {
        const int b = t & 1;
        const bool il = t + 1 < g.n0;
        const Pl cur = nxt;
        if (t < i1) load_plane(t + 1, nxt);  // uniform
        const float exn = cur.exn, eyn = cur.eyn, ezn = cur.ezn, hx0 = cur.hx0, hy0 = cur.hy0, hz0 = cur.hz0;
        const uint32_t c0 = cur.c0, c1 = cur.c1, c2 = cur.c2;
        const float *xv = xa[t - tstart];
        const float idx_ = xv[0], iddx = xv[1];
        float bhx = 0, chx = 0, rdx = 0, bex = 0, cex = 0, rddx = 0;
        float hxy = cur.hxy, hxz = cur.hxz, hyz = cur.hyz, hyx = cur.hyx, hzx = cur.hzx, hzy = cur.hzy;
        float exy = cur.exy, exz = cur.exz, eyz = cur.eyz, eyx = cur.eyx, ezx = cur.ezx, ezy = cur.ezy;
        if constexpr (PXY) {
            bhx = xv[2], chx = xv[3], rdx = xv[4];
            bex = xv[5], cex = xv[6], rddx = xv[7];
        }
        // ---- step n, H at plane t (the +1 neighbours of E0: j+1 from LDS, k+1 from the next lane)
        const float ex_j = xr[0][b ^ 1][w + 2][0][lane], ez_j = xr[0][b ^ 1][w + 2][1][lane];
        const float ex_k = lane_next(ex0), ey_k = lane_next(ey0);
        float nhx, nhy, nhz;
        {
            const float ev = ez_j - ez0, eu = ey_k - ey0;
            float t1 = ev * idy, t2 = eu * idz;
            if constexpr (PXY) hxy = bhy * hxy + chy * ev * rdy, t1 += hxy;
            if constexpr (PZ) hxz = bhz * hxz + chz * eu * rdz, t2 += hxz;
            nhx = jl && kl ? hx0 - g.k_mu * (t1 - t2) : hx0;
        }
        {
            const float ev = ex_k - ex0, eu = ezn - ez0;
            float t1 = ev * idz, t2 = eu * idx_;
            if constexpr (PZ) hyz = bhz * hyz + chz * ev * rdz, t1 += hyz;
            if constexpr (PXY) hyx = bhx * hyx + chx * eu * rdx, t2 += hyx;
            nhy = kl && il ? hy0 - g.k_mu * (t1 - t2) : hy0;
        }
        {
            const float ev = eyn - ey0, eu = ex_j - ex0;
            float t1 = ev * idx_, t2 = eu * idy;
            if constexpr (PXY) hzx = bhx * hzx + chx * ev * rdx, t1 += hzx;
            if constexpr (PXY) hzy = bhy * hzy + chy * eu * rdy, t2 += hzy;
            nhz = il && jl ? hz0 - g.k_mu * (t1 - t2) : hz0;
        }
        xr[1][b][w + 1][0][lane] = nhx, xr[1][b][w + 1][1][lane] = nhz;
        xr[0][b][w + 1][0][lane] = exn, xr[0][b][w + 1][1][lane] = ezn;
        __syncthreads();
        // ---- step n, E at plane t (the -1 neighbours of step-n H: j-1 from LDS, k-1 from the
        // previous lane, i-1 from the previous iteration)
        const float hx_j = xr[1][b][w][0][lane], hz_j = xr[1][b][w][1][lane];
        const float hx_k = lane_prev(nhx), hy_k = lane_prev(nhy);
        const float2 f0 = coef(c0), f1 = coef(c1), f2 = coef(c2);
        float nex, ney, nez;
        {
            const float hv = nhz - hz_j, hu = nhy - hy_k;
            float t1 = hv * iddy, t2 = hu * iddz;
            if constexpr (PXY) exy = bey * exy + cey * hv * rddy, t1 += exy;
            if constexpr (PZ) exz = bez * exz + cez * hu * rddz, t2 += exz;
            nex = f0.y != 0.f ? f0.x * ex0 + f0.y * (t1 - t2) : ex0;
        }
        {
            const float hv = nhx - hx_k, hu = nhz - h1z;
            float t1 = hv * iddz, t2 = hu * iddx;
            if constexpr (PZ) eyz = bez * eyz + cez * hv * rddz, t1 += eyz;
            if constexpr (PXY) eyx = bex * eyx + cex * hu * rddx, t2 += eyx;
            ney = f1.y != 0.f ? f1.x * ey0 + f1.y * (t1 - t2) : ey0;
        }
        {
            const float hv = nhy - h1y, hu = nhx - hx_j;
            float t1 = hv * iddx, t2 = hu * iddy;
            if constexpr (PXY) ezx = bex * ezx + cex * hv * rddx, t1 += ezx;
            if constexpr (PXY) ezy = bey * ezy + cey * hu * rddy, t2 += ezy;
            nez = f2.y != 0.f ? f2.x * ez0 + f2.y * (t1 - t2) : ez0;
        }
        // step n's Debye ADE, then its port sources, on this plane (every copy of the edge,
        // owner or halo), in the order of cpu.rs:284-297
        const int drow = SIDE ? xdeb[(t - tstart) * ROWS + w] : -1;  // wave-uniform
        int32_t dt3[3] = {-1, -1, -1};
        if (SIDE && drow >= 0) {
            const int32_t *dm = db.map + (uint64_t(uint32_t(drow)) * g.n2 + uk) * 3;
            for (int c = 0; c < 3; c++) dt3[c] = kin ? dm[c] : -1;
            for (int c = 0; c < 3; c++) {
                const int32_t di = dt3[c];
                if (di < 0) continue;
                const fg_debye_edge d = fx.debye[di];
                const float cb = (c == 0 ? f0 : c == 1 ? f1 : f2).y;
                const float *tab = fx.dtab + size_t(d.set) * fx.np * 2;
                float s0 = 0.f;
                for (uint32_t q = 0; q < fx.np; q++) s0 += 0.5f * (1.0f + tab[2 * q]) * fx.dstate[q * fx.n_debye + di];
                float &e = c == 0 ? nex : c == 1 ? ney : nez;
                const float next = e - cb * s0;
                e = next;
                const float de = next - (c == 0 ? ex0 : c == 1 ? ey0 : ez0);
                if (own && t >= i0 && t < i1)
                    for (uint32_t q = 0; q < fx.np; q++)
                        db.jn[q * fx.n_debye + di] = tab[2 * q] * fx.dstate[q * fx.n_debye + di] + d.dd * tab[2 * q + 1] * de;
            }
        }
        if (t < g.n0) {
            const uint32_t s0 = xsrc[t - tstart], s1 = xsrc[t - tstart + 1];
            for (uint32_t q = s0; q < s1; q++) {
                const uint4 e = sr.ent[q];
                if (int(e.x) == j && int(e.y) == k) {
                    const float d = __uint_as_float(e.w) * pl;
                    if (e.z == 0) nex -= d;
                    else if (e.z == 1) ney -= d;
                    else nez -= d;
                }
            }
        }
        if (SIDE && drow >= 0 && own && t >= i0 && t < i1)
            for (int c = 0; c < 3; c++)
                if (dt3[c] >= 0) db.eo[dt3[c]] = c == 0 ? nex : c == 1 ? ney : nez;
        // step n's final values at the observed points of this row, by their owner
        if (SIDE && cp.start && t >= i0 && t < i1 && w >= 2 && w < 2 + TJ && jin) {  // wave-uniform
            const uint32_t *hp = xcap + (t - i0) * (TJ + 1);
            const uint32_t q0 = hp[w - 2], q1 = hp[w - 1], at0 = xcoff[t - i0] - hp[0];
            const bool mine = own;
            for (uint32_t q = q0; q < q1; q++) {
                const uint4 sg = seg_lds ? xseg[q + at0] : cp.seg[q];
                if (mine && uk >= sg.x && uk < sg.y) {
                    uint32_t at = sg.w + (uk - sg.x) * uint32_t(__popc(sg.z));
                    const float v[6] = {nex, ney, nez, nhx, nhy, nhz};
                    for (int c = 0; c < 6; c++)
                        if ((sg.z >> c) & 1u) cp.buf[at++] = v[c];
                }
            }
        }
        // ---- step n+1 at plane t-1: H from the step-n E of planes t-1 (carried) and t (+x)
        const bool two = t >= 1 && t - 1 + 1 >= (i0 > 0 ? i0 : 1);
        const uint32_t s = t - 1;  // plane of step n+1
        const bool sl = t < g.n0;  // s+1 inside the grid (the i+1 neighbour exists)
        float mhy = 0.f, mhz = 0.f;
        if (two) {
            const float *xs = xa[s - tstart];
            const float sidx = xs[0], siddx = xs[1];
            float sbhx = 0, schx = 0, srdx = 0, sbex = 0, scex = 0, srddx = 0;
            uint32_t soxh = OOB, soxe = OOB;
            uint64_t szrow = 0, syrow = 0, sxhb = 0, sxeb = 0;
            if constexpr (PZ) szrow = uint64_t(s) * g.n1 * p2;
            if constexpr (PXY) {
                const int shx = __float_as_int(xs[8]), sex = __float_as_int(xs[9]);
                sbhx = xs[2], schx = xs[3], srdx = xs[4];
                sbex = xs[5], scex = xs[6], srddx = xs[7];
                syrow = uint64_t(s) * p2 * g.n2;
                sxhb = uint64_t(max(shx, 0)) * plane, sxeb = uint64_t(max(sex, 0)) * plane;
                soxh = shx >= 0 ? ob4 : OOB, soxe = sex >= 0 ? ob4 : OOB;
            }
            const bool wr = own && s >= i0;  // this thread stores plane s
            const uint32_t sto = wr ? stO : OOB;
            const float ex_j2 = xr[2][b ^ 1][w + 2][0][lane], ez_j2 = xr[2][b ^ 1][w + 2][1][lane];
            const float ex_k2 = lane_next(e1x), ey_k2 = lane_next(e1y);
            float ohx, ohy, ohz;
            {
                const float ev = ez_j2 - e1z, eu = ey_k2 - e1y;
                float t1 = ev * idy, t2 = eu * idz;
                const bool v = jl && kl;
                if constexpr (PXY) {
                    q_hxy = bhy * q_hxy + chy * ev * rdy, t1 += q_hxy;
                    st(psin + PSI(1, 0, 1) + syrow, v ? sto == OOB ? OOB : oyh : OOB, ybytes, q_hxy);
                }
                if constexpr (PZ) {
                    q_hxz = bhz * q_hxz + chz * eu * rdz, t2 += q_hxz;
                    st(psin + PSI(1, 0, 2) + szrow, v ? sto == OOB ? OOB : ozh : OOB, zbytes, q_hxz);
                }
                ohx = v ? h1x - g.k_mu * (t1 - t2) : h1x;
            }
            {
                const float ev = ex_k2 - e1x, eu = nez - e1z;
                float t1 = ev * idz, t2 = eu * sidx;
                const bool v = kl && sl;
                if constexpr (PZ) {
                    q_hyz = bhz * q_hyz + chz * ev * rdz, t1 += q_hyz;
                    st(psin + PSI(1, 1, 2) + szrow, v ? sto == OOB ? OOB : ozh : OOB, zbytes, q_hyz);
                }
                if constexpr (PXY) {
                    q_hyx = sbhx * q_hyx + schx * eu * srdx, t2 += q_hyx;
                    st(psin + PSI(1, 1, 0) + sxhb, v ? sto == OOB ? OOB : soxh : OOB, pl4, q_hyx);
                }
                ohy = v ? h1y - g.k_mu * (t1 - t2) : h1y;
            }
            {
                const float ev = ney - e1y, eu = ex_j2 - e1x;
                float t1 = ev * sidx, t2 = eu * idy;
                const bool v = sl && jl;
                if constexpr (PXY) {
                    q_hzx = sbhx * q_hzx + schx * ev * srdx, t1 += q_hzx;
                    st(psin + PSI(1, 2, 0) + sxhb, v ? sto == OOB ? OOB : soxh : OOB, pl4, q_hzx);
                    q_hzy = bhy * q_hzy + chy * eu * rdy, t2 += q_hzy;
                    st(psin + PSI(1, 2, 1) + syrow, v ? sto == OOB ? OOB : oyh : OOB, ybytes, q_hzy);
                }
                ohz = v ? h1z - g.k_mu * (t1 - t2) : h1z;
            }
            {
                float *q = Hn + uint64_t(s) * plane;
                st(q, sto, pl4, ohx), st(q + nn, sto, pl4, ohy), st(q + 2 * nn, sto, pl4, ohz);
            }
            xr[3][b][w + 1][0][lane] = ohx, xr[3][b][w + 1][1][lane] = ohz;
            xr[2][b][w + 1][0][lane] = nex, xr[2][b][w + 1][1][lane] = nez;
            __syncthreads();
            const float hx_j2 = xr[3][b][w][0][lane], hz_j2 = xr[3][b][w][1][lane];
            const float hx_k2 = lane_prev(ohx), hy_k2 = lane_prev(ohy);
            float *q = En + uint64_t(s) * plane;
            {
                const float hv = ohz - hz_j2, hu = ohy - hy_k2;
                float t1 = hv * iddy, t2 = hu * iddz;
                const bool on = cf0.y != 0.f;
                if constexpr (PXY) {
                    q_exy = bey * q_exy + cey * hv * rddy, t1 += q_exy;
                    st(psin + PSI(0, 0, 1) + syrow, on ? sto == OOB ? OOB : oye : OOB, ybytes, q_exy);
                }
                if constexpr (PZ) {
                    q_exz = bez * q_exz + cez * hu * rddz, t2 += q_exz;
                    st(psin + PSI(0, 0, 2) + szrow, on ? sto == OOB ? OOB : oze : OOB, zbytes, q_exz);
                }
                st(q, sto, pl4, on ? cf0.x * e1x + cf0.y * (t1 - t2) : e1x);
            }
            {
                const float hv = ohx - hx_k2, hu = ohz - h2z;
                float t1 = hv * iddz, t2 = hu * siddx;
                const bool on = cf1.y != 0.f;
                if constexpr (PZ) {
                    q_eyz = bez * q_eyz + cez * hv * rddz, t1 += q_eyz;
                    st(psin + PSI(0, 1, 2) + szrow, on ? sto == OOB ? OOB : oze : OOB, zbytes, q_eyz);
                }
                if constexpr (PXY) {
                    q_eyx = sbex * q_eyx + scex * hu * srddx, t2 += q_eyx;
                    st(psin + PSI(0, 1, 0) + sxeb, on ? sto == OOB ? OOB : soxe : OOB, pl4, q_eyx);
                }
                st(q + nn, sto, pl4, on ? cf1.x * e1y + cf1.y * (t1 - t2) : e1y);
            }
            {
                const float hv = ohy - h2y, hu = ohx - hx_j2;
                float t1 = hv * siddx, t2 = hu * iddy;
                const bool on = cf2.y != 0.f;
                if constexpr (PXY) {
                    q_ezx = sbex * q_ezx + scex * hv * srddx, t1 += q_ezx;
                    st(psin + PSI(0, 2, 0) + sxeb, on ? sto == OOB ? OOB : soxe : OOB, pl4, q_ezx);
                    q_ezy = bey * q_ezy + cey * hu * rddy, t2 += q_ezy;
                    st(psin + PSI(0, 2, 1) + syrow, on ? sto == OOB ? OOB : oye : OOB, ybytes, q_ezy);
                }
                st(q + 2 * nn, sto, pl4, on ? cf2.x * e1z + cf2.y * (t1 - t2) : e1z);
            }
            mhy = ohy, mhz = ohz;
        } else {
            xr[2][b][w + 1][0][lane] = nex, xr[2][b][w + 1][1][lane] = nez;
            __syncthreads();
        }
        // rotate: plane t becomes the previous plane
        h2y = mhy, h2z = mhz;
        h1x = nhx, h1y = nhy, h1z = nhz;
        e1x = nex, e1y = ney, e1z = nez;
        cf0 = f0, cf1 = f1, cf2 = f2;
        q_hxy = hxy, q_hxz = hxz, q_hyz = hyz, q_hyx = hyx, q_hzx = hzx, q_hzy = hzy;
        q_exy = exy, q_exz = exz, q_eyz = eyz, q_eyx = eyx, q_ezx = ezx, q_ezy = ezy;
        ex0 = exn, ey0 = eyn, ez0 = ezn;
    }
#undef PSI
}

// march2_ahead computes exactly what march2 computes (bit for bit), with the plane loop built for
// the memory pipeline rather than for register and scalar economy:
//  - the field and coefficient-index planes come through resources built once per block in the
//    interior class without side work (block-wide resources, 60 scalar registers), per plane
//    elsewhere;
//  - the interior classes issue each plane's loads two planes ahead (two register sets, the loop
//    unrolled by two), the CPML classes one;
//  - every load and store is issued unconditionally (out of range where it does not apply), and
//    each plane's loaded values are moved out where the plane begins, so the compiler's waits
//    stay exact and never drain the loads in flight; the second step also runs on the
//    segment's first planes, storing nothing there;
//  - the E updates are selects.
// Measured A/B on one box: it takes the 768x768x384 layered interior launch from 4.656 to 4.612
// ms per two steps, and the batched sweep from 5.59 to 5.43 s; march2 stays ahead where side
// work and short segments dominate (the 256^3 boxes run: 0.333 against 0.349 ms per step; the
// sweep's single runs: 10.23 against 10.79 s). The engine runs march2_ahead for the interior
// class without side work, the batch for every class; march2 runs everything else.
// SIDE: the first step has side work (Debye ADE through db, captures through cp; either may be
// absent, NULL tables)
template <int TJ, int PML, bool WIDE, bool LDS_SETS, bool SIDE>
__device__ __forceinline__ void march2_ahead(March2Lds<TJ, LDS_SETS, SIDE> &lds, const Geo &g, const MarchBlock mb,
                                       const float *__restrict__ Eo,
                                       float *__restrict__ En, const float *__restrict__ Ho, float *__restrict__ Hn,
                                       const float *__restrict__ psio, float *__restrict__ psin, uint32_t seg,
                                       const Src2 &sr, float pl, const FixArgs &fx, const Deb2 &db,
                                       const Cap2 &cp) {
    constexpr int ROWS = TJ + 4;
    constexpr bool PZ = PML >= 1, PXY = PML == 2;
    auto &xr = lds.xr;
    auto &sets = lds.sets;
    auto &xa = lds.xa;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k = int(mb.bx) * OWN_K2 + lane - 2;
    const int j = int(mb.by) * TJ + w - 2;
    const uint32_t i0 = mb.bz * seg, i1 = min(i0 + seg, g.n0);
    const bool kin = k >= 0 && k < int(g.n2), jin = j >= 0 && j < int(g.n1);
    const bool act = kin && jin;
    const bool own = act && lane >= 2 && lane < 2 + OWN_K2 && w >= 2 && w < 2 + TJ;
    const uint32_t uk = kin ? uint32_t(k) : 0u, uj = jin ? uint32_t(j) : 0u;
    const uint64_t nn = g.nn, plane = uint64_t(g.n1) * g.n2;
    const uint32_t off = uj * g.n2 + uk;
    const bool kl = uk + 1 < g.n2, jl = uj + 1 < g.n1;
    const float idz = g.ax[2].inv_d[uk], iddz = g.ax[2].inv_dd[uk];
    const float idy = g.ax[1].inv_d[uj], iddy = g.ax[1].inv_dd[uj];
    const uint32_t ob4 = off * 4u;
    const uint32_t ldA = act ? ob4 : OOB2, stO = own ? ob4 : OOB2;
    const uint32_t ldC = act ? off * (WIDE ? 4u : 2u) : OOB2;
    const uint32_t pl4 = uint32_t(plane) * 4u, plc = uint32_t(plane) * (WIDE ? 4u : 2u);
    float bhz = 0, chz = 0, rdz = 0, bez = 0, cez = 0, rddz = 0, bhy = 0, chy = 0, rdy = 0, bey = 0, cey = 0, rddy = 0;
    uint32_t ozh = OOB2, oze = OOB2, oyh = OOB2, oye = OOB2;
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
            ozh = shz >= 0 ? (uj * p2 + uint32_t(shz)) * 4u : OOB2, oze = sez >= 0 ? (uj * p2 + uint32_t(sez)) * 4u : OOB2;
            oyh = shy >= 0 ? (uint32_t(shy) * g.n2 + uk) * 4u : OOB2, oye = sey >= 0 ? (uint32_t(sey) * g.n2 + uk) * 4u : OOB2;
        }
    }
    const uint32_t *cix = g.coef_index;
    const uint16_t *cix16 = reinterpret_cast<const uint16_t *>(g.coef_index);
    auto coef = [&](uint32_t c) {
        if constexpr (LDS_SETS) return sets[c];
        else return g.coef_sets[c];
    };
    if constexpr (LDS_SETS)
        for (uint32_t t = w * 64 + lane; t < g.n_sets; t += 64 * ROWS) sets[t] = g.coef_sets[t];
    // step n runs on planes tstart..i1, step n+1 on planes max(i0,1)-1 .. i1-1 (the plane
    // before i0 feeds the -x neighbour of i0's E update)
    const uint32_t tstart = i0 >= 2 ? i0 - 2 : 0;
    for (uint32_t t = w * 64 + lane; t + tstart <= i1; t += 64 * ROWS) {
        const uint32_t q = min(tstart + t, g.n0 - 1);
        xa[t][0] = g.ax[0].inv_d[q], xa[t][1] = g.ax[0].inv_dd[q];
        if constexpr (PXY) {
            xa[t][2] = g.ax[0].bh[q], xa[t][3] = g.ax[0].ch[q], xa[t][4] = g.ax[0].raw_d[q];
            xa[t][5] = g.ax[0].be[q], xa[t][6] = g.ax[0].ce[q], xa[t][7] = g.ax[0].raw_dd[q];
            xa[t][8] = __int_as_float(g.ax[0].slot_h[q]), xa[t][9] = __int_as_float(g.ax[0].slot_e[q]);
        }
    }
    if (w < 2)
        for (int x = 0; x < 4; x++)
            for (int q = 0; q < 2; q++) xr[x][q][w * (ROWS + 1)][0][lane] = 0.f, xr[x][q][w * (ROWS + 1)][1][lane] = 0.f;
    // The per-plane entries of the source, Debye-row and capture tables, staged in LDS before the
    // march: the plane loop then reads them from LDS. Read from global memory inside the loop, a
    // table entry's wait (vmcnt counts every vector memory access in issue order) would also
    // wait for the plane loads issued one iteration ahead, once per plane.
    //   xsrc[t - tstart]: sr.start[t] for the planes tstart .. i1 + 1 (clamped to n0);
    //   xdeb[(t - tstart) * ROWS + w]: the Debye row slot of (t, row w), -1 outside the grid;
    //   xcap[(t - i0) * (TJ + 1) + r]: cp.start of owned row r of owned plane t (r = TJ: the end),
    //   xseg: the owned planes' capture segments, plane p's at xcoff[p] on, when they all fit.
    constexpr uint32_t CAP_SEGS = March2Lds<TJ, LDS_SETS, SIDE>::CAP_SEGS;
    auto &xsrc = lds.xsrc;
    auto &xdeb = lds.xdeb;
    auto &xcap = lds.xcap;
    auto &xcoff = lds.xcoff;
    auto &xseg = lds.xseg;
    for (uint32_t t = w * 64 + lane; t + tstart <= i1 + 1; t += 64 * ROWS) xsrc[t] = sr.start[min(tstart + t, g.n0)];
    bool seg_lds = false;
    if constexpr (SIDE) {
        const uint32_t np = i1 - tstart + 1, no = i1 - i0, j0 = mb.by * TJ;
        for (uint32_t x = w * 64 + lane; x < np * ROWS; x += 64 * ROWS) {
            const uint32_t t = tstart + x / ROWS;
            const int jj = int(j0) + int(x % ROWS) - 2;
            xdeb[x] = db.row && t < g.n0 && jj >= 0 && jj < int(g.n1) ? db.row[uint64_t(t) * g.n1 + uint32_t(jj)] : -1;
        }
        if (cp.start)
            for (uint32_t x = w * 64 + lane; x < no * (TJ + 1); x += 64 * ROWS)
                xcap[x] = cp.start[uint64_t(i0 + x / (TJ + 1)) * g.n1 + min(j0 + x % (TJ + 1), g.n1)];
        __syncthreads();
        if (cp.start && w == 0 && lane == 0) {
            uint32_t o = 0;
            for (uint32_t q = 0; q < no; q++) xcoff[q] = o, o += xcap[q * (TJ + 1) + TJ] - xcap[q * (TJ + 1)];
            xcoff[no] = o;
        }
        __syncthreads();
        seg_lds = cp.start && xcoff[no] <= CAP_SEGS;  // block-uniform
        if (seg_lds)
            for (uint32_t x = w * 64 + lane; x < xcoff[no]; x += 64 * ROWS) {
                uint32_t lo = 0, hi = no;  // the owned plane of entry x: xcoff[lo] <= x < xcoff[hi]
                while (hi - lo > 1) {
                    const uint32_t m = (lo + hi) / 2;
                    if (xcoff[m] <= x) lo = m;
                    else hi = m;
                }
                xseg[x] = cp.seg[xcap[lo * (TJ + 1)] + (x - xcoff[lo])];
            }
    }
    // The block's field and coefficient-index resources, one per component array, based at plane
    // tstart and ranged over its planes inside the grid (tstart .. min(i1 + 2, n0) - 1). Plane p
    // sits (p - tstart) plane strides past a lane's in-plane offset, so the plane loop adds a
    // uniform stride instead of rebuilding a resource per plane and array.
    const uint32_t npl = min(i1 + 2, g.n0) - tstart;
    __amdgpu_buffer_rsrc_t rEo[3], rHo[3], rEn[3], rHn[3], rC[3];
    for (int c = 0; c < 3; c++) {
        const uint64_t o = uint64_t(c) * nn + uint64_t(tstart) * plane;
        rEo[c] = rsrc(Eo + o, npl * pl4), rHo[c] = rsrc(Ho + o, npl * pl4);
        rEn[c] = rsrc(En + o, npl * pl4), rHn[c] = rsrc(Hn + o, npl * pl4);
        if constexpr (WIDE) rC[c] = rsrc(cix + o, npl * plc);
        else rC[c] = rsrc(cix16 + o, npl * plc);
    }
    auto cidx = [&](int c, uint32_t off_) {
        if constexpr (WIDE) return ldr32(rC[c], off_);
        else return ldr16(rC[c], off_);
    };
#define PSI(e, c, a) g.psi_off[(e) * 9 + (c) * 3 + (a)]
    // carried state: E0 of the current plane, step-n values of the previous plane, step-n+1 H
    // of the plane before, coefficient pairs and step-n psi of the previous plane
    float ex0 = ldr(rEo[0], ldA), ey0 = ldr(rEo[1], ldA), ez0 = ldr(rEo[2], ldA);
    {
        const int pb = (tstart + 1) & 1;  // the buffer the first iteration reads
        xr[0][pb][w + 1][0][lane] = ex0, xr[0][pb][w + 1][1][lane] = ez0;
    }
    __syncthreads();
    float h1x = 0, h1y = 0, h1z = 0;           // step-n H at the previous plane
    float e1x = 0, e1y = 0, e1z = 0;           // step-n E at the previous plane
    float h2y = 0, h2z = 0;                    // step-n+1 Hy, Hz at the plane before that
    float2 cf0 = {0, 0}, cf1 = {0, 0}, cf2 = {0, 0};  // (ca, cb) of the previous plane
    // step-n psi of the previous plane (component, axis)
    float q_hxy = 0, q_hxz = 0, q_hyz = 0, q_hyx = 0, q_hzx = 0, q_hzy = 0;
    float q_exy = 0, q_exz = 0, q_eyz = 0, q_eyx = 0, q_ezx = 0, q_ezy = 0;
    // the loads of plane t: E0(t+1), H0(t), coefficient indices and step n-1 psi, issued ahead
    // of the plane that uses them (two planes ahead in the interior class, one in the CPML
    // classes), so their latency overlaps whole iterations instead of stalling between the
    // barriers. The vector memory counter counts loads and stores in issue order, so a wait for
    // a plane's loads also waits for the stores issued before them, never for those after.
    struct Pl {
        float exn, eyn, ezn, hx0, hy0, hz0;
        uint32_t c0, c1, c2;
        float hxy, hxz, hyz, hyx, hzx, hzy, exy, exz, eyz, eyx, ezx, ezy;
    };
    auto load_plane = [&](uint32_t t, Pl &P) {
        // plane t's H and indices and plane t+1's E. The plane loop issues these loads whether
        // or not the plane is in the segment (a branch around them would blur the wait counts);
        // a plane past the segment (t > i1) or past the grid reads from past every resource's
        // range and gets 0.
        const bool live = t <= i1;  // uniform
        const uint32_t rel = live ? t - tstart : npl, oh = ldA + rel * pl4, oe = oh + pl4;
        P.exn = ldr(rEo[0], oe), P.eyn = ldr(rEo[1], oe), P.ezn = ldr(rEo[2], oe);
        P.hx0 = ldr(rHo[0], oh), P.hy0 = ldr(rHo[1], oh), P.hz0 = ldr(rHo[2], oh);
        const uint32_t oc = ldC + rel * plc;
        P.c0 = cidx(0, oc), P.c1 = cidx(1, oc), P.c2 = cidx(2, oc);
        P.hxy = P.hxz = P.hyz = P.hyx = P.hzx = P.hzy = 0.f;
        P.exy = P.exz = P.eyz = P.eyx = P.ezx = P.ezy = 0.f;
        const bool in = live && t < g.n0;
        const uint32_t zb = in ? zbytes : 0u, yb = in ? ybytes : 0u, xb = in ? pl4 : 0u;
        if constexpr (PZ) {
            const uint64_t zrow = in ? uint64_t(t) * g.n1 * p2 : 0;
            P.hxz = ld(psio + PSI(1, 0, 2) + zrow, ozh, zb), P.hyz = ld(psio + PSI(1, 1, 2) + zrow, ozh, zb);
            P.exz = ld(psio + PSI(0, 0, 2) + zrow, oze, zb), P.eyz = ld(psio + PSI(0, 1, 2) + zrow, oze, zb);
        }
        if constexpr (PXY) {
            const float *xv = xa[(live ? t : i1) - tstart];  // xa holds the segment's planes only
            const int shx = __float_as_int(xv[8]), sex = __float_as_int(xv[9]);
            const uint64_t yrow = in ? uint64_t(t) * p2 * g.n2 : 0;
            const uint64_t xhb = uint64_t(max(shx, 0)) * plane, xeb = uint64_t(max(sex, 0)) * plane;
            const uint32_t oxh = shx >= 0 ? ldA : OOB2, oxe = sex >= 0 ? ldA : OOB2;
            P.hxy = ld(psio + PSI(1, 0, 1) + yrow, oyh, yb);
            P.hyx = ld(psio + PSI(1, 1, 0) + xhb, oxh, xb);
            P.hzx = ld(psio + PSI(1, 2, 0) + xhb, oxh, xb), P.hzy = ld(psio + PSI(1, 2, 1) + yrow, oyh, yb);
            P.exy = ld(psio + PSI(0, 0, 1) + yrow, oye, yb);
            P.eyx = ld(psio + PSI(0, 1, 0) + xeb, oxe, xb);
            P.ezx = ld(psio + PSI(0, 2, 0) + xeb, oxe, xb), P.ezy = ld(psio + PSI(0, 2, 1) + yrow, oye, yb);
        }
    };
    // one plane of the march: step n at plane t and step n+1 at plane t-1, from plane t's loads
    auto plane_step = [&](const uint32_t t, const Pl &cur) __attribute__((always_inline)) {
        const int b = t & 1;
        const bool il = t + 1 < g.n0;
        const float exn = cur.exn, eyn = cur.eyn, ezn = cur.ezn, hx0 = cur.hx0, hy0 = cur.hy0, hz0 = cur.hz0;
        const uint32_t c0 = cur.c0, c1 = cur.c1, c2 = cur.c2;
        const float *xv = xa[t - tstart];
        const float idx_ = xv[0], iddx = xv[1];
        float bhx = 0, chx = 0, rdx = 0, bex = 0, cex = 0, rddx = 0;
        float hxy = cur.hxy, hxz = cur.hxz, hyz = cur.hyz, hyx = cur.hyx, hzx = cur.hzx, hzy = cur.hzy;
        float exy = cur.exy, exz = cur.exz, eyz = cur.eyz, eyx = cur.eyx, ezx = cur.ezx, ezy = cur.ezy;
        if constexpr (PXY) {
            bhx = xv[2], chx = xv[3], rdx = xv[4];
            bex = xv[5], cex = xv[6], rddx = xv[7];
        }
        // ---- step n, H at plane t (the +1 neighbours of E0: j+1 from LDS, k+1 from the next lane)
        const float ex_j = xr[0][b ^ 1][w + 2][0][lane], ez_j = xr[0][b ^ 1][w + 2][1][lane];
        const float ex_k = lane_next(ex0), ey_k = lane_next(ey0);
        float nhx, nhy, nhz;
        {
            const float ev = ez_j - ez0, eu = ey_k - ey0;
            float t1 = ev * idy, t2 = eu * idz;
            if constexpr (PXY) hxy = bhy * hxy + chy * ev * rdy, t1 += hxy;
            if constexpr (PZ) hxz = bhz * hxz + chz * eu * rdz, t2 += hxz;
            nhx = jl && kl ? hx0 - g.k_mu * (t1 - t2) : hx0;
        }
        {
            const float ev = ex_k - ex0, eu = ezn - ez0;
            float t1 = ev * idz, t2 = eu * idx_;
            if constexpr (PZ) hyz = bhz * hyz + chz * ev * rdz, t1 += hyz;
            if constexpr (PXY) hyx = bhx * hyx + chx * eu * rdx, t2 += hyx;
            nhy = kl && il ? hy0 - g.k_mu * (t1 - t2) : hy0;
        }
        {
            const float ev = eyn - ey0, eu = ex_j - ex0;
            float t1 = ev * idx_, t2 = eu * idy;
            if constexpr (PXY) hzx = bhx * hzx + chx * ev * rdx, t1 += hzx;
            if constexpr (PXY) hzy = bhy * hzy + chy * eu * rdy, t2 += hzy;
            nhz = il && jl ? hz0 - g.k_mu * (t1 - t2) : hz0;
        }
        xr[1][b][w + 1][0][lane] = nhx, xr[1][b][w + 1][1][lane] = nhz;
        xr[0][b][w + 1][0][lane] = exn, xr[0][b][w + 1][1][lane] = ezn;
        __syncthreads();
        // ---- step n, E at plane t (the -1 neighbours of step-n H: j-1 from LDS, k-1 from the
        // previous lane, i-1 from the previous iteration)
        const float hx_j = xr[1][b][w][0][lane], hz_j = xr[1][b][w][1][lane];
        const float hx_k = lane_prev(nhx), hy_k = lane_prev(nhy);
        const float2 f0 = coef(c0), f1 = coef(c1), f2 = coef(c2);
        float nex, ney, nez;
        {
            const float hv = nhz - hz_j, hu = nhy - hy_k;
            float t1 = hv * iddy, t2 = hu * iddz;
            if constexpr (PXY) exy = bey * exy + cey * hv * rddy, t1 += exy;
            if constexpr (PZ) exz = bez * exz + cez * hu * rddz, t2 += exz;
            const float u = f0.x * ex0 + f0.y * (t1 - t2);  // computed for every lane: a select, not a branch
            nex = f0.y != 0.f ? u : ex0;
        }
        {
            const float hv = nhx - hx_k, hu = nhz - h1z;
            float t1 = hv * iddz, t2 = hu * iddx;
            if constexpr (PZ) eyz = bez * eyz + cez * hv * rddz, t1 += eyz;
            if constexpr (PXY) eyx = bex * eyx + cex * hu * rddx, t2 += eyx;
            const float u = f1.x * ey0 + f1.y * (t1 - t2);
            ney = f1.y != 0.f ? u : ey0;
        }
        {
            const float hv = nhy - h1y, hu = nhx - hx_j;
            float t1 = hv * iddx, t2 = hu * iddy;
            if constexpr (PXY) ezx = bex * ezx + cex * hv * rddx, t1 += ezx;
            if constexpr (PXY) ezy = bey * ezy + cey * hu * rddy, t2 += ezy;
            const float u = f2.x * ez0 + f2.y * (t1 - t2);
            nez = f2.y != 0.f ? u : ez0;
        }
        // step n's Debye ADE, then its port sources, on this plane (every copy of the edge,
        // owner or halo), in the order of cpu.rs:284-297
        const int drow = SIDE ? xdeb[(t - tstart) * ROWS + w] : -1;  // wave-uniform
        int32_t dt3[3] = {-1, -1, -1};
        if (SIDE && drow >= 0) {
            const int32_t *dm = db.map + (uint64_t(uint32_t(drow)) * g.n2 + uk) * 3;
            for (int c = 0; c < 3; c++) dt3[c] = kin ? dm[c] : -1;
            for (int c = 0; c < 3; c++) {
                const int32_t di = dt3[c];
                if (di < 0) continue;
                const fg_debye_edge d = fx.debye[di];
                const float cb = (c == 0 ? f0 : c == 1 ? f1 : f2).y;
                const float *tab = fx.dtab + size_t(d.set) * fx.np * 2;
                float s0 = 0.f;
                for (uint32_t q = 0; q < fx.np; q++) s0 += 0.5f * (1.0f + tab[2 * q]) * fx.dstate[q * fx.n_debye + di];
                float &e = c == 0 ? nex : c == 1 ? ney : nez;
                const float next = e - cb * s0;
                e = next;
                const float de = next - (c == 0 ? ex0 : c == 1 ? ey0 : ez0);
                if (own && t >= i0 && t < i1)
                    for (uint32_t q = 0; q < fx.np; q++)
                        db.jn[q * fx.n_debye + di] = tab[2 * q] * fx.dstate[q * fx.n_debye + di] + d.dd * tab[2 * q + 1] * de;
            }
        }
        if (t < g.n0) {
            const uint32_t s0 = xsrc[t - tstart], s1 = xsrc[t - tstart + 1];
            for (uint32_t q = s0; q < s1; q++) {
                const uint4 e = sr.ent[q];
                if (int(e.x) == j && int(e.y) == k) {
                    const float d = __uint_as_float(e.w) * pl;
                    if (e.z == 0) nex -= d;
                    else if (e.z == 1) ney -= d;
                    else nez -= d;
                }
            }
        }
        if (SIDE && drow >= 0 && own && t >= i0 && t < i1)
            for (int c = 0; c < 3; c++)
                if (dt3[c] >= 0) db.eo[dt3[c]] = c == 0 ? nex : c == 1 ? ney : nez;
        // step n's final values at the observed points of this row, by their owner
        if (SIDE && cp.start && t >= i0 && t < i1 && w >= 2 && w < 2 + TJ && jin) {  // wave-uniform
            const uint32_t *hp = xcap + (t - i0) * (TJ + 1);
            const uint32_t q0 = hp[w - 2], q1 = hp[w - 1], at0 = xcoff[t - i0] - hp[0];
            const bool mine = own;
            for (uint32_t q = q0; q < q1; q++) {
                const uint4 sg = seg_lds ? xseg[q + at0] : cp.seg[q];
                if (mine && uk >= sg.x && uk < sg.y) {
                    uint32_t at = sg.w + (uk - sg.x) * uint32_t(__popc(sg.z));
                    const float v[6] = {nex, ney, nez, nhx, nhy, nhz};
                    for (int c = 0; c < 6; c++)
                        if ((sg.z >> c) & 1u) cp.buf[at++] = v[c];
                }
            }
        }
        // ---- step n+1 at plane t-1: H from the step-n E of planes t-1 (carried) and t (+x)
        // It runs on every plane but stores only where `two` holds (a branch around its stores
        // would blur the wait counts of the loads in flight); elsewhere s is a plane the tables
        // hold and the results go nowhere.
        const bool two = t >= 1 && t - 1 + 1 >= (i0 > 0 ? i0 : 1);
        const uint32_t s = two ? t - 1 : t;  // plane of step n+1
        const bool sl = t < g.n0;  // s+1 inside the grid (the i+1 neighbour exists)
        float mhy, mhz;
        {
            const float *xs = xa[s - tstart];
            const float sidx = xs[0], siddx = xs[1];
            float sbhx = 0, schx = 0, srdx = 0, sbex = 0, scex = 0, srddx = 0;
            uint32_t soxh = OOB2, soxe = OOB2;
            uint64_t szrow = 0, syrow = 0, sxhb = 0, sxeb = 0;
            if constexpr (PZ) szrow = uint64_t(s) * g.n1 * p2;
            if constexpr (PXY) {
                const int shx = __float_as_int(xs[8]), sex = __float_as_int(xs[9]);
                sbhx = xs[2], schx = xs[3], srdx = xs[4];
                sbex = xs[5], scex = xs[6], srddx = xs[7];
                syrow = uint64_t(s) * p2 * g.n2;
                sxhb = uint64_t(max(shx, 0)) * plane, sxeb = uint64_t(max(sex, 0)) * plane;
                soxh = shx >= 0 ? ob4 : OOB2, soxe = sex >= 0 ? ob4 : OOB2;
            }
            const bool wr = own && two && s >= i0;  // this thread stores plane s
            const uint32_t sto = wr ? stO : OOB2;
            const uint32_t sfo = sto + (s - tstart) * pl4;  // plane s of the block's field resources
            const float ex_j2 = xr[2][b ^ 1][w + 2][0][lane], ez_j2 = xr[2][b ^ 1][w + 2][1][lane];
            const float ex_k2 = lane_next(e1x), ey_k2 = lane_next(e1y);
            float ohx, ohy, ohz;
            {
                const float ev = ez_j2 - e1z, eu = ey_k2 - e1y;
                float t1 = ev * idy, t2 = eu * idz;
                const bool v = jl && kl;
                if constexpr (PXY) {
                    q_hxy = bhy * q_hxy + chy * ev * rdy, t1 += q_hxy;
                    st(psin + PSI(1, 0, 1) + syrow, v ? sto == OOB2 ? OOB2 : oyh : OOB2, ybytes, q_hxy);
                }
                if constexpr (PZ) {
                    q_hxz = bhz * q_hxz + chz * eu * rdz, t2 += q_hxz;
                    st(psin + PSI(1, 0, 2) + szrow, v ? sto == OOB2 ? OOB2 : ozh : OOB2, zbytes, q_hxz);
                }
                ohx = v ? h1x - g.k_mu * (t1 - t2) : h1x;
            }
            {
                const float ev = ex_k2 - e1x, eu = nez - e1z;
                float t1 = ev * idz, t2 = eu * sidx;
                const bool v = kl && sl;
                if constexpr (PZ) {
                    q_hyz = bhz * q_hyz + chz * ev * rdz, t1 += q_hyz;
                    st(psin + PSI(1, 1, 2) + szrow, v ? sto == OOB2 ? OOB2 : ozh : OOB2, zbytes, q_hyz);
                }
                if constexpr (PXY) {
                    q_hyx = sbhx * q_hyx + schx * eu * srdx, t2 += q_hyx;
                    st(psin + PSI(1, 1, 0) + sxhb, v ? sto == OOB2 ? OOB2 : soxh : OOB2, pl4, q_hyx);
                }
                ohy = v ? h1y - g.k_mu * (t1 - t2) : h1y;
            }
            {
                const float ev = ney - e1y, eu = ex_j2 - e1x;
                float t1 = ev * sidx, t2 = eu * idy;
                const bool v = sl && jl;
                if constexpr (PXY) {
                    q_hzx = sbhx * q_hzx + schx * ev * srdx, t1 += q_hzx;
                    st(psin + PSI(1, 2, 0) + sxhb, v ? sto == OOB2 ? OOB2 : soxh : OOB2, pl4, q_hzx);
                    q_hzy = bhy * q_hzy + chy * eu * rdy, t2 += q_hzy;
                    st(psin + PSI(1, 2, 1) + syrow, v ? sto == OOB2 ? OOB2 : oyh : OOB2, ybytes, q_hzy);
                }
                ohz = v ? h1z - g.k_mu * (t1 - t2) : h1z;
            }
            str(rHn[0], sfo, ohx), str(rHn[1], sfo, ohy), str(rHn[2], sfo, ohz);
            xr[3][b][w + 1][0][lane] = ohx, xr[3][b][w + 1][1][lane] = ohz;
            xr[2][b][w + 1][0][lane] = nex, xr[2][b][w + 1][1][lane] = nez;
            __syncthreads();
            const float hx_j2 = xr[3][b][w][0][lane], hz_j2 = xr[3][b][w][1][lane];
            const float hx_k2 = lane_prev(ohx), hy_k2 = lane_prev(ohy);
            {
                const float hv = ohz - hz_j2, hu = ohy - hy_k2;
                float t1 = hv * iddy, t2 = hu * iddz;
                const bool on = cf0.y != 0.f;
                if constexpr (PXY) {
                    q_exy = bey * q_exy + cey * hv * rddy, t1 += q_exy;
                    st(psin + PSI(0, 0, 1) + syrow, on ? sto == OOB2 ? OOB2 : oye : OOB2, ybytes, q_exy);
                }
                if constexpr (PZ) {
                    q_exz = bez * q_exz + cez * hu * rddz, t2 += q_exz;
                    st(psin + PSI(0, 0, 2) + szrow, on ? sto == OOB2 ? OOB2 : oze : OOB2, zbytes, q_exz);
                }
                const float u = cf0.x * e1x + cf0.y * (t1 - t2);
                str(rEn[0], sfo, on ? u : e1x);
            }
            {
                const float hv = ohx - hx_k2, hu = ohz - h2z;
                float t1 = hv * iddz, t2 = hu * siddx;
                const bool on = cf1.y != 0.f;
                if constexpr (PZ) {
                    q_eyz = bez * q_eyz + cez * hv * rddz, t1 += q_eyz;
                    st(psin + PSI(0, 1, 2) + szrow, on ? sto == OOB2 ? OOB2 : oze : OOB2, zbytes, q_eyz);
                }
                if constexpr (PXY) {
                    q_eyx = sbex * q_eyx + scex * hu * srddx, t2 += q_eyx;
                    st(psin + PSI(0, 1, 0) + sxeb, on ? sto == OOB2 ? OOB2 : soxe : OOB2, pl4, q_eyx);
                }
                const float u = cf1.x * e1y + cf1.y * (t1 - t2);
                str(rEn[1], sfo, on ? u : e1y);
            }
            {
                const float hv = ohy - h2y, hu = ohx - hx_j2;
                float t1 = hv * siddx, t2 = hu * iddy;
                const bool on = cf2.y != 0.f;
                if constexpr (PXY) {
                    q_ezx = sbex * q_ezx + scex * hv * srddx, t1 += q_ezx;
                    st(psin + PSI(0, 2, 0) + sxeb, on ? sto == OOB2 ? OOB2 : soxe : OOB2, pl4, q_ezx);
                    q_ezy = bey * q_ezy + cey * hu * rddy, t2 += q_ezy;
                    st(psin + PSI(0, 2, 1) + syrow, on ? sto == OOB2 ? OOB2 : oye : OOB2, ybytes, q_ezy);
                }
                const float u = cf2.x * e1z + cf2.y * (t1 - t2);
                str(rEn[2], sfo, on ? u : e1z);
            }
            mhy = two ? ohy : 0.f, mhz = two ? ohz : 0.f;
        }
        // rotate: plane t becomes the previous plane
        h2y = mhy, h2z = mhz;
        h1x = nhx, h1y = nhy, h1z = nhz;
        e1x = nex, e1y = ney, e1z = nez;
        cf0 = f0, cf1 = f1, cf2 = f2;
        q_hxy = hxy, q_hxz = hxz, q_hyz = hyz, q_hyx = hyx, q_hzx = hzx, q_hzy = hzy;
        q_exy = exy, q_exz = exz, q_eyz = eyz, q_eyx = eyx, q_ezx = ezx, q_ezy = ezy;
        ex0 = exn, ey0 = eyn, ez0 = ezn;
    };
    // The interior class keeps two register sets of plane loads, each consumed two planes after
    // it was issued. The loop is unrolled by two so that no set is copied into the other (a copy
    // waits for its loads a plane early). The first two planes' loads are waited for before the
    // loop, so the wait counts inside it follow the loop's own order of loads and stores.
    // A set's values are moved out where its plane begins (vcopy: a move the compiler keeps in
    // place), so the wait for its loads stands there and its registers are free for the loads
    // two planes on; left to itself the compiler moves the newest set at the loop's end instead,
    // which waits for loads issued a plane earlier.
    auto settle = [&](const Pl &P) {
        Pl c = P;
        c.exn = vcopy(P.exn), c.eyn = vcopy(P.eyn), c.ezn = vcopy(P.ezn);
        c.hx0 = vcopy(P.hx0), c.hy0 = vcopy(P.hy0), c.hz0 = vcopy(P.hz0);
        c.c0 = vcopy(P.c0), c.c1 = vcopy(P.c1), c.c2 = vcopy(P.c2);
        if constexpr (PZ) c.hxz = vcopy(P.hxz), c.hyz = vcopy(P.hyz), c.exz = vcopy(P.exz), c.eyz = vcopy(P.eyz);
        if constexpr (PXY) {
            c.hxy = vcopy(P.hxy), c.hyx = vcopy(P.hyx), c.hzx = vcopy(P.hzx), c.hzy = vcopy(P.hzy);
            c.exy = vcopy(P.exy), c.eyx = vcopy(P.eyx), c.ezx = vcopy(P.ezx), c.ezy = vcopy(P.ezy);
        }
        return c;
    };
    if constexpr (PML == 0) {
        Pl pa, pb;
        load_plane(tstart, pa);
        load_plane(tstart + 1, pb);  // i1 > tstart
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt at their maxima (no wait)
        uint32_t t = tstart;
        for (; t + 1 <= i1; t += 2) {
            {
                const Pl cur = settle(pa);
                load_plane(t + 2, pa);
                plane_step(t, cur);
            }
            {
                const Pl cur = settle(pb);
                load_plane(t + 3, pb);
                plane_step(t + 1, cur);
            }
        }
        if (t <= i1) plane_step(t, pa);  // an odd plane count's last plane
    } else {
        // the CPML classes carry up to 12 psi values more per plane: a second set would not fit
        // in the registers a 1024-thread block leaves each lane, so they run one plane ahead
        Pl nx;
        load_plane(tstart, nx);
        for (uint32_t t = tstart; t <= i1; t++) {
            const Pl cur = nx;
            load_plane(t + 1, nx);
            plane_step(t, cur);
        }
    }
#undef PSI
}

// The first step of a two-step launch exists only inside it. What reads it (the observers of
// that step, the adjoint's tape) recomputes the few values it needs from the untouched previous
// copy, with the march's arithmetic: H by h_next, E by the same update, then the driven port's
// sources.
template <int C>
__device__ float h_mid(const Geo &g, const float *__restrict__ E, const float *__restrict__ H,
                       const float *__restrict__ psi, uint32_t i, uint32_t j, uint32_t k) {
    return h_next<C>(g, E, H, psi, nullptr, i, j, k);
}

template <int C>
__device__ float e_mid(const Geo &g, const FixArgs &f, float pl, const float *__restrict__ E,
                       const float *__restrict__ H, const float *__restrict__ psi, uint32_t i, uint32_t j, uint32_t k) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const float eold = E[C * g.nn + id];
    const float2 cf = coef_at(g, C * g.nn + id);
    float e = eold;
    if (cf.y != 0.f) {
        uint32_t um[3] = {i, j, k}, vm[3] = {i, j, k};
        um[U] -= 1;
        vm[V] -= 1;
        const float hv = h_mid<V>(g, E, H, psi, i, j, k) - h_mid<V>(g, E, H, psi, um[0], um[1], um[2]);
        const float hu = h_mid<U>(g, E, H, psi, i, j, k) - h_mid<U>(g, E, H, psi, vm[0], vm[1], vm[2]);
        const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
        float t1 = hv * g.ax[U].inv_dd[pu], t2 = hu * g.ax[V].inv_dd[pv];
        if (g.pml) {
            const int su = g.ax[U].slot_e[pu], sv = g.ax[V].slot_e[pv];
            if (su >= 0)
                t1 += g.ax[U].be[pu] * psi[psi_at<U>(g, 0, C, i, j, k, uint32_t(su))] + g.ax[U].ce[pu] * hv * g.ax[U].raw_dd[pu];
            if (sv >= 0)
                t2 += g.ax[V].be[pv] * psi[psi_at<V>(g, 0, C, i, j, k, uint32_t(sv))] + g.ax[V].ce[pv] * hu * g.ax[V].raw_dd[pv];
        }
        e = cf.x * eold + cf.y * (t1 - t2);
    }
    for (uint32_t q = 0; q < f.n_src; q++)
        if (f.src[q].comp == C && f.src[q].id == id) e -= f.src[q].src * pl;
    return e;
}

__device__ inline float field_mid(const Geo &g, const FixArgs &f, float pl, const float *__restrict__ E,
                           const float *__restrict__ H, const float *__restrict__ psi, uint64_t flat) {
    const uint32_t c = uint32_t(flat / g.nn);
    const uint64_t id = flat - c * g.nn;
    const uint32_t k = uint32_t(id % g.n2), j = uint32_t((id / g.n2) % g.n1), i = uint32_t(id / (uint64_t(g.n1) * g.n2));
    switch (c) {
        case 0: return e_mid<0>(g, f, pl, E, H, psi, i, j, k);
        case 1: return e_mid<1>(g, f, pl, E, H, psi, i, j, k);
        case 2: return e_mid<2>(g, f, pl, E, H, psi, i, j, k);
        case 3: return h_mid<0>(g, E, H, psi, i, j, k);
        case 4: return h_mid<1>(g, E, H, psi, i, j, k);
        default: return h_mid<2>(g, E, H, psi, i, j, k);
    }
}

// The first step of a two-step launch exists only inside it. Its observers (port V/I, point
// probes) recompute the few values they read from the untouched previous copy (h_mid, e_mid,
// field_mid above), block `blk` of the observer launch as in observe().
__device__ inline void observe_mid(const Geo &g, const ObsArgs &a, const FixArgs &f, const float *__restrict__ E,
                                   const float *__restrict__ H, const float *__restrict__ psi, uint32_t s_off,
                                   uint32_t blk, const Cap2 &cp, uint32_t tid) {
    const uint32_t n = a.step[0] + s_off;
    const float pl = n < f.n_pulse ? f.pulse[n] : pulse_at(f, (float(n) + 0.5f) * f.dt);
    const uint64_t nn = g.nn;
    if (blk < a.blk_ports) {  // as observe(): one wave per port, f32 products summed in f64
        const uint32_t p = blk * 4 + tid / 64, lane = tid % 64;
        if (p >= a.n_ports) return;
        const fg_port d = a.ports[p];
        double v = 0.0, cur = 0.0;
        for (uint32_t q = lane; q < d.n_edges; q += 64) {
            const fg_port_edge pe = a.edges[d.first_edge + q];
            v -= double(field_mid(g, f, pl, E, H, psi, d.comp * nn + pe.id) * pe.w);
        }
        for (uint32_t q = lane; q < d.n_loops; q += 64) {
            const fg_loop_term lt = a.loops[d.first_loop + q];
            cur += double(field_mid(g, f, pl, E, H, psi, (3 + lt.comp) * nn + lt.id) * lt.w);
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
        const uint32_t q = blk * 256 + tid;
        if (q < a.n_probes && n < a.cap) a.probes[uint64_t(n) * a.n_probes + q] = field_mid(g, f, pl, E, H, psi, a.probe_f[q]);
        return;
    }
    // plane DFT and NTFF from the values the pass captured (Cap2)
    blk -= a.blk_probes;
    if (blk < a.blk_plane) {
        const uint32_t q = blk * 256 + tid;
        if (q >= a.plane_n) return;
        const uint32_t *ix = cp.plane + uint64_t(q) * 5;
        const float fv[5] = {cp.buf[ix[0]], cp.buf[ix[1]], cp.buf[ix[2]], cp.buf[ix[3]], cp.buf[ix[4]]};
        plane_acc(a, q, n, fv);
        return;
    }
    blk -= a.blk_plane;
    if (blk < a.blk_ntff) {
        const uint64_t gi = uint64_t(blk) * 256 + tid;
        if (gi >= a.n_patches) return;
        const uint32_t *ix = cp.ntff + gi * 12;
        const float e[4] = {cp.buf[ix[0]], cp.buf[ix[1]], cp.buf[ix[2]], cp.buf[ix[3]]};
        const float h[8] = {cp.buf[ix[4]], cp.buf[ix[5]], cp.buf[ix[6]], cp.buf[ix[7]],
                            cp.buf[ix[8]], cp.buf[ix[9]], cp.buf[ix[10]], cp.buf[ix[11]]};
        ntff_acc(a, gi, n, a.patch[gi * 13], e, h);
    }
}

// observe_mid's port V/I and probes from the values the pass captured at those points (Cap2::obs)
// instead of recomputing them: the same values (the capture is step n's final field, as the
// recomputation gives it) summed in the same order, so the series are bit-identical.
__device__ inline void observe_mid_cap(const ObsArgs &a, uint32_t s_off, uint32_t blk, const Cap2 &cp, uint32_t tid) {
    const uint32_t n = a.step[0] + s_off;
    if (blk < a.blk_ports) {
        const uint32_t p = blk * 4 + tid / 64, lane = tid % 64;
        if (p >= a.n_ports) return;
        const fg_port d = a.ports[p];
        double v = 0.0, cur = 0.0;
        for (uint32_t q = lane; q < d.n_edges; q += 64) {
            const fg_port_edge pe = a.edges[d.first_edge + q];
            v -= double(cp.buf[cp.obs[d.first_edge + q]] * pe.w);
        }
        for (uint32_t q = lane; q < d.n_loops; q += 64) {
            const fg_loop_term lt = a.loops[d.first_loop + q];
            cur += double(cp.buf[cp.obs[cp.obs_loop0 + d.first_loop + q]] * lt.w);
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
        const uint32_t q = blk * 256 + tid;
        if (q < a.n_probes && n < a.cap) a.probes[uint64_t(n) * a.n_probes + q] = cp.buf[cp.obs[cp.obs_probe0 + q]];
    }
}

// Both steps' observers of a two-step pass in one launch: port V/I and probes as observe_mid
// (step n, from the input copy) followed by observe (step n+1, from the output copy); the plane
// DFT and NTFF accumulators read once and written once, with step n's term added before step
// n+1's, the order the two separate observer launches add them in, so every accumulator is
// bit-identical to theirs and its traffic is halved.
__device__ __forceinline__ void plane_acc2(const ObsArgs &a, uint32_t q, uint32_t n, const float f0[5],
                                           const float f1[5]) {
    const float dt = a.dt;
    float *out = a.plane + uint64_t(q) * a.nf * 10;
    for (uint32_t kf = 0; kf < a.nf; kf++) {
        const float w = a.omega[kf];
        float ce[2][2], ch[2][2];
        for (int s = 0; s < 2; s++) {
            const uint32_t m = n + uint32_t(s);
            const float te = (float(m) + 1.0f) * dt, th = (float(m) + 0.5f) * dt;
            const float *ph = a.phase + (uint64_t(m) * a.nf + kf) * 4;
            const bool tab = m < a.n_phase;
            ce[s][0] = tab ? ph[0] : cosf(w * te) * dt, ce[s][1] = tab ? ph[1] : -sinf(w * te) * dt;
            ch[s][0] = tab ? ph[2] : cosf(w * th) * dt, ch[s][1] = tab ? ph[3] : -sinf(w * th) * dt;
        }
        for (int c = 0; c < 5; c++)
            for (int r = 0; r < 2; r++) {
                float x = out[kf * 10 + 2 * c + r];
                x += f0[c] * (c >= 3 ? ch[0][r] : ce[0][r]);
                x += f1[c] * (c >= 3 ? ch[1][r] : ce[1][r]);
                out[kf * 10 + 2 * c + r] = x;
            }
    }
}

// J = n x H and M = E x n of one patch (ntff_acc's terms)
__device__ __forceinline__ void ntff_terms(uint32_t meta, const float e[4], const float h[8], float jv[3], float mv[3]) {
    const uint32_t ax = meta & 3u;
    const float side = (meta & 4u) ? 1.f : -1.f;
    const uint32_t u = (ax + 1) % 3, v = (ax + 2) % 3;
    float ef[3] = {0.f, 0.f, 0.f}, hf[3] = {0.f, 0.f, 0.f};
    ef[u] = 0.5f * (e[0] + e[1]);
    ef[v] = 0.5f * (e[2] + e[3]);
    hf[u] = 0.25f * (h[0] + h[1] + h[2] + h[3]);
    hf[v] = 0.25f * (h[4] + h[5] + h[6] + h[7]);
    float nrm[3] = {0.f, 0.f, 0.f};
    nrm[ax] = side;
    jv[0] = nrm[1] * hf[2] - nrm[2] * hf[1], jv[1] = nrm[2] * hf[0] - nrm[0] * hf[2], jv[2] = nrm[0] * hf[1] - nrm[1] * hf[0];
    mv[0] = ef[1] * nrm[2] - ef[2] * nrm[1], mv[1] = ef[2] * nrm[0] - ef[0] * nrm[2], mv[2] = ef[0] * nrm[1] - ef[1] * nrm[0];
}

__device__ __forceinline__ void ntff_acc2(const ObsArgs &a, uint64_t gi, uint32_t n, uint32_t meta, const float e0[4],
                                          const float h0[8], const float e1[4], const float h1[8]) {
    const float dt = a.dt;
    float jv[2][3], mv[2][3];
    ntff_terms(meta, e0, h0, jv[0], mv[0]);
    ntff_terms(meta, e1, h1, jv[1], mv[1]);
    float *out = a.ntff + gi * a.nf * 12;
    for (uint32_t kf = 0; kf < a.nf; kf++) {
        const float wf = a.omega[kf];
        float pe[2][2], ph[2][2];
        for (int s = 0; s < 2; s++) {
            const uint32_t m = n + uint32_t(s);
            const float te = (float(m) + 1.0f) * dt, th = (float(m) + 0.5f) * dt;
            const float *p = a.phase + (uint64_t(m) * a.nf + kf) * 4;
            const bool tab = m < a.n_phase;
            pe[s][0] = tab ? p[0] : cosf(-wf * te) * dt, pe[s][1] = tab ? p[1] : sinf(-wf * te) * dt;
            ph[s][0] = tab ? p[2] : cosf(-wf * th) * dt, ph[s][1] = tab ? p[3] : sinf(-wf * th) * dt;
        }
        float *o = out + kf * 12;
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 2; r++) {
                float x = o[2 * c + r];
                x += ph[0][r] * jv[0][c];
                x += ph[1][r] * jv[1][c];
                o[2 * c + r] = x;
                float y = o[6 + 2 * c + r];
                y += pe[0][r] * mv[0][c];
                y += pe[1][r] * mv[1][c];
                o[6 + 2 * c + r] = y;
            }
    }
}

// The plane-DFT and NTFF blocks of observe_pair (blk counted from the first plane block).
__device__ __forceinline__ void observe_pair_fields(const Geo &g, const ObsArgs &a, const float *__restrict__ En,
                                                    const float *__restrict__ Hn, uint32_t s_off, uint32_t blk,
                                                    const Cap2 &cp, uint32_t tid);

__device__ inline void observe_pair(const Geo &g, const ObsArgs &a, const FixArgs &f, const float *__restrict__ Eo,
                                    const float *__restrict__ Ho, const float *__restrict__ psio,
                                    const float *__restrict__ En, const float *__restrict__ Hn, uint32_t s_off,
                                    uint32_t blk, const Cap2 &cp, uint32_t tid) {
    if (blk < a.blk_ports + a.blk_probes) {
        observe_mid(g, a, f, Eo, Ho, psio, s_off, blk, cp, tid);
        observe(g, a, En, Hn, s_off + 1, blk, tid);
        return;
    }
    observe_pair_fields(g, a, En, Hn, s_off, blk - (a.blk_ports + a.blk_probes), cp, tid);
}

// Both steps' port V/I and probes of a two-step pass with no plane DFT or NTFF, step n's from
// the pass's captures (observe_mid_cap): no recomputation and no field-DFT code, so a kernel
// that runs it next to the march keeps the march's registers.
__device__ inline void observe_pair_ports_cap(const Geo &g, const ObsArgs &a, const float *__restrict__ En,
                                              const float *__restrict__ Hn, uint32_t s_off, uint32_t blk,
                                              const Cap2 &cp, uint32_t tid) {
    if (blk >= a.blk_ports + a.blk_probes) return;
    observe_mid_cap(a, s_off, blk, cp, tid);
    observe_ports(g, a, En, Hn, s_off + 1, blk, tid);
}

__device__ __forceinline__ void observe_pair_fields(const Geo &g, const ObsArgs &a, const float *__restrict__ En,
                                                    const float *__restrict__ Hn, uint32_t s_off, uint32_t blk,
                                                    const Cap2 &cp, uint32_t tid) {
    const uint32_t n = a.step[0] + s_off;
    const uint64_t nn = g.nn;
    if (blk < a.blk_plane) {
        const uint32_t q = blk * 256 + tid;
        if (q >= a.plane_n) return;
        const uint32_t *ix = cp.plane + uint64_t(q) * 5;
        const float f0[5] = {cp.buf[ix[0]], cp.buf[ix[1]], cp.buf[ix[2]], cp.buf[ix[3]], cp.buf[ix[4]]};
        const uint32_t i = q / g.n1, j = q % g.n1;
        const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + a.plane_k;
        const float f1[5] = {En[id], En[nn + id], En[2 * nn + id], Hn[id], Hn[nn + id]};
        plane_acc2(a, q, n, f0, f1);
        return;
    }
    blk -= a.blk_plane;
    if (blk < a.blk_ntff) {
        const uint64_t gi = uint64_t(blk) * 256 + tid;
        if (gi >= a.n_patches) return;
        const uint32_t *cx = cp.ntff + gi * 12;
        const float e0[4] = {cp.buf[cx[0]], cp.buf[cx[1]], cp.buf[cx[2]], cp.buf[cx[3]]};
        const float h0[8] = {cp.buf[cx[4]], cp.buf[cx[5]], cp.buf[cx[6]], cp.buf[cx[7]],
                             cp.buf[cx[8]], cp.buf[cx[9]], cp.buf[cx[10]], cp.buf[cx[11]]};
        const uint32_t *ix = a.patch + gi * 13;
        const uint32_t meta = ix[0], ax = meta & 3u, u = (ax + 1) % 3, v = (ax + 2) % 3;
        const float e1[4] = {En[u * nn + ix[1]], En[u * nn + ix[2]], En[v * nn + ix[3]], En[v * nn + ix[4]]};
        const float h1[8] = {Hn[u * nn + ix[5]], Hn[u * nn + ix[6]], Hn[u * nn + ix[7]], Hn[u * nn + ix[8]],
                             Hn[v * nn + ix[9]], Hn[v * nn + ix[10]], Hn[v * nn + ix[11]], Hn[v * nn + ix[12]]};
        ntff_acc2(a, gi, n, meta, e0, h0, e1, h1);
    }
}

}  // namespace fg
