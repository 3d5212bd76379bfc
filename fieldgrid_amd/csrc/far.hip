// Far field behind include/fieldgrid_far.h: ntff.rs:66-137 on the device.
//
// One block of 256 threads takes DIRS directions; its threads stride over the patches, keeping
// the per-direction sums N = sum J a e^{i k d.p} and L = sum M a e^{i k d.p} (three complex
// components each) in f64 registers, so every patch (centre, area, J, M: 16 doubles, structure
// of arrays) is read once per DIRS directions. The sums are reduced over the wave with
// shuffles and over the block through LDS; thread 0 finishes ntff.rs:107-114. It is compute-
// bound on the f64 sincos and complex products, not on HBM: the patch arrays (128 B per
// patch) sit in L2 / MALL across the blocks.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/fieldgrid_far.h"
#include "host.h"

namespace fg {

constexpr double PI_D = 3.14159265358979323846;
constexpr double C0_D = 299792458.0;      // sim.rs:1
constexpr double ETA0_D = 376.730313668;  // sim.rs:4
constexpr int FAR_DIRS = 4;

struct FarDev {
    uint64_t n;
    const double *px, *py, *pz, *area;
    const double2 *j[3], *m[3];
    double k;
};

template <int D>
__global__ __launch_bounds__(256) void k_far_field(FarDev f, const double *__restrict__ dirs, uint64_t nd,
                                                   double *__restrict__ out, double *__restrict__ inten) {
    __shared__ double red[4][FAR_DIRS * 12];
    const uint64_t d0 = uint64_t(blockIdx.x) * D;
    double dx[D], dy[D], dz[D];
    for (int a = 0; a < D; a++) {
        const uint64_t d = d0 + a < nd ? d0 + a : nd - 1;
        dx[a] = dirs[3 * d], dy[a] = dirs[3 * d + 1], dz[a] = dirs[3 * d + 2];
    }
    double acc[D][12];
    for (int a = 0; a < D; a++)
        for (int c = 0; c < 12; c++) acc[a][c] = 0.0;
    for (uint64_t q = threadIdx.x; q < f.n; q += 256) {
        const double x = f.px[q], y = f.py[q], z = f.pz[q], ar = f.area[q];
        const double2 j0 = f.j[0][q], j1 = f.j[1][q], j2 = f.j[2][q], m0 = f.m[0][q], m1 = f.m[1][q], m2 = f.m[2][q];
        for (int a = 0; a < D; a++) {
            const double arg = f.k * (dx[a] * x + dy[a] * y + dz[a] * z);
            double s, c;
            sincos(arg, &s, &c);
            const double pr = ar * c, pi = ar * s;
            // mul(J, ph) and mul(M, ph), ntff.rs:62-64
            acc[a][0] += j0.x * pr - j0.y * pi, acc[a][1] += j0.x * pi + j0.y * pr;
            acc[a][2] += j1.x * pr - j1.y * pi, acc[a][3] += j1.x * pi + j1.y * pr;
            acc[a][4] += j2.x * pr - j2.y * pi, acc[a][5] += j2.x * pi + j2.y * pr;
            acc[a][6] += m0.x * pr - m0.y * pi, acc[a][7] += m0.x * pi + m0.y * pr;
            acc[a][8] += m1.x * pr - m1.y * pi, acc[a][9] += m1.x * pi + m1.y * pr;
            acc[a][10] += m2.x * pr - m2.y * pi, acc[a][11] += m2.x * pi + m2.y * pr;
        }
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int a = 0; a < D; a++)
        for (int c = 0; c < 12; c++) {
            double v = acc[a][c];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[wave][a * 12 + c] = v;
        }
    __syncthreads();
    if (threadIdx.x >= D) return;
    const int a = threadIdx.x;
    const uint64_t d = d0 + a;
    if (d >= nd) return;
    double s[12];
    for (int c = 0; c < 12; c++) s[c] = ((red[0][a * 12 + c] + red[1][a * 12 + c]) + red[2][a * 12 + c]) + red[3][a * 12 + c];
    const double dir[3] = {dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]};  // (no runtime index into registers)
    // ntff.rs:107-114: E = (d x L - (N - (N.d) d) eta0) k / (4 pi)
    const double rnr = s[0] * dir[0] + s[2] * dir[1] + s[4] * dir[2], rni = s[1] * dir[0] + s[3] * dir[1] + s[5] * dir[2];
    double pr[3], pim[3];
    for (int c = 0; c < 3; c++) pr[c] = s[2 * c] - rnr * dir[c], pim[c] = s[2 * c + 1] - rni * dir[c];
    const double *L = s + 6;
    const double rlr[3] = {dir[1] * L[4] - dir[2] * L[2], dir[2] * L[0] - dir[0] * L[4], dir[0] * L[2] - dir[1] * L[0]};
    const double rli[3] = {dir[1] * L[5] - dir[2] * L[3], dir[2] * L[1] - dir[0] * L[5], dir[0] * L[3] - dir[1] * L[1]};
    const double scale = f.k / (4.0 * PI_D);
    double u = 0.0;
    for (int c = 0; c < 3; c++) {
        const double er = (rlr[c] - pr[c] * ETA0_D) * scale, ei = (rli[c] - pim[c] * ETA0_D) * scale;
        if (out) out[6 * d + 2 * c] = er, out[6 * d + 2 * c + 1] = ei;
        u += er * er + ei * ei;
    }
    if (inten) inten[d] = u / (2.0 * ETA0_D);
}

}  // namespace fg

struct fg_far {
    fg_device *dev = nullptr;
    hipStream_t stream = nullptr;
    fg::DevBuf<double> px, py, pz, area;
    fg::DevBuf<double2> j[3], m[3];
    uint64_t n = 0;
    double freq = 0.0;
    ~fg_far() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace fg {

static FarDev far_dev(const fg_far &F) {
    FarDev f{};
    f.n = F.n;
    f.px = F.px.p, f.py = F.py.p, f.pz = F.pz.p, f.area = F.area.p;
    for (int c = 0; c < 3; c++) f.j[c] = F.j[c].p, f.m[c] = F.m[c].p;
    f.k = 2.0 * PI_D * F.freq / C0_D;
    return f;
}

// field and/or intensity of n directions
static void far_eval(fg_far &F, uint64_t n, const double *dirs, double *field, double *inten) {
    if (!n) return;
    if (!dirs) throw HipError{"fg_far: directions are NULL", FG_INVALID};
    hipStream_t st = F.stream;
    DevBuf<double> dd, dout, dint;
    dd.upload(dirs, 3 * n, st);
    if (field) dout.alloc(6 * n);
    if (inten) dint.alloc(n);
    const uint64_t blocks = (n + FAR_DIRS - 1) / FAR_DIRS;
    if (blocks > 0x7fffffffull) throw HipError{"fg_far: too many directions", FG_INVALID};
    hipLaunchKernelGGL(k_far_field<FAR_DIRS>, dim3(uint32_t(blocks)), dim3(256), 0, st, far_dev(F), dd.p, n,
                       field ? dout.p : nullptr, inten ? dint.p : nullptr);
    FG_CHECK(hipGetLastError());
    if (field) FG_CHECK(hipMemcpyAsync(field, dout.p, 6 * n * 8, hipMemcpyDeviceToHost, st));
    if (inten) FG_CHECK(hipMemcpyAsync(inten, dint.p, n * 8, hipMemcpyDeviceToHost, st));
    FG_CHECK(hipStreamSynchronize(st));
}

}  // namespace fg

extern "C" int fg_far_new(fg_device *dev, uint64_t count, const double *pos, const double *area, const float *acc,
                          uint32_t nf, uint32_t k, double freq, double scale_re, double scale_im, fg_far **out) {
    using namespace fg;
    return guarded([&] {
        if (!dev || !out || (count && (!pos || !area || !acc))) throw HipError{"fg_far_new: NULL argument", FG_INVALID};
        if (k >= nf) throw HipError{"fg_far_new: frequency slot out of range", FG_INVALID};
        auto F = std::make_unique<fg_far>();
        F->dev = dev;
        F->n = count;
        F->freq = freq;
        FG_CHECK(hipSetDevice(dev->ordinal));
        FG_CHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
        // Far::load (ntff.rs:76-88): J, M of slot k times scale, as f64 complex
        std::vector<double> x(count), y(count), z(count), a(count);
        std::vector<double2> jm[6];
        for (auto &v : jm) v.resize(count);
        for (uint64_t q = 0; q < count; q++) {
            x[q] = pos[3 * q], y[q] = pos[3 * q + 1], z[q] = pos[3 * q + 2], a[q] = area[q];
            const uint64_t o = (q * nf + k) * 12;
            for (int c = 0; c < 6; c++) {
                const double re = double(acc[o + 2 * c]), im = double(acc[o + 2 * c + 1]);
                jm[c][q] = make_double2(re * scale_re - im * scale_im, re * scale_im + im * scale_re);
            }
        }
        hipStream_t st = F->stream;
        F->px.upload(x.data(), count, st), F->py.upload(y.data(), count, st), F->pz.upload(z.data(), count, st);
        F->area.upload(a.data(), count, st);
        for (int c = 0; c < 3; c++) F->j[c].upload(jm[c].data(), count, st), F->m[c].upload(jm[3 + c].data(), count, st);
        FG_CHECK(hipStreamSynchronize(st));
        *out = F.release();
        return int(FG_OK);
    });
}

extern "C" int fg_far_field(fg_far *far, uint64_t n, const double *dirs, double *out) {
    using namespace fg;
    return guarded([&] {
        if (!far || (n && !out)) throw HipError{"fg_far_field: NULL argument", FG_INVALID};
        FG_CHECK(hipSetDevice(far->dev->ordinal));
        far_eval(*far, n, dirs, out, nullptr);
        return int(FG_OK);
    });
}

extern "C" int fg_far_intensity(fg_far *far, uint64_t n, const double *dirs, double *out) {
    using namespace fg;
    return guarded([&] {
        if (!far || (n && !out)) throw HipError{"fg_far_intensity: NULL argument", FG_INVALID};
        FG_CHECK(hipSetDevice(far->dev->ordinal));
        far_eval(*far, n, dirs, nullptr, out);
        return int(FG_OK);
    });
}

extern "C" int fg_far_radiated(fg_far *far, uint32_t steps, double *total, double *peak) {
    using namespace fg;
    return guarded([&] {
        if (!far || !total || !peak) throw HipError{"fg_far_radiated: NULL argument", FG_INVALID};
        FG_CHECK(hipSetDevice(far->dev->ordinal));
        // ntff.rs:121-136: theta cell centres, phi from 0, the sums in the reference's order
        const uint64_t n = uint64_t(steps) * 2 * steps;
        std::vector<double> dirs(3 * n), u(n);
        for (uint32_t a = 0; a < steps; a++) {
            const double th = PI_D * (a + 0.5) / steps;
            for (uint32_t b = 0; b < 2 * steps; b++) {
                const double ph = PI_D * b / steps;
                double *d = &dirs[3 * (uint64_t(a) * 2 * steps + b)];
                d[0] = std::sin(th) * std::cos(ph), d[1] = std::sin(th) * std::sin(ph), d[2] = std::cos(th);
            }
        }
        far_eval(*far, n, dirs.data(), nullptr, u.data());
        double t = 0.0, p = 0.0;
        for (uint32_t a = 0; a < steps; a++) {
            const double th = PI_D * (a + 0.5) / steps;
            for (uint32_t b = 0; b < 2 * steps; b++) {
                const double v = u[uint64_t(a) * 2 * steps + b];
                p = std::max(p, v);
                t += v * std::sin(th) * (PI_D / steps) * (PI_D / steps);
            }
        }
        *total = t, *peak = p;
        return int(FG_OK);
    });
}

extern "C" void fg_far_free(fg_far *far) { delete far; }
