// psi_capi.hip -- context management, PCG drivers and the C ABI declared in
// include/psisolve.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/psisolve.h"
#include "psi_device.h"

// The kernels are compiled in this translation unit.
#include "psi_kernels.hip"
#include "psi_poly.hip"
#include "psi_levels.hip"
#include "psi_mg.hip"

using namespace psi;

struct psi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Geom g{};
    size_t plane = 0;
    std::vector<double> h_dxt, h_dxu, h_dyt, h_dyu, h_cost, h_cosu;
    double *dxt = nullptr, *dxu = nullptr, *dyt = nullptr, *dyu = nullptr, *cost = nullptr,
           *cosu = nullptr;
    const double *hur = nullptr, *hvr = nullptr, *maskU = nullptr, *maskV = nullptr;
    // stored operator
    double *E = nullptr, *N = nullptr, *rx = nullptr, *ry = nullptr, *cC = nullptr;
    float *E32 = nullptr, *N32 = nullptr, *rx32 = nullptr, *ry32 = nullptr;
    uint8_t *act = nullptr, *dirichlet = nullptr;
    // islands
    int nisle = 0;
    std::vector<int32_t> offs;
    int32_t *bi = nullptr, *bj = nullptr, *di = nullptr, *dj = nullptr;
    // CG scratch
    double *u64[2] = {nullptr, nullptr}, *p64[2] = {nullptr, nullptr};
    float *u32[2] = {nullptr, nullptr}, *p32[2] = {nullptr, nullptr}, *e32 = nullptr;
    float* rp32[2] = {nullptr, nullptr};   // residuals of the polynomial-preconditioned path
    int Gpoly[PMAXK + 1] = {0};            // resident grid of k_poly_pcg<K> (0: not yet queried)
    double *r64 = nullptr, *tmp64 = nullptr, *tmp64b = nullptr;
    // 3-D parts of solve_streamfunction: fpx/fpy of the whole routine and the
    // layer thicknesses (dzt, dzw) as last uploaded
    double *fpx3 = nullptr, *fpy3 = nullptr, *dz = nullptr;
    std::vector<double> h_dz;
    int vl_nz = 0;
    VelocityLaunch vl;
    MgState* mg = nullptr;           // multigrid hierarchy (psi_mg_setup)
    float *r32 = nullptr, *z32 = nullptr, *q32 = nullptr, *pf32 = nullptr;   // flexible CG
    double* fcg = nullptr;        // flexible-CG device scalars
    int fcg_restart = 1;
    int fcg_mg_ready = 0;         // the multigrid scalars of psi_fcg_step_mg are set
    double *part = nullptr, *result = nullptr;
    double* part2 = nullptr;      // solver partials, 2 x NRED x G
    int G = 0, G32 = 0;           // resident workgroups of the FP64 / FP32 solver kernels
    int march = 0;                // wave-column sweep shape (SEG * 10 + D), 0: LDS-tile sweep
    int Gmarch = 0;               // its grid
    int Gpart = 0;                // partials per slot the part2 halves hold
    int ncu = 1;                  // compute units of the device
    unsigned* counter = nullptr;
    CgState* st = nullptr;        // 2 states (double-buffered by launch)
    CgState* h_st = nullptr;      // pinned: [0..1] poll slots, [2] init
    double* h_res = nullptr;      // pinned scalars
    hipEvent_t ev[2] = {nullptr, nullptr}, t0 = nullptr, t1 = nullptr;
    hipEvent_t ts0 = nullptr, ts1[2] = {nullptr, nullptr};   // sweep-span timing
    int64_t bytes = 0;
    std::vector<void*> allocs;
    std::string err;
};


// ---------------------------------------------------------------------------
// Flexible CG with a learned preconditioner (no reference counterpart).
// FP32 vectors on the solver grid, FP64 atomically accumulated inner
// products with the area weights W = dxu cosu dyu (S = W A is symmetric), FP64
// residual replacement in psi_fcg_flush.  Device scalars (c->fcg):
// 0 (z, W q_old)  1 (p, W q)  2 (p, W q)_old  3 (r, W p)  4 |r|^2 active
// 5 input scale  6 |Z r|^2 active  7 active count  8 scale (r32 = r64*scale)
// 9 |forc|^2 active.
// ---------------------------------------------------------------------------
namespace {

// The scalars are accumulated with device-scope atomics, which complete
// beyond the per-XCD L2 caches; every other access to them must therefore
// be a device-scope atomic load or store as well, never a plain access that
// could sit in (or read from) one XCD's L2.
__device__ __forceinline__ double ld_dev(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double area_w(const Geom& g, const double* dxu, const double* cosu,
                                         const double* dyu, int r, int c) {
    return dxu[c] * cosu[r] * dyu[r];
}

__global__ void k_fcg_prepare(Geom g, const double* __restrict__ cC, const uint8_t* __restrict__ act,
                              const double* __restrict__ hur, const double* __restrict__ hvr,
                              float* __restrict__ Z32, float* __restrict__ hinv32) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    float z = 0.f, h = 0.f;
    if (r >= 1 && r <= g.ny + 2 && c >= 1 && c <= g.nx + 2) {
        if (act[idx]) z = (float)(1.0 / cC[idx]);
        // inverse depth at the psi point: mean of the four faces around it
        h = (float)(0.25 * (hur[idx] + hur[idx + g.pitch] + hvr[idx] + hvr[idx + 1]));
    }
    Z32[idx] = z;
    hinv32[idx] = h;
}

__global__ void k_fcg_begin(Geom g, const double* __restrict__ r64, const uint8_t* __restrict__ act,
                            const double* sc, float* __restrict__ r32,
                            float* __restrict__ e32, float* __restrict__ p32) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    r32[idx] = act[idx] ? (float)(r64[idx] * ld_dev(sc + 8)) : 0.f;
    e32[idx] = 0.f;
    p32[idx] = 0.f;
}

__global__ void k_fcg_zr2(Geom g, const float* __restrict__ r32, const float* __restrict__ Z32,
                          double* __restrict__ sc) {
    __shared__ double sh[2 * NW];
    double acc[2] = {0.0, 0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        const double v = (double)Z32[idx] * (double)r32[idx];
        acc[0] += v * v;
        acc[1] += Z32[idx] != 0.f ? 1.0 : 0.0;
    }
    block_reduce<2>(acc, 0u, sh);
    if (threadIdx.x == 0) {
        atomicAdd(&sc[6], acc[0]);
        atomicAdd(&sc[7], acc[1]);
    }
}

__global__ void k_fcg_scalar(double* sc, int what) {
    if (what == 0) {          // finish the input scale, reset its accumulators
        const double s6 = ld_dev(sc + 6), s7 = ld_dev(sc + 7);
        st_dev(sc + 5, s6 > 0 ? 1.0 / sqrt(s6 / fmax(s7, 1.0)) : 1.0);
        st_dev(sc + 6, 0.0);
        st_dev(sc + 7, 0.0);
    } else {                  // start of a step
        st_dev(sc + 2, ld_dev(sc + 1));
        st_dev(sc + 0, 0.0);
        st_dev(sc + 1, 0.0);
        st_dev(sc + 3, 0.0);
        st_dev(sc + 4, 0.0);
        st_dev(sc + 10, 0.0);
        st_dev(sc + 11, 0.0);
    }
}

// The network's value at solver point idx = (r, c) (network layout H x W),
// zero off active points; both loads are issued whatever the mask (at a
// clamped in-range position), so neither waits for the other.
__device__ __forceinline__ float fcg_net(const float* __restrict__ znet, int H, int W,
                                         const uint8_t* __restrict__ act, size_t idx, int r,
                                         int c) {
    const uint8_t a = act[idx];
    const float z = znet[(size_t)min(r, H - 1) * W + min(c, W - 1)];
    return (a && r < H && c < W) ? z : 0.f;
}

// z (network layout H x W) -> zs on the solver grid; (zs, W q_old)
__global__ void k_fcg_scatter(Geom g, const float* __restrict__ znet, int H, int W,
                              const uint8_t* __restrict__ act, const float* __restrict__ q32,
                              const double* __restrict__ dxu, const double* __restrict__ cosu,
                              const double* __restrict__ dyu, float* __restrict__ zs,
                              double* __restrict__ sc) {
    __shared__ double sh[NW];
    double acc[1] = {0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        const int r = (int)((unsigned)idx / (unsigned)g.pitch), c = (int)idx - r * g.pitch;
        const float zv = fcg_net(znet, H, W, act, idx, r, c);
        const float qv = q32[idx];
        const double wa = area_w(g, dxu, cosu, dyu, r, c);
        zs[idx] = zv;
        acc[0] += (double)zv * (double)qv * wa;
    }
    block_reduce<1>(acc, 0u, sh);
    if (threadIdx.x == 0) atomicAdd(&sc[0], acc[0]);
}

__global__ void k_fcg_dir(Geom g, const float* __restrict__ zs, float* __restrict__ p32,
                          const double* sc, int restart) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const float beta = restart ? 0.f : (float)(ld_dev(sc + 0) / ld_dev(sc + 2));
    p32[idx] = zs[idx] - beta * p32[idx];
}

// q = A p, (p, W q), (r, W p).  A is the FP64 operator (q stored in FP32): with
// the FP32 one the recursion would follow a system whose difference from
// A x = b grows with |A||x| rather than with the residual, which holds the
// FP32 legs between residual replacements to small reductions.
__global__ void k_fcg_apply(Geom g, const double* __restrict__ E, const double* __restrict__ N,
                            const double* __restrict__ rx, const double* __restrict__ ry,
                            const float* __restrict__ p, const float* __restrict__ r32,
                            const double* __restrict__ dxu, const double* __restrict__ cosu,
                            const double* __restrict__ dyu, float* __restrict__ q,
                            double* __restrict__ sc) {
    __shared__ double sh[2 * NW];
    double acc[2] = {0.0, 0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const int P = g.pitch;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        int r = (int)((unsigned)idx / (unsigned)P), c = (int)idx - r * P;
        // halo points compute at the first interior point and keep nothing,
        // so every load is issued without a branch
        const bool in = r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1;
        if (!in) r = 2, c = 2;
        const size_t ic = (size_t)r * P + c;
        const int cl = map_col(c - 1, g), cr = map_col(c + 1, g);
        const size_t il = (size_t)r * P + cl, ir = (size_t)r * P + cr;
        const double e = E[ic], w = E[il] * rx[c], n = N[ic], s = N[ic - P] * ry[r];
        const float ps = p[ic - P], pw = p[il], pc = p[ic], pe = p[ir], pn = p[ic + P];
        const float rv = r32[ic];
        const double wa = area_w(g, dxu, cosu, dyu, r, c);
        const double cc = (((-e) - w) - n) - s;
        const double vd = s * (double)ps + w * (double)pw + cc * (double)pc + e * (double)pe +
                          n * (double)pn;
        q[idx] = in ? (float)vd : 0.f;
        if (in) {   // (arithmetic only: the loads are above)
            acc[0] += (double)pc * vd * wa;
            acc[1] += (double)rv * (double)pc * wa;
        }
    }
    block_reduce<2>(acc, 0u, sh);
    if (threadIdx.x == 0) {
        atomicAdd(&sc[1], acc[0]);
        atomicAdd(&sc[3], acc[1]);
    }
}

__global__ void k_fcg_update(Geom g, const float* __restrict__ p, const float* __restrict__ q,
                             const uint8_t* __restrict__ act, float* __restrict__ e32,
                             float* __restrict__ r32, double* sc) {
    __shared__ double sh[NW];
    double acc[1] = {0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const float alpha = (float)(ld_dev(sc + 3) / ld_dev(sc + 1));
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        int a = act[idx];
        float ev = e32[idx], pv = p[idx], rv0 = r32[idx], qv = q[idx];
        pin_loaded(a), pin_loaded(ev), pin_loaded(pv), pin_loaded(rv0), pin_loaded(qv);
        if (!a) continue;
        e32[idx] = ev + alpha * pv;
        const float rv = rv0 - alpha * qv;
        r32[idx] = rv;
        acc[0] += (double)rv * (double)rv;
    }
    block_reduce<1>(acc, 0u, sh);
    if (threadIdx.x == 0) atomicAdd(&sc[4], acc[0]);
}

// ---- flexible CG with the network composed with the multigrid V-cycle ----
// z = c n + V(r - c A n): n the network's estimate (network layout H x W,
// masked to active points), c = (n, W r) / (n, S n) its A-optimal scale
// (S = W A), V the aggregation-multigrid V-cycle on what the network leaves.
// Scalars sc[10] = (n, W r), sc[11] = (n, S n).

// A n (FP32 operator) and the two products of the scale.
__global__ void k_fcgmg_scale(Geom g, const float* __restrict__ znet, int H, int W,
                              const uint8_t* __restrict__ act, const float* __restrict__ E,
                              const float* __restrict__ N, const float* __restrict__ rx,
                              const float* __restrict__ ry, const float* __restrict__ r32,
                              const double* __restrict__ dxu, const double* __restrict__ cosu,
                              const double* __restrict__ dyu, float* __restrict__ an,
                              double* __restrict__ sc) {
    __shared__ double sh[2 * NW];
    double acc[2] = {0.0, 0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const int P = g.pitch;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        int r = (int)((unsigned)idx / (unsigned)P), c = (int)idx - r * P;
        // halo points compute at the first interior point and keep nothing,
        // so every load is issued without a branch
        const bool in = r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1;
        if (!in) r = 2, c = 2;
        const size_t ic = (size_t)r * P + c;
        const int cl = map_col(c - 1, g), cr = map_col(c + 1, g);
        const size_t il = (size_t)r * P + cl, ir = (size_t)r * P + cr;
        const float e = E[ic], w = E[il] * rx[c], n = N[ic], s = N[ic - P] * ry[r];
        const float nc = fcg_net(znet, H, W, act, ic, r, c);
        const float ns = fcg_net(znet, H, W, act, ic - P, r - 1, c);
        const float nw = fcg_net(znet, H, W, act, il, r, cl);
        const float ne = fcg_net(znet, H, W, act, ir, r, cr);
        const float nn = fcg_net(znet, H, W, act, ic + P, r + 1, c);
        const float rv = r32[ic];
        const double wa = area_w(g, dxu, cosu, dyu, r, c);
        const float cc = (((-e) - w) - n) - s;
        const float v = s * ns + w * nw + cc * nc + e * ne + n * nn;
        an[idx] = in ? v : 0.f;
        if (in) {   // (arithmetic only: the loads are above)
            acc[0] += (double)nc * (double)rv * wa;
            acc[1] += (double)nc * (double)v * wa;
        }
    }
    block_reduce<2>(acc, 0u, sh);
    if (threadIdx.x == 0) {
        atomicAdd(&sc[10], acc[0]);
        atomicAdd(&sc[11], acc[1]);
    }
}

// c minimises the energy norm of e - c n (e = A^-1 r); S is negative definite,
// so (n, S n) < 0 for any n != 0 on active points
__device__ __forceinline__ float fcgmg_c(const double* sc) {
    const double a = ld_dev(sc + 10), q = ld_dev(sc + 11);
    const double cv = q != 0.0 ? a / q : 0.0;
    return cv == cv ? (float)cv : 0.f;
}

// The V-cycle's input on level 0: (w / w_mean) (r - c A n) on active cells.
__global__ void k_fcgmg_input(Geom g, MgLev L0, const uint8_t* __restrict__ act,
                              const float* __restrict__ r32, const float* __restrict__ an,
                              const double* __restrict__ dxu, const double* __restrict__ cosu,
                              const double* __restrict__ dyu, double inv_wmean,
                              const double* sc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int i = t % g.nx, j = t / g.nx;
    const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
    const int ql = (j + 1) * L0.pitch + i + MG_XOFF;
    // the point's loads first, the device scalars beside them
    int a = act[o];
    float rv = r32[o], av = an[o];
    double wa = area_w(g, dxu, cosu, dyu, j + 2, i + 2);
    pin_loaded(a), pin_loaded(rv), pin_loaded(av), pin_loaded(wa);
    const float cn = fcgmg_c(sc);
    L0.b[ql] = a ? (float)(wa * inv_wmean * (double)(rv - cn * av)) : 0.0f;
}

// zs = c n + w_mean y on active points; (zs, W q_old)
__global__ void k_fcgmg_scatter(Geom g, MgLev L0, const float* __restrict__ znet, int H, int W,
                                const uint8_t* __restrict__ act, const float* __restrict__ q32,
                                const double* __restrict__ dxu, const double* __restrict__ cosu,
                                const double* __restrict__ dyu, float w_mean,
                                float* __restrict__ zs, double* __restrict__ sc) {
    __shared__ double sh[NW];
    double acc[1] = {0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const float cn = fcgmg_c(sc);
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        const int r = (int)((unsigned)idx / (unsigned)g.pitch), c = (int)idx - r * g.pitch;
        // every load issued whatever the mask (y at an in-range cell)
        const bool in = r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1;
        const uint8_t a = act[idx];
        const float y = L0.y[in ? (r - 1) * L0.pitch + (c - 2) + MG_XOFF : L0.pitch + MG_XOFF];
        const float nv = fcg_net(znet, H, W, act, idx, r, c);
        const float z = a ? cn * nv + w_mean * y : 0.f;
        zs[idx] = z;
        acc[0] += (double)z * (double)q32[idx] * area_w(g, dxu, cosu, dyu, r, c);
    }
    block_reduce<1>(acc, 0u, sh);
    if (threadIdx.x == 0) atomicAdd(&sc[0], acc[0]);
}

}  // namespace

namespace {
// A-norm cosine loss of the learned preconditioner.  For sample b with
// residual r (FP64, solver layout), true error e (through ER = e . W r) and
// network output z (fp32, network layout H x W, masked to active points):
//   a = z . W r,  q = z . S z (S = W A),  L = 1 - a^2 / (q ER)
//   dL/dz = -(2a/(q ER)) W r + (2a^2/(q^2 ER)) S z
__global__ void k_nn_loss1(Geom g, const float* __restrict__ znet, int H, int W,
                           const double* __restrict__ R, const uint8_t* __restrict__ act,
                           const float* __restrict__ E, const float* __restrict__ N,
                           const float* __restrict__ rx, const float* __restrict__ ry,
                           const double* __restrict__ dxu, const double* __restrict__ cosu,
                           const double* __restrict__ dyu, float* __restrict__ Sz,
                           double* __restrict__ work) {
    __shared__ double sh[2 * NW];
    const int b = blockIdx.y;
    const float* z = znet + (size_t)b * H * W;
    const double* r = R + (size_t)b * g.pitch * (g.ny + 4);
    float* out = Sz + (size_t)b * H * W;
    double acc[2] = {0.0, 0.0};
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const int P = g.pitch;
    auto zm = [&](size_t idx) -> float {
        const int rr = (int)(idx / P), cc = (int)(idx % P);
        return act[idx] ? z[(size_t)rr * W + cc] : 0.f;
    };
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < plane; idx += (size_t)gridDim.x * NT) {
        const int rr = (int)(idx / P), cc = (int)(idx % P);
        if (!act[idx]) continue;
        const float e = E[idx], w = E[idx - 1] * rx[cc], n = N[idx], s = N[idx - P] * ry[rr];
        const float c0 = (((-e) - w) - n) - s;
        const float zc = zm(idx);
        const float az = s * zm(idx - P) + w * zm(idx - 1) + c0 * zc + e * zm(idx + 1) + n * zm(idx + P);
        const double wa = dxu[cc] * cosu[rr] * dyu[rr];
        const double szv = wa * (double)az;
        out[(size_t)rr * W + cc] = (float)szv;
        acc[0] += (double)zc * wa * r[idx];
        acc[1] += (double)zc * szv;
    }
    block_reduce<2>(acc, 0u, sh);
    if (threadIdx.x == 0) {
        atomicAdd(&work[2 * b], acc[0]);
        atomicAdd(&work[2 * b + 1], acc[1]);
    }
}

__global__ void k_nn_loss2(Geom g, int H, int W, int B, const double* __restrict__ R,
                           const uint8_t* __restrict__ act, const double* __restrict__ ER,
                           const double* __restrict__ dxu, const double* __restrict__ cosu,
                           const double* __restrict__ dyu, const double* work,
                           float* __restrict__ Sz_to_grad, double* __restrict__ loss) {
    const int b = blockIdx.y;
    const double a = ld_dev(work + 2 * b), q = ld_dev(work + 2 * b + 1), er = ER[b];
    const double ca = -2.0 * a / (q * er) / B, cs = 2.0 * a * a / (q * q * er) / B;
    const double* r = R + (size_t)b * g.pitch * (g.ny + 4);
    float* gz = Sz_to_grad + (size_t)b * H * W;
    const long n = (long)H * W;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        const int y = (int)(p / W), x = (int)(p % W);
        float gv = 0.f;
        if (y < g.ny + 4 && x < g.pitch) {
            const size_t idx = (size_t)y * g.pitch + x;
            if (act[idx]) {
                const double wa = dxu[x] * cosu[y] * dyu[y];
                gv = (float)(ca * wa * r[idx] + cs * (double)gz[p]);
            }
        }
        gz[p] = gv;
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        double L = 0.0;
        for (int k = 0; k < B; ++k) {
            const double ak = ld_dev(work + 2 * k), qk = ld_dev(work + 2 * k + 1);
            L += 1.0 - ak * ak / (qk * ER[k]);
        }
        *loss = L / B;
    }
}
}  // namespace

namespace {

thread_local std::string g_create_err;

int fail(psi_ctx* c, int code, const char* fmt, const char* what = nullptr) {
    char buf[512];
    snprintf(buf, sizeof buf, fmt, what ? what : "");
    if (c) c->err = buf;
    else g_create_err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return fail((c), PSI_HIP_ERROR, "HIP error: %s", hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(psi_ctx* c, T** p, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(T) + 64);
    if (e != hipSuccess) return fail(c, PSI_HIP_ERROR, "hipMalloc: %s", hipGetErrorString(e));
    e = hipMemsetAsync(v, 0, n * sizeof(T) + 64, c->stream);
    if (e != hipSuccess) return fail(c, PSI_HIP_ERROR, "hipMemset: %s", hipGetErrorString(e));
    c->allocs.push_back(v);
    c->bytes += (int64_t)(n * sizeof(T) + 64);
    *p = (T*)v;
    return PSI_OK;
}

inline unsigned nblk(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

int check_launch(psi_ctx* c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, PSI_HIP_ERROR, "kernel launch: %s", hipGetErrorString(e));
    return PSI_OK;
}

int build_operator(psi_ctx* c) {
    const Geom& g = c->g;
    hipLaunchKernelGGL(k_build_EN, dim3(nblk(c->plane)), dim3(256), 0, c->stream, g, c->hur,
                       c->hvr, c->dxt, c->dxu, c->dyt, c->dyu, c->cost, c->cosu, c->E, c->N);
    hipLaunchKernelGGL(k_build_active, dim3(nblk(c->plane)), dim3(256), 0, c->stream, g, c->E,
                       c->N, c->rx, c->ry, c->dirichlet, c->act, c->cC);
    hipLaunchKernelGGL(k_convert<float>, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->plane,
                       c->E, c->E32);
    hipLaunchKernelGGL(k_convert<float>, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->plane,
                       c->N, c->N32);
    return check_launch(c);
}

struct CgRun {
    CgState fin{};
    double sweep_ms = 0;
    long launches = 0;
};

template <class T>
struct CgPath {
    const T *E, *N, *rx, *ry;
    T *u[2], *p[2];
    int G;
    int poly = 0;                  // > 0: Chebyshev-preconditioned FP32 kernels of degree poly
    int march = 0;                 // > 0: Jacobi sweeps by k_cg_march of this shape
    ChebCoef ch{};
    T* rb[2] = {nullptr, nullptr}; // their residual buffers (u holds z = M r)
};

// Launch of the polynomial-preconditioned kernel of mode MODE and the
// path's degree.
template <int MODE>
int launch_poly(psi_ctx* c, const CgPath<float>& P, float* x, const double* r64, double rscale,
                const double* true_res, double thresh, int new_max_it, const CgState* si,
                CgState* so, const double* pi, double* po) {
#define PSI_POLY_LAUNCH(K)                                                                    \
    hipLaunchKernelGGL((k_poly_pcg<K, MODE>), dim3(P.G), dim3(NT), 0, c->stream,            \
                       poly_geom(c->g, K), P.ch,                                              \
                       P.E, P.N, P.rx, P.ry, c->act, x, P.u[0], P.u[1], P.p[0], P.p[1],       \
                       P.rb[0], P.rb[1], r64, rscale, true_res, thresh, new_max_it, si, so,  \
                       pi, po)
    switch (P.poly) {
        case 1: PSI_POLY_LAUNCH(1); break;
        case 2: PSI_POLY_LAUNCH(2); break;
        case 3: PSI_POLY_LAUNCH(3); break;
        case 4: PSI_POLY_LAUNCH(4); break;
        case 5: PSI_POLY_LAUNCH(5); break;
        case 6: PSI_POLY_LAUNCH(6); break;
        default: return PSI_BAD_ARG;
    }
#undef PSI_POLY_LAUNCH
    return PSI_OK;
}

template <class T>
int launch_poly_checked(psi_ctx* c, int mode, const CgPath<T>& P, T* x, const double* r64,
                        double rscale, const double* true_res, double thresh, int new_max_it,
                        const CgState* si, CgState* so, const double* pi, double* po) {
    if constexpr (std::is_same<T, float>::value) {
        int rc = mode == PM_SETUP
                     ? launch_poly<PM_SETUP>(c, P, x, r64, rscale, true_res, thresh, new_max_it,
                                             si, so, pi, po)
                 : mode == PM_SWEEP
                     ? launch_poly<PM_SWEEP>(c, P, x, r64, rscale, true_res, thresh, new_max_it,
                                             si, so, pi, po)
                     : launch_poly<PM_REFRESH>(c, P, x, r64, rscale, true_res, thresh,
                                               new_max_it, si, so, pi, po);
        return rc;
    } else {
        return PSI_BAD_ARG;
    }
}

// Launch bookkeeping of one solve: launch L reads state/partials (L-1) & 1
// and writes L & 1 (the setup is launch 0 and reads the host-written state 0).
// Each half of part2 holds G partials per slot; Gprev is how many the last
// launch wrote (its grid).
struct Launches {
    long L = 0;
    int G = 0;
    int Gprev = 0;
    CgState* st_in(psi_ctx* c) const { return c->st + ((L - 1) & 1); }
    CgState* st_out(psi_ctx* c) const { return c->st + (L & 1); }
    double* part_in(psi_ctx* c) const { return c->part2 + ((L - 1) & 1) * (size_t)NRED * G; }
    double* part_out(psi_ctx* c) const { return c->part2 + (L & 1) * (size_t)NRED * G; }
};

// Shapes of the wave-column sweep (k_cg_march<T, SEG, D>): SEG rows per
// wave, D rows of loads in flight; code SEG * 10 + D.  Per launch on the
// 2048^2 two-island basin, FP64 / FP32: 24 x 6 52.4 / 32.2 us, 32 x 8
// 55.1 / 33.4 us, 48 x 8 54.3 / 33.2 us; the LDS-tile sweep 54.6 / 35.3 us.
#define PSI_MARCH_SHAPES(X) X(24, 6) X(32, 8) X(48, 8)
#ifndef PSI_MARCH_DEFAULT
#define PSI_MARCH_DEFAULT 246
#endif

// Grid of march shape `code` (0: not a shape), from the resident workgroups
// per CU of its FP64 kernel; *one_round: every wave has at most one item.
inline int march_grid(const Geom& g, int ncu, int code, bool* one_round = nullptr) {
    int occ = 0;
#define PSI_MARCH_GRID(S, D)                                                                  \
    if (code == S * 10 + D) {                                                                 \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_cg_march<double, S, D>, NT,  \
                                                         0) != hipSuccess)                    \
            occ = 1;                                                                          \
        if (const char* ov = getenv("PSI_GRID_PER_CU"))                                       \
            if (atoi(ov) > 0) occ = std::min(occ, atoi(ov));                                  \
        const int G = march_groups<S>(g, ncu, occ);                                           \
        if (one_round) *one_round = march_items<S>(g) <= (long)NW * G;                        \
        return G;                                                                             \
    }
    PSI_MARCH_SHAPES(PSI_MARCH_GRID)
#undef PSI_MARCH_GRID
    return 0;
}

// The Jacobi sweeps' kernel and the partials per slot every solver launch
// fits in; needs c->g, c->G and c->G32.  The march when one round of waves
// covers the grid (2048^2 FP64 52.5 us against 54.6 us for the LDS-tile
// sweep), else the LDS-tile sweep: with several items per wave the march
// streams HBM worse (4096^2: 316 us against 250 us).  PSI_MARCH=SEG*10+D
// forces a march shape, PSI_MARCH=0 the LDS-tile sweep.
void choose_march(psi_ctx* c) {
    bool one = false;
    c->march = march_grid(c->g, c->ncu, PSI_MARCH_DEFAULT, &one) > 0 && one ? PSI_MARCH_DEFAULT : 0;
    if (const char* mv = getenv("PSI_MARCH")) {
        const int m = atoi(mv);
        if (m == 0 || march_grid(c->g, c->ncu, m) > 0) c->march = m;
    }
    c->Gpart = std::max(c->G, c->G32);
    c->Gmarch = c->march ? march_grid(c->g, c->ncu, c->march) : 0;
    c->Gpart = std::max(c->Gpart, c->Gmarch);
}

template <class T>
void launch_march(psi_ctx* c, const CgPath<T>& P, T* x, const CgState* si, CgState* so,
                  const double* pi, int Gprev, double* po) {
#define PSI_MARCH_LAUNCH(S, D)                                                                \
    if (P.march == S * 10 + D) {                                                              \
        hipLaunchKernelGGL((k_cg_march<T, S, D>), dim3(c->Gmarch), dim3(NT), 0,              \
                           c->stream, c->g, P.E, P.N, P.rx, P.ry, c->act, x, P.u[0], P.u[1],  \
                           P.p[0], P.p[1], si, so, pi, Gprev, po);                            \
        return;                                                                               \
    }
    PSI_MARCH_SHAPES(PSI_MARCH_LAUNCH)
#undef PSI_MARCH_LAUNCH
}

template <class T, class TB>
int cg_setup(psi_ctx* c, const CgPath<T>& P, const TB* b, double bscale, T* x, bool zero_guess,
             const CgState& init, Launches& Ln) {
    c->h_st[2] = init;
    HIPCHK(c, hipMemcpyAsync(c->st, &c->h_st[2], sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    Ln.L = 0;
    Ln.G = c->Gpart;
    Ln.Gprev = P.G;
    if (P.poly > 0) {
        // the polynomial path starts from a zero first guess on the residual b
        if constexpr (std::is_same<TB, double>::value) {
            if (!zero_guess) return fail(c, PSI_BAD_ARG, "internal: poly setup%s");
            if (launch_poly_checked<T>(c, PM_SETUP, P, x, b, bscale, nullptr, 0.0, 0, nullptr,
                                       nullptr, nullptr, Ln.part_out(c)))
                return fail(c, PSI_BAD_ARG, "polynomial degree%s");
        } else {
            return fail(c, PSI_BAD_ARG, "internal: poly setup%s");
        }
    } else {
        hipLaunchKernelGGL((k_cg_setup<T, TB>), dim3(P.G), dim3(NT), 0, c->stream, c->g, P.E,
                           P.N, P.rx, P.ry, c->act, b, bscale,
                           zero_guess ? (const T*)nullptr : (const T*)x,
                           zero_guess ? x : (T*)nullptr, P.u[0], P.p[0], Ln.part_out(c));
    }
    Ln.L = 1;
    return check_launch(c);
}

// Issue fused sweeps in batches until a poll of the device state shows the
// solve (or this leg of it) finished.  The host keeps one batch queued ahead
// of the state it inspects, so the GPU never idles on a poll; sweeps issued
// after completion see the finished state, forward it and return at once.
template <class T>
int cg_iterate(psi_ctx* c, const CgPath<T>& P, T* x, Launches& Ln, int batch, long max_launch,
               CgRun* out) {
    int slot = 0, queue[2], qn = 0;
    long issued_at[2] = {0, 0};
    const long L0 = Ln.L;
    // a context without events (psi_congrad_noalloc) polls with stream
    // synchronisations and reports no sweep time
    const bool events = c->ev[0] != nullptr;
    if (events) HIPCHK(c, hipEventRecord(c->ts0, c->stream));
    auto issue = [&]() -> int {
        for (int k = 0; k < batch && Ln.L - L0 < max_launch; ++k, ++Ln.L) {
            if (P.poly > 0) {
                (void)launch_poly_checked<T>(c, PM_SWEEP, P, x, nullptr, 0.0, nullptr, 0.0, 0,
                                             Ln.st_in(c), Ln.st_out(c), Ln.part_in(c),
                                             Ln.part_out(c));
            } else if (P.march > 0) {
                launch_march<T>(c, P, x, Ln.st_in(c), Ln.st_out(c), Ln.part_in(c), Ln.Gprev,
                                Ln.part_out(c));
                Ln.Gprev = c->Gmarch;
            } else {
                hipLaunchKernelGGL((k_cg_sweep<T>), dim3(P.G), dim3(NT), 0, c->stream, c->g,
                                   P.E, P.N, P.rx, P.ry, c->act, x, P.u[0], P.u[1], P.p[0],
                                   P.p[1], Ln.st_in(c), Ln.st_out(c), Ln.part_in(c),
                                   Ln.part_out(c));
            }
        }
        int r = check_launch(c);
        if (r) return r;
        if (events) HIPCHK(c, hipEventRecord(c->ts1[slot], c->stream));
        issued_at[slot] = Ln.L - L0;
        // the latest state is the one the last launch wrote
        HIPCHK(c, hipMemcpyAsync(&c->h_st[slot], c->st + ((Ln.L - 1) & 1), sizeof(CgState),
                                 hipMemcpyDeviceToHost, c->stream));
        if (events) HIPCHK(c, hipEventRecord(c->ev[slot], c->stream));
        queue[qn++] = slot;
        slot ^= 1;
        return PSI_OK;
    };
    for (;;) {
        if (qn < 2 && Ln.L - L0 < max_launch) {
            int rc = issue();
            if (rc) return rc;
        }
        if (qn == 0) return fail(c, PSI_HIP_ERROR, "solver launch budget exhausted%s");
        const int s = queue[0];
        queue[0] = queue[1];
        --qn;
        if (events) HIPCHK(c, hipEventSynchronize(c->ev[s]));
        else HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->h_st[s].done) {
            out->fin = c->h_st[s];
            float ms = 0.f;
            if (events) HIPCHK(c, hipEventElapsedTime(&ms, c->ts0, c->ts1[s]));
            out->sweep_ms += ms;
            out->launches += issued_at[s];
            // drain the batch queued behind this poll so the next leg starts clean
            if (qn) {
                if (events) HIPCHK(c, hipEventSynchronize(c->ev[queue[0]]));
                out->launches += issued_at[queue[0]] - issued_at[s];
            }
            return PSI_OK;
        }
    }
}

// FP64 residual of (b, x): returns |r|^2, |b|^2 (active points), max|r|, count.
// The FP64 residual's norms into c->result on stream s (r_out: r as well).
void residual_launch(psi_ctx* c, hipStream_t s, const double* b, const double* x, double* r_out) {
    const Geom& g = c->g;
    // 4 workgroups per CU, a multiple of 8; part holds NRED x (ntiles + 1024)
    const int Gr = std::max(8, std::min(4 * c->ncu, g.ntiles + 1024) & ~7);
    hipLaunchKernelGGL(k_residual, dim3(Gr), dim3(NT), 0, s, g, c->E, c->N, c->rx, c->ry, c->act,
                       b, x, r_out, c->part);
    hipLaunchKernelGGL(k_residual_fold, dim3(1), dim3(NT), 0, s, (const double*)c->part, Gr,
                       c->result);
}

int residual(psi_ctx* c, const double* b, const double* x, double* r_out, double res[4]) {
    residual_launch(c, c->stream, b, x, r_out);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->h_res, c->result, 4 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) res[k] = c->h_res[k];
    return PSI_OK;
}

template <class T>
int cg_refresh(psi_ctx* c, const CgPath<T>& P, double rscale, double rr_thresh, int new_max_it,
               Launches& Ln) {
    if (P.poly > 0) {
        if (launch_poly_checked<T>(c, PM_REFRESH, P, (T*)nullptr, c->r64, rscale, c->result,
                                   rr_thresh, new_max_it, Ln.st_in(c), Ln.st_out(c), nullptr,
                                   Ln.part_out(c)))
            return fail(c, PSI_BAD_ARG, "polynomial degree%s");
    } else {
        hipLaunchKernelGGL((k_cg_refresh<T>), dim3(P.G), dim3(NT), 0, c->stream, c->g, P.E, P.N,
                           P.rx, P.ry, c->act, c->r64, rscale, P.u[0], P.u[1], (const T*)P.p[0],
                           (const T*)P.p[1], c->result, rr_thresh, new_max_it, Ln.st_in(c),
                           Ln.st_out(c), Ln.part_out(c));
    }
    Ln.Gprev = P.G;
    ++Ln.L;
    return check_launch(c);
}

int status_of(int st) {
    switch (st) {
        case ST_CONVERGED: return PSI_OK;
        case ST_MAXIT: return PSI_NOT_CONVERGED;
        case ST_DIVERGED: return PSI_DIVERGED;
        case ST_NAN: return PSI_NAN;
        default: return PSI_NOT_CONVERGED;
    }
}

CgState make_init(const psi_opts* o, int criterion, double rtarget, int max_it) {
    CgState s{};
    s.eps = o->congr_epsilon;
    s.rtarget2 = rtarget * rtarget;
    s.max_it = max_it;
    s.criterion = criterion;
    s.phase = PHASE_SETUP;
    return s;
}

// One Jacobi-PCG solve in precision T on the correction x (FP64: x is the
// solution itself; FP32: x is a correction flushed into sol).  With the
// relres criterion the recursive residual is confirmed on the FP64 true
// residual; a miss replaces the recursive residual by the true one and
// continues (no restart).  refresh_every > 0 also replaces it every that
// many iterations (needed in FP32, where the two drift apart).
template <class T>
int solve_pcg(psi_ctx* c, const psi_opts* o, const CgPath<T>& P, const double* forc, double* sol,
              T* xw, double bscale_in, bool fp64, int refresh_every, double leg_reduction,
              psi_info* info) {
    const int batch = o->check_every > 0 ? o->check_every : 32;
    const bool relres = fp64 ? (o->criterion == PSI_CRIT_RELRES) : true;
    const double target = o->relres_target > 0 ? o->relres_target : 1e-10;
    const int abs_max = o->max_iterations;
    double res[4] = {0, 0, 0, 0};
    double scale = 1.0, bb_forc = 0.0;
    Launches Ln;
    CgRun run;
    int rc;
    if (!fp64) {
        rc = residual(c, forc, sol, c->r64, res);
        if (rc) return rc;
        info->setup_launches += 1;
        bb_forc = res[1] > 0 ? res[1] : res[0];
        if (res[0] <= target * target * bb_forc) {
            info->iterations = 0;
            info->status = PSI_OK;
            info->converged = 1;
            info->relres_true = bb_forc > 0 ? sqrt(res[0] / bb_forc) : 0.0;
            return PSI_OK;
        }
        const double rms = sqrt(res[0] / std::max(res[3], 1.0));
        if (!(rms > 0) || rms != rms) return fail(c, PSI_NAN, "NaN residual%s");
        scale = 1.0 / rms;
    }
    const int first_max = refresh_every > 0 ? std::min(refresh_every, abs_max) : abs_max;
    CgState init = make_init(o, relres ? CRIT_RELRES : CRIT_REFERENCE, target, first_max);
    if (!fp64) init.bb_fixed = bb_forc * scale * scale;
    init.red2 = leg_reduction > 0 ? leg_reduction * leg_reduction : 0.0;
    if (fp64)
        rc = cg_setup<T, double>(c, P, forc, bscale_in, xw, false, init, Ln);
    else
        rc = cg_setup<T, double>(c, P, c->r64, scale, xw, true, init, Ln);
    if (rc) return rc;
    info->setup_launches += 1;
    int refreshes = 0;
    for (;;) {
        rc = cg_iterate<T>(c, P, xw, Ln, batch, (long)abs_max + 4, &run);
        if (rc) return rc;
        const CgState& f = run.fin;
        const bool leg_end = (f.status == ST_MAXIT && f.n < abs_max) || f.status == ST_LEG;
        if (!relres || !(f.status == ST_CONVERGED || leg_end)) {
            // iteration budget spent or divergence: sol keeps the current
            // iterate, as the reference's does (a NaN correction is dropped)
            if (!fp64 && (f.status == ST_MAXIT || f.status == ST_DIVERGED)) {
                hipLaunchKernelGGL((k_add_correction<T>), dim3(nblk(c->plane)), dim3(256), 0,
                                   c->stream, c->g, sol, xw, 1.0 / scale);
                info->setup_launches += 1;
                res[1] = 0.0;   // final residual below
            }
            break;
        }
        // confirm on the FP64 true residual
        if (!fp64) {
            hipLaunchKernelGGL((k_add_correction<T>), dim3(nblk(c->plane)), dim3(256), 0, c->stream,
                               c->g, sol, xw, 1.0 / scale);
            info->setup_launches += 1;
        }
        rc = residual(c, forc, sol, c->r64, res);
        if (rc) return rc;
        info->setup_launches += 1;
        if (bb_forc == 0.0) bb_forc = res[1] > 0 ? res[1] : f.rr0;
        const double thresh = target * target * bb_forc;
        if (res[0] <= thresh) {
            run.fin.status = ST_CONVERGED;
            break;
        }
        if (f.n >= abs_max) {
            run.fin.status = ST_MAXIT;
            break;
        }
        const int next = refresh_every > 0 ? std::min(f.n + refresh_every, abs_max) : abs_max;
        rc = cg_refresh<T>(c, P, fp64 ? 1.0 : scale, thresh, next, Ln);
        if (rc) return rc;
        info->setup_launches += 1;
        ++refreshes;
    }
    if (!fp64 && !relres) return fail(c, PSI_BAD_ARG, "internal%s");
    const CgState& f = run.fin;
    info->iterations = f.status == ST_MAXIT ? abs_max + 1 : f.n;
    info->status = status_of(f.status);
    info->converged = f.status == ST_CONVERGED;
    info->outer_iterations = refreshes;
    info->est_error = f.est_error;
    info->sweep_ms += run.sweep_ms;
    info->sweep_launches += (int32_t)run.launches;
    info->relres_recursive = f.bb > 0 ? sqrt(f.rr / f.bb) : 0.0;
    if (!relres || res[1] == 0.0) {
        rc = residual(c, forc, sol, nullptr, res);
        if (rc) return rc;
        info->setup_launches += 1;
    }
    const double bbn = bb_forc > 0 ? bb_forc : res[1];
    info->relres_true = bbn > 0 ? sqrt(res[0] / bbn) : sqrt(res[0]);
    return info->status;
}

int solve_plain(psi_ctx* c, const psi_opts* o, const double* forc, double* sol, psi_info* info) {
    CgPath<double> P{c->E, c->N, c->rx, c->ry, {c->u64[0], c->u64[1]}, {c->p64[0], c->p64[1]}, c->G};
    P.march = c->march;
    return solve_pcg<double>(c, o, P, forc, sol, sol, 1.0, true, 0, 0.0, info);
}

// Chebyshev iteration scalars for ZA on [a, b] = [lower * 2, 2] (degree k).
ChebCoef make_cheb(int k, double lower) {
    ChebCoef ch{};
    const double b = 2.0, a = lower * b;
    const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
    double rho = 1.0 / sigma;
    ch.inv_theta = (float)(1.0 / theta);
    for (int j = 0; j < k && j < PMAXK; ++j) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        ch.c1[j] = (float)(rn * rho);
        ch.c2[j] = (float)(2.0 * rn / delta);
        rho = rn;
    }
    return ch;
}

// Resident grid of k_poly_pcg<K, PM_SWEEP> (queried once per degree).
int poly_grid(psi_ctx* c, int k) {
    if (c->Gpoly[k]) return c->Gpoly[k];
    int ncu = 0, occ = 0;
    const void* fn = nullptr;
    switch (k) {
        case 1: fn = reinterpret_cast<const void*>(&k_poly_pcg<1, PM_SWEEP>); break;
        case 2: fn = reinterpret_cast<const void*>(&k_poly_pcg<2, PM_SWEEP>); break;
        case 3: fn = reinterpret_cast<const void*>(&k_poly_pcg<3, PM_SWEEP>); break;
        case 4: fn = reinterpret_cast<const void*>(&k_poly_pcg<4, PM_SWEEP>); break;
        case 5: fn = reinterpret_cast<const void*>(&k_poly_pcg<5, PM_SWEEP>); break;
        default: fn = reinterpret_cast<const void*>(&k_poly_pcg<6, PM_SWEEP>); break;
    }
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, NT, 0) != hipSuccess)
        return 0;
    const char* ov = getenv("PSI_GRID_PER_CU");
    if (ov && atoi(ov) > 0) occ = std::min(occ, atoi(ov));
    int G = std::max(1, ncu * std::max(occ, 1));
    G = std::min(G, c->Gpart);   // the partials buffer holds Gpart per slot
    G = std::min(G, poly_geom(c->g, k).ntiles);
    if (G >= 8) G &= ~7;
    c->Gpoly[k] = G;
    return G;
}

int solve_poly(psi_ctx* c, const psi_opts* o, const double* forc, double* sol, psi_info* info) {
    // defaults measured on the 2048^2 two-island basin (ms per solve to
    // relres 1e-10): k = 4 with the interval [0.01, 2] 143.6; k = 3 154.8;
    // k = 6 187 (two waves per SIMD); mixed Jacobi 434
    const int k = o->poly_degree > 0 ? o->poly_degree : 4;
    if (k > PMAXK) return fail(c, PSI_BAD_ARG, "poly_degree must be 1..6%s");
    const double lower = o->poly_lower > 0 ? o->poly_lower : 0.005;
    if (!(lower < 1.0)) return fail(c, PSI_BAD_ARG, "poly_lower must be in (0, 1)%s");
    // The polynomial is built for a spectrum of Z A in [0, 2], which holds when
    // every off-diagonal coefficient is >= 0 (weak diagonal dominance): hur,
    // hvr >= 0 and positive metrics.  Refuse grids whose metrics are not.
    const int nxh = c->g.nx + 4, nyh = c->g.ny + 4;
    for (int q = 1; q < nxh - 1; ++q)
        if (!(c->h_dxu[q] > 0 && c->h_dxt[q] > 0))
            return fail(c, PSI_BAD_ARG, "poly mode needs positive dxt/dxu%s");
    for (int q = 1; q < nyh - 1; ++q)
        if (!(c->h_dyu[q] > 0 && c->h_dyt[q] > 0 && c->h_cosu[q] > 0 && c->h_cost[q] > 0))
            return fail(c, PSI_BAD_ARG, "poly mode needs positive dyt/dyu/cost/cosu%s");
    CgPath<float> P{c->E32, c->N32, c->rx32, c->ry32, {c->u32[0], c->u32[1]}, {c->p32[0], c->p32[1]},
                    poly_grid(c, k)};
    if (P.G <= 0) return fail(c, PSI_HIP_ERROR, "occupancy query failed%s");
    P.poly = k;
    P.ch = make_cheb(k, lower);
    P.rb[0] = c->rp32[0];
    P.rb[1] = c->rp32[1];
    const double red = o->inner_reduction > 0 ? o->inner_reduction : 0.25;
    return solve_pcg<float>(c, o, P, forc, sol, c->e32, 1.0, false, o->leg_max_iterations, red,
                            info);
}

int solve_mixed(psi_ctx* c, const psi_opts* o, const double* forc, double* sol, psi_info* info) {
    CgPath<float> P{c->E32, c->N32, c->rx32, c->ry32, {c->u32[0], c->u32[1]}, {c->p32[0], c->p32[1]},
                    c->G32};
    P.march = c->march;
    // FP32 legs end after each `inner_reduction` drop of the residual (the
    // recursive FP32 residual is then replaced by the FP64 true one).  The
    // default 0.25 was measured on the 2048^2 two-island basin: legs of 0.2-0.5
    // need ~11.9k iterations to 1e-10, 0.05 needs 13.6k, 0.01 needs 30.6k (the
    // FP32 recursion drifts from the true residual over long legs).
    const double red = o->inner_reduction > 0 ? o->inner_reduction : 0.25;
    return solve_pcg<float>(c, o, P, forc, sol, c->e32, 1.0, false, o->leg_max_iterations, red,
                            info);
}

bool valid_field_ctx(psi_ctx* c) { return c && c->E; }

// ---- multigrid-preconditioned flexible CG ----------------------------------
void mg_release(psi_ctx* c) {
    MgState* m = c->mg;
    c->fcg_mg_ready = 0;
    if (!m) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (m->gs) (void)hipStreamSynchronize(m->gs);
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    if (m->exec_t) (void)hipGraphExecDestroy(m->exec_t);
    if (m->graph_t) (void)hipGraphDestroy(m->graph_t);
    for (hipEvent_t& e : m->kev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->h_S) (void)hipHostFree(m->h_S);
    if (m->h_res) (void)hipHostFree(m->h_res);
    if (m->e_in) (void)hipEventDestroy(m->e_in);
    if (m->e_out) (void)hipEventDestroy(m->e_out);
    if (m->t0) (void)hipEventDestroy(m->t0);
    if (m->t1) (void)hipEventDestroy(m->t1);
    if (m->gs) (void)hipStreamDestroy(m->gs);
    delete m;
    c->mg = nullptr;
}

template <class T>
int mg_alloc(psi_ctx* c, MgState* m, T** p, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(T) + 64);
    if (e != hipSuccess) return fail(c, PSI_HIP_ERROR, "hipMalloc: %s", hipGetErrorString(e));
    m->allocs.push_back(v);
    m->bytes += (int64_t)(n * sizeof(T) + 64);
    e = hipMemsetAsync(v, 0, n * sizeof(T) + 64, c->stream);
    if (e != hipSuccess) return fail(c, PSI_HIP_ERROR, "hipMemset: %s", hipGetErrorString(e));
    *p = (T*)v;
    return PSI_OK;
}

// Level 0 of the hierarchy from the stored operator: S = W A over active
// points (W the psi-cell area), couplings to inactive points dropped.
MgHostLev mg_level0(psi_ctx* c, const std::vector<double>& E, const std::vector<double>& N,
                    const std::vector<double>& cC, const std::vector<uint8_t>& act,
                    double& w_mean) {
    const Geom& g = c->g;
    MgHostLev L;
    L.nx = g.nx;
    L.ny = g.ny;
    L.cyc = g.cyclic;
    const size_t n = (size_t)g.nx * g.ny;
    L.E.assign(n, 0.0);
    L.N.assign(n, 0.0);
    L.C.assign(n, 1.0);
    L.act.assign(n, 0);
    double wsum = 0.0;
    auto o_of = [&](int i, int j) { return (size_t)(j + 2) * g.pitch + (size_t)(i + 2); };
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const size_t o = o_of(i, j), p = (size_t)i + (size_t)j * g.nx;
            if (!act[o]) continue;
            const double w = c->h_dxu[i + 2] * c->h_cosu[j + 2] * c->h_dyu[j + 2];
            L.act[p] = 1;
            ++L.nact;
            wsum += w;
            L.C[p] = w * cC[o];
            const int ie = (i + 1 < g.nx) ? i + 1 : (g.cyclic ? 0 : -1);
            if (ie >= 0 && act[o_of(ie, j)]) L.E[p] = w * E[o];
            if (j + 1 < g.ny && act[o_of(i, j + 1)]) L.N[p] = w * N[o];
        }
    w_mean = L.nact ? wsum / L.nact : 1.0;
    return L;
}

int mg_build(psi_ctx* c, std::vector<MgHostLev>& H);

int mg_setup(psi_ctx* c) {
    mg_release(c);
    const Geom& g = c->g;
    std::vector<double> E(c->plane), N(c->plane), cC(c->plane);
    std::vector<uint8_t> act(c->plane);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(E.data(), c->E, c->plane * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(N.data(), c->N, c->plane * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(cC.data(), c->cC, c->plane * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(act.data(), c->act, c->plane, hipMemcpyDeviceToHost));
    MgState* m = new MgState();
    c->mg = m;
    std::vector<MgHostLev> H;
    H.push_back(mg_level0(c, E, N, cC, act, m->w_mean));
    return mg_build(c, H);
}

// The rest of the hierarchy from its level 0 (H[0], m->w_mean set): 2 x 2
// coarsening down to at most MG_MAX_DENSE active cells, the coarsest level's
// dense inverse, device copies of every level, the outer-iteration vectors.
// Shared by the streamfunction and the surface-pressure contexts.
int mg_build(psi_ctx* c, std::vector<MgHostLev>& H) {
    MgState* m = c->mg;
    const Geom& g = c->g;
    m->nact0 = H[0].nact;
    if (m->nact0 == 0) {
        mg_release(c);
        return fail(c, PSI_BAD_ARG, "multigrid: no active points%s");
    }
    while (H.back().nact > MG_MAX_DENSE && (int)H.size() < MG_MAX_LEVELS) {
        MgHostLev next = mg_coarsen(H.back());
        if ((size_t)next.nx * next.ny >= (size_t)H.back().nx * H.back().ny) break;
        H.push_back(std::move(next));
    }
    if (H.back().nact > MG_MAX_DENSE * 4) {
        mg_release(c);
        return fail(c, PSI_BAD_ARG, "multigrid: coarsest level too large%s");
    }
    std::vector<int> cells;
    std::vector<float> inv;
    if (!mg_dense_inverse(H.back(), cells, inv)) {
        mg_release(c);
        return fail(c, PSI_BAD_ARG, "multigrid: singular coarsest operator%s");
    }
    m->L = (int)H.size() - 1;
    m->ncoarse = (int)cells.size();
    int rc = PSI_OK;
#define MGTRY(x)          \
    do {                  \
        rc = (x);         \
        if (rc) {         \
            mg_release(c); \
            return rc;    \
        }                 \
    } while (0)
    for (const MgHostLev& h : H) {
        MgLev d{};
        d.nx = h.nx;
        d.ny = h.ny;
        d.pitch = mg_pitch(h.nx);
        d.cyc = h.cyc;
        const size_t np = (size_t)d.pitch * (h.ny + 2);
        std::vector<float> hE(np, 0.0f), hN(np, 0.0f), hC(np, 1.0f);
        std::vector<uint8_t> hA(np, 0);
        for (int j = 0; j < h.ny; ++j)
            for (int i = 0; i < h.nx; ++i) {
                const size_t p = (size_t)i + (size_t)j * h.nx;
                const size_t q = (size_t)(j + 1) * d.pitch + i + MG_XOFF;
                hE[q] = (float)h.E[p];
                hN[q] = (float)h.N[p];
                hC[q] = (float)h.C[p];
                hA[q] = h.act[p];
            }
        float *dE, *dN, *dC;
        uint8_t* dA;
        MGTRY(mg_alloc(c, m, &dE, np));
        MGTRY(mg_alloc(c, m, &dN, np));
        MGTRY(mg_alloc(c, m, &dC, np));
        MGTRY(mg_alloc(c, m, &dA, np));
        HIPCHK(c, hipMemcpy(dE, hE.data(), np * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dN, hN.data(), np * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dC, hC.data(), np * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dA, hA.data(), np, hipMemcpyHostToDevice));
        d.E = dE;
        d.N = dN;
        d.C = dC;
        d.act = dA;
        MGTRY(mg_alloc(c, m, &d.b, np));
        MGTRY(mg_alloc(c, m, &d.x, np));
        MGTRY(mg_alloc(c, m, &d.y, np));
        m->lev.push_back(d);
        m->level_nx.push_back(h.nx);
        m->level_ny.push_back(h.ny);
        m->level_nact.push_back(h.nact);
    }
    // rows padded to a multiple of 4 floats (16-byte loads), zeros past n
    const int nc = m->ncoarse, ld = (nc + 3) & ~3;
    std::vector<float> padded((size_t)nc * ld, 0.0f);
    for (int i = 0; i < nc; ++i)
        std::copy(inv.begin() + (size_t)i * nc, inv.begin() + (size_t)(i + 1) * nc,
                  padded.begin() + (size_t)i * ld);
    m->dense_ld = ld;
    MGTRY(mg_alloc(c, m, &m->Ainv, padded.size() + 4));
    MGTRY(mg_alloc(c, m, &m->cells, cells.size() + 1));
    if (!inv.empty()) {
        HIPCHK(c, hipMemcpy(m->Ainv, padded.data(), padded.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(m->cells, cells.data(), cells.size() * 4, hipMemcpyHostToDevice));
    }
    if (m->L > 0 && nc > 0) {
        // the restriction into the coarsest level also writes b in row order
        MgLev& cl = m->lev[m->L];
        const size_t np = (size_t)cl.pitch * (cl.ny + 2);
        std::vector<int> hidx(np, -1);
        for (int k = 0; k < nc; ++k) hidx[(size_t)cells[k]] = k;
        int* didx = nullptr;
        float* bd = nullptr;
        MGTRY(mg_alloc(c, m, &didx, np));
        MGTRY(mg_alloc(c, m, &bd, (size_t)ld + 4));
        HIPCHK(c, hipStreamSynchronize(c->stream));   // the allocation's clearing first
        HIPCHK(c, hipMemcpy(didx, hidx.data(), np * 4, hipMemcpyHostToDevice));
        cl.didx = didx;
        cl.bd = bd;
    }
    MGTRY(mg_alloc(c, m, &m->x, c->plane));
    MGTRY(mg_alloc(c, m, &m->r, c->plane));
    MGTRY(mg_alloc(c, m, &m->p, c->plane));
    MGTRY(mg_alloc(c, m, &m->q, c->plane));
    MGTRY(mg_alloc(c, m, &m->az, c->plane));
    MGTRY(mg_alloc(c, m, &m->S, 1));
    MGTRY(mg_alloc(c, m, &m->S_time, 1));
    MGTRY(mg_alloc(c, m, &m->zscale, 1));
    m->nblk = (int)std::min<size_t>(nblk((size_t)g.nx * g.ny), PSI_MG_OUTER_BLOCKS);
    MGTRY(mg_alloc(c, m, &m->part, (size_t)MG_NRED * m->nblk));
    {
        // pass A on pairs: two 1024-thread workgroups per CU, at most one per
        // row segment
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
                hipSuccess ||
            ncu < 1)
            ncu = 1;
        const long segs = (long)g.ny * ((g.nx + 2 * MG_PASS_A_THREADS - 1) / (2 * MG_PASS_A_THREADS));
        long ga = std::min<long>((long)PSI_MG_PASS_A_PER_CU * ncu, segs);
        if (ga >= 8) ga &= ~7L;
        m->nblk_a = (int)std::max<long>(ga, 1);
        MGTRY(mg_alloc(c, m, &m->part_a, (size_t)MG_NRED * m->nblk_a));
    }
#undef MGTRY
    HIPCHK(c, hipHostMalloc((void**)&m->h_S, sizeof(MgScal), hipHostMallocDefault));
    HIPCHK(c, hipHostMalloc((void**)&m->h_res, 4 * sizeof(double), hipHostMallocDefault));
    HIPCHK(c, hipStreamCreateWithFlags(&m->gs, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&m->e_in, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&m->e_out, hipEventDisableTiming));
    HIPCHK(c, hipEventCreate(&m->t0));
    HIPCHK(c, hipEventCreate(&m->t1));
    for (hipEvent_t& e : m->kev) HIPCHK(c, hipEventCreate(&e));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

// One outer iteration on stream s: V-cycle, pass A, fold, pass B, fold.
// Timing events of a captured outer iteration.  hipEventRecord on a
// capturing stream adds no node to the graph, so each event goes in as an
// explicit event-record node after the stream's current capture
// dependencies, which then become that node.
struct MgEvents {
    const hipEvent_t* ev;
    hipError_t err = hipSuccess;
    void rec(int k, hipStream_t s) {
        if (err != hipSuccess) return;
        hipStreamCaptureStatus st;
        unsigned long long id = 0;
        hipGraph_t g = nullptr;
        const hipGraphNode_t* deps = nullptr;
        size_t nd = 0;
        err = hipStreamGetCaptureInfo_v2(s, &st, &id, &g, &deps, &nd);
        if (err != hipSuccess) return;
        if (st != hipStreamCaptureStatusActive) {
            err = hipErrorInvalidValue;   // only used while capturing
            return;
        }
        hipGraphNode_t node;
        err = hipGraphAddEventRecordNode(&node, g, deps, nd, ev[k]);
        if (err == hipSuccess)
            err = hipStreamUpdateCaptureDependencies(s, &node, 1, hipStreamSetCaptureDependencies);
    }
};

// The level kernels of level v at degree K (waves and rows per lane by the
// level's size).
template <int K, int NW, int R>
void mg_launch_pre_r(const MgLev& v, const MgLev& vc, const MgCheb& ch, const MgScal* S,
                     hipStream_t s) {
    hipLaunchKernelGGL((k_mg_pre<K, NW, R>), dim3(mg_tiles(v, K, NW, R)), dim3(64 * NW), 0, s,
                       v, ch, vc, S);
}

template <int K, int NW, int R>
void mg_launch_post_r(const MgLev& v, const MgLev& vc, const MgCheb& ch, const MgScal* S,
                      hipStream_t s) {
    hipLaunchKernelGGL((k_mg_post<K, NW, R>), dim3(mg_tiles(v, K, NW, R)), dim3(64 * NW), 0, s,
                       v, vc, ch, S);
}

template <int K>
void mg_launch_pre(const MgLev& v, const MgLev& vc, const MgCheb& ch, const MgScal* S,
                   hipStream_t s) {
    if (mg_large_level(v, K))
        mg_launch_pre_r<K, MG_NW_LARGE, MG_RPT_LARGE>(v, vc, ch, S, s);
    else
        mg_launch_pre_r<K, MG_NW_SMALL, mg_rpt_small(K)>(v, vc, ch, S, s);
}

template <int K>
void mg_launch_post(const MgLev& v, const MgLev& vc, const MgCheb& ch, const MgScal* S,
                    hipStream_t s) {
    if (mg_large_level(v, K))
        mg_launch_post_r<K, MG_NW_LARGE, MG_RPT_LARGE>(v, vc, ch, S, s);
    else
        mg_launch_post_r<K, MG_NW_SMALL, mg_rpt_small(K)>(v, vc, ch, S, s);
}

// A coarse level's kernel at degree Kc.
template <int Kc>
void mg_launch_coarse(bool pre, const MgLev& v, const MgLev& vc, const MgCheb& ch,
                      const MgScal* S, hipStream_t s) {
    if (pre)
        mg_launch_pre<Kc>(v, vc, ch, S, s);
    else
        mg_launch_post<Kc>(v, vc, ch, S, s);
}

void mg_launch_coarse(int Kc, bool pre, const MgLev& v, const MgLev& vc, const MgCheb& ch,
                      const MgScal* S, hipStream_t s) {
    switch (Kc) {
        case 1: mg_launch_coarse<1>(pre, v, vc, ch, S, s); break;
        case 2: mg_launch_coarse<2>(pre, v, vc, ch, S, s); break;
        case 3: mg_launch_coarse<3>(pre, v, vc, ch, S, s); break;
        case 4: mg_launch_coarse<4>(pre, v, vc, ch, S, s); break;
        case 5: mg_launch_coarse<5>(pre, v, vc, ch, S, s); break;
        case 6: mg_launch_coarse<6>(pre, v, vc, ch, S, s); break;
        case 7: mg_launch_coarse<7>(pre, v, vc, ch, S, s); break;
        case 8: mg_launch_coarse<8>(pre, v, vc, ch, S, s); break;
        case 9: mg_launch_coarse<9>(pre, v, vc, ch, S, s); break;
        default: mg_launch_coarse<10>(pre, v, vc, ch, S, s); break;
    }
}

// Coarse smoothing of a V-cycle: degree Kc with scalars chc on the coarse
// levels (mg_coarse_level); Kc <= 0 keeps the finest levels' degree.
struct MgCoarse {
    int Kc = 0;
    MgCheb chc{};
    const MgScal* S = nullptr;   // scalars the kernels test (null: the solve's, m->S)
};

template <int K>
void mg_vcycle_fused(MgState* m, hipStream_t s, const MgCheb& ch, const MgCoarse& co,
                     MgEvents* ev) {
    const int L = m->L;
    if (ev && L == 0) {   // no smoothing level: empty intervals
        ev->rec(0, s);
        ev->rec(1, s);
    }
    auto coarse = [&](int l) { return co.Kc > 0 && l > 0 && mg_coarse_level(m->lev[l]); };
    const MgScal* S = co.S ? co.S : (const MgScal*)m->S;
    for (int l = 0; l < L; ++l) {
        if (ev && l == 0) ev->rec(0, s);
        if (coarse(l))
            mg_launch_coarse(co.Kc, true, m->lev[l], m->lev[l + 1], co.chc,
                             S, s);
        else
            mg_launch_pre<K>(m->lev[l], m->lev[l + 1], ch, S, s);
        if (ev && l == 0) ev->rec(1, s);
    }
    if (m->ncoarse > 0)
        hipLaunchKernelGGL(k_mg_dense, dim3((m->ncoarse + 3) / 4), dim3(256), 0, s, m->lev[L],
                           (const float*)m->Ainv, m->dense_ld, (const int*)m->cells, m->ncoarse,
                           S);
    if (ev && L == 0) ev->rec(2, s);
    for (int l = L - 1; l >= 0; --l) {
        if (ev && l == 0) ev->rec(2, s);
        if (coarse(l))
            mg_launch_coarse(co.Kc, false, m->lev[l], m->lev[l + 1], co.chc,
                             S, s);
        else
            mg_launch_post<K>(m->lev[l], m->lev[l + 1], ch, S, s);
    }
    if (ev) ev->rec(3, s);
}

// ev: null, or the 7 events of MgState::kev (the first 4 are recorded here)
void mg_vcycle(MgState* m, hipStream_t s, int K, const MgCheb& ch, MgEvents* ev = nullptr,
               const MgCoarse& co = MgCoarse{}) {
    switch (K) {
        case 1: mg_vcycle_fused<1>(m, s, ch, co, ev); break;
        case 2: mg_vcycle_fused<2>(m, s, ch, co, ev); break;
        case 3: mg_vcycle_fused<3>(m, s, ch, co, ev); break;
        case 4: mg_vcycle_fused<4>(m, s, ch, co, ev); break;
        case 5: mg_vcycle_fused<5>(m, s, ch, co, ev); break;
        default: mg_vcycle_fused<6>(m, s, ch, co, ev); break;
    }
}

// One outer iteration on stream s: V-cycle, pass A, (fold,) pass B, fold;
// with ev, the kernel-timing events of MgState::kev recorded around them.
// On even-width grids pass B folds pass A's partials itself.
void mg_iteration(psi_ctx* c, hipStream_t s, int K, const MgCheb& ch, MgEvents* ev = nullptr,
                  const MgCoarse& co = MgCoarse{}) {
    MgState* m = c->mg;
    mg_vcycle(m, s, K, ch, ev, co);
    const Geom& g = c->g;
    if (!(g.nx & 1)) {
        hipLaunchKernelGGL(k_mg_fcg_a, dim3(m->nblk_a), dim3(MG_PASS_A_THREADS), 0, s, g,
                           m->lev[0], (const double*)c->E, (const double*)c->N,
                           (const double*)c->rx, (const double*)c->ry, (const double*)c->dxu,
                           (const double*)c->cosu, (const double*)c->dyu,
                           (const uint8_t*)c->act, (const double*)m->q, (const double*)m->r,
                           m->az, (const double*)m->zscale, (const MgScal*)m->S, m->part_a);
        if (ev) ev->rec(4, s);
        if (ev) ev->rec(5, s);
        hipLaunchKernelGGL(k_mg_fcg_b<true>, dim3(m->nblk), dim3(256), 0, s, g, m->lev[0],
                           (const double*)c->dxu, (const double*)c->cosu, (const double*)c->dyu,
                           (const uint8_t*)c->act, m->p, m->q, m->r, m->x, (const double*)m->az,
                           (const double*)m->zscale, m->S, (const double*)m->part_a, m->nblk_a,
                           m->part);
    } else {
        hipLaunchKernelGGL(k_mg_fcg_a1, dim3(m->nblk), dim3(256), 0, s, g, m->lev[0],
                           (const double*)c->E, (const double*)c->N, (const double*)c->rx,
                           (const double*)c->ry, (const double*)c->dxu, (const double*)c->cosu,
                           (const double*)c->dyu, (const uint8_t*)c->act,
                           (const double*)m->q, (const double*)m->r, m->az,
                           (const double*)m->zscale, (const MgScal*)m->S, m->part);
        if (ev) ev->rec(4, s);
        hipLaunchKernelGGL(k_mg_fold_a, dim3(1), dim3(1024), 0, s, (const double*)m->part,
                           m->nblk, m->S);
        if (ev) ev->rec(5, s);
        mg_launch_pass_b(g, m->nblk, s, m->lev[0], c->dxu, c->cosu, c->dyu, c->act, m->p, m->q,
                         m->r, m->x, m->az, m->zscale, m->S, m->part);
    }
    if (ev) ev->rec(6, s);
    hipLaunchKernelGGL(k_mg_fold_b, dim3(1), dim3(1024), 0, s, (const double*)m->part, m->nblk,
                       m->nact0, m->S, m->zscale);
}

// The captured outer iteration (timed: with the kernel-timing events).
// Capture one outer iteration, iter(stream, cheb, events), into the graph
// (timed: the copy with the kernel-timing events) unless it is current.
template <class Iter>
int mg_capture(psi_ctx* c, MgState* m, int K, double ratio, bool timed, Iter iter, int Kc = 0) {
    hipGraph_t& graph = timed ? m->graph_t : m->graph;
    hipGraphExec_t& exec = timed ? m->exec_t : m->exec;
    int& gK = timed ? m->graph_t_K : m->graph_K;
    double& gratio = timed ? m->graph_t_ratio : m->graph_ratio;
    int& gKc = timed ? m->graph_t_Kc : m->graph_Kc;
    if (exec && gK == K && gratio == ratio && gKc == Kc) return PSI_OK;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    const MgCheb ch = mg_cheb(K, ratio, m->cscale);
    HIPCHK(c, hipStreamBeginCapture(m->gs, hipStreamCaptureModeThreadLocal));
    MgEvents ev{m->kev};
    iter(m->gs, ch, timed ? &ev : nullptr);
    hipGraph_t cap = nullptr;
    HIPCHK(c, hipStreamEndCapture(m->gs, &cap));
    graph = cap;
    HIPCHK(c, ev.err);
    HIPCHK(c, hipGraphInstantiate(&exec, cap, nullptr, nullptr, 0));
    gK = K;
    gratio = ratio;
    gKc = Kc;
    return PSI_OK;
}

// Kc: the coarse levels' smoothing degree (0: K)
int mg_graph(psi_ctx* c, int K, double ratio, bool timed = false, int Kc = 0) {
    MgCoarse co;
    co.Kc = Kc;
    if (Kc > 0) co.chc = mg_cheb(Kc, ratio, c->mg->cscale);
    return mg_capture(
        c, c->mg, K, ratio, timed,
        [&](hipStream_t s, const MgCheb& ch, MgEvents* ev) { mg_iteration(c, s, K, ch, ev, co); },
        Kc);
}

// What the legs of a multigrid solve need from its operator: where its
// FP64 residual kernels leave the norms {|r|^2, |b|^2, max|r|, count}, and
// the psi-cell areas that weight the V-cycle input.
struct MgOperator {
    const double* res_dev;
    const double *dxu, *cosu, *dyu;
};

// Replay the outer-iteration graph on m->gs from iteration n_start until the
// device flags the end, polling the scalars and the leg residual's norms
// after every batch (the leg's starting |r|^2 is read at the first poll).
// Replays after the end still launch their kernels (they test the flag
// beside their first loads), so a batch is cut to the iterations the
// observed contraction of |r|^2 says remain, and such a batch, expected to be
// the leg's last, is followed by resid(m->gs, nullptr) (the FP64 true
// residual of x) and k_mg_confirm, so the poll that sees the end also sees
// it confirmed.  timed: the last replay of the first batch is the timed
// graph, whose kernel durations are added to info (its event nodes cost a
// few us each, so later batches replay the plain graph).
template <class Resid>
int mg_replay(psi_ctx* c, MgState* m, const MgOperator& op, Resid resid, int batch, bool timed,
              int n_start, psi_info* info) {
    int n_poll = n_start, next = batch;
    double rr_poll = -1.0;
    for (;; timed = false) {
        for (int k = 0; k < next; ++k)
            HIPCHK(c, hipGraphLaunch(timed && k == next - 1 ? m->exec_t : m->exec, m->gs));
        const bool last = next < batch;
        if (last) {
            resid(m->gs, nullptr);
            hipLaunchKernelGGL(k_mg_confirm, dim3(1), dim3(64), 0, m->gs, op.res_dev, m->S);
            HIPCHK(c, hipGetLastError());
            info->setup_launches += 1;
        }
        HIPCHK(c, hipMemcpyAsync(m->h_S, m->S, sizeof(MgScal), hipMemcpyDeviceToHost, m->gs));
        HIPCHK(c, hipMemcpyAsync(m->h_res, op.res_dev, 4 * sizeof(double),
                                 hipMemcpyDeviceToHost, m->gs));
        HIPCHK(c, hipStreamSynchronize(m->gs));
        const MgScal& h = *m->h_S;
        if (rr_poll < 0.0) rr_poll = m->h_res[0];   // first poll: never after a confirmation
        if (timed && h.n - n_poll == next) {
            // every replay of the batch iterated, so the events hold the last
            // one's kernels
            const int pairs[4][2] = {{0, 1}, {2, 3}, {3, 4}, {5, 6}};
            for (int k = 0; k < 4; ++k) {
                float ms = 0.0f;
                HIPCHK(c, hipEventElapsedTime(&ms, m->kev[pairs[k][0]], m->kev[pairs[k][1]]));
                info->kernel_us[k] += 1e3 * ms;
            }
            info->kernel_samples += 1;
        }
        if (h.done) return PSI_OK;
        next = batch;
        if (h.n > n_poll && rr_poll > 0.0 && h.rr > 0.0 && h.rr < rr_poll) {
            const double rate = std::pow(h.rr / rr_poll, 1.0 / (h.n - n_poll));
            const double need = std::log(h.rtarget2 * h.bb / h.rr) / std::log(rate);
            if (need > 0.0 && need < batch) next = std::max(1, (int)std::ceil(need));
        }
        n_poll = h.n;
        rr_poll = h.rr;
    }
}

// The legs of a multigrid solve from m->x (which holds the first guess).  A
// leg starts on the device: resid(stream, m->r) (the FP64 residual into r and
// op.res_dev), k_mg_init (state and end tests) and k_mg_start (V-cycle
// input), so its first batch of replays queues behind them without a host
// round trip.  A leg that ends on the recursive residual is confirmed by the
// FP64 true residual, after its last batch or here when the batch that ended
// it was not expected to; else a new leg starts from the true residual.
// Out: res (the true residual's norms), bb, the iterations n, the status and
// the legs that iterated.
template <class Resid>
int mg_run_legs(psi_ctx* c, MgState* m, const MgOperator& op, Resid resid, int batch,
                bool timed, double target, int max_it, psi_info* info, double res[4], double& bb,
                int& n, int& status, int& legs) {
    const Geom& g = c->g;
    n = 0;
    legs = 0;
    status = ST_RUNNING;
    auto sync_resid = [&]() -> int {
        resid(c->stream, nullptr);
        int rc = check_launch(c);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(m->h_res, op.res_dev, 4 * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < 4; ++k) res[k] = m->h_res[k];
        info->setup_launches += 1;
        return PSI_OK;
    };
    for (int leg = 0;; ++leg) {
        resid(c->stream, m->r);
        info->setup_launches += 1;
        hipLaunchKernelGGL(k_mg_init, dim3(1), dim3(64), 0, c->stream, op.res_dev, m->S,
                           m->zscale, leg == 0 ? 1 : 0, n, max_it, target * target, m->w_mean);
        hipLaunchKernelGGL(k_mg_start, dim3(nblk((size_t)g.nx * g.ny)), dim3(256), 0, c->stream,
                           g, m->lev[0], op.dxu, op.cosu, op.dyu, (const uint8_t*)c->act,
                           (const double*)m->r, m->p, m->q, (const double*)m->zscale,
                           (const MgScal*)m->S);
        int rc = check_launch(c);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(m->e_in, c->stream));
        HIPCHK(c, hipStreamWaitEvent(m->gs, m->e_in, 0));
        const int n_start = n;
        HIPCHK(c, hipEventRecord(m->t0, m->gs));
        rc = mg_replay(c, m, op, resid, batch, timed && leg == 0, n_start, info);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(m->t1, m->gs));
        HIPCHK(c, hipEventSynchronize(m->t1));
        float span = 0.0f;
        (void)hipEventElapsedTime(&span, m->t0, m->t1);
        info->sweep_ms += span;                          // graph replays of this leg
        HIPCHK(c, hipEventRecord(m->e_out, m->gs));
        HIPCHK(c, hipStreamWaitEvent(c->stream, m->e_out, 0));
        const MgScal h = *m->h_S;
        info->sweep_launches += h.n - n_start;           // outer iterations they ran
        if (h.n > n_start) ++legs;
        n = h.n;
        status = h.status;
        bb = h.bb;
        if (status != ST_CONVERGED) return sync_resid();   // the final iterate's
        if (h.confirmed == 1) {   // by the norms in op.res_dev (leg start or confirmation)
            for (int k = 0; k < 4; ++k) res[k] = m->h_res[k];
            return PSI_OK;
        }
        if (h.confirmed == 0) {
            rc = sync_resid();
            if (rc) return rc;
            if (res[0] <= target * target * bb) return PSI_OK;
        }
        // the true residual is above the target: a new leg from it
    }
}

// Average duration (us) of `reps` back-to-back launches queued by fn on the
// multigrid stream, timed with HIP events.
template <class Fn>
double mg_time(MgState* m, int reps, Fn fn) {
    if (hipEventRecord(m->t0, m->gs) != hipSuccess) return -1.0;
    for (int k = 0; k < reps; ++k) fn();
    if (hipEventRecord(m->t1, m->gs) != hipSuccess || hipEventSynchronize(m->t1) != hipSuccess)
        return -1.0;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->t0, m->t1);
    return 1e3 * ms / reps;
}

template <int K>
void mg_time_levels(MgState* m, int reps, double* us) {
    const MgCheb ch = mg_cheb(K, 15.0, m->cscale);
    const MgLev& v = m->lev[0];
    const MgScal* S = m->S_time;
    us[0] = mg_time(m, reps, [&] { mg_launch_post<K>(v, m->lev[1], ch, S, m->gs); });
    us[1] = mg_time(m, reps, [&] { mg_launch_pre<K>(v, m->lev[1], ch, S, m->gs); });
}

int solve_mg(psi_ctx* c, const psi_opts* o, const double* forc, double* sol, psi_info* info) {
    MgState* m = c->mg;
    if (!m) return fail(c, PSI_BAD_ARG, "multigrid mode: call psi_mg_setup first%s");
    if (o->criterion != PSI_CRIT_RELRES)
        return fail(c, PSI_BAD_ARG, "multigrid mode needs the relres criterion%s");
    // defaults: 5 Chebyshev steps on [2/15, 2] (with the coarse-correction
    // weight 1.9, a sweep of 3-6 steps and intervals 1/10-1/30 at 2048^2 put
    // this ahead)
    const int K = o->poly_degree > 0 ? o->poly_degree : 5;
    if (K > MG_MAXK) return fail(c, PSI_BAD_ARG, "multigrid smoothing degree 1..6%s");
    const int Kc = o->mg_coarse_degree > 0 ? o->mg_coarse_degree : PSI_MG_COARSE_DEGREE;
    if (Kc > MG_MAXKC) return fail(c, PSI_BAD_ARG, "multigrid coarse smoothing degree 1..10%s");
    const double lower = o->poly_lower > 0 ? o->poly_lower : 1.0 / 15.0;
    if (!(lower < 1.0)) return fail(c, PSI_BAD_ARG, "multigrid interval: lower < 1%s");
    const double ratio = 1.0 / lower;
    const double target = o->relres_target > 0 ? o->relres_target : 1e-10;
    // replays per poll: 8 measured fastest at 2048^2 (each poll idles the GPU
    // ~20 us; batches are cut near the end, so a long one wastes little)
    const int batch = o->check_every > 0 ? o->check_every : 8;
    // with time_kernels the last replay of the first poll batch runs the
    // timed graph (its event nodes cost about 3 us each)
    const bool timed = o->time_kernels != 0;
    int rc = mg_graph(c, K, ratio, false, Kc == K ? 0 : Kc);
    if (rc == PSI_OK && timed) rc = mg_graph(c, K, ratio, true, Kc == K ? 0 : Kc);
    if (rc) return rc;
    const Geom& g = c->g;
    HIPCHK(c, hipMemcpyAsync(m->x, sol, c->plane * 8, hipMemcpyDeviceToDevice, c->stream));
    double res[4] = {0.0, 0.0, 0.0, 0.0};
    double bb = 0.0;
    int n = 0, status = ST_RUNNING, legs = 0;
    const MgOperator op{c->result, c->dxu, c->cosu, c->dyu};
    rc = mg_run_legs(
        c, m, op,
        [&](hipStream_t s, double* r_out) { residual_launch(c, s, forc, m->x, r_out); }, batch,
        timed, target, o->max_iterations, info, res, bb, n, status, legs);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(sol, m->x, c->plane * 8, hipMemcpyDeviceToDevice, c->stream));
    if (info->kernel_samples > 0)
        for (int k = 0; k < 4; ++k) info->kernel_us[k] /= info->kernel_samples;
    info->iterations = status == ST_MAXIT ? o->max_iterations + 1 : n;
    info->status = status_of(status);
    info->converged = status == ST_CONVERGED;
    info->outer_iterations = legs;
    info->relres_true = bb > 0 ? sqrt(res[0] / bb) : sqrt(res[0]);
    info->relres_recursive = bb > 0 ? sqrt(m->h_S->rr / bb) : 0.0;
    return info->status;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* psi_ctx_last_error(const psi_ctx* c) {
    return c ? c->err.c_str() : g_create_err.c_str();
}

int psi_ctx_nisle(const psi_ctx* c) { return c ? c->nisle : 0; }
int64_t psi_ctx_device_bytes(const psi_ctx* c) {
    return c ? c->bytes + (c->mg ? c->mg->bytes : 0) : 0;
}

int psi_mg_setup(psi_ctx* c) {
    if (!valid_field_ctx(c)) return PSI_BAD_ARG;
    return mg_setup(c);
}

int psi_mg_kernel_times(psi_ctx* c, int32_t degree, int32_t reps, double* us, int32_t n) {
    if (!valid_field_ctx(c) || !us || n < 4 || reps < 1) return PSI_BAD_ARG;
    MgState* m = c->mg;
    if (!m) return fail(c, PSI_BAD_ARG, "call psi_mg_setup first%s");
    if (m->L < 1) return fail(c, PSI_BAD_ARG, "multigrid: a single level has no V-cycle kernels%s");
    const int K = degree > 0 ? degree : 5;   // the solve's default
    if (K > MG_MAXK) return fail(c, PSI_BAD_ARG, "multigrid smoothing degree 1..6%s");
    MgScal z{};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(m->S_time, &z, sizeof z, hipMemcpyHostToDevice));
    switch (K) {
        case 1: mg_time_levels<1>(m, reps, us); break;
        case 2: mg_time_levels<2>(m, reps, us); break;
        case 3: mg_time_levels<3>(m, reps, us); break;
        case 4: mg_time_levels<4>(m, reps, us); break;
        case 5: mg_time_levels<5>(m, reps, us); break;
        default: mg_time_levels<6>(m, reps, us); break;
    }
    const Geom& g = c->g;
    us[2] = mg_time(m, reps, [&] {
        if (g.nx & 1)
            hipLaunchKernelGGL(k_mg_fcg_a1, dim3(m->nblk), dim3(256), 0, m->gs, g, m->lev[0],
                               (const double*)c->E, (const double*)c->N, (const double*)c->rx,
                               (const double*)c->ry, (const double*)c->dxu,
                               (const double*)c->cosu, (const double*)c->dyu,
                               (const uint8_t*)c->act, (const double*)m->q, (const double*)m->r,
                               m->az, (const double*)m->zscale, (const MgScal*)m->S_time,
                               m->part);
        else
            hipLaunchKernelGGL(k_mg_fcg_a, dim3(m->nblk_a), dim3(MG_PASS_A_THREADS), 0, m->gs, g,
                               m->lev[0], (const double*)c->E, (const double*)c->N,
                               (const double*)c->rx, (const double*)c->ry,
                               (const double*)c->dxu, (const double*)c->cosu,
                               (const double*)c->dyu, (const uint8_t*)c->act,
                               (const double*)m->q, (const double*)m->r, m->az,
                               (const double*)m->zscale, (const MgScal*)m->S_time, m->part_a);
    });
    // pass B with its scalars from S_time (the in-solve pass B also folds pass A's partials)
    us[3] = mg_time(m, reps, [&] {
        mg_launch_pass_b(g, m->nblk, m->gs, m->lev[0], c->dxu, c->cosu, c->dyu, c->act, m->p, m->q,
                         m->r, m->x, m->az, m->zscale, m->S_time, m->part);
    });
    HIPCHK(c, hipGetLastError());
    for (int k = 0; k < 4; ++k)
        if (us[k] < 0) return fail(c, PSI_HIP_ERROR, "multigrid kernel timing failed%s");
    return PSI_OK;
}

int psi_mg_levels(const psi_ctx* c, int32_t* nlevels, int32_t* nx, int32_t* ny, int32_t* nact,
                  int32_t cap) {
    if (!c || !nlevels) return PSI_BAD_ARG;
    if (!c->mg) {
        *nlevels = 0;
        return PSI_OK;
    }
    const int n = (int)c->mg->level_nx.size();
    *nlevels = n;
    for (int k = 0; k < n && k < cap; ++k) {
        if (nx) nx[k] = c->mg->level_nx[k];
        if (ny) ny[k] = c->mg->level_ny[k];
        if (nact) nact[k] = c->mg->level_nact[k];
    }
    return PSI_OK;
}

// Frees what a context holds (device memory, pinned buffers, events), not the
// object itself.
static void ctx_release(psi_ctx* c) {
    mg_release(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->h_st) (void)hipHostFree(c->h_st);
    if (c->h_res) (void)hipHostFree(c->h_res);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->t0) (void)hipEventDestroy(c->t0);
    if (c->t1) (void)hipEventDestroy(c->t1);
    if (c->ts0) (void)hipEventDestroy(c->ts0);
    for (auto& e : c->ts1)
        if (e) (void)hipEventDestroy(e);
    c->allocs.clear();
}

void psi_ctx_destroy(psi_ctx* c) {
    if (!c) return;
    ctx_release(c);
    delete c;
}

int psi_ctx_create(psi_ctx** out, const psi_grid* grid, const psi_fields* f,
                   const psi_islands* isl, void* stream) {
    if (!out || !grid || !f || !f->hur || !f->hvr) return fail(nullptr, PSI_BAD_ARG, "null argument%s");
    *out = nullptr;
    if (grid->nx < 2 || grid->ny < 2 || !grid->dxt || !grid->dxu || !grid->dyt || !grid->dyu ||
        !grid->cost || !grid->cosu)
        return fail(nullptr, PSI_BAD_ARG, "bad grid%s");
    psi_ctx* c = new psi_ctx();
    c->stream = (hipStream_t)stream;
    (void)hipGetDevice(&c->device);
    Geom& g = c->g;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.pitch = grid->nx + 4;
    g.cyclic = grid->cyclic_x ? 1 : 0;
    g.tiles_x = (g.nx + TX - 1) / TX;
    g.tiles_y = (g.ny + TY - 1) / TY;
    g.ntiles = g.tiles_x * g.tiles_y;
    c->plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t nxh = g.nx + 4, nyh = g.ny + 4;
    c->h_dxt.assign(grid->dxt, grid->dxt + nxh);
    c->h_dxu.assign(grid->dxu, grid->dxu + nxh);
    c->h_dyt.assign(grid->dyt, grid->dyt + nyh);
    c->h_dyu.assign(grid->dyu, grid->dyu + nyh);
    c->h_cost.assign(grid->cost, grid->cost + nyh);
    c->h_cosu.assign(grid->cosu, grid->cosu + nyh);
    c->hur = f->hur;
    c->hvr = f->hvr;
    c->maskU = f->maskU;
    c->maskV = f->maskV;
    int rc = PSI_OK;
#define TRY(x)                 \
    do {                       \
        rc = (x);              \
        if (rc) goto failed;   \
    } while (0)
    {
        TRY(dalloc(c, &c->dxt, nxh));
        TRY(dalloc(c, &c->dxu, nxh));
        TRY(dalloc(c, &c->dyt, nyh));
        TRY(dalloc(c, &c->dyu, nyh));
        TRY(dalloc(c, &c->cost, nyh));
        TRY(dalloc(c, &c->cosu, nyh));
        TRY(dalloc(c, &c->rx, nxh));
        TRY(dalloc(c, &c->ry, nyh));
        TRY(dalloc(c, &c->rx32, nxh));
        TRY(dalloc(c, &c->ry32, nyh));
        TRY(dalloc(c, &c->E, c->plane));
        TRY(dalloc(c, &c->N, c->plane));
        TRY(dalloc(c, &c->cC, c->plane));
        TRY(dalloc(c, &c->E32, c->plane));
        TRY(dalloc(c, &c->N32, c->plane));
        TRY(dalloc(c, &c->act, c->plane));
        TRY(dalloc(c, &c->dirichlet, c->plane));
        for (int k = 0; k < 2; ++k) {
            TRY(dalloc(c, &c->u64[k], c->plane));
            TRY(dalloc(c, &c->p64[k], c->plane));
            TRY(dalloc(c, &c->u32[k], c->plane));
            TRY(dalloc(c, &c->p32[k], c->plane));
        }
        TRY(dalloc(c, &c->e32, c->plane));
        TRY(dalloc(c, &c->rp32[0], c->plane));
        TRY(dalloc(c, &c->rp32[1], c->plane));
        TRY(dalloc(c, &c->r64, c->plane));
        TRY(dalloc(c, &c->r32, c->plane));
        TRY(dalloc(c, &c->z32, c->plane));
        TRY(dalloc(c, &c->q32, c->plane));
        TRY(dalloc(c, &c->pf32, c->plane));
        TRY(dalloc(c, &c->fcg, 16));
        TRY(dalloc(c, &c->tmp64, c->plane));
        TRY(dalloc(c, &c->tmp64b, c->plane));
        TRY(dalloc(c, &c->fpx3, c->plane));
        TRY(dalloc(c, &c->fpy3, c->plane));
        TRY(dalloc(c, &c->dz, 2 * (size_t)LV_MAX_NZ));
        TRY(dalloc(c, &c->part, (size_t)NRED * (g.ntiles + 1024)));
        {
            // resident grid: CUs x workgroups per CU of the FP64 sweep (the
            // FP32 kernels fit at least as many), a multiple of 8 (XCDs)
            int ncu = 0, occ64 = 0, occ32 = 0;
            if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
                    hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ64, k_cg_sweep<double>, NT, 0) !=
                    hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ32, k_cg_sweep<float>, NT, 0) !=
                    hipSuccess) {
                rc = fail(c, PSI_HIP_ERROR, "device query failed%s");
                goto failed;
            }
            auto grid_for = [&](int occ) {
                const char* ov = getenv("PSI_GRID_PER_CU");
                if (ov && atoi(ov) > 0) occ = std::min(occ, atoi(ov));
                int G = std::max(1, ncu * std::max(occ, 1));
                if (G > g.ntiles) G = g.ntiles;
                if (G >= 8) G &= ~7;
                return G;
            };
            c->ncu = ncu;
            c->G = grid_for(occ64);
            c->G32 = grid_for(occ32);
            choose_march(c);
        }
        TRY(dalloc(c, &c->part2, (size_t)2 * NRED * c->Gpart));
        TRY(dalloc(c, &c->result, 8));
        TRY(dalloc(c, &c->counter, 4));
        TRY(dalloc(c, &c->st, 2));
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(c, PSI_HIP_ERROR, "allocation failed%s");
            goto failed;
        }
        if (hipHostMalloc((void**)&c->h_st, 4 * sizeof(CgState), 0) != hipSuccess ||
            hipHostMalloc((void**)&c->h_res, 8 * sizeof(double), 0) != hipSuccess) {
            rc = fail(c, PSI_HIP_ERROR, "hipHostMalloc failed%s");
            goto failed;
        }
        memset(c->h_st, 0, 4 * sizeof(CgState));
        for (auto& e : c->ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                rc = fail(c, PSI_HIP_ERROR, "event create failed%s");
                goto failed;
            }
        if (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess ||
            hipEventCreate(&c->ts0) != hipSuccess || hipEventCreate(&c->ts1[0]) != hipSuccess ||
            hipEventCreate(&c->ts1[1]) != hipSuccess) {
            rc = fail(c, PSI_HIP_ERROR, "event create failed%s");
            goto failed;
        }
        // 1-D metric ratios: W(i,j) = E(i-1,j) * dxu(i-1)/dxu(i);
        // S(i,j) = N(i,j-1) * dyu(j-1) cosu(j-1) / (dyu(j) cosu(j)).
        std::vector<double> rx(nxh, 0.0), ry(nyh, 0.0);
        std::vector<float> rxf(nxh), ryf(nyh);
        for (size_t k = 1; k < nxh; ++k) rx[k] = c->h_dxu[k - 1] / c->h_dxu[k];
        for (size_t k = 1; k < nyh; ++k)
            ry[k] = c->h_dyu[k - 1] * c->h_cosu[k - 1] / (c->h_dyu[k] * c->h_cosu[k]);
        for (size_t k = 0; k < nxh; ++k) rxf[k] = (float)rx[k];
        for (size_t k = 0; k < nyh; ++k) ryf[k] = (float)ry[k];
        if (hipMemcpy(c->dxt, grid->dxt, nxh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->dxu, grid->dxu, nxh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->dyt, grid->dyt, nyh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->dyu, grid->dyu, nyh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->cost, grid->cost, nyh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->cosu, grid->cosu, nyh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->rx, rx.data(), nxh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->ry, ry.data(), nyh * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->rx32, rxf.data(), nxh * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->ry32, ryf.data(), nyh * 4, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(c, PSI_HIP_ERROR, "metric upload failed%s");
            goto failed;
        }
        // islands
        if (isl && isl->nisle > 0) {
            if (!isl->offsets || !isl->bi || !isl->bj || !isl->di || !isl->dj) {
                rc = fail(c, PSI_BAD_ARG, "island lists incomplete%s");
                goto failed;
            }
            c->nisle = isl->nisle;
            c->offs.assign(isl->offsets, isl->offsets + isl->nisle + 1);
            const int np = c->offs.back();
            if (np > 0) {
                TRY(dalloc(c, &c->bi, np));
                TRY(dalloc(c, &c->bj, np));
                TRY(dalloc(c, &c->di, np));
                TRY(dalloc(c, &c->dj, np));
                if (hipMemcpy(c->bi, isl->bi, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(c->bj, isl->bj, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(c->di, isl->di, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(c->dj, isl->dj, np * 4, hipMemcpyHostToDevice) != hipSuccess) {
                    rc = fail(c, PSI_HIP_ERROR, "island upload failed%s");
                    goto failed;
                }
                hipLaunchKernelGGL(k_mark_points, dim3(nblk(np)), dim3(256), 0, c->stream, g, np,
                                   c->bi, c->bj, c->dirichlet);
                TRY(check_launch(c));
            }
        }
        TRY(build_operator(c));
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(c, PSI_HIP_ERROR, "setup failed%s");
            goto failed;
        }
        *out = c;
        return PSI_OK;
    }
failed:
    g_create_err = c->err;
    psi_ctx_destroy(c);
    return rc;
#undef TRY
}

int psi_ctx_update_coeffs(psi_ctx* c, const double* hur, const double* hvr) {
    if (!valid_field_ctx(c) || !hur || !hvr) return PSI_BAD_ARG;
    c->hur = hur;
    c->hvr = hvr;
    mg_release(c);   // the hierarchy belongs to the old operator
    int rc = build_operator(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_ctx_set_stream(psi_ctx* c, void* stream) {
    if (!c) return PSI_BAD_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)stream;
    return PSI_OK;
}

int psi_congrad(psi_ctx* c, const psi_opts* o, const double* forc, double* sol, psi_info* info) {
    if (!valid_field_ctx(c) || !o || !forc || !sol || !info) return PSI_BAD_ARG;
    if (o->max_iterations < 1) return fail(c, PSI_BAD_ARG, "max_iterations < 1%s");
    memset(info, 0, sizeof *info);
    HIPCHK(c, hipEventRecord(c->t0, c->stream));
    if (c->g.cyclic)
        hipLaunchKernelGGL(k_setcyclic, dim3(nblk(c->g.ny + 4)), dim3(256), 0, c->stream, c->g, sol);
    int rc;
    switch (o->mode) {
        case PSI_MODE_PLAIN: rc = solve_plain(c, o, forc, sol, info); break;
        case PSI_MODE_MIXED: rc = solve_mixed(c, o, forc, sol, info); break;
        case PSI_MODE_POLY: rc = solve_poly(c, o, forc, sol, info); break;
        case PSI_MODE_MG: rc = solve_mg(c, o, forc, sol, info); break;
        case PSI_MODE_LEARNED: return fail(c, PSI_NO_LEARNED_MODEL, "no learned model loaded%s");
        default: return fail(c, PSI_BAD_ARG, "unknown mode%s");
    }
    if (rc == PSI_HIP_ERROR || rc == PSI_BAD_ARG) return rc;
    HIPCHK(c, hipEventRecord(c->t1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->t1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    info->elapsed_ms = ms;
    return info->status;
}

int psi_apply_op(psi_ctx* c, const double* p, double* res) {
    if (!valid_field_ctx(c) || !p || !res) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_zero_halo, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g, res);
    hipLaunchKernelGGL(k_apply_op, dim3(c->g.ntiles), dim3(NT), 0, c->stream, c->g, c->E, c->N,
                       c->rx, c->ry, p, res);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_make_coeff(psi_ctx* c, double* cf) {
    if (!valid_field_ctx(c) || !cf) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_make_coeff_ref, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g, c->hur,
                       c->hvr, c->dxt, c->dxu, c->dyt, c->dyu, c->cost, c->cosu, cf);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_make_inv(psi_ctx* c, double* Z) {
    if (!valid_field_ctx(c) || !Z) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_make_inv, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g, c->cC,
                       c->act, Z);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

static int dot_absmax(psi_ctx* c, const double* a, const double* b, double* out2) {
    const int blocks = 1024;
    hipLaunchKernelGGL(k_dot_absmax, dim3(blocks), dim3(NT), 0, c->stream, c->g, a, b, c->part,
                       c->counter, c->result);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->h_res, c->result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out2[0] = c->h_res[0];
    out2[1] = c->h_res[1];
    return PSI_OK;
}

int psi_dot(psi_ctx* c, const double* a, const double* b, double* out) {
    if (!valid_field_ctx(c) || !a || !b || !out) return PSI_BAD_ARG;
    double r[2];
    int rc = dot_absmax(c, a, b, r);
    if (rc) return rc;
    *out = r[0];
    return PSI_OK;
}

int psi_absmax(psi_ctx* c, const double* a, double* out) {
    if (!valid_field_ctx(c) || !a || !out) return PSI_BAD_ARG;
    double r[2];
    int rc = dot_absmax(c, a, nullptr, r);
    if (rc) return rc;
    *out = r[1];
    return PSI_OK;
}

int psi_setcyclic(psi_ctx* c, double* a) {
    if (!valid_field_ctx(c) || !a) return PSI_BAD_ARG;
    if (!c->g.cyclic) return PSI_OK;
    hipLaunchKernelGGL(k_setcyclic, dim3(nblk(c->g.ny + 4)), dim3(256), 0, c->stream, c->g, a);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

static int line_integral_async(psi_ctx* c, int isle, const double* uloc, const double* vloc,
                               double* dev_out2) {
    const int k = isle - 1;
    const int s = c->offs[k], n = c->offs[k + 1] - s;
    hipLaunchKernelGGL(k_line_integral, dim3(1), dim3(NT), 0, c->stream, c->g, n, c->bi + s,
                       c->bj + s, c->di + s, c->dj + s, c->dxu, c->dyu, c->cost, uloc, vloc,
                       dev_out2);
    return check_launch(c);
}

int psi_line_integral(psi_ctx* c, int32_t isle, const double* uloc, const double* vloc,
                      double* line) {
    if (!valid_field_ctx(c) || !uloc || !vloc || !line || isle < 1 || isle > c->nisle)
        return PSI_BAD_ARG;
    int rc = line_integral_async(c, isle, uloc, vloc, c->result);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->h_res, c->result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->h_res[1] != 0.0) return fail(c, PSI_BAD_ARG, "missing line_dir in line integral%s");
    *line = c->h_res[0];
    return PSI_OK;
}

int psi_psi_velocity(psi_ctx* c, const double* psi_f, double* fpx, double* fpy) {
    if (!valid_field_ctx(c) || !psi_f || !fpx || !fpy) return PSI_BAD_ARG;
    if (!c->maskU || !c->maskV) return fail(c, PSI_BAD_ARG, "maskU/maskV not given%s");
    hipLaunchKernelGGL(k_psi_velocity, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g, psi_f,
                       c->maskU, c->maskV, c->hur, c->hvr, c->dxt, c->dyt, c->cosu, fpx, fpy);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_true_relres(psi_ctx* c, const double* forc, const double* sol, double* relres) {
    if (!valid_field_ctx(c) || !forc || !sol || !relres) return PSI_BAD_ARG;
    double res[4];
    int rc = residual(c, forc, sol, nullptr, res);
    if (rc) return rc;
    *relres = res[1] > 0 ? sqrt(res[0] / res[1]) : sqrt(res[0]);
    return PSI_OK;
}

int psi_streamfunction_init(psi_ctx* c, const psi_opts* o, double* psin, double* line_psin,
                            int32_t* iterations) {
    if (!valid_field_ctx(c) || !o || !psin || !line_psin) return PSI_BAD_ARG;
    if (c->nisle > 0 && (!c->maskU || !c->maskV))
        return fail(c, PSI_BAD_ARG, "maskU/maskV needed for line_psin%s");
    const int n = c->nisle;
    double* zero = c->tmp64;
    HIPCHK(c, hipMemsetAsync(zero, 0, c->plane * 8, c->stream));
    for (int k = 0; k < n; ++k) {
        double* pk = psin + (size_t)k * c->plane;
        HIPCHK(c, hipMemsetAsync(pk, 0, c->plane * 8, c->stream));
        const int s = c->offs[k], np = c->offs[k + 1] - s;
        // psin = 1 on the island's perimeter (streamfct_init.f90:274-279)
        std::vector<int32_t> hbi(np), hbj(np);
        HIPCHK(c, hipMemcpy(hbi.data(), c->bi + s, np * 4, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(hbj.data(), c->bj + s, np * 4, hipMemcpyDeviceToHost));
        std::vector<double> host(c->plane, 0.0);
        for (int q = 0; q < np; ++q) {
            const int i = hbi[q], j = hbj[q];
            if (i >= -1 && i <= c->g.nx + 2 && j >= -1 && j <= c->g.ny + 2)
                host[(size_t)(j + 1) * c->g.pitch + (i + 1)] = 1.0;
        }
        HIPCHK(c, hipMemcpyAsync(pk, host.data(), c->plane * 8, hipMemcpyHostToDevice, c->stream));
        psi_info info;
        const int rc = psi_congrad(c, o, zero, pk, &info);
        if (rc == PSI_HIP_ERROR || rc == PSI_BAD_ARG) return rc;
        if (c->g.cyclic)
            hipLaunchKernelGGL(k_setcyclic, dim3(nblk(c->g.ny + 4)), dim3(256), 0, c->stream, c->g,
                               pk);
        if (iterations) iterations[k] = info.iterations;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    // line_psin(n, isle) (streamfct_init.f90:292-303)
    for (int isle = 0; isle < n; ++isle) {
        hipLaunchKernelGGL(k_psi_velocity, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g,
                           psin + (size_t)isle * c->plane, c->maskU, c->maskV, c->hur, c->hvr,
                           c->dxt, c->dyt, c->cosu, c->tmp64, c->tmp64b);
        for (int m = 0; m < n; ++m) {
            int rc = line_integral_async(c, m + 1, c->tmp64, c->tmp64b, c->result + 2 + 0);
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(c->h_res, c->result + 2, 2 * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            line_psin[(size_t)m * n + isle] = c->h_res[0];
        }
    }
    return PSI_OK;
}

int psi_solve_streamfunction(psi_ctx* c, const psi_opts* o, const double* fpx, const double* fpy,
                             double* psi_f, double* dpsi, double* dpsin, const double* psin,
                             const double* line_psin, int32_t taum1, int32_t tau, int32_t taup1,
                             double dt_mom, double AB_eps, double* forc_out, psi_info* info) {
    if (!valid_field_ctx(c) || !o || !fpx || !fpy || !psi_f || !dpsi || !info) return PSI_BAD_ARG;
    if (taum1 < 0 || taum1 > 2 || tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2)
        return fail(c, PSI_BAD_ARG, "time level out of range%s");
    const int n = c->nisle;
    if (n > 1 && (!dpsin || !psin || !line_psin || !c->maskU || !c->maskV))
        return fail(c, PSI_BAD_ARG, "island arrays missing%s");
    const size_t P = c->plane;
    double* forc = forc_out ? forc_out : c->tmp64;
    hipLaunchKernelGGL(k_curl_forcing, dim3(nblk(P)), dim3(256), 0, c->stream, c->g, fpx, fpy,
                       c->dxu, c->dyu, c->cost, c->cosu, forc);
    hipLaunchKernelGGL(k_warm_start, dim3(nblk(P)), dim3(256), 0, c->stream, P, dpsi + tau * P,
                       dpsi + taum1 * P, dpsi + taup1 * P);
    int rc = check_launch(c);
    if (rc) return rc;
    rc = psi_congrad(c, o, forc, dpsi + taup1 * P, info);
    if (rc == PSI_HIP_ERROR || rc == PSI_BAD_ARG || rc == PSI_NAN) return rc;
    const int solve_rc = rc;
    if (c->g.cyclic)
        hipLaunchKernelGGL(k_setcyclic, dim3(nblk(c->g.ny + 4)), dim3(256), 0, c->stream, c->g,
                           dpsi + taup1 * P);
    std::vector<double> coef((size_t)std::max(n, 1), 0.0);
    const double w1 = 1.5 + AB_eps, w0 = 0.5 + AB_eps;
    if (n > 1) {
        // line_forc(k) - line_psi0(k) for k = 2..nisle (solve_streamfct.f90:63-77)
        std::vector<double> rhs((size_t)n, 0.0);
        for (int k = 2; k <= n; ++k) {
            rc = line_integral_async(c, k, fpx, fpy, c->result);
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(c->h_res, c->result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            rhs[k - 1] = c->h_res[0];
        }
        hipLaunchKernelGGL(k_psi_velocity, dim3(nblk(P)), dim3(256), 0, c->stream, c->g,
                           dpsi + taup1 * P, c->maskU, c->maskV, c->hur, c->hvr, c->dxt, c->dyt,
                           c->cosu, c->tmp64, c->tmp64b);
        for (int k = 2; k <= n; ++k) {
            rc = line_integral_async(c, k, c->tmp64, c->tmp64b, c->result);
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(c->h_res, c->result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            rhs[k - 1] = rhs[k - 1] - c->h_res[0];
        }
        // DGESV on line_psin(2:nisle, 2:nisle) (solve_streamfct.f90:79-87):
        // LU with partial pivoting, FP64, on the host (at most a few unknowns).
        const int m = n - 1;
        std::vector<double> A((size_t)m * m), x(m);
        for (int r = 0; r < m; ++r) {
            for (int q = 0; q < m; ++q) A[(size_t)r * m + q] = line_psin[(size_t)(r + 1) * n + (q + 1)];
            x[r] = rhs[r + 1];
        }
        for (int k = 0; k < m; ++k) {
            int piv = k;
            double amax = fabs(A[(size_t)k * m + k]);
            for (int r = k + 1; r < m; ++r)
                if (fabs(A[(size_t)r * m + k]) > amax) {
                    amax = fabs(A[(size_t)r * m + k]);
                    piv = r;
                }
            if (A[(size_t)piv * m + k] == 0.0)
                return fail(c, PSI_SINGULAR_ISLAND_SYSTEM, "island system singular%s");
            if (piv != k) {
                for (int q = 0; q < m; ++q) std::swap(A[(size_t)k * m + q], A[(size_t)piv * m + q]);
                std::swap(x[k], x[piv]);
            }
            for (int r = k + 1; r < m; ++r) {
                const double l = A[(size_t)r * m + k] / A[(size_t)k * m + k];
                A[(size_t)r * m + k] = l;
                for (int q = k + 1; q < m; ++q) A[(size_t)r * m + q] -= l * A[(size_t)k * m + q];
                x[r] -= l * x[k];
            }
        }
        for (int k = m - 1; k >= 0; --k) {
            double s = x[k];
            for (int q = k + 1; q < m; ++q) s -= A[(size_t)k * m + q] * x[q];
            x[k] = s / A[(size_t)k * m + k];
        }
        for (int k = 2; k <= n; ++k) dpsin[(size_t)(k - 1) * 3 + tau] = x[k - 2];
        for (int k = 2; k <= n; ++k)
            coef[k - 1] = dt_mom * (w1 * dpsin[(size_t)(k - 1) * 3 + tau] -
                                    w0 * dpsin[(size_t)(k - 1) * 3 + taum1]);
    }
    double* dcoef = c->result + 4;   // up to 4 islands via scratch; more via tmp buffer
    const double* psin_arg = psin ? psin : c->tmp64;
    if (n > 1) {
        if (n > 4) dcoef = c->tmp64b;   // velocity scratch is free again here
        HIPCHK(c, hipMemcpyAsync(dcoef, coef.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    }
    hipLaunchKernelGGL(k_ab_update, dim3(nblk(P)), dim3(256), 0, c->stream, P, psi_f + tau * P,
                       dpsi + taup1 * P, dpsi + tau * P, dt_mom, w1, w0, psin_arg, dcoef,
                       n > 1 ? n : 0, psi_f + taup1 * P);
    rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return solve_rc;
}

// ---- the 3-D parts of solve_streamfunction ---------------------------------
static int levels_check(psi_ctx* c, const psi_levels* L) {
    if (!L || L->nz < 1 || L->nz > LV_MAX_NZ || !L->dzt || !L->dzw || !L->maskT || !L->maskU ||
        !L->maskV)
        return fail(c, PSI_BAD_ARG, "levels: nz must be 1..128 and every array given%s");
    // dzt, dzw to the device when they changed (host copy of the last upload)
    const size_t nz = (size_t)L->nz;
    bool same = c->h_dz.size() == 2 * nz;
    for (size_t k = 0; same && k < nz; ++k)
        same = c->h_dz[k] == L->dzt[k] && c->h_dz[nz + k] == L->dzw[k];
    if (!same) {
        c->h_dz.assign(L->dzt, L->dzt + nz);
        c->h_dz.insert(c->h_dz.end(), L->dzw, L->dzw + nz);
        HIPCHK(c, hipMemcpyAsync(c->dz, c->h_dz.data(), 2 * nz * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));   // h_dz may change on the next call
    }
    return PSI_OK;
}

static int forcing_async(psi_ctx* c, const psi_levels* L, double grav, double rho_0,
                         const double* rho, double* du, double* dv, const double* du_mix,
                         const double* dv_mix, double* p_hydro, double* fpx, double* fpy) {
    const double fxa = grav / rho_0;   // solve_streamfct.f90:23
    const int nz = L->nz;
    hipLaunchKernelGGL(k_hydrostatic, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->plane, nz,
                       fxa, c->dz + nz, rho, L->maskT, p_hydro);
    hipLaunchKernelGGL(k_streamfunction_forcing, dim3(nblk(c->plane)), dim3(256), 0, c->stream,
                       c->g, nz,
                       c->dxu, c->dyu, c->cost, c->dz, L->maskU, L->maskV, c->hur, c->hvr,
                       p_hydro, du, dv, du_mix, dv_mix, fpx, fpy);
    if (c->g.cyclic) {
        hipLaunchKernelGGL(k_setcyclic, dim3(nblk(c->g.ny + 4)), dim3(256), 0, c->stream, c->g,
                           fpx);
        hipLaunchKernelGGL(k_setcyclic, dim3(nblk(c->g.ny + 4)), dim3(256), 0, c->stream, c->g,
                           fpy);
    }
    return check_launch(c);
}

static int velocity_async(psi_ctx* c, const psi_levels* L, double dt_mom, double AB_eps,
                          const double* u_tau, const double* v_tau, const double* du_tau,
                          const double* dv_tau, const double* du_taum1, const double* dv_taum1,
                          const double* du_mix, const double* dv_mix, const double* psi_f,
                          double* u_out, double* v_out) {
    const double w1 = 1.5 + AB_eps, w0 = 0.5 + AB_eps;   // solve_streamfct.f90:96-97
    const int nz = L->nz;
    if (c->vl_nz != nz) {   // launch shape per level count, kept until nz changes
        c->vl = velocity_launch(nz);
        c->vl_nz = nz;
    }
    if (!c->vl.kern) return fail(c, PSI_BAD_ARG, "levels: nz out of range%s");
    hipLaunchKernelGGL(c->vl.kern, dim3((unsigned)((c->plane + 63) / 64)), dim3(64, c->vl.nw), 0,
                       c->stream, c->g, nz, dt_mom, w1, w0, c->dxt, c->dyt, c->cosu, c->dz,
                       L->maskU, L->maskV, c->hur, c->hvr, u_tau, v_tau, du_tau, dv_tau,
                       du_taum1, dv_taum1, du_mix, dv_mix, psi_f, u_out, v_out);
    return check_launch(c);
}

int psi_streamfunction_forcing(psi_ctx* c, const psi_levels* L, double grav, double rho_0,
                               const double* rho, double* du, double* dv, const double* du_mix,
                               const double* dv_mix, double* p_hydro, double* fpx, double* fpy) {
    if (!valid_field_ctx(c) || !rho || !du || !dv || !du_mix || !dv_mix || !p_hydro || !fpx ||
        !fpy || !(rho_0 != 0.0))
        return PSI_BAD_ARG;
    int rc = levels_check(c, L);
    if (rc) return rc;
    rc = forcing_async(c, L, grav, rho_0, rho, du, dv, du_mix, dv_mix, p_hydro, fpx, fpy);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_velocity_update(psi_ctx* c, const psi_levels* L, double dt_mom, double AB_eps,
                        const double* u_tau, const double* v_tau, const double* du_tau,
                        const double* dv_tau, const double* du_taum1, const double* dv_taum1,
                        const double* du_mix, const double* dv_mix, const double* psi_taup1,
                        double* u_taup1, double* v_taup1) {
    if (!valid_field_ctx(c) || !u_tau || !v_tau || !du_tau || !dv_tau || !du_taum1 ||
        !dv_taum1 || !du_mix || !dv_mix || !psi_taup1 || !u_taup1 || !v_taup1)
        return PSI_BAD_ARG;
    int rc = levels_check(c, L);
    if (rc) return rc;
    rc = velocity_async(c, L, dt_mom, AB_eps, u_tau, v_tau, du_tau, dv_tau, du_taum1, dv_taum1,
                        du_mix, dv_mix, psi_taup1, u_taup1, v_taup1);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_solve_streamfunction_3d(psi_ctx* c, const psi_opts* o, const psi_levels* L,
                                const psi_momentum* m, double grav, double rho_0, double* psi_f,
                                double* dpsi, double* dpsin, const double* psin,
                                const double* line_psin, int32_t taum1, int32_t tau,
                                int32_t taup1, double dt_mom, double AB_eps, double* forc_out,
                                psi_info* info) {
    if (!valid_field_ctx(c) || !o || !m || !m->rho || !m->u || !m->v || !m->du || !m->dv ||
        !m->du_mix || !m->dv_mix || !m->p_hydro || !psi_f || !dpsi || !info || !(rho_0 != 0.0))
        return PSI_BAD_ARG;
    if (taum1 < 0 || taum1 > 2 || tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2)
        return fail(c, PSI_BAD_ARG, "time level out of range%s");
    int rc = levels_check(c, L);
    if (rc) return rc;
    const size_t V = (size_t)L->nz * c->plane;   // one time level of a 3-D field
    rc = forcing_async(c, L, grav, rho_0, m->rho + tau * V, m->du + tau * V, m->dv + tau * V,
                       m->du_mix, m->dv_mix, m->p_hydro, c->fpx3, c->fpy3);
    if (rc) return rc;
    rc = psi_solve_streamfunction(c, o, c->fpx3, c->fpy3, psi_f, dpsi, dpsin, psin, line_psin,
                                  taum1, tau, taup1, dt_mom, AB_eps, forc_out, info);
    if (rc == PSI_HIP_ERROR || rc == PSI_BAD_ARG || rc == PSI_NAN ||
        rc == PSI_SINGULAR_ISLAND_SYSTEM)
        return rc;
    const int solve_rc = rc;
    rc = velocity_async(c, L, dt_mom, AB_eps, m->u + tau * V, m->v + tau * V, m->du + tau * V,
                        m->dv + tau * V, m->du + taum1 * V, m->dv + taum1 * V, m->du_mix,
                        m->dv_mix, psi_f + taup1 * c->plane, m->u + taup1 * V,
                        m->v + taup1 * V);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return solve_rc;
}

// ---- flexible CG with the learned preconditioner --------------------------
int psi_fcg_prepare(psi_ctx* c, float* Z32, float* hinv32) {
    if (!valid_field_ctx(c) || !Z32 || !hinv32) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_fcg_prepare, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g, c->cC,
                       c->act, c->hur, c->hvr, Z32, hinv32);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

static int fcg_residual(psi_ctx* c, const double* forc, const double* sol, bool set_scale,
                        double* relres) {
    double res[4];
    int rc = residual(c, forc, sol, c->r64, res);
    if (rc) return rc;
    double h[10];
    HIPCHK(c, hipMemcpy(h, c->fcg, 10 * sizeof(double), hipMemcpyDeviceToHost));
    if (set_scale) {
        const double rms = sqrt(res[0] / std::max(res[3], 1.0));
        h[8] = rms > 0 ? 1.0 / rms : 1.0;
        h[9] = res[1] > 0 ? res[1] : res[0];
        HIPCHK(c, hipMemcpy(c->fcg + 8, h + 8, 2 * sizeof(double), hipMemcpyHostToDevice));
    }
    *relres = h[9] > 0 ? sqrt(res[0] / h[9]) : 0.0;
    hipLaunchKernelGGL(k_fcg_begin, dim3(nblk(c->plane)), dim3(256), 0, c->stream, c->g, c->r64,
                       c->act, c->fcg, c->r32, c->e32, c->pf32);
    c->fcg_restart = 1;
    return check_launch(c);
}

int psi_fcg_begin(psi_ctx* c, const double* forc, const double* sol, double* relres) {
    if (!valid_field_ctx(c) || !forc || !sol || !relres) return PSI_BAD_ARG;
    HIPCHK(c, hipMemsetAsync(c->fcg, 0, 16 * sizeof(double), c->stream));
    return fcg_residual(c, forc, sol, true, relres);
}

int psi_fcg_input_scale(psi_ctx* c, const float* Z32, const float** r32, const double** scale) {
    if (!valid_field_ctx(c) || !Z32) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_fcg_zr2, dim3(1024), dim3(NT), 0, c->stream, c->g, c->r32, Z32, c->fcg);
    hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(1), 0, c->stream, c->fcg, 0);
    if (r32) *r32 = c->r32;
    if (scale) *scale = c->fcg + 5;
    return check_launch(c);
}

int psi_fcg_step(psi_ctx* c, const float* znet, int H, int W) {
    if (!valid_field_ctx(c) || !znet) return PSI_BAD_ARG;
    const Geom& g = c->g;
    hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(1), 0, c->stream, c->fcg, 1);
    hipLaunchKernelGGL(k_fcg_scatter, dim3(1024), dim3(NT), 0, c->stream, g, znet, H, W, c->act,
                       c->q32, c->dxu, c->cosu, c->dyu, c->z32, c->fcg);
    hipLaunchKernelGGL(k_fcg_dir, dim3(nblk(c->plane)), dim3(256), 0, c->stream, g, c->z32, c->pf32,
                       c->fcg, c->fcg_restart);
    hipLaunchKernelGGL(k_fcg_apply, dim3(1024), dim3(NT), 0, c->stream, g, c->E, c->N, c->rx,
                       c->ry, c->pf32, c->r32, c->dxu, c->cosu, c->dyu, c->q32, c->fcg);
    hipLaunchKernelGGL(k_fcg_update, dim3(1024), dim3(NT), 0, c->stream, g, c->pf32, c->q32, c->act,
                       c->e32, c->r32, c->fcg);
    c->fcg_restart = 0;
    return check_launch(c);
}

int psi_fcg_step_mg(psi_ctx* c, const float* znet, int H, int W) {
    if (!valid_field_ctx(c) || !znet || H < 1 || W < 1) return PSI_BAD_ARG;
    MgState* m = c->mg;
    if (!m) return fail(c, PSI_BAD_ARG, "psi_fcg_step_mg: call psi_mg_setup first%s");
    const Geom& g = c->g;
    if (!c->fcg_mg_ready) {
        // the V-cycle kernels test S_time's done flag, which stays clear
        MgScal z{};
        HIPCHK(c, hipMemcpyAsync(m->S_time, &z, sizeof z, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->fcg_mg_ready = 1;
    }
    hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(1), 0, c->stream, c->fcg, 1);
    float* an = c->rp32[0];
    hipLaunchKernelGGL(k_fcgmg_scale, dim3(1024), dim3(NT), 0, c->stream, g, znet, H, W, c->act,
                       c->E32, c->N32, c->rx32, c->ry32, c->r32, c->dxu, c->cosu, c->dyu, an,
                       c->fcg);
    hipLaunchKernelGGL(k_fcgmg_input, dim3(nblk((size_t)g.nx * g.ny)), dim3(256), 0, c->stream, g,
                       m->lev[0], c->act, c->r32, (const float*)an, c->dxu, c->cosu, c->dyu,
                       1.0 / m->w_mean, c->fcg);
    MgCoarse co;
    co.S = m->S_time;
    mg_vcycle(m, c->stream, 5, mg_cheb(5, 15.0, m->cscale), nullptr, co);
    hipLaunchKernelGGL(k_fcgmg_scatter, dim3(1024), dim3(NT), 0, c->stream, g, m->lev[0], znet, H,
                       W, c->act, c->q32, c->dxu, c->cosu, c->dyu, (float)m->w_mean, c->z32,
                       c->fcg);
    hipLaunchKernelGGL(k_fcg_dir, dim3(nblk(c->plane)), dim3(256), 0, c->stream, g, c->z32, c->pf32,
                       c->fcg, c->fcg_restart);
    hipLaunchKernelGGL(k_fcg_apply, dim3(1024), dim3(NT), 0, c->stream, g, c->E, c->N, c->rx,
                       c->ry, c->pf32, c->r32, c->dxu, c->cosu, c->dyu, c->q32, c->fcg);
    hipLaunchKernelGGL(k_fcg_update, dim3(1024), dim3(NT), 0, c->stream, g, c->pf32, c->q32, c->act,
                       c->e32, c->r32, c->fcg);
    c->fcg_restart = 0;
    return check_launch(c);
}

int psi_fcg_read(psi_ctx* c, double* relres_recursive) {
    if (!valid_field_ctx(c) || !relres_recursive) return PSI_BAD_ARG;
    HIPCHK(c, hipMemcpyAsync(c->h_res, c->fcg + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_res + 1, c->fcg + 8, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double bbs = c->h_res[2] * c->h_res[1] * c->h_res[1];
    *relres_recursive = bbs > 0 ? sqrt(c->h_res[0] / bbs) : 0.0;
    return PSI_OK;
}

int psi_fcg_flush(psi_ctx* c, const double* forc, double* sol, double* relres) {
    if (!valid_field_ctx(c) || !forc || !sol || !relres) return PSI_BAD_ARG;
    double sc[10];
    HIPCHK(c, hipMemcpy(sc, c->fcg, 10 * sizeof(double), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL((k_add_correction<float>), dim3(nblk(c->plane)), dim3(256), 0, c->stream,
                       c->g, sol, c->e32, 1.0 / sc[8]);
    int rc = check_launch(c);
    if (rc) return rc;
    return fcg_residual(c, forc, sol, false, relres);
}

// A-norm cosine training loss (see k_nn_loss1); work: 2*B doubles, loss: 1
// double, both device.  gz (B x H x W fp32) receives dL/dz.
int psi_nn_anorm_loss(psi_ctx* c, const float* znet, int H, int W, int B, const double* R,
                      const double* ER, double* work, float* gz, double* loss) {
    if (!valid_field_ctx(c) || !znet || !R || !ER || !work || !gz || !loss || B < 1 ||
        H < c->g.ny + 4 || W < c->g.pitch)
        return PSI_BAD_ARG;
    HIPCHK(c, hipMemsetAsync(work, 0, 2 * B * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_nn_loss1, dim3(256, B), dim3(NT), 0, c->stream, c->g, znet, H, W, R, c->act,
                       c->E32, c->N32, c->rx32, c->ry32, c->dxu, c->cosu, c->dyu, gz, work);
    hipLaunchKernelGGL(k_nn_loss2, dim3(256, B), dim3(NT), 0, c->stream, c->g, H, W, B, R, c->act, ER,
                       c->dxu, c->cosu, c->dyu, work, gz, loss);
    return check_launch(c);
}

// Products with the area weights for training data: ER_b = e_b . W r_b.
int psi_nn_weighted_dot(psi_ctx* c, const double* a, const double* b, double* out) {
    if (!valid_field_ctx(c) || !a || !b || !out) return PSI_BAD_ARG;
    std::vector<double> ha(c->plane), hb(c->plane);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(ha.data(), a, c->plane * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hb.data(), b, c->plane * 8, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (int r = 2; r <= c->g.ny + 1; ++r)
        for (int col = 2; col <= c->g.nx + 1; ++col) {
            const size_t i = (size_t)r * c->g.pitch + col;
            s += ha[i] * hb[i] * c->h_dxu[col] * c->h_cosu[r] * c->h_dyu[r];
        }
    *out = s;
    return PSI_OK;
}

int psi_poly_coefficients(int32_t degree, double lower, float* out) {
    if (degree < 1 || degree > PMAXK || !(lower > 0 && lower < 1) || !out) return PSI_BAD_ARG;
    const ChebCoef ch = make_cheb(degree, lower);
    out[0] = ch.inv_theta;
    for (int j = 0; j < degree; ++j) {
        out[1 + 2 * j] = ch.c1[j];
        out[2 + 2 * j] = ch.c2[j];
    }
    return PSI_OK;
}

int psi_fcg_scalars(psi_ctx* c, double* out10) {
    if (!valid_field_ctx(c) || !out10) return PSI_BAD_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out10, c->fcg, 10 * sizeof(double), hipMemcpyDeviceToHost));
    return PSI_OK;
}

}  // extern "C"

// Batched ensembles (include/psibatch.h) share this translation unit's
// kernels and helpers.
#include "../../include/psibatch.h"
#include "psi_batch.hip"

// The surface-pressure solver (psi_press_*) and the non-hydrostatic one
// (psi_nh_*) likewise.
#include "psi_press.hip"
#include "psi_nonhydro.hip"

// The allocation-free plain solve (psi_congrad_noalloc) and its static scratch.
#include "psi_noalloc.hip"
