// psi_kernels.hip -- HIP kernels of the streamfunction solver for gfx950.
//
// The PCG iteration runs as one fused sweep per iteration (k_cg_sweep):
// each workgroup stages a (TX+4)x(TY+4) frame of u = Z r, p, and the two
// stored coefficient planes in LDS, and computes
//     p_k     = u_k + beta_k p_{k-1}                    (radius 2)
//     u_{k+1} = u_k - alpha_k Z (A p_k)                 (radius 1)
//     x_{k+1} = x_k + alpha_k p_k,  w = A u_{k+1}       (radius 0)
// together with the reductions (r, Z r), max|p|, |r|^2 and the three inner
// products of u, w = A u, p and s = A p that give the next (p, A p) exactly,
// so one reduction point per iteration yields alpha_{k+1}, beta_{k+1} of the
// reference recurrence (solve_streamfct.f90:223-299) also for the
// non-symmetric operator of spherical grids.  The last
// workgroup to finish folds the partials in a fixed order and runs the
// reference's divergence / degenerate / convergence tests on the device.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "psi_device.h"

namespace psi {

// ---------------------------------------------------------------------------
// Operator construction
// ---------------------------------------------------------------------------

// Reference coefficient planes (make_coeff_streamfunction,
// solve_streamfct.f90:123-148) in the reference's exact operation order;
// 9 planes, plane (ii+1)+3*(jj+1).  Interior only; everything else 0.
__global__ void k_make_coeff_ref(Geom g, const double* __restrict__ hur,
                                 const double* __restrict__ hvr, const double* __restrict__ dxt,
                                 const double* __restrict__ dxu, const double* __restrict__ dyt,
                                 const double* __restrict__ dyu, const double* __restrict__ cost,
                                 const double* __restrict__ cosu, double* __restrict__ cf) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    double vC = 0, vE = 0, vW = 0, vN = 0, vS = 0;
    if (r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1) {
        const double cu2 = cosu[r] * cosu[r];
        double t;
        t = hvr[idx + 1] / dxu[c] / dxt[c + 1] / cu2;
        vC = vC - t; vE = vE + t;
        t = hvr[idx] / dxu[c] / dxt[c] / cu2;
        vC = vC - t; vW = vW + t;
        t = hur[idx + g.pitch] / dyu[r] / dyt[r + 1] * cost[r + 1] / cosu[r];
        vC = vC - t; vN = vN + t;
        t = hur[idx] / dyu[r] / dyt[r] * cost[r] / cosu[r];
        vC = vC - t; vS = vS + t;
    }
    for (int k = 0; k < 9; ++k) cf[(size_t)k * plane + idx] = 0.0;
    cf[4 * plane + idx] = vC;
    cf[5 * plane + idx] = vE;
    cf[3 * plane + idx] = vW;
    cf[7 * plane + idx] = vN;
    cf[1 * plane + idx] = vS;
}

// Stored operator: E(i,j) for i in [0,nx], j in [1,ny] and N(i,j) for
// i in [1,nx], j in [0,ny] (the values make_coeff gives the E and N planes
// at interior points, extended by one column/row so W and S of interior
// points follow from their neighbours).  Zero elsewhere.
__global__ void k_build_EN(Geom g, const double* __restrict__ hur, const double* __restrict__ hvr,
                           const double* __restrict__ dxt, const double* __restrict__ dxu,
                           const double* __restrict__ dyt, const double* __restrict__ dyu,
                           const double* __restrict__ cost, const double* __restrict__ cosu,
                           double* __restrict__ E, double* __restrict__ N) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    double e = 0.0, n = 0.0;
    // Fortran i = c-1, j = r-1
    if (c >= 1 && c <= g.nx + 1 && r >= 2 && r <= g.ny + 1) {
        const double cu2 = cosu[r] * cosu[r];
        e = hvr[idx + 1] / dxu[c] / dxt[c + 1] / cu2;
    }
    if (c >= 2 && c <= g.nx + 1 && r >= 1 && r <= g.ny + 1) {
        n = hur[idx + g.pitch] / dyu[r] / dyt[r + 1] * cost[r + 1] / cosu[r];
    }
    E[idx] = e;
    N[idx] = n;
}

// Active mask: 1 where Z != 0, i.e. interior points with a non-zero
// diagonal that are not on any island perimeter (make_inv_sfc,
// solve_streamfct.f90:394-421).  `dirichlet` holds 1 on perimeter points.
__global__ void k_build_active(Geom g, const double* __restrict__ E, const double* __restrict__ N,
                               const double* __restrict__ rx, const double* __restrict__ ry,
                               const uint8_t* __restrict__ dirichlet, uint8_t* __restrict__ act,
                               double* __restrict__ cC_out) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    uint8_t a = 0;
    double cc = 0.0;
    if (r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1) {
        const double e = E[idx], w = E[idx - 1] * rx[c], n = N[idx], s = N[idx - g.pitch] * ry[r];
        cc = (((-e) - w) - n) - s;
        a = (cc != 0.0 && !dirichlet[idx]) ? 1 : 0;
    }
    act[idx] = a;
    if (cC_out) cC_out[idx] = cc;
}

__global__ void k_mark_points(Geom g, int npts, const int* __restrict__ bi,
                              const int* __restrict__ bj, uint8_t* __restrict__ mark) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    const int i = bi[k], j = bj[k];
    if (i >= -1 && i <= g.nx + 2 && j >= -1 && j <= g.ny + 2)
        mark[(size_t)(j + 1) * g.pitch + (i + 1)] = 1;
}

// Z (make_inv_sfc) exported: 1/C on active points, 0 elsewhere.
__global__ void k_make_inv(Geom g, const double* __restrict__ cC, const uint8_t* __restrict__ act,
                           double* __restrict__ Z) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    Z[idx] = act[idx] ? 1.0 / cC[idx] : 0.0;
}

template <class T>
__global__ void k_convert(size_t n, const double* __restrict__ a, T* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

// ---------------------------------------------------------------------------
// Frame loading helpers
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void load_metrics_1d(const Geom& g, int c0, int r0, const T* rx,
                                                const T* ry, T* sRX, T* sRY) {
    if (threadIdx.x < HX) {
        const int c = map_col(c0 + threadIdx.x, g);
        sRX[threadIdx.x] = (c >= 0 && c < g.nx + 4) ? rx[c] : T(0);
    } else if (threadIdx.x < HX + HY) {
        const int k = threadIdx.x - HX, r = r0 + k;
        sRY[k] = (r >= 0 && r < g.ny + 4) ? ry[r] : T(0);
    }
}

// ---------------------------------------------------------------------------
// Work schedule.  The solver kernels run on a fixed grid of G resident
// workgroups; every workgroup sweeps a static set of tiles and publishes one
// partial per reduction slot.  Each XCD (workgroup b runs on XCD b % 8) owns
// a contiguous range of tiles, so neighbouring tiles, which share halo rows,
// are swept through the same L2.
// ---------------------------------------------------------------------------
// Pins a loaded value in registers at this point of the program.  The
// compiler moves a load that only a data-dependent branch uses below that
// branch, so the load would be issued only after the branch's own operand
// had arrived: two memory round trips instead of one.  An empty asm
// statement that reads and writes the value keeps its load above it, issued
// with the others.
__device__ __forceinline__ void pin_loaded(double& v) { asm volatile("" : "+v"(v)); }
typedef double pin_d2 __attribute__((ext_vector_type(2)));
typedef float pin_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void pin_loaded(double2& v) {   // one 128-bit operand
    pin_d2 t = {v.x, v.y};
    asm volatile("" : "+v"(t));
    v.x = t.x, v.y = t.y;
}
__device__ __forceinline__ void pin_loaded(float& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin_loaded(float2& v) {
    pin_f2 t = {v.x, v.y};
    asm volatile("" : "+v"(t));
    v.x = t.x, v.y = t.y;
}
__device__ __forceinline__ void pin_loaded(int& v) { asm volatile("" : "+v"(v)); }

__device__ __forceinline__ void range_schedule(int n, int& first, int& stride, int& end) {
    const int b = blockIdx.x, G = gridDim.x;
    if (G < 8 || (G & 7)) {
        first = b;
        stride = G;
        end = n;
        return;
    }
    const int x = b & 7, k = b >> 3;
    const int q = n >> 3, rem = n & 7;
    const int start = x * q + (x < rem ? x : rem);
    first = start + k;
    stride = G >> 3;
    end = start + q + (x < rem ? 1 : 0);
}

__device__ __forceinline__ void tile_schedule(const Geom& g, int& first, int& stride, int& end) {
    range_schedule(g.ntiles, first, stride, end);
}

// Fold the previous kernel's partials (slot-major, G per slot) in a fixed
// order.  Every workgroup does this and gets the bit-identical result, so
// no grid-wide synchronisation (and no cross-XCD cache maintenance) is
// needed inside a kernel.  Result valid in thread 0.
// The loads are unconditional (slots past G read partial 0 and are dropped
// in fold_add), so all 4 NR of them are in flight at once instead of one
// round of loads per predicated group.
template <int NR>
__device__ __forceinline__ void fold_load(const double* __restrict__ part, int G, int b0,
                                          double (&x)[4][NR]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int b = b0 + q * NT < G ? b0 + q * NT : 0;
#pragma unroll
        for (int k = 0; k < NR; ++k) x[q][k] = part[(size_t)k * G + b];
    }
}
template <int NR>
__device__ __forceinline__ void fold_add(const double (&x)[4][NR], int G, int b0, unsigned maxmask,
                                         double (&v)[NR]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool in = b0 + q * NT < G;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double y = in ? x[q][k] : 0.0;
            v[k] = ((maxmask >> k) & 1u) ? fmax(v[k], y) : v[k] + y;
        }
    }
}
template <int NR>
__device__ __forceinline__ void fold_prev(const double* __restrict__ part, int G, unsigned maxmask,
                                          double (&v)[NR], double* sh) {
#pragma unroll
    for (int k = 0; k < NR; ++k) v[k] = 0.0;
    for (int b0 = threadIdx.x; b0 < G; b0 += 4 * NT) {
        double x[4][NR];
        fold_load<NR>(part, G, b0, x);
        fold_add<NR>(x, G, b0, maxmask, v);
    }
    block_reduce<NR>(v, maxmask, sh);
}

template <int NR>
__device__ __forceinline__ void write_partials(const double (&v)[NR], double* __restrict__ part) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = v[k];
    }
}

// ---------------------------------------------------------------------------
// CG state transitions, evaluated identically by every workgroup from the
// folded partials.  Slots: setup {gamma, d1, max|u|, |r|^2, |b|^2, -};
// sweep and refresh {gamma, d1, max|p|, |r|^2, d2, d3}.
// ---------------------------------------------------------------------------

// Prologue of iteration m = n+1 given gamma = (r, Z r) and the pieces of
// (p_{n+1}, A p_{n+1}) = d1 + beta d2 + beta^2 d3: beta, the divergence test
// (solve_streamfct.f90:232-240), the degenerate test (:262-266), alpha (:270).
__device__ void cg_prologue(CgState& s, double gam, double d1, double d2, double d3) {
    const int m = s.n + 1;
    const double beta = gam / s.gamma;
    if (m > 2) {
        s.betak_min = fmin(s.betak_min, fabs(gam));
        if (fabs(gam) > 100.0 * s.betak_min) {
            s.n = m;
            s.status = ST_DIVERGED;
            s.done = 1;
            return;
        }
    }
    const double pAp = d1 + beta * d2 + beta * beta * d3;
    s.beta = beta;
    s.gamma = gam;
    if (fabs(pAp) < fabs(gam) * (double)1.e-10f) {
        s.degenerate = 1;
        s.alpha = 0.0;
        return;
    }
    s.alpha = gam / pAp;
}

// After the setup sweep: the reference's entry test (:206-213) and the
// prologue of iteration 1 (ss = 0, so p_1 = Z r_0 and (p, A p) = d1).
__device__ void cg_from_setup(CgState& s, const double* red) {
    const double gam = red[0], d1 = red[1], zmax = red[2], rr = red[3];
    double bb = s.bb_fixed > 0 ? s.bb_fixed : red[4];
    s.rr0 = rr;
    s.rr = rr;
    s.rr_ref = rr;
    if (bb == 0.0) bb = rr;
    s.bb = bb;
    s.zmax0 = zmax;
    s.n = 0;
    s.degenerate = 0;
    s.est_error = 100.0 * zmax;
    const bool conv = (s.criterion == CRIT_REFERENCE) ? (100.0 * zmax < s.eps)
                                                       : (rr <= s.rtarget2 * bb);
    if (conv) {
        s.status = ST_CONVERGED;
        s.done = 1;
        return;
    }
    if (gam != gam || d1 != d1) {
        s.status = ST_NAN;
        s.done = 1;
        return;
    }
    s.betak_min = fabs(gam);
    s.beta = 0.0;
    s.gamma = gam;
    if (fabs(d1) < fabs(gam) * (double)1.e-10f) {
        s.degenerate = 1;
        s.alpha = 0.0;
        return;
    }
    s.alpha = gam / d1;
}

// After sweep n: the step / extrapolation test of iteration n (:283-297),
// then the prologue of iteration n+1.
__device__ void cg_from_sweep(CgState& s, const double* red) {
    const double gam = red[0], d1 = red[1], pmax = red[2], rr = red[3], d2 = red[4], d3 = red[5];
    const int n = s.n + 1;
    const double alpha = s.alpha;
    s.rr = rr;
    s.pmax = pmax;
    if (s.degenerate) {
        s.est_error = 100.0 * pmax;
        s.n = n;
        s.status = ST_CONVERGED;
        s.done = 1;
        return;
    }
    const double step = fabs(alpha) * pmax;
    bool conv = false;
    if (n == 1) {
        s.step1 = step;
        s.est_error = step;
        conv = step < s.eps;
    } else if (step < s.eps) {
        const double rate = exp(log(step / s.step1) / (double)(n - 1));
        s.est_error = step * rate / (1.0 - rate);
        conv = s.est_error < s.eps;
    }
    if (s.criterion == CRIT_RELRES) conv = (rr <= s.rtarget2 * s.bb);
    s.n = n;
    if (conv) {
        s.status = ST_CONVERGED;
        s.done = 1;
        return;
    }
    if (n >= s.max_it) {
        s.status = ST_MAXIT;
        s.done = 1;
        return;
    }
    if (s.red2 > 0 && rr <= s.red2 * s.rr_ref) {   // time to replace the residual
        s.status = ST_LEG;
        s.done = 1;
        return;
    }
    if (gam != gam || d1 != d1 || rr != rr) {
        s.status = ST_NAN;
        s.done = 1;
        return;
    }
    cg_prologue(s, gam, d1, d2, d3);
}

// After a residual replacement: the recursively updated residual was
// replaced by the true one; continue with the prologue of iteration n+1.
__device__ void cg_from_refresh(CgState& s, const double* red) {
    const double gam = red[0], d1 = red[1], rr = red[3], d2 = red[4], d3 = red[5];
    s.rr = rr;
    s.rr_ref = rr;
    if (gam != gam || d1 != d1 || rr != rr) {
        s.status = ST_NAN;
        s.done = 1;
        return;
    }
    // the true residual may sit above the drifted recursive one: rebase the
    // divergence test on it
    s.betak_min = fabs(gam);
    cg_prologue(s, gam, d1, d2, d3);
}

// The state update from the folded partials `red` (valid in thread 0):
// workgroup 0 publishes it, every workgroup gets it in LDS.
__device__ __forceinline__ void advance_state_folded(const CgState* __restrict__ st_in,
                                                     CgState* __restrict__ st_out,
                                                     const double (&red)[NRED], int done_in,
                                                     int next_phase, CgState& sst) {
    if (threadIdx.x == 0) {
        CgState s = *st_in;
        if (!done_in) {
            if (s.phase == PHASE_SETUP) cg_from_setup(s, red);
            else if (s.phase == PHASE_REFRESH) cg_from_refresh(s, red);
            else cg_from_sweep(s, red);
            s.phase = next_phase;
        }
        sst = s;
        if (blockIdx.x == 0) *st_out = s;
    }
    __syncthreads();
}

// Every workgroup: fold the previous partials (G_prev of them per slot; by
// default the previous launch ran on this launch's grid), advance the state,
// publish it (workgroup 0) and return it in LDS.
__device__ __forceinline__ void advance_state(const CgState* __restrict__ st_in,
                                              CgState* __restrict__ st_out,
                                              const double* __restrict__ part_in, int next_phase,
                                              CgState& sst, double* sRed, int G_prev = -1) {
    const int done_in = st_in->done;
    double red[NRED];
    if (!done_in)
        fold_prev<NRED>(part_in, G_prev > 0 ? G_prev : (int)gridDim.x, 1u << 2, red, sRed);
    if (threadIdx.x == 0) {
        CgState s = *st_in;
        if (!done_in) {
            if (s.phase == PHASE_SETUP) cg_from_setup(s, red);
            else if (s.phase == PHASE_REFRESH) cg_from_refresh(s, red);
            else cg_from_sweep(s, red);
            s.phase = next_phase;
        }
        sst = s;
        if (blockIdx.x == 0) *st_out = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Setup sweep: r0 = b - A x0, u0 = Z r0, w0 = A u0; partials for
// cg_from_setup.  TB is the type of the right-hand side (the FP32 inner
// solve reads the FP64 residual and scales it on the fly).  x == null means
// a zero first guess; then x_zero_out (if set) is zeroed on the interior.
// ---------------------------------------------------------------------------
template <class T, class TB>
__global__ __launch_bounds__(NT) void k_cg_setup(
    Geom g, const T* __restrict__ E, const T* __restrict__ N, const T* __restrict__ rx,
    const T* __restrict__ ry, const uint8_t* __restrict__ act, const TB* __restrict__ b,
    double bscale, const T* __restrict__ x, T* __restrict__ x_zero_out, T* __restrict__ u_out,
    T* __restrict__ p_out, double* __restrict__ part_out) {
    __shared__ T sX[HX * HY], sU[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ uint8_t sA[HX * HY];
    __shared__ T sRX[HX], sRY[HY];
    __shared__ double sRed[NRED * NW];
    constexpr int W1 = HX - 2, H1 = HY - 2;
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    int t0, ts, te;
    tile_schedule(g, t0, ts, te);
    for (int tile = t0; tile < te; tile += ts) {
        const int c0 = (tile % g.tiles_x) * TX, r0 = (tile / g.tiles_x) * TY;
        for (int l = threadIdx.x; l < HX * HY; l += NT) {
            const int gg = l / HX, f = l - gg * HX;
            const int r = r0 + gg, c = map_col(c0 + f, g);
            T xv = 0, bv = 0, e = 0, n = 0;
            uint8_t a = 0;
            if (in_array(r, c, g)) {
                const size_t o = (size_t)r * g.pitch + c;
                xv = x ? x[o] : T(0);
                bv = (T)((double)b[o] * bscale);
                e = E[o];
                n = N[o];
                a = act[o];
            }
            sX[l] = xv;
            sU[l] = bv;
            sE[l] = e;
            sN[l] = n;
            sA[l] = a;
        }
        load_metrics_1d(g, c0, r0, rx, ry, sRX, sRY);
        __syncthreads();
        // radius 1: u0 = Z (b - A x), in place over b (each point reads only
        // its own b); |b|^2 on the tile's own active points first
        for (int l = threadIdx.x; l < W1 * H1; l += NT) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HX + f;
            T cc;
            const T ax = stencil5(sX, sE, sN, idx, sRX[f], sRY[gg], cc);
            const T bv = sU[idx];
            const bool own = gg >= 2 && gg < HY - 2 && f >= 2 && f < HX - 2 &&
                             (r0 + gg) <= g.ny + 1 && (c0 + f) <= g.nx + 1;
            if (own && sA[idx]) acc[4] += (double)bv * (double)bv;
            sU[idx] = sA[idx] ? (T(1) / cc) * (bv - ax) : T(0);
        }
        __syncthreads();
        for (int k = 0; k < (TX * TY) / NT; ++k) {
            const int l = threadIdx.x + k * NT;
            const int gg = 2 + l / TX, f = 2 + l % TX;
            const int r = r0 + gg, c = c0 + f;
            if (r <= g.ny + 1 && c <= g.nx + 1) {
                const int idx = gg * HX + f;
                T cc;
                const T w = stencil5(sU, sE, sN, idx, sRX[f], sRY[gg], cc);
                const T u0 = sU[idx];
                const size_t o = (size_t)r * g.pitch + c;
                u_out[o] = u0;
                p_out[o] = T(0);
                if (x_zero_out) x_zero_out[o] = T(0);
                const double ud = (double)u0, ccd = (double)cc;
                acc[0] += ud * ud * ccd;
                acc[1] += ud * (double)w;
                acc[2] = fmax(acc[2], fabs(ud));
                if (sA[idx]) {
                    const double rv = ud * ccd;
                    acc[3] += rv * rv;
                }
            }
        }
        __syncthreads();
    }
    block_reduce<NRED>(acc, 1u << 2, sRed);
    write_partials<NRED>(acc, part_out);
}

// ---------------------------------------------------------------------------
// One fused PCG iteration.  u/p are double-buffered by iteration parity:
// after n iterations u_n and p_{n-1} live in buffer n & 1.  The frame of the
// next tile is fetched into registers while the current one is computed;
// the first frame is in flight while the previous partials are folded.
// ---------------------------------------------------------------------------
#ifndef PSI_PREFETCH
#define PSI_PREFETCH 1
#endif
// PSI_PREFETCH_X: the tile's own x values travel with the prefetched frame
// (else x is read where it is updated).  PSI_NT_STORE: the sweep's outputs
// are written with non-temporal stores.
#ifndef PSI_PREFETCH_X
#define PSI_PREFETCH_X 1
#endif
#ifndef PSI_NT_STORE
#define PSI_NT_STORE 0
#endif
template <class T>
__device__ __forceinline__ void put(T* p, T v) {
#if PSI_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

template <class T>
struct FrameRegs {
    static constexpr int NL = (HX * HY + NT - 1) / NT;
    static constexpr int PER = (TX * TY) / NT;
    T u[NL], p[NL], e[NL], n[NL];
    uint8_t a[NL];
#if PSI_PREFETCH_X
    T xo[PER];
#endif

    __device__ __forceinline__ void fetch(const Geom& g, int tile, const T* __restrict__ u_in,
                                          const T* __restrict__ p_in, const T* __restrict__ E,
                                          const T* __restrict__ N,
                                          const uint8_t* __restrict__ act,
                                          const T* __restrict__ x) {
        const int c0 = (tile % g.tiles_x) * TX, r0 = (tile / g.tiles_x) * TY;
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int l = threadIdx.x + q * NT;
            const int gg = l / HX, f = l - gg * HX;
            const int r = r0 + gg, c = map_col(c0 + f, g);
            u[q] = 0;
            p[q] = 0;
            e[q] = 0;
            n[q] = 0;
            a[q] = 0;
            if (l < HX * HY && in_array(r, c, g)) {
                const size_t o = (size_t)r * g.pitch + c;
                u[q] = u_in[o];
                p[q] = p_in[o];
                e[q] = E[o];
                n[q] = N[o];
                a[q] = act[o];
            }
        }
#if PSI_PREFETCH_X
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int l = threadIdx.x + k * NT;
            const int r = r0 + 2 + l / TX, c = c0 + 2 + l % TX;
            xo[k] = (r <= g.ny + 1 && c <= g.nx + 1) ? x[(size_t)r * g.pitch + c] : T(0);
        }
#else
        (void)x;
#endif
    }
    // frame -> LDS, fused with p_k = u_k + beta p_{k-1}
    __device__ __forceinline__ void store(T beta, T* sU, T* sP, T* sE, T* sN, uint8_t* sA) const {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int l = threadIdx.x + q * NT;
            if (l < HX * HY) {
                sU[l] = u[q];
                sP[l] = u[q] + beta * p[q];
                sE[l] = e[q];
                sN[l] = n[q];
                sA[l] = a[q];
            }
        }
    }
};

template <class T>
__global__ __launch_bounds__(NT) void k_cg_sweep(
    Geom g, const T* __restrict__ E, const T* __restrict__ N, const T* __restrict__ rx,
    const T* __restrict__ ry, const uint8_t* __restrict__ act, T* __restrict__ x, T* ub0, T* ub1,
    T* pb0, T* pb1, const CgState* __restrict__ st_in, CgState* __restrict__ st_out,
    const double* __restrict__ part_in, double* __restrict__ part_out) {
    __shared__ T sU[HX * HY], sP[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ uint8_t sA[HX * HY];
    __shared__ T sRX[HX], sRY[HY];
    __shared__ double sRed[NRED * NW];
    __shared__ CgState sst;
    // iteration count after this launch's state update (setup -> 0,
    // sweep -> n+1, refresh -> n): selects the u/p buffers before the fold
    const int phase_in = st_in->phase, n_in = st_in->n;
    const int n_pred = st_in->done ? n_in
                                   : (phase_in == PHASE_SETUP ? 0
                                                              : (phase_in == PHASE_SWEEP ? n_in + 1
                                                                                         : n_in));
    const int odd = n_pred & 1;
    const T* __restrict__ u_in = odd ? ub1 : ub0;
    const T* __restrict__ p_in = odd ? pb1 : pb0;
    T* __restrict__ u_out = odd ? ub0 : ub1;
    T* __restrict__ p_out = odd ? pb0 : pb1;
    int t0, ts, te;
    tile_schedule(g, t0, ts, te);
    FrameRegs<T> fr;
    if (!st_in->done && t0 < te) fr.fetch(g, t0, u_in, p_in, E, N, act, x);
    advance_state(st_in, st_out, part_in, PHASE_SWEEP, sst, sRed);
    if (sst.done) return;
    const T alpha = (T)sst.alpha, beta = (T)sst.beta;
    constexpr int W1 = HX - 2, H1 = HY - 2;
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    for (int tile = t0; tile < te; tile += ts) {
        const int c0 = (tile % g.tiles_x) * TX, r0 = (tile / g.tiles_x) * TY;
#if !PSI_PREFETCH
        if (tile != t0) fr.fetch(g, tile, u_in, p_in, E, N, act, x);
#endif
        fr.store(beta, sU, sP, sE, sN, sA);
#if PSI_PREFETCH_X
        T xo[FrameRegs<T>::PER];
#pragma unroll
        for (int k = 0; k < FrameRegs<T>::PER; ++k) xo[k] = fr.xo[k];
#endif
        load_metrics_1d(g, c0, r0, rx, ry, sRX, sRY);
        __syncthreads();
#if PSI_PREFETCH
        if (tile + ts < te) fr.fetch(g, tile + ts, u_in, p_in, E, N, act, x);
#endif
        // radius 1: u_{k+1} = u_k - alpha Z (A p_k)
        for (int l = threadIdx.x; l < W1 * H1; l += NT) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HX + f;
            if (sA[idx]) {
                T cc;
                const T s = stencil5(sP, sE, sN, idx, sRX[f], sRY[gg], cc);
                sU[idx] = sU[idx] - alpha * ((T(1) / cc) * s);
            }
        }
        __syncthreads();
        // radius 0: x update, w = A u_{k+1}, s = A p_k, reductions
#pragma unroll
        for (int k = 0; k < (TX * TY) / NT; ++k) {
            const int l = threadIdx.x + k * NT;
            const int gg = 2 + l / TX, f = 2 + l % TX;
            const int r = r0 + gg, c = c0 + f;
            if (r <= g.ny + 1 && c <= g.nx + 1) {
                const int idx = gg * HX + f;
                T cc;
                const T w = stencil5(sU, sE, sN, idx, sRX[f], sRY[gg], cc);
                const T sv = stencil5(sP, sE, sN, idx, sRX[f], sRY[gg], cc);
                const T un = sU[idx], pn = sP[idx];
                const size_t o = (size_t)r * g.pitch + c;
#if PSI_PREFETCH_X
                put(x + o, xo[k] + alpha * pn);
#else
                put(x + o, x[o] + alpha * pn);
#endif
                put(u_out + o, un);
                put(p_out + o, pn);
                const double ud = (double)un, pd = (double)pn, ccd = (double)cc;
                const double wd = (double)w, sd = (double)sv;
                acc[0] += ud * ud * ccd;
                acc[1] += ud * wd;
                acc[2] = fmax(acc[2], fabs(pd));
                if (sA[idx]) {
                    const double rv = ud * ccd;
                    acc[3] += rv * rv;
                }
                acc[4] += ud * sd + pd * wd;
                acc[5] += pd * sd;
            }
        }
        __syncthreads();
    }
    block_reduce<NRED>(acc, 1u << 2, sRed);
    write_partials<NRED>(acc, part_out);
}

// ---------------------------------------------------------------------------
// The same fused iteration as a wave-column march.  Each wave owns a strip
// of MOWN = 60 columns and a segment of SEG rows: one frame column per lane
// (the strip with its 2-column margin each side), and it walks down the
// segment row by row with every value in registers.  West/east neighbours
// come from the adjacent lanes by DPP whole-wave shifts; north/south
// neighbours are the rows before and after in the march.  No LDS frame and
// no barrier until the final reduction: each row is loaded once (plus the
// two rows either side of the segment), the loads of the next D rows are in
// flight while a row is computed, and the per-point arithmetic is
// k_cg_sweep's (stencil5's terms and order, the same tests), so both give
// the same iterates up to the order of the partial sums.
//
// Per loaded row r (march step i = r - R0 + 2):
//   p(r) = u(r) + beta p_old(r)
//   row r-1 (i >= 2): A p, u_new = u - alpha Z (A p) on active points
//   row r-2 (i >= 4, an own row): w = A u_new, s = A p, outputs, reductions
// ---------------------------------------------------------------------------
constexpr int MW = 64;            // frame columns per wave (one per lane)
constexpr int MOWN = MW - 4;      // own columns per strip

__device__ __forceinline__ int lane_shift_i(int v, bool west) {
    // wave_shr:1 (0x138): lane l gets lane l-1; wave_shl:1 (0x130): l gets l+1;
    // the lane with no source gets 0
    return west ? __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true)
                : __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double lane_shift(double v, bool west) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = lane_shift_i((int)(unsigned)(b & 0xffffffffull), west);
    const int hi = lane_shift_i((int)(unsigned)(b >> 32), west);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}
__device__ __forceinline__ float lane_shift(float v, bool west) {
    return __int_as_float(lane_shift_i(__float_as_int(v), west));
}

// The 5-point operator of one lane's cell given its coefficients and the
// values south, here and north; west/east from the neighbouring lanes.
// Same terms and order as stencil5.
template <class T>
__device__ __forceinline__ T march_apply(T so, T w, T cc, T e, T n, T vs, T vc, T vn) {
    const T vw = lane_shift(vc, true), ve = lane_shift(vc, false);
    T r = so * vs;
    r = r + w * vw;
    r = r + cc * vc;
    r = r + e * ve;
    r = r + n * vn;
    return r;
}

template <class T>
struct MarchRow {   // one loaded row of a lane's column, as loaded
    T u, po, e, n, x, ry;
    int a;
};

// Work items of the march (strips x segments) and its grid: a multiple of
// the CU count, so that every CU holds the same number of workgroups, and
// the items spread evenly over the workgroups.  A CU's vector memory
// pipeline is shared by its waves: with an uneven grid the CUs holding one
// workgroup more set the launch's length.  Up to NW items per workgroup
// (one per wave) the grid is ceil(items / (NW CUs)) workgroups per CU; past
// the resident limit `per_cu` the waves take several items in turn.
template <int SEG>
__host__ __device__ inline int march_items(const Geom& g) {
    return ((g.nx + MOWN - 1) / MOWN) * ((g.ny + SEG - 1) / SEG);
}
template <int SEG>
inline int march_groups(const Geom& g, int ncu, int per_cu) {
    const long items = march_items<SEG>(g), per = (long)NW * ncu;
    const long k = std::min<long>((items + per - 1) / per, std::max(per_cu, 1));
    return (int)(k * ncu);
}

// One work item's place: the segment's rows R0 .. R0+H-1 of a strip, the
// lane's column, and the direction of its march.
struct MarchItem {
    int R0, H, rlast, c, cs, up;
    bool live, own_col;
};

template <int SEG>
__device__ __forceinline__ MarchItem march_item(const Geom& g, int item, int end) {
    MarchItem m;
    const int lane = threadIdx.x & 63, nseg = (g.ny + SEG - 1) / SEG;
    m.live = item < end;
    const int strip = m.live ? item / nseg : 0, seg = m.live ? item - strip * nseg : 0;
    m.c = strip * MOWN + lane;                      // frame column (array column before wrap)
    // Lanes past the array's last column (closed basins) load column 0
    // instead: an own point's results depend only on lanes up to nx+3, so
    // their values are never used.  Own columns (2 .. nx+1) map to
    // themselves, also when the grid is cyclic.
    const int cm = map_col(m.c, g);
    m.cs = (cm >= 0 && cm < g.nx + 4) ? cm : 0;
    m.own_col = lane >= 2 && lane < MW - 2 && m.c <= g.nx + 1;
    m.R0 = 2 + seg * SEG;
    m.H = m.live ? min(SEG, g.ny + 2 - m.R0) : 0;
    m.rlast = m.R0 + m.H + 1;
    // Odd segments march up, even ones down: two neighbouring segments then
    // reach the two rows either side of their common edge at the same time
    // (both at the start or both at the end of the march), and the second
    // reads them from L2 instead of memory.
    m.up = seg & 1;
    return m;
}

template <class T, int SEG, int D>
__global__ __launch_bounds__(NT) void k_cg_march(
    Geom g, const T* __restrict__ E, const T* __restrict__ N, const T* __restrict__ rx,
    const T* __restrict__ ry, const uint8_t* __restrict__ act, T* __restrict__ x, T* ub0, T* ub1,
    T* pb0, T* pb1, const CgState* __restrict__ st_in, CgState* __restrict__ st_out,
    const double* __restrict__ part_in, int G_prev, double* __restrict__ part_out) {
    __shared__ double sRed[NRED * NW];
    __shared__ CgState sst;
    const int phase_in = st_in->phase, n_in = st_in->n;
    const int n_pred = st_in->done ? n_in
                                   : (phase_in == PHASE_SETUP ? 0
                                                              : (phase_in == PHASE_SWEEP ? n_in + 1
                                                                                         : n_in));
    const int odd = n_pred & 1;
    const T* __restrict__ u_in = odd ? ub1 : ub0;
    const T* __restrict__ p_in = odd ? pb1 : pb0;
    T* __restrict__ u_out = odd ? ub0 : ub1;
    T* __restrict__ p_out = odd ? pb0 : pb1;

    // Work items of this workgroup: a contiguous range of the items (strips
    // in order, the segments of a strip consecutive), each XCD a contiguous
    // range of workgroups; wave w takes items first + w, first + w + NW, ...
    // (The item index is wave-uniform and kept in a scalar register: the
    // march's branches are then scalar, and no load result is copied at a
    // join.)
    const int wv = threadIdx.x >> 6;
    const long nitems = march_items<SEG>(g), wg = xcd_tile(blockIdx.x, gridDim.x);
    const int first = (int)(wg * nitems / gridDim.x), end = (int)((wg + 1) * nitems / gridDim.x);
    int item = __builtin_amdgcn_readfirstlane(first + wv);
    MarchItem cur = march_item<SEG>(g, item, end);
    const size_t P = g.pitch;

    // Loads only, every one in bounds and none under a branch, so they stay
    // in flight until the march consumes the row D steps later.  Row i of an
    // item's march; steps past its end reload its last row (unused); x is
    // read on every loaded row but used only on own rows.  Idle waves load
    // rows of strip 0 and use nothing.
    auto load = [&](MarchRow<T>& m, const MarchItem& it, int i) {
        const int r = it.up ? max(it.rlast - i, it.R0 - 2) : min(it.R0 - 2 + i, it.rlast);
        const size_t o = (size_t)r * P + it.cs;
        m.u = u_in[o];
        m.po = p_in[o];
        m.e = E[o];
        m.n = N[o];
        m.a = act[o];
        m.x = x[o];
        m.ry = ry[r];
    };

    // the previous launch's partials first (up to 4 NT of them per slot), then
    // the first D rows: vector loads complete in order, so the fold then
    // waits for its own loads only
    const int done_in = st_in->done;
    double pre[4][NRED];
    fold_load<NRED>(part_in, G_prev, threadIdx.x, pre);   // (valid memory also when done)
    MarchRow<T> ring[D];
#pragma unroll
    for (int q = 0; q < D; ++q) load(ring[q], cur, q);   // (finished solves return below)
    __builtin_amdgcn_sched_barrier(0);   // keep every load above the fold's first use
    {
        double red[NRED] = {0, 0, 0, 0, 0, 0};
        if (!done_in) {
            fold_add<NRED>(pre, G_prev, threadIdx.x, 1u << 2, red);
            for (int b0 = threadIdx.x + 4 * NT; b0 < G_prev; b0 += 4 * NT) {   // larger grids
                double x[4][NRED];
                fold_load<NRED>(part_in, G_prev, b0, x);
                fold_add<NRED>(x, G_prev, b0, 1u << 2, red);
            }
            block_reduce<NRED>(red, 1u << 2, sRed);
        }
        advance_state_folded(st_in, st_out, red, done_in, PHASE_SWEEP, sst);
    }
    if (sst.done) return;
    const T alpha = (T)sst.alpha, beta = (T)sst.beta;

    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    // (idle waves skip the loop and still join the final reduction)
    for (; item < end; item += NW) {
        // the next item of this wave, whose first D rows the last D steps of
        // this march load
        const int inext = __builtin_amdgcn_readfirstlane(item + NW);
        const MarchItem nxt = march_item<SEG>(g, inext, end);
        const T rxv = rx[cur.cs];
        const int rstart = cur.up ? cur.rlast : cur.R0 - 2, dir = cur.up ? -1 : 1;
        const int nrows = cur.H + 4;                 // rows R0-2 .. R0+H+1
        const bool up = cur.up, own_col = cur.own_col;
        const int c = cur.c;
        // march state, for the step that loads row r (the rows before it in
        // the march are r-1, r-2, r-3, written for a downward march): p of
        // rows r-2, r-1; the loaded row r-1; N of row r-2; the finished row
        // r-2 (coefficients, A p, u_new, p, x, mask); u_new of r-3
        T p_m2 = 0, p_m1 = 0, n_m2 = 0, un_m3 = 0;
        MarchRow<T> L1{};
        T c1_e = 0, c1_w = 0, c1_n = 0, c1_s = 0, c1_cc = 0, c1_ap = 0, c1_un = 0, c1_p = 0,
          c1_x = 0;
        int c1_a = 0;
        T ry_m1 = 0;
        // The march is unrolled whole (SEG + 4 steps), so no load in flight
        // is copied at an edge inside it; a segment cut short by the grid
        // runs the same steps, those past its last row keeping nothing.
#pragma unroll
        for (int i0 = 0; i0 < SEG + 4; i0 += D) {
#pragma unroll
            for (int q = 0; q < D; ++q) {
                const int i = i0 + q;
                if (i < SEG + 4) {
                    const MarchRow<T> rw = ring[q];
                    const int r = rstart + dir * i;
                    const T ry_r = rw.ry;
                    const T p0 = rw.u + beta * rw.po;
                    T c0_e = 0, c0_w = 0, c0_n = 0, c0_s = 0, c0_cc = 0, c0_ap = 0, c0_un = 0;
                    if (i >= 2) {
                        // row r-1: its coefficients, A p and u_new
                        c0_e = L1.e;
                        c0_w = lane_shift(L1.e, true) * rxv;
                        c0_n = L1.n;
                        c0_s = (up ? rw.n : n_m2) * ry_m1;
                        c0_cc = (((-c0_e) - c0_w) - c0_n) - c0_s;
                        c0_ap = march_apply(c0_s, c0_w, c0_cc, c0_e, c0_n, up ? p0 : p_m2, p_m1,
                                            up ? p_m2 : p0);
                        c0_un = L1.a ? L1.u - alpha * ((T(1) / c0_cc) * c0_ap) : L1.u;
                    }
                    if (i >= 4 && i < nrows) {
                        // row r-2, an own row: w = A u_new, s = A p, outputs
                        const T wv2 = march_apply(c1_s, c1_w, c1_cc, c1_e, c1_n,
                                                  up ? c0_un : un_m3, c1_un, up ? un_m3 : c0_un);
                        if (own_col) {
                            const size_t o = (size_t)(r - 2 * dir) * P + c;
                            put(x + o, c1_x + alpha * c1_p);
                            put(u_out + o, c1_un);
                            put(p_out + o, c1_p);
                            const double ud = (double)c1_un, pd = (double)c1_p,
                                         ccd = (double)c1_cc;
                            const double wd = (double)wv2, sd = (double)c1_ap;
                            acc[0] += ud * ud * ccd;
                            acc[1] += ud * wd;
                            acc[2] = fmax(acc[2], fabs(pd));
                            if (c1_a) {
                                const double rv = ud * ccd;
                                acc[3] += rv * rv;
                            }
                            acc[4] += ud * sd + pd * wd;
                            acc[5] += pd * sd;
                        }
                    }
                    // shift the march by one row
                    un_m3 = c1_un;
                    c1_e = c0_e;
                    c1_w = c0_w;
                    c1_n = c0_n;
                    c1_s = c0_s;
                    c1_cc = c0_cc;
                    c1_ap = c0_ap;
                    c1_un = c0_un;
                    c1_p = p_m1;
                    c1_x = L1.x;
                    c1_a = L1.a;
                    n_m2 = L1.n;
                    ry_m1 = ry_r;
                    L1 = rw;
                    p_m2 = p_m1;
                    p_m1 = p0;
                    // the slot's row is consumed: reload it (in the same
                    // registers) with the row D steps on; in the last D steps
                    // with the next item's row that its march reads from
                    // this slot first (row q: step q reads slot q), or, with
                    // no next item, this item's last row again (in L2)
                    if (i + D < SEG + 4) load(ring[q], cur, i + D);
                    else load(ring[q], nxt.live ? nxt : cur, nxt.live ? q : i + D);
                }
            }
        }
        cur = nxt;
    }
    block_reduce<NRED>(acc, 1u << 2, sRed);
    write_partials<NRED>(acc, part_out);
}

// ---------------------------------------------------------------------------
// Residual replacement: u_n = Z r_true * rscale (r_true from k_residual,
// FP64), p_{n-1} kept; partials for cg_from_refresh.  If the true residual
// already meets the target (rr_true <= rr_thresh) the solve is finished.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_cg_refresh(
    Geom g, const T* __restrict__ E, const T* __restrict__ N, const T* __restrict__ rx,
    const T* __restrict__ ry, const uint8_t* __restrict__ act, const double* __restrict__ r_true,
    double rscale, T* ub0, T* ub1, const T* pb0, const T* pb1, const double* __restrict__ true_res,
    double rr_thresh, int new_max_it, const CgState* __restrict__ st_in,
    CgState* __restrict__ st_out, double* __restrict__ part_out) {
    __shared__ T sU[HX * HY], sP[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ uint8_t sA[HX * HY];
    __shared__ T sRX[HX], sRY[HY];
    __shared__ double sRed[NRED * NW];
    const bool finished = true_res[0] <= rr_thresh;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        CgState s = *st_in;
        s.phase = PHASE_REFRESH;
        s.max_it = new_max_it;
        s.degenerate = 0;
        s.done = finished ? 1 : 0;
        s.status = finished ? ST_CONVERGED : ST_RUNNING;
        *st_out = s;
    }
    if (finished) return;
    const int odd = st_in->n & 1;
    T* __restrict__ u_out = odd ? ub1 : ub0;
    const T* __restrict__ p_in = odd ? pb1 : pb0;
    constexpr int W1 = HX - 2, H1 = HY - 2;
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    int t0, ts, te;
    tile_schedule(g, t0, ts, te);
    for (int tile = t0; tile < te; tile += ts) {
        const int c0 = (tile % g.tiles_x) * TX, r0 = (tile / g.tiles_x) * TY;
        for (int l = threadIdx.x; l < HX * HY; l += NT) {
            const int gg = l / HX, f = l - gg * HX;
            const int r = r0 + gg, c = map_col(c0 + f, g);
            T rv = 0, pp = 0, e = 0, n = 0;
            uint8_t a = 0;
            if (in_array(r, c, g)) {
                const size_t o = (size_t)r * g.pitch + c;
                rv = (T)(r_true[o] * rscale);
                pp = p_in[o];
                e = E[o];
                n = N[o];
                a = act[o];
            }
            sU[l] = rv;
            sP[l] = pp;
            sE[l] = e;
            sN[l] = n;
            sA[l] = a;
        }
        load_metrics_1d(g, c0, r0, rx, ry, sRX, sRY);
        __syncthreads();
        // radius 1: u = Z r
        for (int l = threadIdx.x; l < W1 * H1; l += NT) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HX + f;
            T cc;
            (void)stencil5(sP, sE, sN, idx, sRX[f], sRY[gg], cc);
            sU[idx] = sA[idx] ? (T(1) / cc) * sU[idx] : T(0);
        }
        __syncthreads();
        for (int k = 0; k < (TX * TY) / NT; ++k) {
            const int l = threadIdx.x + k * NT;
            const int gg = 2 + l / TX, f = 2 + l % TX;
            const int r = r0 + gg, c = c0 + f;
            if (r <= g.ny + 1 && c <= g.nx + 1) {
                const int idx = gg * HX + f;
                T cc;
                const T w = stencil5(sU, sE, sN, idx, sRX[f], sRY[gg], cc);
                const T sv = stencil5(sP, sE, sN, idx, sRX[f], sRY[gg], cc);
                const T un = sU[idx], pn = sP[idx];
                u_out[(size_t)r * g.pitch + c] = un;
                const double ud = (double)un, pd = (double)pn, ccd = (double)cc;
                const double wd = (double)w, sd = (double)sv;
                acc[0] += ud * ud * ccd;
                acc[1] += ud * wd;
                acc[2] = fmax(acc[2], fabs(pd));
                if (sA[idx]) {
                    const double rr = ud * ccd;
                    acc[3] += rr * rr;
                }
                acc[4] += ud * sd + pd * wd;
                acc[5] += pd * sd;
            }
        }
        __syncthreads();
    }
    block_reduce<NRED>(acc, 1u << 2, sRed);
    write_partials<NRED>(acc, part_out);
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - A x with |r|^2 and |b|^2 over active points and
// max|r| (the true residual of the refinement loop and of every report).
// r may be null (norms only).  One pass over row segments of NT points;
// each XCD takes a contiguous range of rows, so a row's neighbours above and
// below come from the same L2 and the columns either side from the same
// cache lines.  Partials go to k_residual_fold: a last-arriving-workgroup
// fold here would need a device-scope fence per workgroup, which writes
// back the XCD's L2 each time (with one workgroup per tile and an LDS frame
// the kernel ran at a fifth of the stream rate).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_residual(
    Geom g, const double* __restrict__ E, const double* __restrict__ N,
    const double* __restrict__ rx, const double* __restrict__ ry, const uint8_t* __restrict__ act,
    const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r_out,
    double* __restrict__ part) {
    __shared__ double sRed[NRED * NW];
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    const int segs = (g.nx + NT - 1) / NT;
    int first, stride, end;
    range_schedule(g.ny * segs, first, stride, end);
    const size_t P = g.pitch;
    for (int it = first; it < end; it += stride) {
        const int j = it / segs, i = (it - j * segs) * NT + threadIdx.x;
        if (i >= g.nx) continue;
        const int r = j + 2, c = i + 2;
        const size_t o = (size_t)r * P + c;
        const size_t ow = (size_t)r * P + map_col(c - 1, g), oe = (size_t)r * P + map_col(c + 1, g);
        // the 5-point operator in stencil5's order and arithmetic
        const double e = E[o], w = E[ow] * rx[c], n = N[o], so = N[o - P] * ry[r];
        const double cc = (((-e) - w) - n) - so;
        double ax = so * x[o - P];
        ax = ax + w * x[ow];
        ax = ax + cc * x[o];
        ax = ax + e * x[oe];
        ax = ax + n * x[o + P];
        const double bv = b[o];
        const double rv = bv - ax;
        const bool a = act[o] != 0;
        if (a) {
            acc[0] += rv * rv;
            acc[1] += bv * bv;
            acc[2] = fmax(acc[2], fabs(rv));
            acc[3] += 1.0;
        }
        if (r_out) r_out[o] = a ? rv : 0.0;
    }
    block_reduce<NRED>(acc, 1u << 2, sRed);
    write_partials<NRED>(acc, part);
}

// Fold k_residual's G partials (one workgroup): result = {|r|^2, |b|^2,
// max|r|, count}.
__global__ __launch_bounds__(NT) void k_residual_fold(const double* __restrict__ part, int G,
                                                     double* __restrict__ result) {
    __shared__ double sRed[NRED * NW];
    double v[NRED];
    fold_prev<NRED>(part, G, 1u << 2, v, sRed);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) result[k] = v[k];
}

// x += s * e on the interior, then e = 0 (FP64 accumulate of an inner
// FP32 correction).
template <class T>
__global__ void k_add_correction(Geom g, double* __restrict__ x, T* __restrict__ e, double s) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    if (r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1) {
        x[idx] = x[idx] + s * (double)e[idx];
        e[idx] = T(0);
    }
}

// apply_op with the stored operator: res = A p on the interior, 0 elsewhere.
__global__ __launch_bounds__(NT) void k_apply_op(Geom g, const double* __restrict__ E,
                                                 const double* __restrict__ N,
                                                 const double* __restrict__ rx,
                                                 const double* __restrict__ ry,
                                                 const double* __restrict__ p,
                                                 double* __restrict__ res) {
    __shared__ double sX[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ double sRX[HX], sRY[HY];
    const int tile = blockIdx.x;
    const int tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const int c0 = tx * TX, r0 = ty * TY;
    for (int l = threadIdx.x; l < HX * HY; l += NT) {
        const int gg = l / HX, f = l - gg * HX;
        const int r = r0 + gg, c = c0 + f;          // apply_op reads the halo as stored
        double xv = 0, e = 0, n = 0;
        if (in_array(r, c, g)) {
            const size_t o = (size_t)r * g.pitch + c;
            xv = p[o];
            e = E[o];
            n = N[o];
        }
        sX[l] = xv;
        sE[l] = e;
        sN[l] = n;
    }
    if (threadIdx.x < HX) {
        const int c = c0 + threadIdx.x;
        sRX[threadIdx.x] = (c < g.nx + 4) ? rx[c] : 0.0;
    } else if (threadIdx.x < HX + HY) {
        const int k = threadIdx.x - HX, r = r0 + k;
        sRY[k] = (r < g.ny + 4) ? ry[r] : 0.0;
    }
    __syncthreads();
    // every array point of the frame's inner TX x TY block is written
    for (int k = 0; k < (TX * TY) / NT; ++k) {
        const int l = threadIdx.x + k * NT;
        const int gg = 2 + l / TX, f = 2 + l % TX;
        const int r = r0 + gg, c = c0 + f;
        if (r <= g.ny + 1 && c <= g.nx + 1) {
            double cc;
            res[(size_t)r * g.pitch + c] = stencil5(sX, sE, sN, gg * HX + f, sRX[f], sRY[gg], cc);
        }
    }
}

// zero the halo ring of a field (rows 0,1,ny+2,ny+3 and columns 0,1,nx+2,nx+3)
__global__ void k_zero_halo(Geom g, double* __restrict__ a) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    if (r < 2 || r > g.ny + 1 || c < 2 || c > g.nx + 1) a[idx] = 0.0;
}

// dot_sfc / absmax_sfc over the interior (deterministic two-level fold).
__global__ __launch_bounds__(NT) void k_dot_absmax(Geom g, const double* __restrict__ a,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ part,
                                                   unsigned* __restrict__ counter,
                                                   double* __restrict__ result) {
    __shared__ double sRed[2 * NW];
    double acc[2] = {0.0, 0.0};
    const size_t n = (size_t)g.nx * g.ny;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += (size_t)gridDim.x * NT) {
        const int j = (int)(k / g.nx), i = (int)(k % g.nx);
        const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
        const double av = a[o];
        if (b) acc[0] += av * b[o];
        acc[1] = fmax(acc[1], fabs(av));
    }
    block_reduce<2>(acc, 2u, sRed);
    if (publish_partials<2>(acc, part, counter, blockIdx.x, gridDim.x)) {
        fold_partials<2>(acc, 2u, part, gridDim.x, sRed);
        if (threadIdx.x == 0) {
            result[0] = acc[0];
            result[1] = acc[1];
            *counter = 0u;
        }
    }
}

// setcyclic_xy (parallel_none.f90:194-200) for onx=2: p(nx+i)=p(i), p(1-i)=p(nx-i+1).
__global__ void k_setcyclic(Geom g, double* __restrict__ a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;   // array row
    if (j >= g.ny + 4) return;
    double* row = a + (size_t)j * g.pitch;
    for (int i = 1; i <= 2; ++i) {
        row[g.nx + i + 1] = row[i + 1];
        row[1 - i + 1] = row[g.nx - i + 1 + 1];
    }
}

// line_integral (streamfct_init.f90:311-377) around one island: one
// workgroup folds the contributions in a fixed order.
__global__ __launch_bounds__(NT) void k_line_integral(
    Geom g, int npts, const int* __restrict__ bi, const int* __restrict__ bj,
    const int* __restrict__ di, const int* __restrict__ dj, const double* __restrict__ dxu,
    const double* __restrict__ dyu, const double* __restrict__ cost,
    const double* __restrict__ uloc, const double* __restrict__ vloc, double* __restrict__ out) {
    __shared__ double sRed[2 * NW];
    double acc[2] = {0.0, 0.0};
    const int P = g.pitch;
    for (int k = threadIdx.x; k < npts; k += NT) {
        const int i = bi[k], j = bj[k];
        if (!(i >= 0 && i <= g.nx && j >= 0 && j <= g.ny)) continue;
        const size_t o = (size_t)(j + 1) * P + (i + 1);
        const int a = di[k], b = dj[k];
        double v;
        if (a == 1 && b == 0)
            v = vloc[o] * dyu[j + 1] + uloc[o + P] * dxu[i + 1] * cost[j + 2];
        else if (a == -1 && b == 0)
            v = -vloc[o + 1] * dyu[j + 1] - uloc[o] * dxu[i + 1] * cost[j + 1];
        else if (a == 0 && b == 1)
            v = vloc[o] * dyu[j + 1] - uloc[o] * dxu[i + 1] * cost[j + 1];
        else if (a == 0 && b == -1)
            v = uloc[o + P] * dxu[i + 1] * cost[j + 2] - vloc[o + 1] * dyu[j + 1];
        else {
            acc[1] = 1.0;   // missing line_dir
            continue;
        }
        acc[0] += v;
    }
    block_reduce<2>(acc, 2u, sRed);
    if (threadIdx.x == 0) {
        out[0] = acc[0];
        out[1] = acc[1];
    }
}

// Velocity of a streamfunction for the island integrals
// (streamfct_init.f90:294-300): zero, then i,j in [0, n+2].
__global__ void k_psi_velocity(Geom g, const double* __restrict__ psi,
                               const double* __restrict__ maskU, const double* __restrict__ maskV,
                               const double* __restrict__ hur, const double* __restrict__ hvr,
                               const double* __restrict__ dxt, const double* __restrict__ dyt,
                               const double* __restrict__ cosu, double* __restrict__ fpx,
                               double* __restrict__ fpy) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    double u = 0.0, v = 0.0;
    if (r >= 1 && c >= 1) {
        u = -(maskU[idx] * (psi[idx] - psi[idx - g.pitch]) / dyt[r] * hur[idx]);
        v = maskV[idx] * (psi[idx] - psi[idx - 1]) / (cosu[r] * dxt[c]) * hvr[idx];
    }
    fpx[idx] = u;
    fpy[idx] = v;
}

// forc = curl(fpx, fpy) on the interior (solve_streamfct.f90:49-53), 0 elsewhere.
__global__ void k_curl_forcing(Geom g, const double* __restrict__ fpx,
                               const double* __restrict__ fpy, const double* __restrict__ dxu,
                               const double* __restrict__ dyu, const double* __restrict__ cost,
                               const double* __restrict__ cosu, double* __restrict__ forc) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    double f = 0.0;
    if (r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1)
        f = (fpy[idx + 1] - fpy[idx]) / (cosu[r] * dxu[c]) -
            (cost[r + 1] * fpx[idx + g.pitch] - cost[r] * fpx[idx]) / (cosu[r] * dyu[r]);
    forc[idx] = f;
}

// dpsi(taup1) = 2 dpsi(tau) - dpsi(taum1) on the whole array (:56).
__global__ void k_warm_start(size_t n, const double* __restrict__ d_tau,
                             const double* __restrict__ d_taum1, double* __restrict__ d_taup1) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d_taup1[i] = 2 * d_tau[i] - d_taum1[i];
}

// psi(taup1) = psi(tau) + dt ((1.5+AB) dpsi(taup1) - (0.5+AB) dpsi(tau))
// then + sum_isle dt ((1.5+AB) dpsin(isle,tau) - (0.5+AB) dpsin(isle,taum1)) psin
// (solve_streamfct.f90:91-94), whole array.
__global__ void k_ab_update(size_t n, const double* __restrict__ psi_tau,
                            const double* __restrict__ d_taup1, const double* __restrict__ d_tau,
                            double dt, double w1, double w0, const double* __restrict__ psin,
                            const double* __restrict__ coef, int nisle,
                            double* __restrict__ psi_taup1) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = psi_tau[i] + dt * (w1 * d_taup1[i] - w0 * d_tau[i]);
    for (int k = 1; k < nisle; ++k) v = v + coef[k] * psin[(size_t)k * n + i];
    psi_taup1[i] = v;
}

}  // namespace psi
