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
// CG state transitions (run by one thread of the last workgroup)
// ---------------------------------------------------------------------------

// After the setup sweep: the reference's entry test (solve_streamfct.f90:
// 206-213) and the prologue of iteration 1 (:227-266 with ss = 0).
__device__ void cg_after_setup(CgState* st, double gam, double del, double zmax, double rr,
                               double bb) {
    st->rr0 = rr;
    st->rr = rr;
    if (st->rel_to_initial || bb == 0.0) bb = rr;
    st->bb = bb;
    st->zmax0 = zmax;
    st->n = 0;
    st->degenerate = 0;
    st->est_error = 100.0 * zmax;
    bool conv = (st->criterion == CRIT_REFERENCE) ? (100.0 * zmax < st->eps)
                                                   : (rr <= st->rtarget2 * bb);
    if (conv) {
        st->status = ST_CONVERGED;
        st->done = 1;
        return;
    }
    if (gam != gam || del != del) {
        st->status = ST_NAN;
        st->done = 1;
        return;
    }
    st->betak_min = fabs(gam);
    st->beta = 0.0;
    st->gamma = gam;
    if (fabs(del) < fabs(gam) * (double)1.e-10f) {
        st->degenerate = 1;
        st->alpha = 0.0;
        return;
    }
    st->alpha = gam / del;
}

// After sweep n: step/extrapolation test of iteration n (:283-297), then
// the prologue of iteration n+1: beta, divergence test (:232-240),
// degenerate test (:262-266) and alpha (:270).  With u = Z r_{n+1},
// w = A u, p = p_n and s = A p_n, the reference's s_dot_As of iteration
// n+1 is (p_{n+1}, A p_{n+1}) = d1 + beta d2 + beta^2 d3 with
// d1 = (u, w), d2 = (u, s) + (p, w), d3 = (p, s): exact for the
// non-symmetric operator of spherical grids, where the symmetric
// shortcut d1 - beta gamma / alpha would change the iteration.
__device__ void cg_after_sweep(CgState* st, double gam, double d1, double d2, double d3,
                               double pmax, double rr) {
    const int n = st->n + 1;
    const double alpha = st->alpha;
    st->rr = rr;
    st->pmax = pmax;
    if (st->degenerate) {
        st->est_error = 100.0 * pmax;
        st->n = n;
        st->status = ST_CONVERGED;
        st->done = 1;
        return;
    }
    const double step = fabs(alpha) * pmax;
    bool conv = false;
    if (n == 1) {
        st->step1 = step;
        st->est_error = step;
        conv = step < st->eps;
    } else if (step < st->eps) {
        const double rate = exp(log(step / st->step1) / (double)(n - 1));
        st->est_error = step * rate / (1.0 - rate);
        conv = st->est_error < st->eps;
    }
    if (st->criterion == CRIT_RELRES) conv = (rr <= st->rtarget2 * st->bb);
    if (conv) {
        st->n = n;
        st->status = ST_CONVERGED;
        st->done = 1;
        return;
    }
    if (n >= st->max_it) {
        st->n = st->max_it + 1;
        st->status = ST_MAXIT;
        st->done = 1;
        return;
    }
    if (gam != gam || d1 != d1 || rr != rr) {
        st->n = n;
        st->status = ST_NAN;
        st->done = 1;
        return;
    }
    const int m = n + 1;
    const double beta = gam / st->gamma;
    if (m > 2) {
        st->betak_min = fmin(st->betak_min, fabs(gam));
        if (fabs(gam) > 100.0 * st->betak_min) {
            st->n = m;
            st->status = ST_DIVERGED;
            st->done = 1;
            return;
        }
    }
    const double pAp = d1 + beta * d2 + beta * beta * d3;
    st->beta = beta;
    st->gamma = gam;
    st->n = n;
    if (fabs(pAp) < fabs(gam) * (double)1.e-10f) {
        st->degenerate = 1;
        st->alpha = 0.0;
        return;
    }
    st->alpha = gam / pAp;
}

// ---------------------------------------------------------------------------
// Setup sweep: r0 = b - A x0, u0 = Z r0, w0 = A u0.
// TB is the type of the right-hand side (double for the FP64 path; the FP32
// inner solve reads the FP64 residual and scales it on the fly).  x == null
// means a zero first guess; then x_out (if set) is zeroed on the interior.
// ---------------------------------------------------------------------------
template <class T, class TB>
__global__ __launch_bounds__(NT) void k_cg_setup(
    Geom g, const T* __restrict__ E, const T* __restrict__ N, const T* __restrict__ rx,
    const T* __restrict__ ry, const uint8_t* __restrict__ act, const TB* __restrict__ b,
    double bscale, const T* __restrict__ x, T* __restrict__ x_zero_out, T* __restrict__ u_out,
    T* __restrict__ p_out, double* __restrict__ part, unsigned* __restrict__ counter,
    CgState* __restrict__ st) {
    __shared__ T sX[HX * HY], sU[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ uint8_t sA[HX * HY];
    __shared__ T sRX[HX], sRY[HY];
    __shared__ double sRed[NRED * NW];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const int c0 = tx * TX, r0 = ty * TY;
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
        sU[l] = bv;   // holds b until overwritten by u0 on radius-1 points
        sE[l] = e;
        sN[l] = n;
        sA[l] = a;
    }
    load_metrics_1d(g, c0, r0, rx, ry, sRX, sRY);
    __syncthreads();
    constexpr int W1 = HX - 2, H1 = HY - 2;
    // radius 1: u0 = Z (b - A x)
    T uloc[(W1 * H1 + NT - 1) / NT];
#pragma unroll
    for (int k = 0; k < (W1 * H1 + NT - 1) / NT; ++k) {
        const int l = threadIdx.x + k * NT;
        uloc[k] = 0;
        if (l < W1 * H1) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HX + f;
            T cc;
            const T ax = stencil5(sX, sE, sN, idx, sRX[f], sRY[gg], cc);
            const T res = sU[idx] - ax;
            uloc[k] = sA[idx] ? (T(1) / cc) * res : T(0);
        }
    }
    // keep b at radius 0 for |b|^2 and r0 for |r|^2: compute them before
    // overwriting (radius-0 points are a subset of radius-1 points).
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (W1 * H1 + NT - 1) / NT; ++k) {
        const int l = threadIdx.x + k * NT;
        if (l < W1 * H1) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HX + f;
            const bool inner = gg >= 2 && gg < HY - 2 && f >= 2 && f < HX - 2 &&
                               (r0 + gg) <= g.ny + 1 && (c0 + f) <= g.nx + 1;
            if (inner && sA[idx]) {
                const double bv = (double)sU[idx];
                acc[4] += bv * bv;
            }
            sU[idx] = uloc[k];
        }
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
    block_reduce<NRED>(acc, 1u << 2, sRed);
    if (publish_partials<NRED>(acc, part, counter, blockIdx.x, gridDim.x)) {
        fold_partials<NRED>(acc, 1u << 2, part, gridDim.x, sRed);
        if (threadIdx.x == 0) {
            st->done = 0;
            st->status = ST_RUNNING;
            cg_after_setup(st, acc[0], acc[1], acc[2], acc[3], acc[4]);
            *counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// One fused PCG iteration.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_cg_sweep(
    Geom g, const T* __restrict__ E, const T* __restrict__ N, const T* __restrict__ rx,
    const T* __restrict__ ry, const uint8_t* __restrict__ act, T* __restrict__ x,
    const T* __restrict__ u_in, const T* __restrict__ p_in, T* __restrict__ u_out,
    T* __restrict__ p_out, double* __restrict__ part, unsigned* __restrict__ counter,
    CgState* __restrict__ st) {
    if (st->done) return;
    const T alpha = (T)st->alpha, beta = (T)st->beta;
    __shared__ T sU[HX * HY], sP[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ uint8_t sA[HX * HY];
    __shared__ T sRX[HX], sRY[HY];
    __shared__ double sRed[NRED * NW];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const int c0 = tx * TX, r0 = ty * TY;
    // frame load fused with p_k = u_k + beta p_{k-1} (radius 2)
    for (int l = threadIdx.x; l < HX * HY; l += NT) {
        const int gg = l / HX, f = l - gg * HX;
        const int r = r0 + gg, c = map_col(c0 + f, g);
        T uu = 0, pp = 0, e = 0, n = 0;
        uint8_t a = 0;
        if (in_array(r, c, g)) {
            const size_t o = (size_t)r * g.pitch + c;
            uu = u_in[o];
            pp = p_in[o];
            e = E[o];
            n = N[o];
            a = act[o];
        }
        sU[l] = uu;
        sP[l] = uu + beta * pp;
        sE[l] = e;
        sN[l] = n;
        sA[l] = a;
    }
    load_metrics_1d(g, c0, r0, rx, ry, sRX, sRY);
    __syncthreads();
    // radius 1: u_{k+1} = u_k - alpha Z (A p_k)
    constexpr int W1 = HX - 2, H1 = HY - 2;
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
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
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
            x[o] = x[o] + alpha * pn;
            u_out[o] = un;
            p_out[o] = pn;
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
    block_reduce<NRED>(acc, 1u << 2, sRed);
    if (publish_partials<NRED>(acc, part, counter, blockIdx.x, gridDim.x)) {
        fold_partials<NRED>(acc, 1u << 2, part, gridDim.x, sRed);
        if (threadIdx.x == 0) {
            cg_after_sweep(st, acc[0], acc[1], acc[4], acc[5], acc[2], acc[3]);
            *counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - A x with |r|^2 and |b|^2 over active points and
// max|r| (the true residual of the refinement loop and of every report).
// r may be null (norms only).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_residual(
    Geom g, const double* __restrict__ E, const double* __restrict__ N,
    const double* __restrict__ rx, const double* __restrict__ ry, const uint8_t* __restrict__ act,
    const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r_out,
    double* __restrict__ part, unsigned* __restrict__ counter, double* __restrict__ result) {
    __shared__ double sX[HX * HY], sE[HX * HY], sN[HX * HY];
    __shared__ double sRX[HX], sRY[HY];
    __shared__ double sRed[NRED * NW];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const int c0 = tx * TX, r0 = ty * TY;
    for (int l = threadIdx.x; l < HX * HY; l += NT) {
        const int gg = l / HX, f = l - gg * HX;
        const int r = r0 + gg, c = map_col(c0 + f, g);
        double xv = 0, e = 0, n = 0;
        if (in_array(r, c, g)) {
            const size_t o = (size_t)r * g.pitch + c;
            xv = x[o];
            e = E[o];
            n = N[o];
        }
        sX[l] = xv;
        sE[l] = e;
        sN[l] = n;
    }
    load_metrics_1d(g, c0, r0, rx, ry, sRX, sRY);
    __syncthreads();
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < (TX * TY) / NT; ++k) {
        const int l = threadIdx.x + k * NT;
        const int gg = 2 + l / TX, f = 2 + l % TX;
        const int r = r0 + gg, c = c0 + f;
        if (r <= g.ny + 1 && c <= g.nx + 1) {
            const int idx = gg * HX + f;
            double cc;
            const double ax = stencil5(sX, sE, sN, idx, sRX[f], sRY[gg], cc);
            const size_t o = (size_t)r * g.pitch + c;
            const double bv = b[o];
            const double rv = bv - ax;
            if (act[o]) {
                acc[0] += rv * rv;
                acc[1] += bv * bv;
                acc[2] = fmax(acc[2], fabs(rv));
                acc[3] += 1.0;
            }
            if (r_out) r_out[o] = act[o] ? rv : 0.0;
        }
    }
    block_reduce<NRED>(acc, 1u << 2, sRed);
    if (publish_partials<NRED>(acc, part, counter, blockIdx.x, gridDim.x)) {
        fold_partials<NRED>(acc, 1u << 2, part, gridDim.x, sRed);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 4; ++k) result[k] = acc[k];
            *counter = 0u;
        }
    }
}

// x += s * e on the interior (FP64 accumulate of an inner correction).
template <class T>
__global__ void k_add_correction(Geom g, double* __restrict__ x, const T* __restrict__ e, double s) {
    const size_t plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane) return;
    const int r = (int)(idx / g.pitch), c = (int)(idx % g.pitch);
    if (r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1) x[idx] = x[idx] + s * (double)e[idx];
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

// ---------------------------------------------------------------------------
// Explicit instantiations used by the host code.
// ---------------------------------------------------------------------------
template __global__ void k_convert<float>(size_t, const double*, float*);
template __global__ void k_cg_setup<double, double>(Geom, const double*, const double*,
                                                    const double*, const double*, const uint8_t*,
                                                    const double*, double, const double*, double*,
                                                    double*, double*, double*, unsigned*, CgState*);
template __global__ void k_cg_setup<float, double>(Geom, const float*, const float*, const float*,
                                                   const float*, const uint8_t*, const double*,
                                                   double, const float*, float*, float*, float*,
                                                   double*, unsigned*, CgState*);
template __global__ void k_cg_sweep<double>(Geom, const double*, const double*, const double*,
                                            const double*, const uint8_t*, double*, const double*,
                                            const double*, double*, double*, double*, unsigned*,
                                            CgState*);
template __global__ void k_cg_sweep<float>(Geom, const float*, const float*, const float*,
                                           const float*, const uint8_t*, float*, const float*,
                                           const float*, float*, float*, double*, unsigned*,
                                           CgState*);
template __global__ void k_add_correction<float>(Geom, double*, const float*, double);

}  // namespace psi
