// psi_levels.hip -- the 3-D parts of solve_streamfunction on either side of
// the 2-D solve (for_src/external/solve_streamfct.f90:22-48 and :95-118).
//
// A 3-D field a(i,j,k) is nz consecutive (ny+4) x (nx+4) planes, level k
// (1 = bottom, nz = surface) at offset (k-1) * plane: the Fortran
// a(-1:nx+2, -1:ny+2, nz) in place.  Every kernel streams whole planes, one
// thread per column (i,j), so each wave reads 64 consecutive doubles of one
// plane per access; all three are HBM-bound.
//
// Floating-point contraction is off in these kernels and every expression
// keeps the reference's operation order, so results are bit-identical to
// the Fortran loops (and to the oracle built with -ffp-contract=off).
//
// Included by psi_capi.hip after psi_kernels.hip.

namespace psi {

// Hydrostatic pressure (:23-27), one thread per column, from the surface
// down:  p(nz) = 0.5 rho(nz) fxa dzw(nz) maskT(nz),
//        p(k)  = maskT(k) (p(k+1) + 0.5 (rho(k+1)+rho(k)) fxa dzw(k)).
// 24 bytes per point and level (rho, maskT in; p out).
__global__ void __launch_bounds__(256) k_hydrostatic(size_t plane, int nz, double fxa,
                                                     const double* __restrict__ dzw,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ maskT,
                                                     double* __restrict__ p_hydro) {
#pragma clang fp contract(off)
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= plane) return;
    size_t o = (size_t)(nz - 1) * plane + q;
    double r_up = rho[o];
    double p = 0.5 * r_up * fxa * dzw[nz - 1] * maskT[o];
    p_hydro[o] = p;
#pragma unroll 4
    for (int k = nz - 2; k >= 0; --k) {
        o -= plane;
        const double r = rho[o];
        p = maskT[o] * (p + 0.5 * (r_up + r) * fxa * dzw[k]);
        p_hydro[o] = p;
        r_up = r;
    }
}

// Pressure gradient into du/dv (interior, :30-35) and the depth-integrated
// forcing fpx = hur sum_k (du+du_mix) maskU dzt(k), fpy likewise (:37-42),
// bottom level first as the reference sums.  One thread per column over the
// flat plane, so every workgroup reads and writes whole 128-byte lines (rows
// of nx+4 doubles are not line-aligned); the p(i,j+1) neighbours sit one row
// (about 8 workgroups of 256) further on, which the round-robin dispatch
// places on the same XCD shortly after, so they come from its L2.
// 72 bytes per point and level (p; du, dv read and written; du_mix, dv_mix,
// maskU, maskV) plus 32 per column (fpx, fpy out; hur, hvr in).
__global__ void __launch_bounds__(256) k_streamfunction_forcing(
    Geom g, int nz, const double* __restrict__ dxu, const double* __restrict__ dyu,
    const double* __restrict__ cost, const double* __restrict__ dzt,
    const double* __restrict__ maskU, const double* __restrict__ maskV,
    const double* __restrict__ hur, const double* __restrict__ hvr,
    const double* __restrict__ p_hydro, double* __restrict__ du, double* __restrict__ dv,
    const double* __restrict__ du_mix, const double* __restrict__ dv_mix,
    double* __restrict__ fpx, double* __restrict__ fpy) {
#pragma clang fp contract(off)
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= plane) return;
    const int r = (int)(q / g.pitch), c = (int)(q - (size_t)r * g.pitch);
    const bool inner = r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1;
    const double denx = inner ? dxu[c] * cost[r] : 1.0;
    const double deny = inner ? dyu[r] : 1.0;
    double fx = 0.0, fy = 0.0;
#pragma unroll 2
    for (int k = 0; k < nz; ++k) {
        const size_t o = (size_t)k * plane + q;
        const double mu = maskU[o], mv = maskV[o];
        double a = du[o], b = dv[o];
        if (inner) {
            const double p0 = p_hydro[o];
            a = a - (p_hydro[o + 1] - p0) / denx * mu;
            b = b - (p_hydro[o + g.pitch] - p0) / deny * mv;
            du[o] = a;
            dv[o] = b;
        }
        fx = fx + (a + du_mix[o]) * mu * dzt[k];
        fy = fy + (b + dv_mix[o]) * mv * dzt[k];
    }
    fpx[q] = fx * hur[q];
    fpy[q] = fy * hvr[q];
}

// Velocity update (:95-118 without the open-boundary call):
//   u' = u(tau) + dt (du_mix + w1 du(tau) - w0 du(taum1)) maskU        (:96)
//   u' -= (sum_k u' maskU dzt) maskU hur                             (:104-112)
//   u' -= maskU (psi(i,j)-psi(i,j-1))/dyt(j) hur     on the interior (:114-119)
// and v likewise.  The depth sum needs every level of u' before any level
// can be finished, so a workgroup holds 64 consecutive columns of the flat
// plane (whole 128-byte lines) with their levels split over NW waves, LPT
// levels per thread kept in registers:
// u' is written once, never re-read.  The sum runs through the waves in
// level order (wave w adds its levels to the running sum left by wave w-1),
// which keeps the reference's summation order.  96 bytes per point and
// level (u, v, du, dv at tau; du, dv at taum1; du_mix, dv_mix, maskU, maskV
// in; u, v out).
template <int LPT>
__global__ void __launch_bounds__(1024) k_velocity_update(
    Geom g, int nz, double dt, double w1, double w0, const double* __restrict__ dxt,
    const double* __restrict__ dyt, const double* __restrict__ cosu,
    const double* __restrict__ dzt, const double* __restrict__ maskU,
    const double* __restrict__ maskV, const double* __restrict__ hur,
    const double* __restrict__ hvr, const double* __restrict__ u_tau,
    const double* __restrict__ v_tau, const double* __restrict__ du_tau,
    const double* __restrict__ dv_tau, const double* __restrict__ du_taum1,
    const double* __restrict__ dv_taum1, const double* __restrict__ du_mix,
    const double* __restrict__ dv_mix, const double* __restrict__ psi,
    double* __restrict__ u_out, double* __restrict__ v_out) {
#pragma clang fp contract(off)
    __shared__ double sum_u[64], sum_v[64];
    const int lane = threadIdx.x, w = threadIdx.y, nw = blockDim.y;
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t q0 = (size_t)blockIdx.x * 64 + lane;
    const bool col = q0 < plane;
    const size_t q = col ? q0 : 0;
    const int r = (int)(q / g.pitch), c = (int)(q - (size_t)r * g.pitch);
    const int k0 = w * LPT;
    double un[LPT], vn[LPT], mu[LPT], mv[LPT];
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
        const int k = k0 + l;
        un[l] = vn[l] = mu[l] = mv[l] = 0.0;
        if (col && k < nz) {
            const size_t o = (size_t)k * plane + q;
            mu[l] = maskU[o];
            mv[l] = maskV[o];
            un[l] = u_tau[o] + dt * (du_mix[o] + w1 * du_tau[o] - w0 * du_taum1[o]) * mu[l];
            vn[l] = v_tau[o] + dt * (dv_mix[o] + w1 * dv_tau[o] - w0 * dv_taum1[o]) * mv[l];
        }
    }
    // depth sums in level order, handed from wave to wave
    for (int ww = 0; ww < nw; ++ww) {
        if (w == ww) {
            double sx = ww ? sum_u[lane] : 0.0, sy = ww ? sum_v[lane] : 0.0;
#pragma unroll
            for (int l = 0; l < LPT; ++l)
                if (k0 + l < nz) {
                    sx = sx + un[l] * mu[l] * dzt[k0 + l];
                    sy = sy + vn[l] * mv[l] * dzt[k0 + l];
                }
            sum_u[lane] = sx;
            sum_v[lane] = sy;
        }
        __syncthreads();
    }
    if (!col) return;
    const double fx = sum_u[lane], fy = sum_v[lane];
    const double hu_r = hur[q], hv_r = hvr[q];
    const bool inner = r >= 2 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1;
    double dpy = 0.0, dpx = 0.0, dy = 1.0, dx = 1.0;
    if (inner) {
        dpy = psi[q] - psi[q - g.pitch];
        dpx = psi[q] - psi[q - 1];
        dy = dyt[r];
        dx = cosu[r] * dxt[c];
    }
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
        const int k = k0 + l;
        if (k < nz) {
            const size_t o = (size_t)k * plane + q;
            double a = un[l] - fx * mu[l] * hu_r;
            double b = vn[l] - fy * mv[l] * hv_r;
            if (inner) {
                a = a - mu[l] * dpy / dy * hu_r;
                b = b + mv[l] * dpx / dx * hv_r;
            }
            u_out[o] = a;
            v_out[o] = b;
        }
    }
}

// Launch shape of k_velocity_update for nz levels: among LPT in {2, 4, 8}
// (NW = ceil(nz/LPT) waves, at most 16), the one with the most waves
// resident per CU, from the occupancy calculator (more waves in flight hide more of the HBM latency;
// measured at 2048^2: nz 40 -> LPT 4 x 10 waves, nz 60 -> LPT 8 x 8).  Ties
// go to the smaller LPT.  kern is null when nz is outside 1..LV_MAX_NZ.
constexpr int LV_MAX_NZ = 128;
using VelocityKernel = decltype(&k_velocity_update<2>);
struct VelocityLaunch {
    VelocityKernel kern = nullptr;
    int lpt = 0, nw = 0, resident = 0;
};

inline VelocityLaunch velocity_launch(int nz) {
    VelocityLaunch best;
    if (nz < 1 || nz > LV_MAX_NZ) return best;
    const VelocityKernel kerns[3] = {k_velocity_update<2>, k_velocity_update<4>,
                                     k_velocity_update<8>};
    for (int i = 0; i < 3; ++i) {
        const int lpt = 2 << i, nw = (nz + lpt - 1) / lpt;
        if (nw > 16) continue;
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kerns[i], 64 * nw, 0) !=
            hipSuccess)
            occ = 1;
        if (occ * nw > best.resident) best = {kerns[i], lpt, nw, occ * nw};
    }
    return best;
}

}  // namespace psi
