// psi_nonhydro.hip -- the non-hydrostatic pressure solver
// (for_src/non_hydrostatic/non_hydrostatic.f90), used when enable_hydrostatic
// is off (momentum.f90:167): the w momentum step and forcing, make_coeff_non_hydro,
// congrad_non_hydro on the 7-point 3-D operator and the velocity correction,
// behind the psi_nh_* / psi_*_non_hydro entry points of include/psisolve.h.
//
// 3-D fields are nz consecutive (ny+4) x (nx+4) planes, level 1 (the bottom)
// first.  The operator is not stored: with 0/1 masks every coefficient of
// make_coeff_non_hydro is either zero or a product of 1-D metric factors, so
// the kernels rebuild it from one mask byte per point (bit 0 maskU, 1 maskV,
// 2 maskW, 3 maskT) and 1-D tables, in the reference's operation order
// (bit-identical).  Contraction is off; the CG inner products are parallel
// sums, so iterates agree with the reference to rounding.
//
// Included at the end of psi_capi.hip after psi_press.hip (context helpers,
// SpScal and the fold kernels).

namespace psi {

enum : uint8_t { NH_MU = 1, NH_MV = 2, NH_MW = 4, NH_MT = 8 };

// 1-D factors of make_coeff_non_hydro (non_hydrostatic.f90:87-141) with mp = 1:
// ex_e(i) = (1/dxu(i))/dxt(i), ex_w(i) = (1/dxu(i-1))/dxt(i), c2(j) = cost(j)^2,
// tn(j) = ((1/dyu(j))/dyt(j) * cosu(j))/cost(j), ts(j) likewise with j-1,
// tu(k) = (1/dzw(k))/dzt(k), td(k) = (1/dzw(k-1))/dzt(k).  (Keeping ex/c2 as
// two 2-D planes instead of dividing in the kernel measured no faster.)
struct NhTables {
    const double *ex_e, *ex_w, *c2, *tn, *ts, *tu, *td;
};

// The seven coefficients of point (row, col, k) in apply_op_3D's summation
// order and the centre accumulated as make_coeff_non_hydro does (E, W, N, S,
// then above, then below).
struct NhCoef {
    double d, s, w, c, e, n, u;
};

__device__ __forceinline__ NhCoef nh_coef(const NhTables& t, int row, int col, int k, int nz,
                                          uint8_t m, uint8_t m_w, uint8_t m_s, uint8_t m_d) {
#pragma clang fp contract(off)
    NhCoef f;
    const double c2 = t.c2[row];
    f.e = (m & NH_MU) ? t.ex_e[col] / c2 : 0.0;
    f.w = (m_w & NH_MU) ? t.ex_w[col] / c2 : 0.0;
    f.n = (m & NH_MV) ? t.tn[row] : 0.0;
    f.s = (m_s & NH_MV) ? t.ts[row] : 0.0;
    f.u = (k < nz - 1 && (m & NH_MW)) ? t.tu[k] : 0.0;
    f.d = (k > 0 && (m_d & NH_MW)) ? t.td[k] : 0.0;
    double c0 = (((0.0 - f.e) - f.w) - f.n) - f.s;
    if (k < nz - 1) c0 = c0 - f.u;
    if (k > 0) c0 = c0 - f.d;
    f.c = c0;
    return f;
}

// apply_op_3D (non_hydrostatic.f90:247-270) at an interior point: the seven
// non-zero terms in the reference's order (kk, jj, ii outer to inner); the
// vertical terms at the bottom and top levels have zero coefficients.
__device__ __forceinline__ double nh_apply(const NhCoef& f, double pd, double ps, double pw,
                                           double pc, double pe, double pn, double pu) {
#pragma clang fp contract(off)
    double a = 0.0 + f.d * pd;
    a = a + f.s * ps;
    a = a + f.w * pw;
    a = a + f.c * pc;
    a = a + f.e * pe;
    a = a + f.n * pn;
    a = a + f.u * pu;
    return a;
}

// Packed mask bytes; bad counts mask values other than 0 and 1.
__global__ void __launch_bounds__(256) k_nh_masks(size_t n3, const double* __restrict__ mT,
                                                  const double* __restrict__ mU,
                                                  const double* __restrict__ mV,
                                                  const double* __restrict__ mW,
                                                  uint8_t* __restrict__ m8, int* __restrict__ bad) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const double t = mT[q], u = mU[q], v = mV[q], w = mW[q];
    int nb = 0;
    for (double x : {t, u, v, w}) nb += (x != 0.0 && x != 1.0);
    if (nb) atomicAdd(bad, nb);
    m8[q] = (uint8_t)((u != 0.0 ? NH_MU : 0) | (v != 0.0 ? NH_MV : 0) | (w != 0.0 ? NH_MW : 0) |
                      (t != 0.0 ? NH_MT : 0));
}

// The coefficients as 7 blocks of nz planes (below, S, W, C, E, N, above),
// zero outside the interior.
__global__ void __launch_bounds__(256) k_nh_export(Geom g, int nz, NhTables t,
                                                   const uint8_t* __restrict__ m8,
                                                   double* __restrict__ cf) {
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    int row, col;
    sp_rc(g, q, row, col);
    if (q >= plane) return;
    const size_t o = (size_t)k * plane + q, n3 = (size_t)nz * plane;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (sp_inner(g, row, col)) {
        const NhCoef f = nh_coef(t, row, col, k, nz, m8[o], m8[o - 1], m8[o - g.pitch],
                                 k > 0 ? m8[o - plane] : 0);
        const double c[7] = {f.d, f.s, f.w, f.c, f.e, f.n, f.u};
        for (int l = 0; l < 7; ++l) v[l] = c[l];
    }
    for (int l = 0; l < 7; ++l) cf[(size_t)l * n3 + o] = v[l];
}

// res = A p on the interior of every level, zero elsewhere.
__global__ void __launch_bounds__(256) k_nh_apply(Geom g, int nz, NhTables t,
                                                  const uint8_t* __restrict__ m8,
                                                  const double* __restrict__ p,
                                                  double* __restrict__ res) {
#pragma clang fp contract(off)
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= plane) return;
    const int k = blockIdx.y;
    int row, col;
    sp_rc(g, q, row, col);
    const size_t o = (size_t)k * plane + q, P = g.pitch;
    double a = 0.0;
    if (sp_inner(g, row, col)) {
        const NhCoef f = nh_coef(t, row, col, k, nz, m8[o], m8[o - 1], m8[o - P],
                                 k > 0 ? m8[o - plane] : 0);
        const double pc = p[o];
        a = nh_apply(f, k > 0 ? p[o - plane] : pc, p[o - P], p[o - 1], pc, p[o + 1], p[o + P],
                     k < nz - 1 ? p[o + plane] : pc);
    }
    res[o] = a;
}

// res = forc - A x on the interior, zero elsewhere (non_hydrostatic.f90:169-177),
// and the partials of dot_3D(res, res) (:181).  x's halo is read as it stands;
// p0 (the direction buffer iteration 1 reads as the old one) is zeroed.
__global__ void __launch_bounds__(256) k_nh_init(Geom g, int nz, NhTables t,
                                                 const uint8_t* __restrict__ m8,
                                                 const double* __restrict__ forc,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ r,
                                                 double* __restrict__ p0,
                                                 double* __restrict__ part) {
#pragma clang fp contract(off)
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t P = g.pitch;
    const int k = blockIdx.y;
    double acc[1] = {0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < plane;
         q += (size_t)gridDim.x * blockDim.x) {
        int row, col;
        sp_rc(g, q, row, col);
        const size_t o = (size_t)k * plane + q;
        double rv = 0.0;
        if (sp_inner(g, row, col)) {
            const uint8_t m = m8[o];
            const NhCoef f = nh_coef(t, row, col, k, nz, m, m8[o - 1], m8[o - P],
                                     k > 0 ? m8[o - plane] : 0);
            const double pc = x[o];
            const double a = nh_apply(f, k > 0 ? x[o - plane] : pc, x[o - P], x[o - 1], pc,
                                      x[o + 1], x[o + P], k < nz - 1 ? x[o + plane] : pc);
            rv = forc[o] - a;
            acc[0] = acc[0] + rv * rv * (double)((m >> 3) & 1);
        }
        r[o] = rv;
        p0[o] = 0.0;   // the old direction of iteration 1
    }
    // one partial per workgroup of the (gridDim.x, nz) grid
    const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
    __shared__ double sh[4];
    double v = acc[0];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[b] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// The direction of iteration n, p = res + (rsnew/rsold) p (:190; p = res for
// n = 1, :179) with setcyclic_xyz (:191-192), Ap = A p (:185; zero off the
// interior) and the partials of dot_3D(p, Ap) (:186) and absmax_3D(p)
// (:215), in two passes: rebuilding the seven neighbours' new directions in
// the stencil (one pass, 33 B/pt) made it bound by instruction issue, and the
// split (41 B/pt) measured 18% faster.
// k_nh_p: p = res + beta p_old over the whole array with setcyclic_xyz, and
// the partial maxima of absmax_3D(p) (slot 1).
__global__ void __launch_bounds__(256) k_nh_p(Geom g, const uint8_t* __restrict__ m8,
                                              const double* __restrict__ r,
                                              const double* __restrict__ p_old,
                                              double* __restrict__ p_new,
                                              const SpScal* __restrict__ S,
                                              double* __restrict__ part) {
#pragma clang fp contract(off)
    // the flag is tested once the first point's loads are in flight; every
    // thread of a finished solve returns before the partials' barrier
    const int done = S->done;
    const double beta = S->beta;   // 0 in iteration 1, with the old direction zeroed
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t nx = (size_t)g.nx;
    const size_t base = (size_t)blockIdx.y * plane;
    const bool cyc = g.cyclic != 0;
    double s_max = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < plane;
         q += (size_t)gridDim.x * blockDim.x) {
        int row, col;
        sp_rc(g, q, row, col);
        const size_t o = base + q;
        const size_t src = o + ((cyc && col < 2) ? nx : 0) - ((cyc && col >= g.nx + 2) ? nx : 0);
        double rv = r[src], pv = p_old[src];
        int mm = m8[o];
        pin_loaded(rv), pin_loaded(pv), pin_loaded(mm);
        if (done) return;
        const double v = rv + beta * pv;
        p_new[o] = v;
        const double mt = sp_inner(g, row, col) ? (double)((mm >> 3) & 1) : 0.0;
        s_max = fmax(fabs(v * mt), s_max);
    }
    if (done) return;
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) s_max = fmax(s_max, __shfl_xor(s_max, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s_max;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nb = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
        part[nb + b] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    }
}

// k_nh_ap: Ap = A p on the interior (zero elsewhere; p's halo columns hold
// their cyclic copies) and the partial sums of dot_3D(p, Ap) (slot 0).
// Workgroup row blockIdx.y is level k; within a level, consecutive chunks of
// the plane go to the workgroups of one XCD (gridDim.x is a multiple of 8, so
// the chunk's levels above and below run on that XCD too), and its L2 serves
// the neighbours (a plain grid-stride walk fetched 3.5x the reads).
__global__ void __launch_bounds__(256) k_nh_ap(Geom g, int nz, NhTables t,
                                               const uint8_t* __restrict__ m8,
                                               const double* __restrict__ p,
                                               double* __restrict__ Ap,
                                               const SpScal* __restrict__ S,
                                               double* __restrict__ part) {
#pragma clang fp contract(off)
    if (S->done) return;
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t P = g.pitch;
    const int k = blockIdx.y;
    const size_t base = (size_t)k * plane;
    const size_t dn = k > 0 ? plane : 0, up = k < nz - 1 ? plane : 0;
    const size_t q_in = 2 * P + 2;
    double s_dot = 0.0;
    const size_t chunk = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x);
    for (size_t q = chunk * blockDim.x + threadIdx.x; q < plane;
         q += (size_t)gridDim.x * blockDim.x) {
        int row, col;
        sp_rc(g, q, row, col);
        const bool inner = sp_inner(g, row, col);
        const size_t oc = base + (inner ? q : q_in);
        const uint8_t m = m8[oc];
        const NhCoef f = nh_coef(t, row, col, k, nz, m, m8[oc - 1], m8[oc - P], m8[oc - dn]);
        const double pc = p[oc];
        const double a = nh_apply(f, p[oc - dn], p[oc - P], p[oc - 1], pc, p[oc + 1],
                                  p[oc + P], p[oc + up]);
        const double mt = inner ? (double)((m >> 3) & 1) : 0.0;
        Ap[base + q] = inner ? a : 0.0;
        s_dot = s_dot + pc * (inner ? a : 0.0) * mt;
    }
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) s_dot += __shfl_xor(s_dot, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s_dot;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.y * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// k_nh_ap on LPT consecutive levels per thread (nz a multiple of LPT):
// the point's row, column and the divisions of its horizontal coefficients
// (level-independent: ex_e / c2, ex_w / c2) are formed once for its LPT
// levels, and each level's centre value and mask serve the next level as
// its "below" ones, so a level costs five loads of p and three mask bytes
// instead of seven and four.  The coefficients, the stencil and every
// partial are those of k_nh_ap bit for bit: workgroup (x, y) takes the
// same chunk of each of its levels as workgroup x of that level in k_nh_ap,
// and writes one partial per level, in the same slot.
template <int LPT>
__global__ void __launch_bounds__(256) k_nh_ap_l(Geom g, int nz, NhTables t,
                                                 const uint8_t* __restrict__ m8,
                                                 const double* __restrict__ p,
                                                 double* __restrict__ Ap,
                                                 const SpScal* __restrict__ S,
                                                 double* __restrict__ part) {
#pragma clang fp contract(off)
    if (S->done) return;
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t P = g.pitch;
    const int k0 = blockIdx.y * LPT;
    const size_t q_in = 2 * P + 2;
    double s_dot[LPT];
#pragma unroll
    for (int l = 0; l < LPT; ++l) s_dot[l] = 0.0;
    const size_t chunk = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x);
    for (size_t q = chunk * blockDim.x + threadIdx.x; q < plane;
         q += (size_t)gridDim.x * blockDim.x) {
        int row, col;
        sp_rc(g, q, row, col);
        const bool inner = sp_inner(g, row, col);
        const size_t qc = inner ? q : q_in;
        const int rc = inner ? row : 2, cc = inner ? col : 2;
        // nh_coef's level-independent parts
        const double c2 = t.c2[rc];
        const double e0 = t.ex_e[cc] / c2, w0 = t.ex_w[cc] / c2, n0 = t.tn[rc], s0 = t.ts[rc];
        // the level below the first (k_nh_ap reads the point itself at k = 0)
        size_t oc = (size_t)k0 * plane + qc;
        double pd = p[k0 > 0 ? oc - plane : oc];
        uint8_t m_d = m8[k0 > 0 ? oc - plane : oc];
        double pc = p[oc];
#pragma unroll
        for (int l = 0; l < LPT; ++l) {
            const int k = k0 + l;
            const uint8_t m = m8[oc], m_w = m8[oc - 1], m_s = m8[oc - P];
            const double pu = p[k < nz - 1 ? oc + plane : oc];
            NhCoef f;
            f.e = (m & NH_MU) ? e0 : 0.0;
            f.w = (m_w & NH_MU) ? w0 : 0.0;
            f.n = (m & NH_MV) ? n0 : 0.0;
            f.s = (m_s & NH_MV) ? s0 : 0.0;
            f.u = (k < nz - 1 && (m & NH_MW)) ? t.tu[k] : 0.0;
            f.d = (k > 0 && (m_d & NH_MW)) ? t.td[k] : 0.0;
            double c0 = (((0.0 - f.e) - f.w) - f.n) - f.s;
            if (k < nz - 1) c0 = c0 - f.u;
            if (k > 0) c0 = c0 - f.d;
            f.c = c0;
            const double a = nh_apply(f, pd, p[oc - P], p[oc - 1], pc, p[oc + 1], p[oc + P], pu);
            const double mt = inner ? (double)((m >> 3) & 1) : 0.0;
            Ap[(size_t)k * plane + q] = inner ? a : 0.0;
            s_dot[l] = s_dot[l] + pc * (inner ? a : 0.0) * mt;
            pd = pc;
            pc = pu;
            m_d = m;
            oc += plane;
        }
    }
    __shared__ double sh[LPT][4];
#pragma unroll
    for (int l = 0; l < LPT; ++l) {
        double v = s_dot[l];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) sh[l][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < LPT) {
        const int l = threadIdx.x;
        part[(size_t)(k0 + l) * gridDim.x + blockIdx.x] =
            ((sh[l][0] + sh[l][1]) + sh[l][2]) + sh[l][3];
    }
}

// p_non_hydro += alpha p, res -= alpha Ap over the whole array (:188-189), and
// the partials of dot_3D(res, res) (:190).
__global__ void __launch_bounds__(256) k_nh_update(Geom g, int nz,
                                                   const uint8_t* __restrict__ m8,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ Ap,
                                                   const SpScal* __restrict__ S,
                                                   double* __restrict__ part) {
#pragma clang fp contract(off)
    // the flag is tested once the first point's loads are in flight; every
    // thread of a finished solve returns before the partials' barrier
    const int done = S->done;
    const double alpha = S->alpha;
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t base = (size_t)blockIdx.y * plane;
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < plane;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t o = base + q;
        double xv = x[o], pv = p[o], rv = r[o], av = Ap[o];
        int mm = m8[o];
        pin_loaded(xv), pin_loaded(pv), pin_loaded(rv), pin_loaded(av), pin_loaded(mm);
        if (done) return;
        x[o] = xv + alpha * pv;
        const double rn = rv - alpha * av;
        r[o] = rn;
        int row, col;
        sp_rc(g, q, row, col);
        if (sp_inner(g, row, col)) acc = acc + rn * rn * (double)((mm >> 3) & 1);
    }
    if (done) return;
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.y * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// w(taup1) = w(tau) + dt (dw_mix + w1 dw(tau) - w0 dw(taum1)) maskW for
// k < nz, whole planes (non_hydrostatic.f90:15-17).
__global__ void __launch_bounds__(256) k_nh_wstep(size_t n, double dt, double w1, double w0,
                                                  const double* __restrict__ w_tau,
                                                  const double* __restrict__ dw,
                                                  const double* __restrict__ dw_m1,
                                                  const double* __restrict__ dw_mix,
                                                  const double* __restrict__ maskW,
                                                  double* __restrict__ w_new) {
#pragma clang fp contract(off)
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    w_new[q] = w_tau[q] + dt * (dw_mix[q] + w1 * dw[q] - w0 * dw_m1[q]) * maskW[q];
}

// setcyclic_xyz: every row of every level (gridDim.y = levels).
__global__ void k_setcyclic3(Geom g, double* __restrict__ a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.ny + 4) return;
    double* row = a + (size_t)blockIdx.y * g.pitch * (g.ny + 4) + (size_t)j * g.pitch;
    for (int i = 1; i <= 2; ++i) {
        row[g.nx + i + 1] = row[i + 1];
        row[1 - i + 1] = row[g.nx - i + 1 + 1];
    }
}

// forc = ((u(i)-u(i-1))/(cost dxt) + (cosu(j) v(j) - cosu(j-1) v(j-1))/(cost dyt)
//         + (w(k) - w(k-1))/dzt(k)) / dt on the interior (w(0) = 0 term absent
// at k = 1), zero elsewhere (:26-38); with p_new given, also the first guess
// p(taup1) = 2 p(tau) - p(taum1) everywhere (:43).
__global__ void __launch_bounds__(256) k_nh_rhs(Geom g, int nz, double dt,
                                                const double* __restrict__ dxt,
                                                const double* __restrict__ dyt,
                                                const double* __restrict__ cost,
                                                const double* __restrict__ cosu,
                                                const double* __restrict__ dzt,
                                                const double* __restrict__ u,
                                                const double* __restrict__ v,
                                                const double* __restrict__ w,
                                                const double* __restrict__ p_tau,
                                                const double* __restrict__ p_m1,
                                                double* __restrict__ p_new,
                                                double* __restrict__ forc) {
#pragma clang fp contract(off)
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= plane) return;
    const int k = blockIdx.y;
    int row, col;
    sp_rc(g, q, row, col);
    const size_t o = (size_t)k * plane + q;
    if (p_new) p_new[o] = 2.0 * p_tau[o] - p_m1[o];
    double f = 0.0;
    if (sp_inner(g, row, col)) {
        f = (u[o] - u[o - 1]) / (cost[row] * dxt[col]) +
            (cosu[row] * v[o] - cosu[row - 1] * v[o - g.pitch]) / (cost[row] * dyt[row]);
        if (k == 0) f = f + w[o] / dzt[0];
        else f = f + (w[o] - w[o - plane]) / dzt[k];
        f = f / dt;
    }
    forc[o] = f;
}

// The pressure gradient removed from u, v (interior, every level) and w
// (whole planes, k < nz) at taup1 (:53-63); p after setcyclic_xyz.
__global__ void __launch_bounds__(256) k_nh_velocity(Geom g, int nz, double dt,
                                                     const double* __restrict__ dxu,
                                                     const double* __restrict__ dyu,
                                                     const double* __restrict__ cost,
                                                     const double* __restrict__ dzw,
                                                     const double* __restrict__ maskU,
                                                     const double* __restrict__ maskV,
                                                     const double* __restrict__ maskW,
                                                     const double* __restrict__ p,
                                                     double* __restrict__ u,
                                                     double* __restrict__ v,
                                                     double* __restrict__ w) {
#pragma clang fp contract(off)
    const size_t plane = (size_t)g.pitch * (g.ny + 4);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= plane) return;
    const int k = blockIdx.y;
    int row, col;
    sp_rc(g, q, row, col);
    const size_t o = (size_t)k * plane + q;
    if (sp_inner(g, row, col)) {
        u[o] = u[o] - dt * (p[o + 1] - p[o]) / (dxu[col] * cost[row]) * maskU[o];
        v[o] = v[o] - dt * (p[o + g.pitch] - p[o]) / dyu[row] * maskV[o];
    }
    if (k < nz - 1) w[o] = w[o] - dt * (p[o + plane] - p[o]) / dzw[k] * maskW[o];
}

}  // namespace psi

// ---------------------------------------------------------------------------
// The non-hydrostatic context: grid, levels, packed masks and 1-D tables, the
// CG vectors (nz planes each) and a captured graph of two iterations per
// solution buffer.
// ---------------------------------------------------------------------------
struct psi_nh : psi_ctx {
    int nz = 0;
    size_t n3 = 0;
    int gx = 0;                      // workgroups per level of the grid-stride passes
    uint8_t* m8 = nullptr;
    const double *maskT = nullptr, *maskU = nullptr, *maskV = nullptr, *maskW = nullptr;
    double* tables = nullptr;        // ex_e, ex_w (nx+4 each), c2, tn, ts (ny+4 each),
                                     // tu, td (nz each), dzt, dzw (nz each)
    psi::NhTables T{};
    const double *d_dzt = nullptr, *d_dzw = nullptr;
    std::vector<double> h_dzt, h_dzw;
    double *r = nullptr, *pv[2] = {nullptr, nullptr}, *Ap = nullptr, *part = nullptr;
    double* forc = nullptr;
    psi::SpScal *S = nullptr, *h_S = nullptr, *S_time = nullptr;
    int* d_bad = nullptr;
    struct Graph {
        double* x = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    hipStream_t gs = nullptr;
    hipEvent_t e_in = nullptr, e_out = nullptr, g0 = nullptr, g1 = nullptr;
};

namespace {

void nh_release_graphs(psi_nh* c) {
    for (auto& G : c->graphs) {
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        if (G.graph) (void)hipGraphDestroy(G.graph);
    }
    c->graphs.clear();
}

bool valid_nh(const psi_nh* c) { return c && c->m8 && c->S; }

// A p with its dot partials: several levels per thread where nz allows
// (k_nh_ap_l, the same results), else one (k_nh_ap); PSI_NH_LEVELS=1 keeps
// one level per thread (tests compare the two).
void nh_launch_ap(psi_nh* c, hipStream_t s, const double* pin, const SpScal* S) {
    using namespace psi;
    const int nz = c->nz;
    const uint8_t* m8 = c->m8;
    const char* ev = getenv("PSI_NH_LEVELS");
    const bool one = ev && atoi(ev) == 1;
    if (one)
        hipLaunchKernelGGL(k_nh_ap, dim3((unsigned)c->gx, (unsigned)nz), dim3(256), 0, s, c->g,
                           nz, c->T, m8, pin, c->Ap, S, c->part);
    else if (nz % 4 == 0)
        hipLaunchKernelGGL(k_nh_ap_l<4>, dim3((unsigned)c->gx, (unsigned)(nz / 4)), dim3(256), 0,
                           s, c->g, nz, c->T, m8, pin, c->Ap, S, c->part);
    else if (nz % 2 == 0)
        hipLaunchKernelGGL(k_nh_ap_l<2>, dim3((unsigned)c->gx, (unsigned)(nz / 2)), dim3(256), 0,
                           s, c->g, nz, c->T, m8, pin, c->Ap, S, c->part);
    else
        hipLaunchKernelGGL(k_nh_ap, dim3((unsigned)c->gx, (unsigned)nz), dim3(256), 0, s, c->g,
                           nz, c->T, m8, pin, c->Ap, S, c->part);
}

void nh_iterations(psi_nh* c, hipStream_t s, double* x) {
    using namespace psi;
    const dim3 grid((unsigned)c->gx, (unsigned)c->nz);
    const int nb = c->gx * c->nz;
    for (int h = 0; h < 2; ++h) {
        hipLaunchKernelGGL(k_nh_p, grid, dim3(256), 0, s, c->g, (const uint8_t*)c->m8,
                           (const double*)c->r, (const double*)c->pv[h], c->pv[1 - h],
                           (const SpScal*)c->S, c->part);
        nh_launch_ap(c, s, (const double*)c->pv[1 - h], (const SpScal*)c->S);
        hipLaunchKernelGGL(k_sp_fold_a, dim3(1), dim3(1024), 0, s, (const double*)c->part, nb,
                           c->S);
        hipLaunchKernelGGL(k_nh_update, grid, dim3(256), 0, s, c->g, c->nz,
                           (const uint8_t*)c->m8, x, c->r, (const double*)c->pv[1 - h],
                           (const double*)c->Ap, (const SpScal*)c->S, c->part);
        hipLaunchKernelGGL(k_sp_fold_b, dim3(1), dim3(1024), 0, s, (const double*)c->part, nb,
                           c->S);
    }
}

int nh_graph(psi_nh* c, double* x, hipGraphExec_t* exec) {
    for (auto& G : c->graphs)
        if (G.x == x) {
            *exec = G.exec;
            return PSI_OK;
        }
    if (c->graphs.size() >= 4) nh_release_graphs(c);
    psi_nh::Graph G;
    G.x = x;
    HIPCHK(c, hipStreamBeginCapture(c->gs, hipStreamCaptureModeThreadLocal));
    nh_iterations(c, c->gs, x);
    HIPCHK(c, hipStreamEndCapture(c->gs, &G.graph));
    hipError_t e = hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(G.graph);
        return fail(c, PSI_HIP_ERROR, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    c->graphs.push_back(G);
    *exec = G.exec;
    return PSI_OK;
}

// congrad_non_hydro on x (device, nz planes, in place) with the reference's
// stopping rule (no NaN test in this solver).
int nh_congrad(psi_nh* c, const psi_opts* o, const double* forc, double* x, psi_info* info) {
    using namespace psi;
    if (o->mode != PSI_MODE_PLAIN || o->criterion != PSI_CRIT_REFERENCE)
        return fail(c, PSI_BAD_ARG,
                    "non-hydrostatic pressure: mode plain with the reference criterion only%s");
    const int batch = o->check_every > 0 ? (o->check_every + 1) / 2 : 16;
    hipGraphExec_t exec = nullptr;
    int rc = nh_graph(c, x, &exec);
    if (rc) return rc;
    const int nb = c->gx * c->nz;
    HIPCHK(c, hipEventRecord(c->t0, c->stream));
    SpScal S0{};
    S0.eps = o->congr_epsilon;
    S0.max_it = o->max_iterations;
    S0.n = 1;
    S0.nan_test = 0;
    *c->h_S = S0;
    HIPCHK(c, hipMemcpyAsync(c->S, c->h_S, sizeof(SpScal), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_nh_init, dim3((unsigned)c->gx, (unsigned)c->nz), dim3(256), 0,
                       c->stream, c->g, c->nz, c->T, (const uint8_t*)c->m8, forc,
                       (const double*)x, c->r, c->pv[0], c->part);
    hipLaunchKernelGGL(k_sp_fold_init, dim3(1), dim3(1024), 0, c->stream, (const double*)c->part,
                       nb, c->S);
    rc = check_launch(c);
    if (rc) return rc;
    info->setup_launches = 2;
    HIPCHK(c, hipEventRecord(c->e_in, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->gs, c->e_in, 0));
    HIPCHK(c, hipEventRecord(c->g0, c->gs));
    int replays = 0;
    for (;;) {
        for (int k = 0; k < batch; ++k) HIPCHK(c, hipGraphLaunch(exec, c->gs));
        replays += batch;
        HIPCHK(c, hipMemcpyAsync(c->h_S, c->S, sizeof(SpScal), hipMemcpyDeviceToHost, c->gs));
        HIPCHK(c, hipStreamSynchronize(c->gs));
        if (c->h_S->done) break;
    }
    HIPCHK(c, hipEventRecord(c->g1, c->gs));
    HIPCHK(c, hipEventRecord(c->e_out, c->gs));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->e_out, 0));
    HIPCHK(c, hipEventRecord(c->t1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->t1));
    float ms = 0.0f, span = 0.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    (void)hipEventElapsedTime(&span, c->g0, c->g1);
    const SpScal& S = *c->h_S;
    info->iterations = S.iterations;
    info->status = status_of(S.status);
    info->converged = S.status == ST_CONVERGED;
    info->outer_iterations = 0;
    info->est_error = S.est;
    info->relres_recursive = 0.0;
    info->relres_true = -1.0;
    info->elapsed_ms = ms;
    info->sweep_ms = span;
    info->sweep_launches = 8 * replays;
    if (info->status == PSI_NOT_CONVERGED)
        fail(c, PSI_NOT_CONVERGED, "max iterations exceeded (non-hydrostatic pressure)%s");
    else if (info->status == PSI_DIVERGED)
        fail(c, PSI_DIVERGED, "non hydrostatic solver diverging%s");
    return info->status;
}

}  // namespace

extern "C" {

const char* psi_nh_last_error(const psi_nh* c) {
    return c ? c->err.c_str() : g_create_err.c_str();
}

void psi_nh_destroy(psi_nh* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gs) (void)hipStreamSynchronize(c->gs);
    nh_release_graphs(c);
    if (c->gs) (void)hipStreamDestroy(c->gs);
    for (hipEvent_t e : {c->e_in, c->e_out, c->g0, c->g1})
        if (e) (void)hipEventDestroy(e);
    if (c->h_S) (void)hipHostFree(c->h_S);
    c->h_S = nullptr;
    ctx_release(c);
    delete c;
}

int psi_nh_create(psi_nh** out, const psi_grid* grid, const psi_levels* lev,
                  const double* maskW, void* stream) {
    using namespace psi;
    if (!out || !grid || !lev || !maskW) return fail(nullptr, PSI_BAD_ARG, "null argument%s");
    *out = nullptr;
    if (grid->nx < 2 || grid->ny < 2 || !grid->dxt || !grid->dxu || !grid->dyt || !grid->dyu ||
        !grid->cost || !grid->cosu)
        return fail(nullptr, PSI_BAD_ARG, "bad grid%s");
    if (lev->nz < 2 || lev->nz > 65535 || !lev->dzt || !lev->dzw || !lev->maskT ||
        !lev->maskU || !lev->maskV)
        return fail(nullptr, PSI_BAD_ARG,
                    "levels: nz must be at least 2 (make_coeff_non_hydro reads maskW(:,:,k-1) "
                    "and dzw(k-1) at k = nz) and every array given%s");
    if ((double)(grid->nx + 4) * (double)(grid->ny + 4) >= 2147483648.0)
        return fail(nullptr, PSI_BAD_ARG, "a plane must hold fewer than 2^31 points%s");
    psi_nh* c = new psi_nh();
    c->stream = (hipStream_t)stream;
    (void)hipGetDevice(&c->device);
    Geom& g = c->g;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.pitch = grid->nx + 4;
    g.cyclic = grid->cyclic_x ? 1 : 0;
    c->plane = (size_t)(g.nx + 4) * (g.ny + 4);
    c->nz = lev->nz;
    c->n3 = (size_t)c->nz * c->plane;
    // workgroups per level: a multiple of 8 (the XCD chunk mapping), about
    // 8192 in all, at most one per 256 points of a level
    c->gx = std::max(8, (8192 / c->nz) & ~7);
    c->gx = std::min<int>(c->gx, std::max<int>(8, (int)nblk(c->plane) & ~7));
    const size_t nxh = g.nx + 4, nyh = g.ny + 4, nz = (size_t)c->nz;
    c->h_dxt.assign(grid->dxt, grid->dxt + nxh);
    c->h_dxu.assign(grid->dxu, grid->dxu + nxh);
    c->h_dyt.assign(grid->dyt, grid->dyt + nyh);
    c->h_dyu.assign(grid->dyu, grid->dyu + nyh);
    c->h_cost.assign(grid->cost, grid->cost + nyh);
    c->h_cosu.assign(grid->cosu, grid->cosu + nyh);
    c->h_dzt.assign(lev->dzt, lev->dzt + nz);
    c->h_dzw.assign(lev->dzw, lev->dzw + nz);
    c->maskT = lev->maskT;
    c->maskU = lev->maskU;
    c->maskV = lev->maskV;
    c->maskW = maskW;
    // 1-D factors of make_coeff_non_hydro with unit masks, in its operation order
    const size_t ntab = 2 * nxh + 3 * nyh + 4 * nz;
    std::vector<double> tab(ntab, 0.0);
    double *ex_e = tab.data(), *ex_w = ex_e + nxh, *c2 = ex_w + nxh, *tn = c2 + nyh,
           *ts = tn + nyh, *tu = ts + nyh, *td = tu + nz, *dzt = td + nz, *dzw = dzt + nz;
    {
#pragma clang fp contract(off)
        for (size_t i = 1; i < nxh; ++i) {
            ex_e[i] = 1.0 / c->h_dxu[i] / c->h_dxt[i];
            ex_w[i] = 1.0 / c->h_dxu[i - 1] / c->h_dxt[i];
        }
        for (size_t j = 1; j < nyh; ++j) {
            c2[j] = c->h_cost[j] * c->h_cost[j];
            tn[j] = 1.0 / c->h_dyu[j] / c->h_dyt[j] * c->h_cosu[j] / c->h_cost[j];
            ts[j] = 1.0 / c->h_dyu[j - 1] / c->h_dyt[j] * c->h_cosu[j - 1] / c->h_cost[j];
        }
        for (size_t k = 0; k < nz; ++k) {
            tu[k] = 1.0 / c->h_dzw[k] / c->h_dzt[k];
            td[k] = k > 0 ? 1.0 / c->h_dzw[k - 1] / c->h_dzt[k] : 0.0;
            dzt[k] = c->h_dzt[k];
            dzw[k] = c->h_dzw[k];
        }
    }
    int rc = PSI_OK;
    int bad = 0;
#define TRY(x)               \
    do {                     \
        rc = (x);            \
        if (rc) goto failed; \
    } while (0)
    {
        TRY(dalloc(c, &c->dxt, nxh));
        TRY(dalloc(c, &c->dxu, nxh));
        TRY(dalloc(c, &c->dyt, nyh));
        TRY(dalloc(c, &c->dyu, nyh));
        TRY(dalloc(c, &c->cost, nyh));
        TRY(dalloc(c, &c->cosu, nyh));
        TRY(dalloc(c, &c->tables, ntab));
        TRY(dalloc(c, &c->m8, c->n3));
        TRY(dalloc(c, &c->r, c->n3));
        TRY(dalloc(c, &c->pv[0], c->n3));
        TRY(dalloc(c, &c->pv[1], c->n3));
        TRY(dalloc(c, &c->Ap, c->n3));
        TRY(dalloc(c, &c->forc, c->n3));
        TRY(dalloc(c, &c->part, 2 * (size_t)c->gx * nz));
        TRY(dalloc(c, &c->S, 1));
        TRY(dalloc(c, &c->S_time, 1));
        TRY(dalloc(c, &c->d_bad, 1));
        if (hipHostMalloc((void**)&c->h_S, sizeof(SpScal), hipHostMallocDefault) != hipSuccess ||
            hipStreamCreateWithFlags(&c->gs, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess ||
            hipEventCreateWithFlags(&c->e_in, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->e_out, hipEventDisableTiming) != hipSuccess ||
            hipEventCreate(&c->g0) != hipSuccess || hipEventCreate(&c->g1) != hipSuccess) {
            rc = fail(c, PSI_HIP_ERROR, "stream/event/pinned allocation failed%s");
            goto failed;
        }
        if (hipMemcpyAsync(c->dxt, grid->dxt, nxh * 8, hipMemcpyHostToDevice, c->stream) ||
            hipMemcpyAsync(c->dxu, grid->dxu, nxh * 8, hipMemcpyHostToDevice, c->stream) ||
            hipMemcpyAsync(c->dyt, grid->dyt, nyh * 8, hipMemcpyHostToDevice, c->stream) ||
            hipMemcpyAsync(c->dyu, grid->dyu, nyh * 8, hipMemcpyHostToDevice, c->stream) ||
            hipMemcpyAsync(c->cost, grid->cost, nyh * 8, hipMemcpyHostToDevice, c->stream) ||
            hipMemcpyAsync(c->cosu, grid->cosu, nyh * 8, hipMemcpyHostToDevice, c->stream) ||
            hipMemcpyAsync(c->tables, tab.data(), ntab * 8, hipMemcpyHostToDevice, c->stream)) {
            rc = fail(c, PSI_HIP_ERROR, "metric upload failed%s");
            goto failed;
        }
        c->T.ex_e = c->tables;
        c->T.ex_w = c->tables + nxh;
        c->T.c2 = c->tables + 2 * nxh;
        c->T.tn = c->T.c2 + nyh;
        c->T.ts = c->T.tn + nyh;
        c->T.tu = c->T.ts + nyh;
        c->T.td = c->T.tu + nz;
        c->d_dzt = c->T.td + nz;
        c->d_dzw = c->d_dzt + nz;

        hipLaunchKernelGGL(k_nh_masks, dim3(nblk(c->n3)), dim3(256), 0, c->stream, c->n3,
                           lev->maskT, lev->maskU, lev->maskV, maskW, c->m8, c->d_bad);
        TRY(check_launch(c));
        if (hipMemcpyAsync(&bad, c->d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream) ||
            hipStreamSynchronize(c->stream)) {
            rc = fail(c, PSI_HIP_ERROR, "operator setup failed%s");
            goto failed;
        }
        if (bad) {
            rc = fail(c, PSI_BAD_ARG, "masks must hold 0 or 1%s");
            goto failed;
        }
    }
    *out = c;
    return PSI_OK;
failed:
    g_create_err = c->err;
    psi_nh_destroy(c);
    return rc;
#undef TRY
}

int psi_nh_set_stream(psi_nh* c, void* stream) {
    if (!valid_nh(c)) return PSI_BAD_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)stream;
    return PSI_OK;
}

int64_t psi_nh_device_bytes(const psi_nh* c) { return c ? c->bytes : 0; }

int psi_make_coeff_non_hydro(psi_nh* c, double* cf) {
    using namespace psi;
    if (!valid_nh(c) || !cf) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_nh_export, dim3(nblk(c->plane), (unsigned)c->nz), dim3(256), 0,
                       c->stream, c->g, c->nz, c->T, (const uint8_t*)c->m8, cf);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_apply_op_3d(psi_nh* c, const double* p, double* res) {
    using namespace psi;
    if (!valid_nh(c) || !p || !res) return PSI_BAD_ARG;
    hipLaunchKernelGGL(k_nh_apply, dim3(nblk(c->plane), (unsigned)c->nz), dim3(256), 0,
                       c->stream, c->g, c->nz, c->T, (const uint8_t*)c->m8, p, res);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSI_OK;
}

int psi_congrad_non_hydro(psi_nh* c, const psi_opts* o, const double* forc, double* p,
                          psi_info* info) {
    if (!valid_nh(c) || !o || !forc || !p || !info) return PSI_BAD_ARG;
    *info = psi_info{};
    return nh_congrad(c, o, forc, p, info);
}

int psi_solve_non_hydrostatic(psi_nh* c, const psi_opts* o, double* u, double* v, double* w,
                              const double* dw, const double* dw_mix, double* p_nh,
                              int32_t taum1, int32_t tau, int32_t taup1, double dt_mom,
                              double AB_eps, int32_t itt, double* forc_out, psi_info* info) {
    using namespace psi;
    if (!valid_nh(c) || !o || !u || !v || !w || !dw || !dw_mix || !p_nh || !info ||
        !(dt_mom != 0.0))
        return PSI_BAD_ARG;
    if (taum1 < 0 || taum1 > 2 || tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2)
        return fail(c, PSI_BAD_ARG, "time level out of range%s");
    *info = psi_info{};
    const size_t V = c->n3, P = c->plane;
    const int nz = c->nz;
    const double w1 = 1.5 + AB_eps, w0 = 0.5 + AB_eps;
    double* forc = forc_out ? forc_out : c->forc;
    double *ut = u + taup1 * V, *vt = v + taup1 * V, *wt = w + taup1 * V;
    const dim3 lv(nblk(P), (unsigned)nz), rows(nblk(c->g.ny + 4), (unsigned)nz);
    hipLaunchKernelGGL(k_nh_wstep, dim3(nblk((size_t)(nz - 1) * P)), dim3(256), 0, c->stream,
                       (size_t)(nz - 1) * P, dt_mom, w1, w0, (const double*)(w + tau * V),
                       dw + tau * V, dw + taum1 * V, dw_mix, c->maskW, wt);
    if (c->g.cyclic)
        for (double* a : {ut, vt, wt})
            hipLaunchKernelGGL(k_setcyclic3, rows, dim3(256), 0, c->stream, c->g, a);
    hipLaunchKernelGGL(k_nh_rhs, lv, dim3(256), 0, c->stream, c->g, nz, dt_mom,
                       (const double*)c->dxt, (const double*)c->dyt, (const double*)c->cost,
                       (const double*)c->cosu, c->d_dzt, (const double*)ut, (const double*)vt,
                       (const double*)wt, (const double*)(p_nh + tau * V),
                       (const double*)(p_nh + taum1 * V), p_nh + taup1 * V, forc);
    int rc = check_launch(c);
    if (rc) return rc;
    rc = nh_congrad(c, o, forc, p_nh + taup1 * V, info);
    if (rc == PSI_HIP_ERROR || rc == PSI_BAD_ARG) return rc;
    const int solve_rc = rc;
    if (c->g.cyclic)
        hipLaunchKernelGGL(k_setcyclic3, rows, dim3(256), 0, c->stream, c->g, p_nh + taup1 * V);
    if (itt == 0) {   // the first time step starts all three levels there (:47-50)
        HIPCHK(c, hipMemcpyAsync(p_nh + tau * V, p_nh + taup1 * V, V * 8,
                                 hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(p_nh + taum1 * V, p_nh + taup1 * V, V * 8,
                                 hipMemcpyDeviceToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_nh_velocity, lv, dim3(256), 0, c->stream, c->g, nz, dt_mom,
                       (const double*)c->dxu, (const double*)c->dyu, (const double*)c->cost,
                       c->d_dzw, c->maskU, c->maskV, c->maskW, (const double*)(p_nh + taup1 * V),
                       ut, vt, wt);
    rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return solve_rc;
}

// Average durations (us) of `reps` launches of the iteration's kernels: [0]
// the direction pair k_nh_p + k_nh_ap, [1] k_nh_update, [2] k_sp_fold_a, [3]
// k_sp_fold_b, and with n >= 6 [4] k_nh_p and [5] k_nh_ap alone; HIP events
// on the graph stream, on the context's own vectors (call between solves).
int psi_nh_kernel_times(psi_nh* c, int32_t reps, double* us, int32_t n) {
    using namespace psi;
    if (!valid_nh(c) || !us || n < 4 || reps < 1) return PSI_BAD_ARG;
    SpScal z{};
    z.n = 2;
    z.rsold = 1.0;
    z.max_it = 1 << 30;
    z.step1 = 1.0;
    z.rs_min = 1e300;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->S_time, &z, sizeof z, hipMemcpyHostToDevice));
    auto timed = [&](auto fn) -> double {
        if (hipEventRecord(c->g0, c->gs) != hipSuccess) return -1.0;
        for (int k = 0; k < reps; ++k) fn();
        if (hipEventRecord(c->g1, c->gs) != hipSuccess || hipEventSynchronize(c->g1) != hipSuccess)
            return -1.0;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->g0, c->g1);
        return 1e3 * ms / reps;
    };
    const dim3 grid((unsigned)c->gx, (unsigned)c->nz);
    const int nb = c->gx * c->nz;
    const SpScal* St = c->S_time;
    us[0] = timed([&] {
        hipLaunchKernelGGL(k_nh_p, grid, dim3(256), 0, c->gs, c->g, (const uint8_t*)c->m8,
                           (const double*)c->r, (const double*)c->pv[0], c->pv[1], St, c->part);
        nh_launch_ap(c, c->gs, (const double*)c->pv[1], St);
    });
    if (n >= 6) {
        us[4] = timed([&] {
            hipLaunchKernelGGL(k_nh_p, grid, dim3(256), 0, c->gs, c->g, (const uint8_t*)c->m8,
                               (const double*)c->r, (const double*)c->pv[0], c->pv[1], St,
                               c->part);
        });
        us[5] = timed([&] {
            nh_launch_ap(c, c->gs, (const double*)c->pv[1], St);
        });
    }
    us[1] = timed([&] {
        hipLaunchKernelGGL(k_nh_update, grid, dim3(256), 0, c->gs, c->g, c->nz,
                           (const uint8_t*)c->m8, c->pv[0], c->r, (const double*)c->pv[1],
                           (const double*)c->Ap, St, c->part);
    });
    us[2] = timed([&] {
        hipLaunchKernelGGL(k_sp_fold_a, dim3(1), dim3(1024), 0, c->gs, (const double*)c->part,
                           nb, c->S_time);
    });
    us[3] = timed([&] {
        hipLaunchKernelGGL(k_sp_fold_b, dim3(1), dim3(1024), 0, c->gs, (const double*)c->part,
                           nb, c->S_time);
    });
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->gs));
    for (int k = 0; k < 4; ++k)
        if (us[k] < 0) return fail(c, PSI_HIP_ERROR, "event timing failed%s");
    return PSI_OK;
}

}  // extern "C"
