// psi_mg.hip -- multigrid-preconditioned flexible CG for the streamfunction
// solve (no counterpart in the reference, whose congrad_streamfunction,
// solve_streamfct.f90:153-327, is Jacobi-preconditioned CG; this is a
// faster preconditioner for the same system, reached through psi_congrad
// with PSI_MODE_MG and the relative-residual criterion).
//
// The hierarchy is built on the host once per context (psi_mg_setup) from
// the stored operator.  Every level holds an explicit symmetric 5-point
// operator S = W A in FP32 (W the psi-cell areas dxu cosu dyu, so S is
// symmetric where A is not on spherical grids): east and north couplings
// E, N (the west and south couplings of a cell are its neighbours' E and N)
// and the diagonal C.  Level 0 is the solver grid's active points; level
// l+1 aggregates 2 x 2 cells of level l (active cells only) with the
// Galerkin product P^T S P of piecewise-constant prolongation, which stays
// 5-point.  The coarsest level (at most MG_MAX_DENSE active cells) is solved
// by a dense inverse computed on the host.
//
// One V-cycle: on each level K steps of Chebyshev smoothing of D^-1 S on
// [2/ratio, 2] from zero (Gershgorin bounds D^-1 S by 2 on every level),
// the residual restricted to the next level, the coarse correction added
// with weight MG_SCALE, and K Chebyshev steps again.  Vectors are FP32 in a
// 1-cell-halo layout per level; each kernel applies one stencil to a vector
// its predecessor wrote.
//
// The outer iteration is flexible CG (FCG(1)) on A x = b in FP64 with
// area-weighted inner products, the V-cycle applied to the area-weighted
// residual.  Scalars live on the device; one outer iteration (V-cycle, two
// passes, two folds) is a captured hipGraph replayed between host polls,
// and kernels return at once after convergence.
//
// Included by psi_capi.hip after psi_kernels.hip.

namespace psi {

constexpr int MG_MAX_DENSE = 1024;   // active cells of the coarsest level at most
constexpr int MG_MAX_LEVELS = 16;
constexpr int MG_MAXK = 6;           // Chebyshev steps per smoothing at most
constexpr float MG_SCALE = 1.7f;     // weight of the coarse-grid correction
constexpr int MG_NRED = 6;

struct MgLev {
    int nx, ny, pitch, cyc;          // interior cells; pitch = nx + 2
    const float *E, *N, *C;
    const uint8_t* act;
    float *b, *x, *y, *r, *d0, *d1;
};

struct MgCheb {
    float c0;               // 1/theta
    float c1[MG_MAXK];      // rho_k rho_{k-1}
    float c2[MG_MAXK];      // 2 rho_k / delta
};

// Device scalars of the outer iteration.
struct MgScal {
    double beta, alpha, pq, pq_old;
    double s_r, s_pending;           // scale of the FP32 residual fed to the V-cycle
    double rr, bb, rtarget2;
    int n, done, status, max_it;
};

__device__ __forceinline__ void mg_cell(const MgLev& L, int t, int& I, int& J, int& q) {
    I = t % L.nx;
    J = t / L.nx;
    q = (J + 1) * L.pitch + (I + 1);
}

__device__ __forceinline__ void mg_nbrs(const MgLev& L, int I, int q, int& qE, int& qW) {
    qE = (L.cyc && I == L.nx - 1) ? q - (L.nx - 1) : q + 1;
    qW = (L.cyc && I == 0) ? q + (L.nx - 1) : q - 1;
}

// S v at cell q
__device__ __forceinline__ float mg_apply(const MgLev& L, const float* v, int q, int qE, int qW) {
    return L.C[q] * v[q] + L.E[q] * v[qE] + L.E[qW] * v[qW] + L.N[q] * v[q + L.pitch] +
           L.N[q - L.pitch] * v[q - L.pitch];
}

// Pre-smoothing, steps 0 and 1 from x = 0: d0 = D^-1 b / theta at the cell
// and its neighbours, r1 = b - S d0, d1 = c1 d0 + c2 D^-1 r1, x = d0 + d1.
// With K = 1 only d0 is formed (x = d0, r = b).
__global__ void __launch_bounds__(256) k_mg_pre1(MgLev L, MgCheb ch, int K,
                                                 const MgScal* __restrict__ S) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    int I, J, q, qE, qW;
    mg_cell(L, t, I, J, q);
    mg_nbrs(L, I, q, qE, qW);
    const float d0 = L.b[q] * ch.c0 / L.C[q];
    if (K == 1) {
        L.x[q] = d0;
        L.r[q] = L.b[q];
        L.d0[q] = d0;
        return;
    }
    const float sd = L.C[q] * d0 + L.E[q] * (L.b[qE] * ch.c0 / L.C[qE]) +
                     L.E[qW] * (L.b[qW] * ch.c0 / L.C[qW]) +
                     L.N[q] * (L.b[q + L.pitch] * ch.c0 / L.C[q + L.pitch]) +
                     L.N[q - L.pitch] * (L.b[q - L.pitch] * ch.c0 / L.C[q - L.pitch]);
    const float r1 = L.b[q] - sd;
    const float d1 = ch.c1[1] * d0 + ch.c2[1] * (r1 / L.C[q]);
    L.r[q] = r1;
    L.d1[q] = d1;
    L.x[q] = d0 + d1;
}

// Chebyshev step k >= 1 (k = 1 only after post1): r -= S d_{k-1};
// d_k = c1 d_{k-1} + c2 D^-1 r; v += d_k, v = x (pre) or y (post).
// d_{k-1} is din, d_k goes to dout.
__global__ void __launch_bounds__(256) k_mg_step(MgLev L, MgCheb ch, int k, int post,
                                                 const float* __restrict__ din, float* dout,
                                                 const MgScal* __restrict__ S) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    int I, J, q, qE, qW;
    mg_cell(L, t, I, J, q);
    mg_nbrs(L, I, q, qE, qW);
    const float r = L.r[q] - mg_apply(L, din, q, qE, qW);
    const float d = ch.c1[k] * din[q] + ch.c2[k] * (r / L.C[q]);
    L.r[q] = r;
    dout[q] = d;
    float* v = post ? L.y : L.x;
    v[q] = v[q] + d;
}

// Residual after pre-smoothing, r - S d_last, restricted to the coarse
// level: one thread per coarse cell sums its (up to) 2 x 2 fine cells.
__global__ void __launch_bounds__(256) k_mg_restrict(MgLev L, const float* __restrict__ dlast,
                                                     MgLev Lc, const MgScal* __restrict__ S) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Lc.nx * Lc.ny) return;
    int IC, JC, Q;
    mg_cell(Lc, t, IC, JC, Q);
    float s = 0.0f;
    for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
            const int I = 2 * IC + di, J = 2 * JC + dj;
            if (I >= L.nx || J >= L.ny) continue;
            const int q = (J + 1) * L.pitch + (I + 1);
            if (!L.act[q]) continue;
            int qE, qW;
            mg_nbrs(L, I, q, qE, qW);
            s += L.r[q] - mg_apply(L, dlast, q, qE, qW);
        }
    Lc.b[Q] = Lc.act[Q] ? s : 0.0f;
}

// Post-smoothing step 0: y' = x + MG_SCALE * P e_c at the cell and its
// neighbours (active cells only), r = b - S y', d0 = D^-1 r / theta,
// y = y' + d0.
__device__ __forceinline__ float mg_prolonged(const MgLev& L, const MgLev& Lc, int q, int I,
                                              int J) {
    if (!L.act[q]) return 0.0f;
    return L.x[q] + MG_SCALE * Lc.y[(J / 2 + 1) * Lc.pitch + (I / 2 + 1)];
}

__global__ void __launch_bounds__(256) k_mg_post1(MgLev L, MgLev Lc, MgCheb ch,
                                                  const MgScal* __restrict__ S) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.nx * L.ny) return;
    int I, J, q, qE, qW;
    mg_cell(L, t, I, J, q);
    mg_nbrs(L, I, q, qE, qW);
    const int IE = (I == L.nx - 1) ? (L.cyc ? 0 : -1) : I + 1;
    const int IW = (I == 0) ? (L.cyc ? L.nx - 1 : -1) : I - 1;
    const float v = mg_prolonged(L, Lc, q, I, J);
    const float vE = IE >= 0 ? mg_prolonged(L, Lc, qE, IE, J) : 0.0f;
    const float vW = IW >= 0 ? mg_prolonged(L, Lc, qW, IW, J) : 0.0f;
    const float vN = J + 1 < L.ny ? mg_prolonged(L, Lc, q + L.pitch, I, J + 1) : 0.0f;
    const float vS = J > 0 ? mg_prolonged(L, Lc, q - L.pitch, I, J - 1) : 0.0f;
    const float sv = L.C[q] * v + L.E[q] * vE + L.E[qW] * vW + L.N[q] * vN +
                     L.N[q - L.pitch] * vS;
    const float r = L.b[q] - sv;
    const float d = ch.c0 * (r / L.C[q]);
    L.r[q] = r;
    L.d0[q] = d;
    L.y[q] = v + d;
}

// Coarsest level: y = Ainv b over its active cells (one wave per row).
__global__ void __launch_bounds__(256) k_mg_dense(MgLev L, const float* __restrict__ Ainv,
                                                  const int* __restrict__ cells, int n,
                                                  const MgScal* __restrict__ S) {
    if (S->done) return;
    const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float s = 0.0f;
    for (int k = lane; k < n; k += 64) s += Ainv[(size_t)row * n + k] * L.b[cells[k]];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) L.y[cells[row]] = s;
}

// ---- outer flexible CG (FP64, solver grid layout) ---------------------------
// z(i,j) = y0(i,j) / s_r with wrap in x (cyclic) and 0 outside the interior.
__device__ __forceinline__ double mg_z(const Geom& g, const MgLev& L0, int i, int j, double inv) {
    if (j < 0 || j >= g.ny) return 0.0;
    if (i < 0 || i >= g.nx) {
        if (!g.cyclic) return 0.0;
        i = (i + g.nx) % g.nx;
    }
    return (double)L0.y[(j + 1) * L0.pitch + (i + 1)] * inv;
}

// A z at interior point (i,j) (ctx index o): the stored operator in FP64.
__device__ __forceinline__ double mg_Az(const Geom& g, const MgLev& L0,
                                        const double* __restrict__ E,
                                        const double* __restrict__ N,
                                        const double* __restrict__ cC,
                                        const double* __restrict__ rx,
                                        const double* __restrict__ ry, int i, int j, size_t o,
                                        double inv, double& z) {
    z = mg_z(g, L0, i, j, inv);
    const int c = i + 2, r = j + 2;
    return cC[o] * z + E[o] * mg_z(g, L0, i + 1, j, inv) +
           E[o - 1] * rx[c] * mg_z(g, L0, i - 1, j, inv) + N[o] * mg_z(g, L0, i, j + 1, inv) +
           N[o - g.pitch] * ry[r] * mg_z(g, L0, i, j - 1, inv);
}

__device__ __forceinline__ double mg_area(const double* dxu, const double* cosu,
                                          const double* dyu, int i, int j) {
    return dxu[i + 2] * cosu[j + 2] * dyu[j + 2];
}

template <int NR>
__device__ __forceinline__ void mg_block_partials(double (&v)[NR], double* part) {
    __shared__ double sh[NR * 4];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NR; ++k) sh[k * 4 + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NR; ++k)
            part[(size_t)k * gridDim.x + blockIdx.x] = ((sh[k * 4] + sh[k * 4 + 1]) + sh[k * 4 + 2]) +
                                                       sh[k * 4 + 3];
}

// Pass A: with z = M r from the V-cycle, the weighted products
// <Az,p>, <z,Az>, <z,q>, <r,z>, <r,p> over active points.
__global__ void __launch_bounds__(256) k_mg_fcg_a(
    Geom g, MgLev L0, const double* __restrict__ E, const double* __restrict__ N,
    const double* __restrict__ cC, const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ dxu, const double* __restrict__ cosu,
    const double* __restrict__ dyu, const uint8_t* __restrict__ act,
    const double* __restrict__ p, const double* __restrict__ q, const double* __restrict__ r,
    const double* __restrict__ zscale, const MgScal* __restrict__ S, double* __restrict__ part) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    if (t < g.nx * g.ny) {
        const int i = t % g.nx, j = t / g.nx;
        const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
        if (act[o]) {
            const double inv = 1.0 / *zscale;   // the scale of the V-cycle's input
            double z;
            const double az = mg_Az(g, L0, E, N, cC, rx, ry, i, j, o, inv, z);
            const double w = mg_area(dxu, cosu, dyu, i, j);
            const double pv = p[o], qv = q[o], rv = r[o];
            acc[0] = w * az * pv;
            acc[1] = w * z * az;
            acc[2] = w * z * qv;
            acc[3] = w * rv * z;
            acc[4] = w * rv * pv;
        }
    }
    mg_block_partials<5>(acc, part);
}

// Fold of a pass's partials by one workgroup (fixed order).
template <int NR>
__device__ __forceinline__ void mg_fold(const double* __restrict__ part, int nblocks,
                                        double (&v)[NR]) {
    __shared__ double sh[NR * 16];
    for (int k = 0; k < NR; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < NR; ++k) v[k] += part[(size_t)k * nblocks + b];
    for (int k = 0; k < NR; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NR; ++k) sh[k * 16 + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NR; ++k) {
            double a = sh[k * 16];
            for (int w = 1; w < nw; ++w) a += sh[k * 16 + w];
            v[k] = a;
        }
}

// beta = -<z,q>/<p,q>_old, <p,q> = <z,Az> + beta(<Az,p> + <z,q>) + beta^2 <p,q>_old,
// alpha = (<r,z> + beta <r,p>) / <p,q>.
__global__ void __launch_bounds__(1024) k_mg_fold_a(const double* __restrict__ part, int nblocks,
                                                    MgScal* S) {
    if (S->done) return;
    double v[5];
    mg_fold<5>(part, nblocks, v);
    if (threadIdx.x == 0) {
        const double beta = S->pq_old != 0.0 ? -v[2] / S->pq_old : 0.0;
        const double pq = v[1] + beta * (v[0] + v[2]) + beta * beta * S->pq_old;
        const double alpha = (v[3] + beta * v[4]) / pq;
        S->beta = beta;
        S->pq = pq;
        S->alpha = alpha;
        S->s_r = S->s_pending;
        if (!(pq != 0.0) || alpha != alpha) {
            S->status = ST_NAN;
            S->done = 1;
        }
    }
}

// Pass B: p = z + beta p, q = Az + beta q, x += alpha p, r -= alpha q; the
// V-cycle's next input w r * s_r on level 0; |r|^2 (active) and |w r|^2.
__global__ void __launch_bounds__(256) k_mg_fcg_b(
    Geom g, MgLev L0, const double* __restrict__ E, const double* __restrict__ N,
    const double* __restrict__ cC, const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ dxu, const double* __restrict__ cosu,
    const double* __restrict__ dyu, const uint8_t* __restrict__ act, double* __restrict__ p,
    double* __restrict__ q, double* __restrict__ r, double* __restrict__ x,
    const double* __restrict__ zscale_src, const MgScal* __restrict__ S,
    double* __restrict__ part) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0, 0};
    if (t < g.nx * g.ny) {
        const int i = t % g.nx, j = t / g.nx;
        const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
        const int ql = (j + 1) * L0.pitch + (i + 1);
        if (act[o]) {
            // z was scaled by the s_r in force before fold_a replaced it
            const double inv = 1.0 / *zscale_src;
            double z;
            const double az = mg_Az(g, L0, E, N, cC, rx, ry, i, j, o, inv, z);
            const double beta = S->beta, alpha = S->alpha;
            const double pn = z + beta * p[o];
            const double qn = az + beta * q[o];
            p[o] = pn;
            q[o] = qn;
            x[o] = x[o] + alpha * pn;
            const double rn = r[o] - alpha * qn;
            r[o] = rn;
            const double w = mg_area(dxu, cosu, dyu, i, j);
            L0.b[ql] = (float)(w * rn * S->s_r);
            acc[0] = rn * rn;
            acc[1] = (w * rn) * (w * rn);
        } else {
            L0.b[ql] = 0.0f;
        }
    }
    mg_block_partials<2>(acc, part);
}

// After pass B: the iteration count, convergence on the recursive residual,
// the scale of the next V-cycle input.
__global__ void __launch_bounds__(1024) k_mg_fold_b(const double* __restrict__ part, int nblocks,
                                                    int nact, MgScal* S, double* zscale) {
    if (S->done) return;
    double v[2];
    mg_fold<2>(part, nblocks, v);
    if (threadIdx.x == 0) {
        *zscale = S->s_r;   // what pass A of the next iteration divides by
        S->rr = v[0];
        S->pq_old = S->pq;
        S->n = S->n + 1;
        const double rms = sqrt(v[1] / (double)(nact > 0 ? nact : 1));
        S->s_pending = rms > 0.0 ? 1.0 / rms : 1.0;
        if (v[0] != v[0]) {
            S->status = ST_NAN;
            S->done = 1;
        } else if (v[0] <= S->rtarget2 * S->bb) {
            S->status = ST_CONVERGED;
            S->done = 1;
        } else if (S->n >= S->max_it) {
            S->status = ST_MAXIT;
            S->done = 1;
        }
    }
}

// Start of a (re)started outer iteration from the residual r: the V-cycle
// input w r s on level 0 with s = 1/rms(w r) (computed by the host).
__global__ void __launch_bounds__(256) k_mg_start(Geom g, MgLev L0, const double* __restrict__ dxu,
                                                  const double* __restrict__ cosu,
                                                  const double* __restrict__ dyu,
                                                  const uint8_t* __restrict__ act,
                                                  const double* __restrict__ r, double s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int i = t % g.nx, j = t / g.nx;
    const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
    const int ql = (j + 1) * L0.pitch + (i + 1);
    L0.b[ql] = act[o] ? (float)(mg_area(dxu, cosu, dyu, i, j) * r[o] * s) : 0.0f;
}

// ---- host: hierarchy ---------------------------------------------------------
struct MgHostLev {
    int nx = 0, ny = 0, cyc = 0, nact = 0;
    std::vector<double> E, N, C;   // interior-indexed, I + J * nx
    std::vector<uint8_t> act;
};

// 2 x 2 aggregation with the Galerkin product of piecewise-constant
// prolongation over active cells.
inline MgHostLev mg_coarsen(const MgHostLev& f) {
    MgHostLev c;
    c.nx = (f.nx + 1) / 2;
    c.ny = (f.ny + 1) / 2;
    c.cyc = f.cyc;
    const size_t nc = (size_t)c.nx * c.ny;
    c.E.assign(nc, 0.0);
    c.N.assign(nc, 0.0);
    c.C.assign(nc, 0.0);
    c.act.assign(nc, 0);
    for (int j = 0; j < f.ny; ++j)
        for (int i = 0; i < f.nx; ++i) {
            const size_t p = (size_t)i + (size_t)j * f.nx;
            if (!f.act[p]) continue;
            const size_t a = (size_t)(i / 2) + (size_t)(j / 2) * c.nx;
            c.act[a] = 1;
            c.C[a] += f.C[p];
            if (f.E[p] != 0.0) {
                const int ie = (i + 1 < f.nx) ? i + 1 : 0;   // nonzero only with a partner
                const size_t ae = (size_t)(ie / 2) + (size_t)(j / 2) * c.nx;
                if (ae == a) c.C[a] += 2.0 * f.E[p];
                else c.E[a] += f.E[p];
            }
            if (f.N[p] != 0.0) {
                const size_t an = (size_t)(i / 2) + (size_t)((j + 1) / 2) * c.nx;
                if (an == a) c.C[a] += 2.0 * f.N[p];
                else c.N[a] += f.N[p];
            }
        }
    for (size_t a = 0; a < nc; ++a) {
        if (c.act[a]) ++c.nact;
        else c.C[a] = 1.0;
    }
    return c;
}

// Dense inverse (Gauss-Jordan with partial pivoting, FP64) of the coarsest
// level's operator over its active cells; false if singular.
inline bool mg_dense_inverse(const MgHostLev& L, std::vector<int>& cells,
                             std::vector<float>& inv) {
    std::vector<int> id((size_t)L.nx * L.ny, -1);
    cells.clear();
    for (int j = 0; j < L.ny; ++j)
        for (int i = 0; i < L.nx; ++i)
            if (L.act[(size_t)i + (size_t)j * L.nx]) {
                id[(size_t)i + (size_t)j * L.nx] = (int)cells.size();
                cells.push_back((j + 1) * (L.nx + 2) + (i + 1));   // level layout index
            }
    const int n = (int)cells.size();
    std::vector<double> M((size_t)n * n, 0.0), X((size_t)n * n, 0.0);
    for (int j = 0; j < L.ny; ++j)
        for (int i = 0; i < L.nx; ++i) {
            const size_t p = (size_t)i + (size_t)j * L.nx;
            const int a = id[p];
            if (a < 0) continue;
            M[(size_t)a * n + a] += L.C[p];
            const int ie = (i + 1 < L.nx) ? i + 1 : (L.cyc ? 0 : -1);
            if (ie >= 0 && L.E[p] != 0.0) {
                const int b = id[(size_t)ie + (size_t)j * L.nx];
                if (b >= 0 && b != a) {
                    M[(size_t)a * n + b] += L.E[p];
                    M[(size_t)b * n + a] += L.E[p];
                }
            }
            if (j + 1 < L.ny && L.N[p] != 0.0) {
                const int b = id[(size_t)i + (size_t)(j + 1) * L.nx];
                if (b >= 0) {
                    M[(size_t)a * n + b] += L.N[p];
                    M[(size_t)b * n + a] += L.N[p];
                }
            }
        }
    for (int k = 0; k < n; ++k) X[(size_t)k * n + k] = 1.0;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int r = k + 1; r < n; ++r)
            if (fabs(M[(size_t)r * n + k]) > fabs(M[(size_t)piv * n + k])) piv = r;
        if (M[(size_t)piv * n + k] == 0.0) return false;
        if (piv != k)
            for (int c2 = 0; c2 < n; ++c2) {
                std::swap(M[(size_t)k * n + c2], M[(size_t)piv * n + c2]);
                std::swap(X[(size_t)k * n + c2], X[(size_t)piv * n + c2]);
            }
        const double d = 1.0 / M[(size_t)k * n + k];
        for (int c2 = 0; c2 < n; ++c2) {
            M[(size_t)k * n + c2] *= d;
            X[(size_t)k * n + c2] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == k) continue;
            const double f = M[(size_t)r * n + k];
            if (f == 0.0) continue;
            for (int c2 = 0; c2 < n; ++c2) {
                M[(size_t)r * n + c2] -= f * M[(size_t)k * n + c2];
                X[(size_t)r * n + c2] -= f * X[(size_t)k * n + c2];
            }
        }
    }
    inv.resize((size_t)n * n);
    for (size_t k = 0; k < inv.size(); ++k) inv[k] = (float)X[k];
    return true;
}

// Device side of a context's hierarchy and the outer iteration's buffers.
struct MgState {
    std::vector<MgLev> lev;          // smoothing levels 0..L-1 and the coarsest L
    int L = 0;                       // index of the coarsest level
    int ncoarse = 0;                 // its active cells
    float* Ainv = nullptr;
    int* cells = nullptr;
    int nact0 = 0;
    double w_mean = 1.0;             // mean psi-cell area over active points
    double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
    MgScal* S = nullptr;
    MgScal* h_S = nullptr;           // pinned
    double* zscale = nullptr;
    double* part = nullptr;
    int nblk = 0;                    // workgroups of the outer passes
    hipStream_t gs = nullptr;        // stream the graph is replayed on
    hipEvent_t e_in = nullptr, e_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_K = 0;
    double graph_ratio = 0.0;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    std::vector<int> level_nx, level_ny, level_nact;
};

// Chebyshev scalars of K steps on [2/ratio, 2].
inline MgCheb mg_cheb(int K, double ratio) {
    MgCheb ch{};
    const double hi = 2.0, lo = hi / ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    ch.c0 = (float)(1.0 / theta);
    for (int k = 1; k < K && k < MG_MAXK; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        ch.c1[k] = (float)(rn * rho);
        ch.c2[k] = (float)(2.0 * rn / delta);
        rho = rn;
    }
    return ch;
}

}  // namespace psi
