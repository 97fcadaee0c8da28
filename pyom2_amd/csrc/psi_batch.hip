// psi_batch.hip -- batched ensembles of independent streamfunction solves
// (include/psibatch.h).  Compiled as part of psi_capi.hip's translation unit:
// it reuses the operator builders and the CG state transitions.
//
// Every member is one congrad_streamfunction solve (solve_streamfct.f90:
// 153-327) on its own grid, run as FP32 inner Jacobi-PCG under FP64
// refinement to a relative FP64 residual.  One kernel advances all members:
// a device item list holds (member, chunk of tiles) work items, resident
// workgroups walk it, and each item first folds its member's partials of
// the previous launch and applies that member's state machine
//
//   START -> RESID0 -> SETUP -> SWEEP* -> CORR -> RESID -> REFRESH -> SWEEP*
//                                            ... -> DONE
//
// so members at different stages (and of different sizes) share every
// launch.  Launches are captured into a hipGraph and replayed; between
// replays the host compacts the item list to the members still running.

namespace psi {
namespace batch {

enum : int {
    BP_START = 0,    // nothing done yet
    BP_RESID0 = 1,   // r = b - A x (FP64), |b|^2, active count
    BP_SETUP = 2,    // u0 = Z r*scale, p = 0, e = 0 (FP32)
    BP_SWEEP = 3,    // fused FP32 PCG iteration on the correction e
    BP_CORR = 4,     // x += e / scale, e = 0 (FP64)
    BP_RESID = 5,    // true FP64 residual of the corrected x
    BP_REFRESH = 6,  // u = Z r*scale (residual replacement), p kept
    BP_DONE = 7
};

struct BMember {
    Geom g;                      // tiles_x/ntiles for the batch tile shape
    int nseg, seg_base;          // work items of this member, first partial record
    long long off;               // plane offset into the packed per-point arrays
    int offx, offy;              // offsets into the packed 1-D metric arrays
    const double* b;             // caller's forcing (device, member plane)
    double* x;                   // caller's solution (device, member plane)
};

struct BArrays {
    const double *E64, *N64, *rx64, *ry64;
    const float *E32, *N32, *rx32, *ry32;
    const uint8_t* act;
    double* r64;
    float* e32;
    float *u0, *u1, *p0, *p1;
    float *r0, *r1;              // residuals of the polynomial path (u holds z = M r)
};

struct BParams {
    double target2;              // relres_target^2
    double red2;                 // inner_reduction^2: FP32 leg length
    double eps;                  // congr_epsilon (estimate only)
    int max_it;                  // inner iterations per member
    int K;                       // tiles per work item
    int grab;                    // work items per grab from the launch's counter
    int frozen;                  // autotuner: every member sweeps every launch
    ChebCoef ch;                 // polynomial path: Chebyshev scalars
};

struct BState {
    CgState c;
    double scale, bbt, rrt;      // FP32 scaling, FP64 |b|^2 and true |r|^2
    int phase, legs, launches, pad;
};

// What the host reads back (members are dropped from the item list once
// done is seen, so this, not the double-buffered state, holds the result).
struct BFinal {
    double rrt, bbt, rr_rec, bb_rec, est_error;
    unsigned long long t_done;
    int done, status, iterations, legs, launches, pad;
};

// Values every lane of the workgroup holds alike (read from LDS after a
// barrier), declared so: the member, its geometry, its phase and scalars
// then live in scalar registers instead of one VGPR per lane each.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float wave_uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ bool needs_fold(int ph) {
    return ph == BP_RESID0 || ph == BP_SETUP || ph == BP_SWEEP || ph == BP_RESID ||
           ph == BP_REFRESH;
}

__device__ inline void b_finish(BState& s, int status) {
    s.c.status = status;
    s.c.done = 1;
    s.phase = BP_DONE;
}

// After an FP32 CG transition: keep sweeping, or go and confirm on the FP64
// residual (the leg ended, the recursive residual met the target or the
// iteration budget is spent), or stop on NaN / divergence.
__device__ inline void b_after_cg(BState& s) {
    CgState& c = s.c;
    if (!c.done) {
        s.phase = BP_SWEEP;
        return;
    }
    if (c.status == ST_CONVERGED || c.status == ST_LEG || c.status == ST_MAXIT) {
        c.done = 0;
        s.phase = BP_CORR;
        return;
    }
    b_finish(s, c.status);
}

// One member's transition from the work of the previous launch (partials
// folded into red) to the work of this launch.
__device__ inline void b_transition(BState& s, const double* red, const BParams& P) {
    CgState& c = s.c;
    if (P.frozen) {
        // timing runs: a full sweep per member per launch, state unchanged
        s.phase = BP_SWEEP;
        c.alpha = 0.0;
        c.beta = 0.0;
        c.n = 0;
        return;
    }
    switch (s.phase) {
        case BP_START:
            s.phase = BP_RESID0;
            return;
        case BP_RESID0: {
            const double rr = red[0], bb = red[1], cnt = red[3];
            s.bbt = bb > 0 ? bb : rr;
            s.rrt = rr;
            if (rr != rr || bb != bb) return b_finish(s, ST_NAN);
            if (rr <= P.target2 * s.bbt) return b_finish(s, ST_CONVERGED);
            s.scale = 1.0 / sqrt(rr / fmax(cnt, 1.0));
            c = CgState{};
            c.eps = P.eps;
            c.rtarget2 = P.target2;
            c.max_it = P.max_it;
            c.criterion = CRIT_RELRES;
            c.bb_fixed = s.bbt * s.scale * s.scale;
            c.red2 = P.red2;
            s.phase = BP_SETUP;
            return;
        }
        case BP_SETUP:
            cg_from_setup(c, red);
            return b_after_cg(s);
        case BP_SWEEP:
            cg_from_sweep(c, red);
            return b_after_cg(s);
        case BP_CORR:
            s.phase = BP_RESID;
            return;
        case BP_RESID: {
            const double rr = red[0];
            s.rrt = rr;
            s.legs += 1;
            if (rr != rr) return b_finish(s, ST_NAN);
            if (rr <= P.target2 * s.bbt) return b_finish(s, ST_CONVERGED);
            if (c.n >= P.max_it) return b_finish(s, ST_MAXIT);
            c.done = 0;
            c.status = ST_RUNNING;
            c.degenerate = 0;
            s.phase = BP_REFRESH;
            return;
        }
        case BP_REFRESH:
            cg_from_refresh(c, red);
            if (c.done) return b_finish(s, c.status);
            s.phase = BP_SWEEP;
            return;
        default:
            return;
    }
}

// Fold a member's partials (records seg_base .. seg_base+nseg-1, NRED
// doubles each) in a fixed order: every item of the member gets the
// bit-identical result.  Valid in thread 0.
__device__ __forceinline__ void fold_member(const double* __restrict__ part, int nseg,
                                            double (&v)[NRED], double* sh) {
#pragma unroll
    for (int k = 0; k < NRED; ++k) v[k] = 0.0;
    for (int j = threadIdx.x; j < nseg; j += NT) {
#pragma unroll
        for (int k = 0; k < NRED; ++k) {
            const double x = part[(size_t)j * NRED + k];
            v[k] = (k == 2) ? fmax(v[k], x) : v[k] + x;
        }
    }
    block_reduce<NRED>(v, 1u << 2, sh);
}

// ---- per-phase tile work ---------------------------------------------------

// FP64 residual r = b - A x on the item's own points (no LDS: the phase runs
// once per leg).  Partials: |r|^2, |b|^2, max|r|, active count.
template <int BX, int BY>
__device__ __forceinline__ void work_resid(const BMember& m, const BArrays& A, int tb, int te,
                                           double (&acc)[NRED]) {
    const Geom& g = m.g;
    const double* __restrict__ E = A.E64 + m.off;
    const double* __restrict__ N = A.N64 + m.off;
    const double* __restrict__ rx = A.rx64 + m.offx;
    const double* __restrict__ ry = A.ry64 + m.offy;
    const uint8_t* __restrict__ act = A.act + m.off;
    double* __restrict__ r64 = A.r64 + m.off;
    const double* __restrict__ b = m.b;
    const double* __restrict__ x = m.x;
    for (int t = tb; t < te; ++t) {
        const int c0 = (t % g.tiles_x) * BX, r0 = (t / g.tiles_x) * BY;
#pragma unroll
        for (int k = 0; k < (BX * BY + NT - 1) / NT; ++k) {
            const int l = threadIdx.x + k * NT;
            const int r = r0 + 2 + l / BX, c = c0 + 2 + l % BX;
            if (l < BX * BY && r <= g.ny + 1 && c <= g.nx + 1) {
                const size_t o = (size_t)r * g.pitch + c;
                const size_t row = (size_t)r * g.pitch;
                const int cm = map_col(c - 1, g), cp = map_col(c + 1, g);
                const double e = E[o], w = E[row + cm] * rx[c], n = N[o],
                             so = N[o - g.pitch] * ry[r];
                const double cc = (((-e) - w) - n) - so;
                double ax = so * x[o - g.pitch];
                ax = ax + w * x[row + cm];
                ax = ax + cc * x[o];
                ax = ax + e * x[row + cp];
                ax = ax + n * x[o + g.pitch];
                const double bv = b[o];
                const double rv = bv - ax;
                const bool a = act[o] != 0;
                if (a) {
                    acc[0] += rv * rv;
                    acc[1] += bv * bv;
                    acc[2] = fmax(acc[2], fabs(rv));
                    acc[3] += 1.0;
                }
                r64[o] = a ? rv : 0.0;
            }
        }
    }
}

// x += e / scale, e = 0 on the item's own points.
template <int BX, int BY>
__device__ __forceinline__ void work_corr(const BMember& m, const BArrays& A, int tb, int te,
                                          double inv_scale) {
    const Geom& g = m.g;
    float* __restrict__ e32 = A.e32 + m.off;
    double* __restrict__ x = m.x;
    for (int t = tb; t < te; ++t) {
        const int c0 = (t % g.tiles_x) * BX, r0 = (t / g.tiles_x) * BY;
#pragma unroll
        for (int k = 0; k < (BX * BY + NT - 1) / NT; ++k) {
            const int l = threadIdx.x + k * NT;
            const int r = r0 + 2 + l / BX, c = c0 + 2 + l % BX;
            if (l < BX * BY && r <= g.ny + 1 && c <= g.nx + 1) {
                const size_t o = (size_t)r * g.pitch + c;
                x[o] = x[o] + (double)e32[o] * inv_scale;
                e32[o] = 0.f;
            }
        }
    }
}

// Shared frame buffers of the FP32 phases.
template <int BX, int BY>
struct Frames {
    static constexpr int HXb = BX + 4, HYb = BY + 4, FR = HXb * HYb;
    float U[FR], P[FR], E[FR], N[FR];
    uint8_t A[FR];
    float RX[HXb], RY[HYb];
};

template <int BX, int BY>
__device__ __forceinline__ void load_metrics_b(const BMember& m, const BArrays& A, int c0, int r0,
                                               Frames<BX, BY>& F) {
    constexpr int HXb = BX + 4, HYb = BY + 4;
    const Geom& g = m.g;
    for (int l = threadIdx.x; l < HXb + HYb; l += NT) {
        if (l < HXb) {
            const int c = map_col(c0 + l, g);
            F.RX[l] = (c >= 0 && c < g.nx + 4) ? A.rx32[m.offx + c] : 0.f;
        } else {
            const int k = l - HXb, r = r0 + k;
            F.RY[k] = (r >= 0 && r < g.ny + 4) ? A.ry32[m.offy + r] : 0.f;
        }
    }
}

// Residual-driven start of an FP32 leg.  refresh = 0 (SETUP): u0 = Z r*scale
// into buffer 0, p = 0, e = 0.  refresh = 1 (REFRESH): u_n = Z r*scale into
// buffer n & 1 next to the kept p_{n-1}.  Partials as the sweep's.
template <int BX, int BY>
__device__ __forceinline__ void work_start(const BMember& m, const BArrays& A, int tb, int te,
                                           bool refresh, int n, float scale, double (&acc)[NRED],
                                           Frames<BX, BY>& F) {
    constexpr int HXb = BX + 4, HYb = BY + 4, FR = HXb * HYb, W1 = HXb - 2, H1 = HYb - 2;
    const Geom& g = m.g;
    const float* __restrict__ E = A.E32 + m.off;
    const float* __restrict__ N = A.N32 + m.off;
    const uint8_t* __restrict__ act = A.act + m.off;
    const double* __restrict__ r64 = A.r64 + m.off;
    const int odd = refresh ? (n & 1) : 0;
    float* __restrict__ u_out = (odd ? A.u1 : A.u0) + m.off;
    float* __restrict__ p_buf = (odd ? A.p1 : A.p0) + m.off;
    float* __restrict__ e32 = A.e32 + m.off;
    for (int t = tb; t < te; ++t) {
        const int c0 = (t % g.tiles_x) * BX, r0 = (t / g.tiles_x) * BY;
        for (int l = threadIdx.x; l < FR; l += NT) {
            const int gg = l / HXb, f = l - gg * HXb;
            const int r = r0 + gg, c = map_col(c0 + f, g);
            float rv = 0.f, pp = 0.f, e = 0.f, nn = 0.f;
            uint8_t a = 0;
            if (in_array(r, c, g)) {
                const size_t o = (size_t)r * g.pitch + c;
                rv = (float)(r64[o] * (double)scale);
                if (refresh) pp = p_buf[o];
                e = E[o];
                nn = N[o];
                a = act[o];
            }
            F.U[l] = rv;
            F.P[l] = pp;
            F.E[l] = e;
            F.N[l] = nn;
            F.A[l] = a;
        }
        load_metrics_b<BX, BY>(m, A, c0, r0, F);
        __syncthreads();
        for (int l = threadIdx.x; l < W1 * H1; l += NT) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HXb + f;
            float cc;
            (void)stencil5<float, HXb>(F.P, F.E, F.N, idx, F.RX[f], F.RY[gg], cc);
            F.U[idx] = F.A[idx] ? (1.f / cc) * F.U[idx] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < (BX * BY) / NT; ++k) {
            const int l = threadIdx.x + k * NT;
            const int gg = 2 + l / BX, f = 2 + l % BX;
            const int r = r0 + gg, c = c0 + f;
            if (r <= g.ny + 1 && c <= g.nx + 1) {
                const int idx = gg * HXb + f;
                float cc;
                const float w = stencil5<float, HXb>(F.U, F.E, F.N, idx, F.RX[f], F.RY[gg], cc);
                const float sv = stencil5<float, HXb>(F.P, F.E, F.N, idx, F.RX[f], F.RY[gg], cc);
                const float un = F.U[idx], pn = F.P[idx];
                const size_t o = (size_t)r * g.pitch + c;
                u_out[o] = un;
                if (!refresh) {
                    p_buf[o] = 0.f;
                    e32[o] = 0.f;
                }
                const double ud = (double)un, pd = (double)pn, ccd = (double)cc;
                const double wd = (double)w, sd = (double)sv;
                acc[0] += ud * ud * ccd;
                acc[1] += ud * wd;
                acc[2] = fmax(acc[2], refresh ? fabs(pd) : fabs(ud));
                if (F.A[idx]) {
                    const double rr = ud * ccd;
                    acc[3] += rr * rr;
                }
                acc[4] += ud * sd + pd * wd;
                acc[5] += pd * sd;
            }
        }
        __syncthreads();
    }
}

// Pointers of one FP32 sweep: after n iterations u_n and p_{n-1} live in
// buffer n & 1.
struct SweepPtrs {
    const float *u_in, *p_in, *E, *N;
    const uint8_t* act;
    float *u_out, *p_out, *x;
};

__device__ __forceinline__ SweepPtrs sweep_ptrs(const BMember& m, const BArrays& A, int n) {
    const bool odd = n & 1;
    SweepPtrs q;
    q.u_in = (odd ? A.u1 : A.u0) + m.off;
    q.p_in = (odd ? A.p1 : A.p0) + m.off;
    q.E = A.E32 + m.off;
    q.N = A.N32 + m.off;
    q.act = A.act + m.off;
    q.u_out = (odd ? A.u0 : A.u1) + m.off;
    q.p_out = (odd ? A.p0 : A.p1) + m.off;
    q.x = A.e32 + m.off;
    return q;
}

// One tile's sweep inputs in registers: u, p, E, N and the mask over the
// (BX+4) x (BY+4) frame, and the correction e at the tile's own points.
// The next tile is fetched while the current one is computed.
template <int BX, int BY>
struct FrameRegsB {
    static constexpr int HXb = BX + 4, HYb = BY + 4, FR = HXb * HYb;
    static constexpr int NL = (FR + NT - 1) / NT, PER = (BX * BY) / NT;
    float u[NL], p[NL], e[NL], n[NL];
    uint8_t a[NL];
    float xo[PER];

    __device__ __forceinline__ void fetch(const Geom& g, int tile, const SweepPtrs& q) {
        const int c0 = (tile % g.tiles_x) * BX, r0 = (tile / g.tiles_x) * BY;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int l = threadIdx.x + k * NT;
            const int gg = l / HXb, f = l - gg * HXb;
            const int r = r0 + gg, c = map_col(c0 + f, g);
            u[k] = 0.f;
            p[k] = 0.f;
            e[k] = 0.f;
            n[k] = 0.f;
            a[k] = 0;
            if (l < FR && in_array(r, c, g)) {
                const unsigned o = (unsigned)(r * g.pitch + c);
                u[k] = q.u_in[o];
                p[k] = q.p_in[o];
                e[k] = q.E[o];
                n[k] = q.N[o];
                a[k] = q.act[o];
            }
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int l = threadIdx.x + k * NT;
            const int r = r0 + 2 + l / BX, c = c0 + 2 + l % BX;
            xo[k] = (r <= g.ny + 1 && c <= g.nx + 1) ? q.x[(unsigned)(r * g.pitch + c)] : 0.f;
        }
    }
    // frame -> LDS, fused with p_k = u_k + beta p_{k-1}
    __device__ __forceinline__ void store(float beta, Frames<BX, BY>& F) const {
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int l = threadIdx.x + k * NT;
            if (l < FR) {
                F.U[l] = u[k];
                F.P[l] = u[k] + beta * p[k];
                F.E[l] = e[k];
                F.N[l] = n[k];
                F.A[l] = a[k];
            }
        }
    }
};

// Fused FP32 PCG iterations over tiles [T0, T1) of one member (the
// single-solve sweep with per-member geometry): p = u + beta p,
// u -= alpha Z A p (radius 1), e += alpha p, w = A u, s = A p and the six
// partials, written per work item (every K tiles).  fr holds tile T0.
template <int BX, int BY>
__device__ __forceinline__ void work_sweep_run(const BMember& m, const SweepPtrs& q, int T0,
                                               int T1, int K, float alpha, float beta,
                                               const BArrays& A, FrameRegsB<BX, BY>& fr,
                                               Frames<BX, BY>& F, double* sRed,
                                               double* __restrict__ part_out) {
    constexpr int HXb = BX + 4, HYb = BY + 4, W1 = HXb - 2, H1 = HYb - 2;
    constexpr int PER = (BX * BY) / NT;
    const Geom& g = m.g;
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    for (int t = T0; t < T1; ++t) {
        const int c0 = (t % g.tiles_x) * BX, r0 = (t / g.tiles_x) * BY;
        fr.store(beta, F);
        float xo[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) xo[k] = fr.xo[k];
        load_metrics_b<BX, BY>(m, A, c0, r0, F);
        __syncthreads();
        if (t + 1 < T1) fr.fetch(g, t + 1, q);
        for (int l = threadIdx.x; l < W1 * H1; l += NT) {
            const int gg = 1 + l / W1, f = 1 + (l - (gg - 1) * W1);
            const int idx = gg * HXb + f;
            if (F.A[idx]) {
                float cc;
                const float s = stencil5<float, HXb>(F.P, F.E, F.N, idx, F.RX[f], F.RY[gg], cc);
                F.U[idx] = F.U[idx] - alpha * ((1.f / cc) * s);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int l = threadIdx.x + k * NT;
            const int gg = 2 + l / BX, f = 2 + l % BX;
            const int r = r0 + gg, c = c0 + f;
            if (r <= g.ny + 1 && c <= g.nx + 1) {
                const int idx = gg * HXb + f;
                float cc;
                const float w = stencil5<float, HXb>(F.U, F.E, F.N, idx, F.RX[f], F.RY[gg], cc);
                const float sv = stencil5<float, HXb>(F.P, F.E, F.N, idx, F.RX[f], F.RY[gg], cc);
                const float un = F.U[idx], pn = F.P[idx];
                const unsigned o = (unsigned)(r * g.pitch + c);
                q.x[o] = xo[k] + alpha * pn;
                q.u_out[o] = un;
                q.p_out[o] = pn;
                const double ud = (double)un, pd = (double)pn, ccd = (double)cc;
                const double wd = (double)w, sd = (double)sv;
                acc[0] += ud * ud * ccd;
                acc[1] += ud * wd;
                acc[2] = fmax(acc[2], fabs(pd));
                if (F.A[idx]) {
                    const double rv = ud * ccd;
                    acc[3] += rv * rv;
                }
                acc[4] += ud * sd + pd * wd;
                acc[5] += pd * sd;
            }
        }
        __syncthreads();
        if ((t + 1) % K == 0 || t + 1 == T1) {    // end of a work item
            block_reduce<NRED>(acc, 1u << 2, sRed);
            if (threadIdx.x == 0) {
#pragma unroll
                for (int k = 0; k < NRED; ++k)
                    part_out[(size_t)(m.seg_base + t / K) * NRED + k] = acc[k];
            }
#pragma unroll
            for (int k = 0; k < NRED; ++k) acc[k] = 0.0;
        }
    }
}

// Fold member s's partials of the previous launch, advance its state (every
// item of the member computes the same), publish it (the item holding the
// member's first segment) and leave it in sst.
__device__ __forceinline__ void member_step(const BMember& m, int s, int seg0, int ph_in,
                                            const BParams& P,
                                            const BState* __restrict__ st_in,
                                            BState* __restrict__ st_out,
                                            const double* __restrict__ part_in,
                                            BFinal* __restrict__ fin, BState& sst, double* sRed) {
    double red[NRED] = {0, 0, 0, 0, 0, 0};
    if (needs_fold(ph_in)) fold_member(part_in + (size_t)m.seg_base * NRED, m.nseg, red, sRed);
    if (threadIdx.x == 0) {
        BState st = st_in[s];
        if (st.phase != BP_DONE) {
            b_transition(st, red, P);
            st.launches += 1;
            if (st.phase == BP_DONE && seg0 == 0) {
                BFinal f;
                f.rrt = st.rrt;
                f.bbt = st.bbt;
                f.rr_rec = st.c.rr;
                f.bb_rec = st.c.bb;
                f.est_error = st.c.est_error;
                f.t_done = wall_clock64();
                f.done = 1;
                f.status = st.c.status;
                f.iterations = st.c.status == ST_MAXIT ? P.max_it + 1 : st.c.n;
                f.legs = st.legs;
                f.launches = st.launches;
                f.pad = 0;
                fin[s] = f;
            }
        }
        sst = st;
        if (seg0 == 0) st_out[s] = st;
    }
    __syncthreads();
}

// ---- the batched step --------------------------------------------------------
// Launch L reads states/partials of parity L & 1 and writes (L + 1) & 1.
// Workgroups grab chunks of `grab` consecutive work items from the launch's
// counter (ctr[L & 1]; workgroup 0 clears ctr[(L + 1) & 1], last used by
// the previous launch, for the next one), so members in the slower
// residual phases do not leave one workgroup behind the rest.  Within a
// chunk a workgroup folds and advances a member's state once per run of
// that member's items.
template <int BX, int BY>
__global__ __launch_bounds__(NT) void k_batch_step(
    const BMember* __restrict__ mem, const int2* __restrict__ items,
    const int* __restrict__ nitems_p, BArrays A, BParams P, const BState* __restrict__ st_in,
    BState* __restrict__ st_out, const double* __restrict__ part_in,
    double* __restrict__ part_out, BFinal* __restrict__ fin, int* ctr, int par) {
    __shared__ Frames<BX, BY> F;
    __shared__ double sRed[NRED * NW];
    __shared__ BState sst;
    __shared__ int s_chunk;
    const int nitems = *nitems_p;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(ctr + (par ^ 1), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    FrameRegsB<BX, BY> fr;
    for (;;) {
    if (threadIdx.x == 0)
        s_chunk = __hip_atomic_fetch_add(ctr + par, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int chunk = wave_uniform(s_chunk);
    if ((long long)chunk * P.grab >= nitems) break;
    const int i_end = min(nitems, (chunk + 1) * P.grab);
    for (int it = chunk * P.grab; it < i_end;) {
        const int2 item = items[it];
        const int s = item.x, seg0 = item.y;
        const BMember m = mem[s];
        const int run = min(i_end - it, m.nseg - seg0);   // this range's items of member s
        const int T0 = seg0 * P.K, T1 = min(m.g.ntiles, (seg0 + run) * P.K);
        const int ph_in = st_in[s].phase, n_in = st_in[s].c.n;
        // the first tile of a sweep is fetched before the fold, with the
        // iteration parity the transition will give if the member keeps sweeping
        const bool may_sweep = ph_in == BP_SWEEP || ph_in == BP_SETUP || ph_in == BP_REFRESH;
        const int n_pred = P.frozen ? 0 : (ph_in == BP_SWEEP ? n_in + 1
                                                             : (ph_in == BP_SETUP ? 0 : n_in));
        if (may_sweep) fr.fetch(m.g, T0, sweep_ptrs(m, A, n_pred));
        member_step(m, s, seg0, ph_in, P, st_in, st_out, part_in, fin, sst, sRed);
        const int ph = wave_uniform(sst.phase);
        if (ph == BP_SWEEP) {
            const int n = wave_uniform(sst.c.n);
            const SweepPtrs q = sweep_ptrs(m, A, n);
            if (!may_sweep || n != n_pred) fr.fetch(m.g, T0, q);
            work_sweep_run<BX, BY>(m, q, T0, T1, P.K, wave_uniform((float)sst.c.alpha),
                                   wave_uniform((float)sst.c.beta), A,
                                   fr, F, sRed, part_out);
        } else if (ph != BP_DONE) {
            for (int j = 0; j < run; ++j) {
                const int seg = seg0 + j;
                const int tb = seg * P.K, te = min(m.g.ntiles, tb + P.K);
                double acc[NRED] = {0, 0, 0, 0, 0, 0};
                switch (ph) {
                    case BP_RESID0:
                    case BP_RESID: work_resid<BX, BY>(m, A, tb, te, acc); break;
                    case BP_SETUP:
                        work_start<BX, BY>(m, A, tb, te, false, 0, (float)sst.scale, acc, F);
                        break;
                    case BP_REFRESH:
                        work_start<BX, BY>(m, A, tb, te, true, sst.c.n, (float)sst.scale, acc,
                                           F);
                        break;
                    case BP_CORR: work_corr<BX, BY>(m, A, tb, te, 1.0 / wave_uniform(sst.scale)); break;
                    default: break;
                }
                if (ph != BP_CORR) {
                    block_reduce<NRED>(acc, 1u << 2, sRed);
                    if (threadIdx.x == 0) {
#pragma unroll
                        for (int k = 0; k < NRED; ++k)
                            part_out[(size_t)(m.seg_base + seg) * NRED + k] = acc[k];
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
        it += run;
    }
    }
}

// Rows per thread of the batched polynomial tile.  The batched kernel runs
// at 2 waves per SIMD whatever the frame height (its member state machine
// holds registers beside the tile), so it takes the tallest frame that fits
// 256 VGPRs without spilling: 9 rows (a 64 x 36 frame, 1.85x its own cells at
// degree 4 against 2.15x with 7 rows; 1,811 -> 1,968 solves/s at 512
// members, 8 rows 1,908).
#ifndef PSI_BATCH_POLY_RPT
#define PSI_BATCH_POLY_RPT 9
#endif
constexpr int batch_poly_rows(int k) { return PSI_POLY_RPT > 0 ? PSI_POLY_RPT : PSI_BATCH_POLY_RPT; }

// Polynomial-preconditioned FP32 legs of one member over tiles [T0, T1) of
// its poly tiling (partials per work item of K tiles).
template <int KP, int MODE>
__device__ __forceinline__ void poly_run(const BMember& m, const BArrays& A, const BParams& P,
                                         const BState& st, int T0, int T1, Exchange& X, int& xbuf,
                                         double* sRed, double* __restrict__ part_out) {
    PolyPtrs q;
    q.E = A.E32 + m.off;
    q.N = A.N32 + m.off;
    q.rx = A.rx32 + m.offx;
    q.ry = A.ry32 + m.offy;
    q.act = A.act + m.off;
    q.x = A.e32 + m.off;
    q.r64 = A.r64 + m.off;
    q.rscale = wave_uniform(st.scale);
    const int n = MODE == PM_SETUP ? 0 : wave_uniform(st.c.n);
    poly_buffers<MODE>(q, n, A.u0 + m.off, A.u1 + m.off, A.p0 + m.off, A.p1 + m.off,
                       A.r0 + m.off, A.r1 + m.off);
    const float alpha = wave_uniform((float)st.c.alpha), beta = wave_uniform((float)st.c.beta);
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    for (int t = T0; t < T1; ++t) {
        poly_tile<KP, MODE, batch_poly_rows(KP)>(m.g, t, q, alpha, beta, P.ch, X, xbuf, acc);
        if ((t + 1) % P.K == 0 || t + 1 == T1) {    // end of a work item
            block_reduce<NRED>(acc, 1u << 2, sRed);
            if (threadIdx.x == 0) {
#pragma unroll
                for (int k = 0; k < NRED; ++k)
                    part_out[(size_t)(m.seg_base + t / P.K) * NRED + k] = acc[k];
            }
#pragma unroll
            for (int k = 0; k < NRED; ++k) acc[k] = 0.0;
        }
    }
}

// The batched step with the degree-KP polynomial preconditioner in the FP32
// legs (the member's tiles are those of the poly tiling, (64 - 2R) x
// (4 RPT - 2R)); work distribution and state machine as k_batch_step.
// Waves per SIMD the register allocation must allow.  The register-resident
// polynomial tile (7 arrays of RPT rows per thread) and the member state
// machine took 223 VGPRs at degree 4 with 7 rows, so 2 waves per SIMD;
// forcing 3 spilled 43 VGPRs to scratch and measured 4% fewer ensemble
// solves per second at 512 members (4 waves spilled 122 and lost 44%).
#ifndef PSI_BATCH_POLY_WAVES
#define PSI_BATCH_POLY_WAVES 1
#endif
template <int KP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PSI_BATCH_POLY_WAVES)))
void k_batch_poly(
    const BMember* __restrict__ mem, const int2* __restrict__ items,
    const int* __restrict__ nitems_p, BArrays A, BParams P, const BState* __restrict__ st_in,
    BState* __restrict__ st_out, const double* __restrict__ part_in,
    double* __restrict__ part_out, BFinal* __restrict__ fin, int* ctr, int par) {
    constexpr int RPT = batch_poly_rows(KP);
    constexpr int TXo = PolyShape<KP, RPT>::TXo, TYo = PolyShape<KP, RPT>::TYo;
    __shared__ Exchange X;
    __shared__ double sRed[NRED * NW];
    __shared__ BState sst;
    __shared__ int s_chunk;
    const int nitems = *nitems_p;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(ctr + (par ^ 1), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int xbuf = 0;
    for (;;) {
        if (threadIdx.x == 0)
            s_chunk =
                __hip_atomic_fetch_add(ctr + par, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int chunk = wave_uniform(s_chunk);
        if ((long long)chunk * P.grab >= nitems) break;
        const int i_end = min(nitems, (chunk + 1) * P.grab);
        for (int it = chunk * P.grab; it < i_end;) {
            const int2 item = items[it];
            const int s = item.x, seg0 = item.y;
            const BMember m = mem[s];
            const int run = min(i_end - it, m.nseg - seg0);
            const int T0 = seg0 * P.K, T1 = min(m.g.ntiles, (seg0 + run) * P.K);
            member_step(m, s, seg0, st_in[s].phase, P, st_in, st_out, part_in, fin, sst, sRed);
            const int ph = wave_uniform(sst.phase);
            if (ph == BP_SWEEP) {
                poly_run<KP, PM_SWEEP>(m, A, P, sst, T0, T1, X, xbuf, sRed, part_out);
            } else if (ph == BP_SETUP) {
                poly_run<KP, PM_SETUP>(m, A, P, sst, T0, T1, X, xbuf, sRed, part_out);
            } else if (ph == BP_REFRESH) {
                poly_run<KP, PM_REFRESH>(m, A, P, sst, T0, T1, X, xbuf, sRed, part_out);
            } else if (ph != BP_DONE) {
                for (int j = 0; j < run; ++j) {
                    const int seg = seg0 + j;
                    const int tb = seg * P.K, te = min(m.g.ntiles, tb + P.K);
                    if (ph == BP_CORR) {
                        work_corr<TXo, TYo>(m, A, tb, te, 1.0 / wave_uniform(sst.scale));
                    } else {
                        double acc[NRED] = {0, 0, 0, 0, 0, 0};
                        work_resid<TXo, TYo>(m, A, tb, te, acc);
                        block_reduce<NRED>(acc, 1u << 2, sRed);
                        if (threadIdx.x == 0) {
#pragma unroll
                            for (int k = 0; k < NRED; ++k)
                                part_out[(size_t)(m.seg_base + seg) * NRED + k] = acc[k];
                        }
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            it += run;
        }
    }
}

__global__ void k_batch_clock(unsigned long long* t) { *t = wall_clock64(); }

// Right-hand side and zero first guess of the autotuner's timing runs.
__global__ void k_batch_tune_fill(size_t n, const uint8_t* __restrict__ act, double* __restrict__ b,
                                  double* __restrict__ x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = act[i] ? 1.0 + 0.25 * (double)((i * 2654435761u) % 7u) : 0.0;
    x[i] = 0.0;
}

}  // namespace batch
}  // namespace psi

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
using namespace psi::batch;

namespace {

struct Shape {
    int bx, by;
};
constexpr Shape kShapes[] = {{64, 16}, {64, 8}, {32, 16}, {32, 8}};
constexpr int kNumShapes = 4;

int shape_index(int bx, int by) {
    for (int k = 0; k < kNumShapes; ++k)
        if (kShapes[k].bx == bx && kShapes[k].by == by) return k;
    return -1;
}

template <int BX, int BY>
const void* step_kernel() {
    return reinterpret_cast<const void*>(&k_batch_step<BX, BY>);
}

const void* kernel_for(int shape) {
    switch (shape) {
        case 0: return step_kernel<64, 16>();
        case 1: return step_kernel<64, 8>();
        case 2: return step_kernel<32, 16>();
        default: return step_kernel<32, 8>();
    }
}

// Polynomial degrees of the batched path and their kernels / own tiles.
constexpr int kPolyMin = 2, kPolyMax = 4;
const void* poly_kernel_for(int k) {
    switch (k) {
        case 2: return reinterpret_cast<const void*>(&k_batch_poly<2>);
        case 3: return reinterpret_cast<const void*>(&k_batch_poly<3>);
        default: return reinterpret_cast<const void*>(&k_batch_poly<4>);
    }
}
void poly_tile_dims(int k, int* bx, int* by) {
    const int R = k + 2, rpt = batch_poly_rows(k);
    *bx = 64 - 2 * R;
    *by = 4 * rpt - 2 * R;
}

}  // namespace

struct psi_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;      // created here: the caller passed the null stream
    hipEvent_t ev_in = nullptr;   // orders an own stream after the null stream
    int count = 0;
    struct HostMember {
        int nx, ny, cyclic;
        size_t off, plane;
        int offx, offy;
    };
    std::vector<HostMember> hm;
    size_t total = 0;
    int totx = 0, toty = 0;
    // packed operator and scratch
    double *E64 = nullptr, *N64 = nullptr, *rx64 = nullptr, *ry64 = nullptr, *r64 = nullptr;
    float *E32 = nullptr, *N32 = nullptr, *rx32 = nullptr, *ry32 = nullptr, *e32 = nullptr;
    float *u[2] = {nullptr, nullptr}, *p[2] = {nullptr, nullptr};
    float* rb[2] = {nullptr, nullptr};   // residuals of the polynomial path
    uint8_t* act = nullptr;
    bool metrics_positive = true;        // the polynomial path's precondition
    int poly_k = 0;                      // > 0: descriptors laid out for k_batch_poly<poly_k>
    int occ_poly[kPolyMax + 1] = {0};
    int64_t active_points = 0;
    // schedule
    psi_batch_tune tune{};
    int shape = 0, G = 0;
    int ncu = 0, occ[kNumShapes] = {0, 0, 0, 0};
    std::vector<BMember> h_mem;
    BMember* d_mem = nullptr;
    int total_seg = 0;
    size_t items_cap = 0, part_cap = 0;
    int2* d_items = nullptr;
    int* d_nitems = nullptr;
    int* d_ctr = nullptr;         // 2 work counters (by launch parity)
    int2* h_items[2] = {nullptr, nullptr};   // pinned, alternating
    int* h_nitems[2] = {nullptr, nullptr};   // pinned, alternating
    double* part = nullptr;       // 2 x part_cap
    BState* st = nullptr;         // 2 x count
    BState* h_st = nullptr;       // pinned, count
    BFinal* fin = nullptr;
    BFinal* h_fin[2] = {nullptr, nullptr};   // pinned poll slots
    unsigned long long* d_t0 = nullptr;
    unsigned long long* h_t0 = nullptr;
    int clock_khz = 0;
    hipEvent_t ev[2] = {nullptr, nullptr}, tb = nullptr, te = nullptr;
    int64_t bytes = 0;
    std::vector<void*> allocs;
    std::string err;
};

namespace {

int bfail(psi_batch* b, int code, const char* fmt, const char* what = nullptr) {
    char buf[512];
    snprintf(buf, sizeof buf, fmt, what ? what : "");
    if (b) b->err = buf;
    else g_create_err = buf;
    return code;
}

#define BHIP(b, expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return bfail((b), PSI_HIP_ERROR, "HIP error: %s", hipGetErrorString(e_)); \
    } while (0)

template <class T>
int balloc(psi_batch* b, T** p, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(T) + 64);
    if (e != hipSuccess) return bfail(b, PSI_HIP_ERROR, "hipMalloc: %s", hipGetErrorString(e));
    e = hipMemsetAsync(v, 0, n * sizeof(T) + 64, b->stream);
    if (e != hipSuccess) return bfail(b, PSI_HIP_ERROR, "hipMemset: %s", hipGetErrorString(e));
    b->allocs.push_back(v);
    b->bytes += (int64_t)(n * sizeof(T) + 64);
    *p = (T*)v;
    return PSI_OK;
}

void batch_free(psi_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    for (void* v : b->allocs) (void)hipFree(v);
    for (int q = 0; q < 2; ++q) {
        if (b->h_items[q]) (void)hipHostFree(b->h_items[q]);
        if (b->h_nitems[q]) (void)hipHostFree(b->h_nitems[q]);
    }
    if (b->h_st) (void)hipHostFree(b->h_st);
    for (auto* f : b->h_fin)
        if (f) (void)hipHostFree(f);
    if (b->h_t0) (void)hipHostFree(b->h_t0);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->tb) (void)hipEventDestroy(b->tb);
    if (b->te) (void)hipEventDestroy(b->te);
    delete b;
}

// Builds one member's stored operator into the packed arrays (the
// single-solve builders, run with the member's geometry).
int build_member(psi_batch* b, int k, const psi_batch_member& mb) {
    const auto& h = b->hm[k];
    Geom g{};
    g.nx = h.nx;
    g.ny = h.ny;
    g.pitch = h.nx + 4;
    g.cyclic = h.cyclic;
    const size_t nxh = h.nx + 4, nyh = h.ny + 4, plane = h.plane;
    const psi_grid& gr = mb.grid;
    std::vector<double> m1(4 * nxh + 4 * nyh);
    double* dxt = m1.data();
    double* dxu = dxt + nxh;
    double* rx = dxu + nxh;
    double* spare = rx + nxh;
    (void)spare;
    double* dyt = m1.data() + 4 * nxh;
    double* dyu = dyt + nyh;
    double* cost = dyu + nyh;
    double* cosu = cost + nyh;
    std::copy(gr.dxt, gr.dxt + nxh, dxt);
    std::copy(gr.dxu, gr.dxu + nxh, dxu);
    std::copy(gr.dyt, gr.dyt + nyh, dyt);
    std::copy(gr.dyu, gr.dyu + nyh, dyu);
    std::copy(gr.cost, gr.cost + nyh, cost);
    std::copy(gr.cosu, gr.cosu + nyh, cosu);
    std::vector<double> hrx(nxh, 0.0), hry(nyh, 0.0);
    std::vector<float> hrxf(nxh), hryf(nyh);
    for (size_t q = 1; q < nxh; ++q) hrx[q] = dxu[q - 1] / dxu[q];
    for (size_t q = 1; q < nyh; ++q) hry[q] = dyu[q - 1] * cosu[q - 1] / (dyu[q] * cosu[q]);
    for (size_t q = 0; q < nxh; ++q) hrxf[q] = (float)hrx[q];
    for (size_t q = 0; q < nyh; ++q) hryf[q] = (float)hry[q];
    // temporaries: metrics, Dirichlet mark plane, boundary points
    const psi_islands& isl = mb.islands;
    const int npts = (isl.nisle > 0 && isl.offsets) ? isl.offsets[isl.nisle] : 0;
    const size_t tmp_bytes = m1.size() * 8 + plane + 2 * (size_t)std::max(npts, 1) * 4 + 256;
    char* tmp = nullptr;
    BHIP(b, hipMalloc((void**)&tmp, tmp_bytes));
    double* d_m1 = (double*)tmp;
    uint8_t* d_dir = (uint8_t*)(tmp + m1.size() * 8);
    int32_t* d_bi = (int32_t*)(tmp + ((m1.size() * 8 + plane + 15) & ~(size_t)15));
    int32_t* d_bj = d_bi + std::max(npts, 1);
    int rc = PSI_OK;
    do {
        if (hipMemcpyAsync(d_m1, m1.data(), m1.size() * 8, hipMemcpyHostToDevice, b->stream) !=
                hipSuccess ||
            hipMemsetAsync(d_dir, 0, plane, b->stream) != hipSuccess ||
            hipMemcpyAsync(b->rx64 + h.offx, hrx.data(), nxh * 8, hipMemcpyHostToDevice,
                           b->stream) != hipSuccess ||
            hipMemcpyAsync(b->ry64 + h.offy, hry.data(), nyh * 8, hipMemcpyHostToDevice,
                           b->stream) != hipSuccess ||
            hipMemcpyAsync(b->rx32 + h.offx, hrxf.data(), nxh * 4, hipMemcpyHostToDevice,
                           b->stream) != hipSuccess ||
            hipMemcpyAsync(b->ry32 + h.offy, hryf.data(), nyh * 4, hipMemcpyHostToDevice,
                           b->stream) != hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "member upload failed%s");
            break;
        }
        if (npts > 0) {
            if (hipMemcpyAsync(d_bi, isl.bi, (size_t)npts * 4, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(d_bj, isl.bj, (size_t)npts * 4, hipMemcpyHostToDevice, b->stream) !=
                    hipSuccess) {
                rc = bfail(b, PSI_HIP_ERROR, "island upload failed%s");
                break;
            }
            hipLaunchKernelGGL(k_mark_points, dim3(nblk(npts)), dim3(256), 0, b->stream, g, npts,
                               d_bi, d_bj, d_dir);
        }
        const double* d_dxt = d_m1;
        const double* d_dxu = d_m1 + nxh;
        const double* d_dyt = d_m1 + 4 * nxh;
        const double* d_dyu = d_dyt + nyh;
        const double* d_cost = d_dyu + nyh;
        const double* d_cosu = d_cost + nyh;
        double* E = b->E64 + h.off;
        double* N = b->N64 + h.off;
        hipLaunchKernelGGL(k_build_EN, dim3(nblk(plane)), dim3(256), 0, b->stream, g, mb.hur,
                           mb.hvr, d_dxt, d_dxu, d_dyt, d_dyu, d_cost, d_cosu, E, N);
        hipLaunchKernelGGL(k_build_active, dim3(nblk(plane)), dim3(256), 0, b->stream, g,
                           (const double*)E, (const double*)N, (const double*)(b->rx64 + h.offx),
                           (const double*)(b->ry64 + h.offy), (const uint8_t*)d_dir,
                           b->act + h.off, (double*)nullptr);
        hipLaunchKernelGGL(k_convert<float>, dim3(nblk(plane)), dim3(256), 0, b->stream, plane,
                           (const double*)E, b->E32 + h.off);
        hipLaunchKernelGGL(k_convert<float>, dim3(nblk(plane)), dim3(256), 0, b->stream, plane,
                           (const double*)N, b->N32 + h.off);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "operator build: %s", hipGetErrorString(e));
            break;
        }
    } while (0);
    // the temporaries are released once the builders have run
    const hipError_t es = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (rc == PSI_OK && es != hipSuccess)
        rc = bfail(b, PSI_HIP_ERROR, "operator build: %s", hipGetErrorString(es));
    return rc;
}

int valid_tune(const psi_batch* b, const psi_batch_tune* t) {
    const int s = shape_index(t->tile_x, t->tile_y);
    if (s < 0) return -1;
    if (t->tiles_per_item < 1 || t->tiles_per_item > 64) return -1;
    if (t->items_per_grab < 1 || t->items_per_grab > 1024) return -1;
    if (t->groups_per_cu < 1 || t->groups_per_cu > std::max(b->occ[s], 1)) return -1;
    if (t->graph_launches < 2 || t->graph_launches > 256 || (t->graph_launches & 1)) return -1;
    return s;
}

// Descriptors, partial-record layout and the full item list for a schedule.
// Descriptors and partial-record layout for tiles bx x by, K tiles per item.
int apply_layout(psi_batch* b, int bx, int by, int K) {
    b->h_mem.resize(b->count);
    int seg_base = 0;
    size_t nitems = 0;
    for (int k = 0; k < b->count; ++k) {
        const auto& h = b->hm[k];
        BMember& m = b->h_mem[k];
        m = BMember{};
        m.g.nx = h.nx;
        m.g.ny = h.ny;
        m.g.pitch = h.nx + 4;
        m.g.cyclic = h.cyclic;
        m.g.tiles_x = (h.nx + bx - 1) / bx;
        m.g.tiles_y = (h.ny + by - 1) / by;
        m.g.ntiles = m.g.tiles_x * m.g.tiles_y;
        m.nseg = (m.g.ntiles + K - 1) / K;
        m.seg_base = seg_base;
        m.off = (long long)h.off;
        m.offx = h.offx;
        m.offy = h.offy;
        seg_base += m.nseg;
        nitems += m.nseg;
    }
    if (nitems > b->items_cap || (size_t)seg_base * NRED > b->part_cap)
        return bfail(b, PSI_BAD_ARG, "schedule exceeds the preallocated work list%s");
    b->total_seg = seg_base;
    return PSI_OK;
}

int grid_of(const psi_batch* b, int groups_per_cu) {
    int G = b->ncu * groups_per_cu;
    if (G >= 8) G &= ~7;
    return std::max(G, 1);
}

// Installs a Jacobi-path schedule.
int apply_tune(psi_batch* b, const psi_batch_tune* t) {
    const int s = valid_tune(b, t);
    if (s < 0) return bfail(b, PSI_BAD_ARG, "invalid batch schedule%s");
    const int rc = apply_layout(b, kShapes[s].bx, kShapes[s].by, t->tiles_per_item);
    if (rc) return rc;
    b->tune = *t;
    b->shape = s;
    b->poly_k = 0;
    b->G = grid_of(b, t->groups_per_cu);
    return PSI_OK;
}

// Lays the descriptors out for the polynomial kernel of degree k, keeping
// the schedule's tiles per item and grab size (occupancy capped by the
// kernel's).
int apply_poly(psi_batch* b, int k) {
    int bx, by;
    poly_tile_dims(k, &bx, &by);
    const int rc = apply_layout(b, bx, by, b->tune.tiles_per_item);
    if (rc) return rc;
    b->poly_k = k;
    b->G = grid_of(b, std::max(1, std::min(b->tune.groups_per_cu, b->occ_poly[k])));
    return PSI_OK;
}

// Item list of the members not yet known to be done (all when done == null).
void fill_items(psi_batch* b, const BFinal* done, int list) {
    size_t n = 0;
    int2* items = b->h_items[list];
    for (int k = 0; k < b->count; ++k) {
        if (done && done[k].done) continue;
        for (int j = 0; j < b->h_mem[k].nseg; ++j) items[n++] = make_int2(k, j);
    }
    *b->h_nitems[list] = (int)n;
}

int upload_schedule(psi_batch* b, const double* const* forc, double* const* sol) {
    for (int k = 0; k < b->count; ++k) {
        b->h_mem[k].b = forc[k];
        b->h_mem[k].x = sol[k];
    }
    BHIP(b, hipMemcpyAsync(b->d_mem, b->h_mem.data(), sizeof(BMember) * b->count,
                           hipMemcpyHostToDevice, b->stream));
    fill_items(b, nullptr, 0);
    BHIP(b, hipMemcpyAsync(b->d_items, b->h_items[0], sizeof(int2) * (size_t)*b->h_nitems[0],
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_nitems, b->h_nitems[0], sizeof(int), hipMemcpyHostToDevice,
                           b->stream));
    // every member starts from START; nothing is done
    memset(b->h_st, 0, sizeof(BState) * b->count);
    BHIP(b, hipMemcpyAsync(b->st, b->h_st, sizeof(BState) * b->count, hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipMemsetAsync(b->fin, 0, sizeof(BFinal) * b->count, b->stream));
    BHIP(b, hipMemsetAsync(b->d_ctr, 0, 2 * sizeof(int), b->stream));
    return PSI_OK;
}

BArrays arrays_of(const psi_batch* b) {
    BArrays A;
    A.E64 = b->E64;
    A.N64 = b->N64;
    A.rx64 = b->rx64;
    A.ry64 = b->ry64;
    A.E32 = b->E32;
    A.N32 = b->N32;
    A.rx32 = b->rx32;
    A.ry32 = b->ry32;
    A.act = b->act;
    A.r64 = b->r64;
    A.e32 = b->e32;
    A.u0 = b->u[0];
    A.u1 = b->u[1];
    A.p0 = b->p[0];
    A.p1 = b->p[1];
    A.r0 = b->rb[0];
    A.r1 = b->rb[1];
    return A;
}

// Launch with parity `par` (reads parity par, writes par ^ 1).
int launch_step(psi_batch* b, int par, const BArrays& A, const BParams& P) {
    const BState* si = b->st + (size_t)par * b->count;
    BState* so = b->st + (size_t)(par ^ 1) * b->count;
    const double* pi = b->part + (size_t)par * b->part_cap;
    double* po = b->part + (size_t)(par ^ 1) * b->part_cap;
    const BMember* mem = b->d_mem;
    const int2* items = b->d_items;
    const int* nitems = b->d_nitems;
    BFinal* fin = b->fin;
    int* ctr = b->d_ctr;
    void* args[] = {(void*)&mem, (void*)&items, (void*)&nitems, (void*)&A,   (void*)&P,
                    (void*)&si,  (void*)&so,    (void*)&pi,     (void*)&po,  (void*)&fin,
                    (void*)&ctr, (void*)&par};
    const void* fn = b->poly_k ? poly_kernel_for(b->poly_k) : kernel_for(b->shape);
    BHIP(b, hipLaunchKernel(fn, dim3(b->G), dim3(NT), args, 0, b->stream));
    return PSI_OK;
}

BParams params_of(const psi_batch* b, const psi_opts* o) {
    BParams P;
    const double target = o->relres_target > 0 ? o->relres_target : 1e-10;
    // default FP32 leg: a 20x residual drop (measured on the 512-member
    // ensemble: 662 ms per pass against 675 ms with 4x legs, whose extra
    // residual phases cost whole launches)
    const double red = o->inner_reduction > 0 ? o->inner_reduction : 0.05;
    P.target2 = target * target;
    P.red2 = red * red;
    P.eps = o->congr_epsilon;
    P.max_it = o->max_iterations;
    P.K = b->tune.tiles_per_item;
    P.grab = b->tune.items_per_grab;
    P.frozen = 0;
    return P;
}

}  // namespace

extern "C" {

const char* psi_batch_last_error(const psi_batch* b) {
    return b ? b->err.c_str() : g_create_err.c_str();
}

int psi_batch_create(psi_batch** out, int32_t count, const psi_batch_member* members,
                     void* stream) {
    if (!out || !members || count < 1) return bfail(nullptr, PSI_BAD_ARG, "null argument%s");
    *out = nullptr;
    for (int k = 0; k < count; ++k) {
        const psi_batch_member& m = members[k];
        const psi_grid& g = m.grid;
        if (g.nx < 2 || g.ny < 2 || !g.dxt || !g.dxu || !g.dyt || !g.dyu || !g.cost || !g.cosu ||
            !m.hur || !m.hvr || m.islands.nisle < 0 ||
            (m.islands.nisle > 0 && (!m.islands.offsets || !m.islands.bi || !m.islands.bj)))
            return bfail(nullptr, PSI_BAD_ARG, "bad member%s");
    }
    psi_batch* b = new psi_batch();
    b->stream = (hipStream_t)stream;
    if (!b->stream) {
        // graphs cannot be captured on the null stream: use a stream of our
        // own (default priority) and order it after the null stream per call
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess) {
            b->err = "stream create failed";
            g_create_err = b->err;
            batch_free(b);
            return PSI_HIP_ERROR;
        }
        b->own_stream = true;
        if (hipEventRecord(b->ev_in, nullptr) != hipSuccess ||
            hipStreamWaitEvent(b->stream, b->ev_in, 0) != hipSuccess) {
            b->err = "stream ordering failed";
            g_create_err = b->err;
            batch_free(b);
            return PSI_HIP_ERROR;
        }
    }
    b->count = count;
    (void)hipGetDevice(&b->device);
    int rc = PSI_OK;
#define BTRY(x)                  \
    do {                         \
        rc = (x);                \
        if (rc) goto failed;     \
    } while (0)
    {
        b->hm.resize(count);
        size_t off = 0;
        int offx = 0, offy = 0;
        size_t max_tiles = 0;
        for (int k = 0; k < count; ++k) {
            auto& h = b->hm[k];
            h.nx = members[k].grid.nx;
            h.ny = members[k].grid.ny;
            h.cyclic = members[k].grid.cyclic_x ? 1 : 0;
            h.plane = (size_t)(h.nx + 4) * (h.ny + 4);
            h.off = off;
            h.offx = offx;
            h.offy = offy;
            off += (h.plane + 63) & ~(size_t)63;
            offx += h.nx + 4;
            offy += h.ny + 4;
            // most work items any schedule can make: the smallest tile, 1 per item
            max_tiles += (size_t)((h.nx + 31) / 32) * ((h.ny + 7) / 8);
            // the polynomial path needs positive metrics (diagonal dominance)
            const psi_grid& gr = members[k].grid;
            for (int q = 1; q < h.nx + 3; ++q)
                if (!(gr.dxt[q] > 0 && gr.dxu[q] > 0)) b->metrics_positive = false;
            for (int q = 1; q < h.ny + 3; ++q)
                if (!(gr.dyt[q] > 0 && gr.dyu[q] > 0 && gr.cost[q] > 0 && gr.cosu[q] > 0))
                    b->metrics_positive = false;
        }
        b->total = off;
        b->totx = offx;
        b->toty = offy;
        BTRY(balloc(b, &b->E64, b->total));
        BTRY(balloc(b, &b->N64, b->total));
        BTRY(balloc(b, &b->r64, b->total));
        BTRY(balloc(b, &b->E32, b->total));
        BTRY(balloc(b, &b->N32, b->total));
        BTRY(balloc(b, &b->e32, b->total));
        for (int q = 0; q < 2; ++q) {
            BTRY(balloc(b, &b->u[q], b->total));
            BTRY(balloc(b, &b->p[q], b->total));
            BTRY(balloc(b, &b->rb[q], b->total));
        }
        BTRY(balloc(b, &b->act, b->total));
        BTRY(balloc(b, &b->rx64, (size_t)offx));
        BTRY(balloc(b, &b->ry64, (size_t)offy));
        BTRY(balloc(b, &b->rx32, (size_t)offx));
        BTRY(balloc(b, &b->ry32, (size_t)offy));
        b->items_cap = max_tiles;
        b->part_cap = max_tiles * NRED;
        BTRY(balloc(b, &b->d_items, b->items_cap));
        BTRY(balloc(b, &b->d_nitems, 1));
        BTRY(balloc(b, &b->d_ctr, 2));
        BTRY(balloc(b, &b->part, 2 * b->part_cap));
        BTRY(balloc(b, &b->st, 2 * (size_t)count));
        BTRY(balloc(b, &b->fin, (size_t)count));
        BTRY(balloc(b, &b->d_mem, (size_t)count));
        BTRY(balloc(b, &b->d_t0, 1));
        if (hipHostMalloc((void**)&b->h_items[0], sizeof(int2) * b->items_cap, 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_items[1], sizeof(int2) * b->items_cap, 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_nitems[0], sizeof(int), 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_nitems[1], sizeof(int), 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_st, sizeof(BState) * count, 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_fin[0], sizeof(BFinal) * count, 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_fin[1], sizeof(BFinal) * count, 0) != hipSuccess ||
            hipHostMalloc((void**)&b->h_t0, sizeof(unsigned long long), 0) != hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "hipHostMalloc failed%s");
            goto failed;
        }
        for (auto& e : b->ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                rc = bfail(b, PSI_HIP_ERROR, "event create failed%s");
                goto failed;
            }
        if (hipEventCreate(&b->tb) != hipSuccess || hipEventCreate(&b->te) != hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "event create failed%s");
            goto failed;
        }
        if (hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device) !=
                hipSuccess ||
            hipDeviceGetAttribute(&b->clock_khz, hipDeviceAttributeWallClockRate, b->device) !=
                hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "device query failed%s");
            goto failed;
        }
        for (int s = 0; s < kNumShapes; ++s)
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->occ[s], kernel_for(s), NT, 0) !=
                hipSuccess) {
                rc = bfail(b, PSI_HIP_ERROR, "occupancy query failed%s");
                goto failed;
            }
        for (int k = kPolyMin; k <= kPolyMax; ++k)
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->occ_poly[k], poly_kernel_for(k),
                                                             NT, 0) != hipSuccess) {
                rc = bfail(b, PSI_HIP_ERROR, "occupancy query failed%s");
                goto failed;
            }
        if (hipStreamSynchronize(b->stream) != hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "allocation failed%s");
            goto failed;
        }
        for (int k = 0; k < count; ++k) BTRY(build_member(b, k, members[k]));
        // active points (the throughput denominator)
        {
            std::vector<uint8_t> a(b->total);
            if (hipMemcpy(a.data(), b->act, b->total, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = bfail(b, PSI_HIP_ERROR, "readback failed%s");
                goto failed;
            }
            int64_t n = 0;
            for (uint8_t v : a) n += v;
            b->active_points = n;
        }
        psi_batch_tune t{};
        t.tile_x = 64;
        t.tile_y = 16;
        t.tiles_per_item = 4;
        t.items_per_grab = 2;
        t.groups_per_cu = std::max(b->occ[0], 1);
        t.graph_launches = 16;
        BTRY(apply_tune(b, &t));
    }
    *out = b;
    return PSI_OK;
failed:
    g_create_err = b->err;
    batch_free(b);
    return rc;
#undef BTRY
}

void psi_batch_destroy(psi_batch* b) { batch_free(b); }

int64_t psi_batch_device_bytes(const psi_batch* b) { return b ? b->bytes : 0; }

int64_t psi_batch_active_points(const psi_batch* b) { return b ? b->active_points : 0; }

int psi_batch_get_tune(const psi_batch* b, psi_batch_tune* t) {
    if (!b || !t) return PSI_BAD_ARG;
    *t = b->tune;
    return PSI_OK;
}

int psi_batch_set_tune(psi_batch* b, const psi_batch_tune* t) {
    if (!b || !t) return bfail(b, PSI_BAD_ARG, "null argument%s");
    return apply_tune(b, t);
}

int psi_batch_max_groups_per_cu(const psi_batch* b, int32_t tile_x, int32_t tile_y) {
    if (!b) return 0;
    const int s = shape_index(tile_x, tile_y);
    return s < 0 ? 0 : b->occ[s];
}

// Times `launches` batched sweeps per candidate schedule on the members'
// operators (right-hand side of ones, zero first guess, internal buffers)
// and keeps the fastest.  Candidates: every tile shape x tiles per item in
// {2, 4, 8} x items per grab in {1, 4} x {full, half} occupancy.
int psi_batch_autotune(psi_batch* b, int32_t launches, psi_batch_tune* best_out,
                       double* table, int32_t table_cap, int32_t* table_n) {
    if (!b) return bfail(nullptr, PSI_BAD_ARG, "null batch%s");
    if (launches < 2) launches = 16;
    const psi_batch_tune keep = b->tune;
    double *tb64 = nullptr, *tx64 = nullptr;
    BHIP(b, hipMalloc((void**)&tb64, b->total * 8));
    if (hipMalloc((void**)&tx64, b->total * 8) != hipSuccess) {
        (void)hipFree(tb64);
        return bfail(b, PSI_HIP_ERROR, "hipMalloc failed%s");
    }
    std::vector<const double*> fb(b->count);
    std::vector<double*> fx(b->count);
    for (int k = 0; k < b->count; ++k) {
        fb[k] = tb64 + b->hm[k].off;
        fx[k] = tx64 + b->hm[k].off;
    }
    psi_opts o{};
    o.mode = PSI_MODE_MIXED;
    o.criterion = PSI_CRIT_RELRES;
    o.relres_target = 1e-30;        // never met in the timing window
    o.max_iterations = 1 << 30;
    o.inner_reduction = 1e-30;      // no leg ends in the timing window
    int rc = PSI_OK;
    psi_batch_tune best = keep;
    double best_ns = 1e300;
    int rows = 0;
    const BArrays A = arrays_of(b);
    for (int s = 0; s < kNumShapes && rc == PSI_OK; ++s) {
        const int occ = std::max(b->occ[s], 1);
        const int gpcs[2] = {occ, std::max(occ / 2, 1)};
        for (int K : {2, 4, 8}) {
            for (int gi = 0; gi < (gpcs[1] == gpcs[0] ? 2 : 4); ++gi) {
                psi_batch_tune t{};
                t.tile_x = kShapes[s].bx;
                t.tile_y = kShapes[s].by;
                t.tiles_per_item = K;
                t.items_per_grab = (gi & 1) ? 4 : 1;
                t.groups_per_cu = gpcs[gi >> 1];
                t.graph_launches = keep.graph_launches;
                if ((rc = apply_tune(b, &t))) break;
                hipLaunchKernelGGL(k_batch_tune_fill, dim3(nblk(b->total)), dim3(256), 0,
                                   b->stream, b->total, (const uint8_t*)b->act, tb64, tx64);
                if ((rc = upload_schedule(b, fb.data(), fx.data()))) break;
                BParams P = params_of(b, &o);
                P.frozen = 1;
                // every member sweeps in every launch: 4 warm-up launches, then the timed ones
                int par = 0;
                for (int w = 0; w < 4 && rc == PSI_OK; ++w, par ^= 1) rc = launch_step(b, par, A, P);
                if (rc) break;
                float ms = 0.f;
                if (hipEventRecord(b->tb, b->stream) != hipSuccess) rc = PSI_HIP_ERROR;
                for (int w = 0; w < launches && rc == PSI_OK; ++w, par ^= 1)
                    rc = launch_step(b, par, A, P);
                if (rc == PSI_OK && (hipEventRecord(b->te, b->stream) != hipSuccess ||
                                     hipEventSynchronize(b->te) != hipSuccess ||
                                     hipEventElapsedTime(&ms, b->tb, b->te) != hipSuccess))
                    rc = bfail(b, PSI_HIP_ERROR, "autotune timing failed%s");
                if (rc) break;
                const double ns = 1e6 * ms / launches / std::max<int64_t>(b->active_points, 1);
                if (table && rows < table_cap) {
                    double* row = table + (size_t)rows * 6;
                    row[0] = t.tile_x;
                    row[1] = t.tile_y;
                    row[2] = t.tiles_per_item;
                    row[3] = t.items_per_grab;
                    row[4] = t.groups_per_cu;
                    row[5] = ns;
                }
                ++rows;
                if (ns < best_ns) {
                    best_ns = ns;
                    best = t;
                    best.ns_per_point = ns;
                }
            }
            if (rc) break;
        }
    }
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(tb64);
    (void)hipFree(tx64);
    if (table_n) *table_n = rows;
    if (rc) {
        (void)apply_tune(b, &keep);
        return rc;
    }
    rc = apply_tune(b, &best);
    if (rc == PSI_OK && best_out) *best_out = best;
    return rc;
}

int psi_batch_congrad(psi_batch* b, const psi_opts* o, const double* const* forc,
                      double* const* sol, psi_info* infos, double* latency_ms,
                      psi_batch_stats* stats) {
    if (!b || !o || !forc || !sol) return bfail(b, PSI_BAD_ARG, "null argument%s");
    if ((o->mode != PSI_MODE_MIXED && o->mode != PSI_MODE_POLY) ||
        o->criterion != PSI_CRIT_RELRES)
        return bfail(b, PSI_BAD_ARG,
                     "batched solves run the mixed or poly path with the relres criterion%s");
    if (o->max_iterations < 1) return bfail(b, PSI_BAD_ARG, "max_iterations < 1%s");
    int poly = 0;
    double lower = 0.0;
    if (o->mode == PSI_MODE_POLY) {
        poly = o->poly_degree > 0 ? o->poly_degree : 4;
        lower = o->poly_lower > 0 ? o->poly_lower : 0.005;
        if (poly < kPolyMin || poly > kPolyMax || !(lower < 1.0))
            return bfail(b, PSI_BAD_ARG, "batched poly path: degree 2..4, poly_lower in (0,1)%s");
        if (!b->metrics_positive)
            return bfail(b, PSI_BAD_ARG, "poly mode needs positive grid metrics%s");
    }
    for (int k = 0; k < b->count; ++k)
        if (!forc[k] || !sol[k]) return bfail(b, PSI_BAD_ARG, "null member field%s");
    if (b->own_stream) {
        BHIP(b, hipEventRecord(b->ev_in, nullptr));
        BHIP(b, hipStreamWaitEvent(b->stream, b->ev_in, 0));
    }
    int rc = poly ? apply_poly(b, poly) : apply_tune(b, &b->tune);
    if (rc) return rc;
    rc = upload_schedule(b, forc, sol);
    if (rc) return rc;
    BParams P = params_of(b, o);
    if (poly) P.ch = make_cheb(poly, lower);
    const BArrays A = arrays_of(b);
    const int KL = b->tune.graph_launches;
    // capture KL launches (parities alternate; KL is even, so every replay
    // starts on parity 0 again)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < KL && rc == PSI_OK; ++k) rc = launch_step(b, k & 1, A, P);
    {
        const hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return bfail(b, PSI_HIP_ERROR, "graph capture: %s", hipGetErrorString(e));
    }
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return bfail(b, PSI_HIP_ERROR, "graph instantiate failed%s");
    }
    // at most ~4 launches per inner iteration (a leg adds CORR, RESID, REFRESH)
    const long max_replays = (4L * o->max_iterations + 64) / KL + 2;
    long replays = 0, polls = 0;
    int queue[2], qn = 0, slot = 0;
    bool all_done = false;
    int ndone_seen = 0;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    };
    auto issue = [&]() -> int {
        if (hipGraphLaunch(exec, b->stream) != hipSuccess)
            return bfail(b, PSI_HIP_ERROR, "graph launch failed%s");
        ++replays;
        if (hipMemcpyAsync(b->h_fin[slot], b->fin, sizeof(BFinal) * b->count,
                           hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
            hipEventRecord(b->ev[slot], b->stream) != hipSuccess)
            return bfail(b, PSI_HIP_ERROR, "poll enqueue failed%s");
        queue[qn++] = slot;
        slot ^= 1;
        return PSI_OK;
    };
    if (hipEventRecord(b->tb, b->stream) != hipSuccess) {
        cleanup();
        return bfail(b, PSI_HIP_ERROR, "event record failed%s");
    }
    hipLaunchKernelGGL(k_batch_clock, dim3(1), dim3(1), 0, b->stream, b->d_t0);
    int list = 0;
    while (!all_done) {
        while (rc == PSI_OK && qn < 2 && replays < max_replays) rc = issue();
        if (rc) break;
        if (qn == 0) {
            rc = bfail(b, PSI_HIP_ERROR, "batched solve launch budget exhausted%s");
            break;
        }
        const int s = queue[0];
        queue[0] = queue[1];
        --qn;
        if (hipEventSynchronize(b->ev[s]) != hipSuccess) {
            rc = bfail(b, PSI_HIP_ERROR, "poll failed%s");
            break;
        }
        ++polls;
        int nd = 0;
        for (int k = 0; k < b->count; ++k) nd += b->h_fin[s][k].done ? 1 : 0;
        if (nd == b->count) {
            all_done = true;
            break;
        }
        if (nd != ndone_seen) {
            // Drop finished members.  The copy runs after the replay already
            // queued.  The two pinned lists alternate: the copy from the list
            // written two rebuilds ago was enqueued before the poll just
            // read (at most one queued poll precedes a copy), so it is done.
            ndone_seen = nd;
            fill_items(b, b->h_fin[s], list);
            if (hipMemcpyAsync(b->d_items, b->h_items[list],
                               sizeof(int2) * (size_t)*b->h_nitems[list], hipMemcpyHostToDevice,
                               b->stream) != hipSuccess ||
                hipMemcpyAsync(b->d_nitems, b->h_nitems[list], sizeof(int),
                               hipMemcpyHostToDevice, b->stream) != hipSuccess) {
                rc = bfail(b, PSI_HIP_ERROR, "item list upload failed%s");
                break;
            }
            list ^= 1;
        }
    }
    if (rc) {
        cleanup();
        return rc;
    }
    if (hipEventRecord(b->te, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->h_t0, b->d_t0, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       b->stream) != hipSuccess) {
        cleanup();
        return bfail(b, PSI_HIP_ERROR, "readback failed%s");
    }
    cleanup();
    // the last poll slot holds every member's final record
    const BFinal* fin = nullptr;
    for (int s = 0; s < 2 && !fin; ++s) {
        bool all = true;
        for (int k = 0; k < b->count && all; ++k) all = b->h_fin[s][k].done != 0;
        if (all) fin = b->h_fin[s];
    }
    if (!fin) return bfail(b, PSI_HIP_ERROR, "batched solve ended without results%s");
    float total_ms = 0.f;
    (void)hipEventElapsedTime(&total_ms, b->tb, b->te);
    int worst = PSI_OK;
    const double khz = b->clock_khz > 0 ? (double)b->clock_khz : 1e5;
    for (int k = 0; k < b->count; ++k) {
        const BFinal& f = fin[k];
        const int st = status_of(f.status);
        if (worst == PSI_OK && st != PSI_OK) worst = st;
        if (infos) {
            psi_info& in = infos[k];
            in = psi_info{};
            in.iterations = f.iterations;
            in.converged = f.status == ST_CONVERGED;
            in.status = st;
            in.outer_iterations = f.legs;
            in.est_error = f.est_error;
            in.relres_recursive = f.bb_rec > 0 ? sqrt(f.rr_rec / f.bb_rec) : 0.0;
            in.relres_true = f.bbt > 0 ? sqrt(f.rrt / f.bbt) : sqrt(f.rrt);
            in.elapsed_ms = (double)(int64_t)(f.t_done - *b->h_t0) / khz;
            in.sweep_launches = f.launches;
        }
        if (latency_ms) latency_ms[k] = (double)(int64_t)(f.t_done - *b->h_t0) / khz;
    }
    if (stats) {
        stats->elapsed_ms = total_ms;
        stats->graph_replays = (int32_t)replays;
        stats->launches = (int32_t)(replays * KL);
        stats->polls = (int32_t)polls;
        stats->active_points = b->active_points;
    }
    return worst;
}

}  // extern "C"
