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
// with a weight (MG_SCALE_*), and K Chebyshev steps again.  Vectors are FP32 in a
// 1-cell-halo layout per level; a level costs two LDS-tiled kernels per
// cycle (pre-smoothing with the restriction, prolongation with the
// post-smoothing).
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
constexpr int MG_MAXK = 6;           // Chebyshev steps per smoothing at most (finest levels)
constexpr int MG_MAXKC = 10;         // ... on the coarse levels (psi_opts.mg_coarse_degree)
#ifndef PSI_MG_COARSE_DEGREE
#define PSI_MG_COARSE_DEGREE 5       // default degree of the coarse levels
#endif
// A level is coarse when it has at most this many 64 x 16 tiles (one per CU
// or fewer): its kernels are bound by launch latency, not by their work.
constexpr long MG_COARSE_TILES = 256;
// weight of the coarse-grid correction, per operator (swept 1.3-2.4 at 2048^2:
// 1.9 for the streamfunction with 5 smoothing steps, 1.5 for the surface
// pressure, whose free-surface term makes the operator better conditioned,
// with 4)
#ifndef PSI_MG_SCALE_PSI
#define PSI_MG_SCALE_PSI 1.9
#endif
constexpr double MG_SCALE_PSI = PSI_MG_SCALE_PSI, MG_SCALE_PRESS = 1.5;
constexpr int MG_NRED = 3;           // reduction slots of the outer passes

// Level layout: cell (I, J) at (J + 1) * pitch + I + MG_XOFF, one halo row
// above and below, MG_XOFF halo columns on the left (so the LDS frames of
// the smoothing kernels, which start MG_XOFF cells left of a 64-cell tile
// boundary at the default degree, begin on a 128-byte line), pitch a
// multiple of 32 floats with at least 4 halo columns on the right.
constexpr int MG_XOFF = 4;
inline int mg_pitch(int nx) { return ((nx + MG_XOFF + 4 + 31) / 32) * 32; }

struct MgLev {
    int nx, ny, pitch, cyc;          // interior cells; pitch = mg_pitch(nx)
    const float *E, *N, *C;
    const uint8_t* act;
    float *b;                        // right-hand side (restricted residual)
    float *x;                        // pre-smoothed solution
    float *y;                        // solution after the coarse correction and post-smoothing
    // the coarsest level below a smoothing level only: each cell's row of the
    // dense inverse (-1: none) and b in row order, written by the restriction
    // so that the dense solve reads it without an index lookup
    const int* didx;
    float* bd;
};

struct MgCheb {
    float c0;               // 1/theta
    float c1[MG_MAXKC];     // rho_k rho_{k-1}
    float c2[MG_MAXKC];     // 2 rho_k / delta
    float cscale;           // weight of the coarse-grid correction
};

// Device scalars of the outer iteration.
struct MgScal {
    double beta, alpha, pq, pq_old;
    double s_r, s_pending;           // scale of the FP32 residual fed to the V-cycle
    double rr, bb, rtarget2;
    int n, done, status, max_it;
    // FP64 true residual of a leg that converged on the recursive one:
    // 0 not checked yet, 1 confirmed, -1 above the target (a new leg)
    int confirmed;
};

// Coarsest level: y = Ainv b over its active cells.  The workgroup gathers b
// into LDS once; each wave takes one row of Ainv (row stride ld, a multiple
// of 4 zero-padded past n) in 16-byte loads, all in flight together.
__global__ void __launch_bounds__(256) k_mg_dense(MgLev L, const float* __restrict__ Ainv, int ld,
                                                  const int* __restrict__ cells, int n,
                                                  const MgScal* __restrict__ S) {
    __shared__ float4 sB[MG_MAX_DENSE];   // 4 MG_MAX_DENSE floats: the largest coarsest level
    const int done = S->done;
    const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int yq = cells[row < n ? row : 0];   // where the row's result goes, loaded up front
    // the row's first MG_MAX_DENSE / 256 16-byte pieces per lane (all of a
    // row of at most MG_MAX_DENSE cells), loaded beside the gather of b; rows
    // past n read row 0 and store nothing
    constexpr int PL = MG_MAX_DENSE / 256;
    const float4* a = reinterpret_cast<const float4*>(Ainv + (size_t)(row < n ? row : 0) * ld);
    float4 av[PL];
#pragma unroll
    for (int t = 0; t < PL; ++t) {
        const int k = lane + 64 * t;
        av[t] = a[k < ld / 4 ? k : 0];
    }
    float* sb = reinterpret_cast<float*>(sB);
    if (L.bd)
        for (int k = threadIdx.x; k < ld; k += blockDim.x) sb[k] = k < n ? L.bd[k] : 0.0f;
    else
        for (int k = threadIdx.x; k < ld; k += blockDim.x) sb[k] = k < n ? L.b[cells[k]] : 0.0f;
    __syncthreads();
    if (done || row >= n) return;
    float s = 0.0f;
#pragma unroll
    for (int t = 0; t < PL; ++t) {
        const int k = lane + 64 * t;
        if (k < ld / 4) {
            const float4 v = av[t], w = sB[k];
            s += ((v.x * w.x + v.y * w.y) + v.z * w.z) + v.w * w.w;
        }
    }
    for (int k = lane + 64 * PL; k < ld / 4; k += 64) {   // coarsest levels above MG_MAX_DENSE
        const float4 v = a[k], w = sB[k];
        s += ((v.x * w.x + v.y * w.y) + v.z * w.z) + v.w * w.w;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) L.y[yq] = s;
}

// ---- fused level kernels ------------------------------------------------------
// One workgroup per tile of a level with a K-cell frame around it: the K
// smoothing steps recompute the frame's shrinking margins instead of writing
// every step to memory, so a level costs two kernels per V-cycle
// (pre-smoothing + restriction, prolongation + post-smoothing).
//
// The frame is 64 columns (one per lane) x NW * RPT rows (NW waves); wave w holds
// rows [w RPT, (w+1) RPT) of every column in registers, with each cell's
// couplings (its own C, E, N, its west neighbour's E and its south
// neighbour's N) and 1/C.  A stencil takes the west and east neighbours from
// the adjacent lanes (DPP whole-wave shifts), the rows inside the band from
// registers, and only the rows just below and above the band from LDS (one
// exchange, one barrier per step).  Every cell is updated at every step:
// a cell at margin m is exact after step m, and garbage from the frame edge
// (where a shift brings in zero) moves in by one cell per step, so the own
// tile (margin >= K) is exact at the end.
// Waves and rows per lane: (MG_NW_LARGE, MG_RPT_LARGE) on levels that keep
// at least MG_LARGE_TILES tiles with it, (MG_NW_SMALL, mg_rpt_small) on the
// others, whose kernels are bound by latency and gain from more, smaller
// workgroups.
constexpr int MG_FW = 64;
#ifndef PSI_MG_NW_LARGE
#define PSI_MG_NW_LARGE 8
#endif
#ifndef PSI_MG_RPT_LARGE
#define PSI_MG_RPT_LARGE 4
#endif
#ifndef PSI_MG_NW_SMALL
#define PSI_MG_NW_SMALL 16
#endif
#ifndef PSI_MG_RPT_SMALL
#define PSI_MG_RPT_SMALL 1
#endif
#ifndef PSI_MG_LARGE_TILES
#define PSI_MG_LARGE_TILES 200
#endif
constexpr int MG_NW_LARGE = PSI_MG_NW_LARGE, MG_RPT_LARGE = PSI_MG_RPT_LARGE;
constexpr int MG_NW_SMALL = PSI_MG_NW_SMALL;
constexpr long MG_LARGE_TILES = PSI_MG_LARGE_TILES;
// the small levels' rows per lane at degree K: PSI_MG_RPT_SMALL where that
// leaves a tile at least 4 rows tall, else the fewest that do
constexpr int mg_rpt_small(int K) {
    return MG_NW_SMALL * PSI_MG_RPT_SMALL - 2 * K >= 4
               ? PSI_MG_RPT_SMALL
               : (MG_NW_SMALL * 2 - 2 * K >= 4 ? 2 : 3);
}

// Tile of a level kernel at degree K, NW waves and RPT rows per lane:
// (64 - 2K) x (NW RPT - 2K), both even.
inline int mg_tile_x(int K) { return MG_FW - 2 * K; }
inline int mg_tile_h(int K, int nw, int rpt) { return nw * rpt - 2 * K; }
inline unsigned mg_tiles(const MgLev& v, int K, int nw, int rpt) {
    return (unsigned)(((v.nx + mg_tile_x(K) - 1) / mg_tile_x(K)) *
                      ((v.ny + mg_tile_h(K, nw, rpt) - 1) / mg_tile_h(K, nw, rpt)));
}
inline bool mg_large_level(const MgLev& v, int K) {
    return mg_tiles(v, K, MG_NW_LARGE, MG_RPT_LARGE) >= MG_LARGE_TILES;
}

// A level is coarse when it has at most MG_COARSE_TILES 64 x 16 blocks of
// cells (one per CU or fewer): its kernels are bound by launch latency, not
// by their work.
inline bool mg_coarse_level(const MgLev& v) {
    return (long)((v.nx + 63) / 64) * ((v.ny + 15) / 16) <= MG_COARSE_TILES;
}

// Level cell of frame cell (fx, fy) of the tile at (I0, J0); -1 outside the
// level (cyclic in x wraps).
__device__ __forceinline__ int mg_frame_cell(const MgLev& L, int I, int J) {
    if (J < 0 || J >= L.ny) return -1;
    if (I < 0 || I >= L.nx) {
        if (!L.cyc) return -1;
        I = (I % L.nx + L.nx) % L.nx;
    }
    return (J + 1) * L.pitch + I + MG_XOFF;
}

// A level cell every load may use in place of one outside the level (row
// 0 of the interior, first column), so the frame's loads need no branch.
__device__ __forceinline__ int mg_safe_cell(const MgLev& L, int q) {
    return q >= 0 ? q : L.pitch + MG_XOFF;
}

// The value of the lane to the west (lane - 1) and to the east (lane + 1);
// the frame's edge lanes get zero.
__device__ __forceinline__ float mg_from_west(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float mg_from_east(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// Rows below and above each wave's band: v[buf][0 = first row | 1 = last
// row][wave][lane].  Buffers alternate, so one is rewritten only after the
// barrier that follows its last read.
template <int NW>
struct MgXch {
    float v[2][2][NW][64];
};

template <int NW>
__device__ __forceinline__ void mg_xch(MgXch<NW>& X, int& buf, float first, float last,
                                       float& below, float& above) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    X.v[buf][0][w][lane] = first;
    X.v[buf][1][w][lane] = last;
    __syncthreads();
    below = w > 0 ? X.v[buf][1][w - 1][lane] : 0.0f;
    above = w < NW - 1 ? X.v[buf][0][w + 1][lane] : 0.0f;
    buf ^= 1;
}

// S v for the band's cells (sum C, E, W, N, S), given the rows below and
// above the band.
template <int R>
__device__ __forceinline__ void mg_band_apply(const float (&v)[R], float below, float above,
                                              const float (&cC)[R], const float (&cE)[R],
                                              const float (&cW)[R], const float (&cN)[R],
                                              const float (&cS)[R], float (&out)[R]) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const float vs = k ? v[k - 1] : below, vn = k < R - 1 ? v[k + 1] : above;
        const float vw = mg_from_west(v[k]), ve = mg_from_east(v[k]);
        out[k] = cC[k] * v[k] + cE[k] * ve + cW[k] * vw + cN[k] * vn + cS[k] * vs;
    }
}

// The cell's west and south couplings (its neighbours' E and N).
template <int NW, int R>
__device__ __forceinline__ void mg_band_couplings(MgXch<NW>& X, int& buf, const float (&cE)[R],
                                                  const float (&cN)[R], float (&cW)[R],
                                                  float (&cS)[R]) {
    float below, above;
    mg_xch(X, buf, cN[0], cN[R - 1], below, above);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        cW[k] = mg_from_west(cE[k]);
        cS[k] = k ? cN[k - 1] : below;
    }
}

// Pre-smoothing from x = 0 (K Chebyshev steps), the residual, and its 2 x 2
// restriction to the next level's right-hand side.
template <int K, int NW, int R>
__global__ void __launch_bounds__(64 * NW) k_mg_pre(MgLev L, MgCheb ch, MgLev Lc,
                                                  const MgScal* __restrict__ S) {
    constexpr int FH = NW * R, TX = MG_FW - 2 * K, TY = FH - 2 * K;
    static_assert(TY >= 2, "tile too short for the degree");
    __shared__ MgXch<NW> X;
    __shared__ float sR[FH][MG_FW];   // the residual, for the restriction
    const int done = S->done;   // read beside the frame loads, checked before any store
    const int tiles_x = (L.nx + TX - 1) / TX;
    // each XCD sweeps a contiguous band of tiles, so the frame margins shared
    // with the neighbouring tiles are read through the same L2
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int I0 = (tile % tiles_x) * TX, J0 = (tile / tiles_x) * TY;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = I0 - K + lane;
    // the tile's coarse cells: this thread's first cell and its mask are
    // loaded beside the frame, so the restriction at the end waits for no
    // load (shapes with more coarse cells than threads load the rest there)
    constexpr int NC = (TX / 2) * (TY / 2);
    const int ccy = threadIdx.x / (TX / 2), ccx = threadIdx.x - ccy * (TX / 2);
    const int IC = I0 / 2 + ccx, JC = J0 / 2 + ccy;
    const bool cown = threadIdx.x < NC && IC < Lc.nx && JC < Lc.ny;
    const int QC = cown ? (JC + 1) * Lc.pitch + IC + MG_XOFF : MG_XOFF;
    const uint8_t cact = Lc.act[QC];
    const int cdq = Lc.didx ? Lc.didx[QC] : -1;
    float r[R], x[R], d[R], cC[R], cE[R], cN[R], cW[R], cS[R], di[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        // loads at a safe cell outside the level (no branch, one round trip)
        const int q = mg_frame_cell(L, I, J0 - K + w * R + k), qs = mg_safe_cell(L, q);
        const float lb = L.b[qs], lc = L.C[qs], le = L.E[qs], ln = L.N[qs];
        const float b = q >= 0 ? lb : 0.0f, c = q >= 0 ? lc : 1.0f;
        cE[k] = q >= 0 ? le : 0.0f;
        cN[k] = q >= 0 ? ln : 0.0f;
        cC[k] = c;
        di[k] = 1.0f / c;
        const float d0 = b * ch.c0 * di[k];
        r[k] = b;
        x[k] = d0;
        d[k] = d0;
    }
    int buf = 0;
    mg_band_couplings(X, buf, cE, cN, cW, cS);   // (uses every loaded row: no load below the test)
    if (done) return;
    // steps 1..K-1; step K: the residual
#pragma unroll
    for (int s = 1; s <= K; ++s) {
        float below, above, sv[R];
        mg_xch(X, buf, d[0], d[R - 1], below, above);
        mg_band_apply(d, below, above, cC, cE, cW, cN, cS, sv);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            r[k] = r[k] - sv[k];
            if (s < K) {
                const float dn = ch.c1[s] * d[k] + ch.c2[s] * (r[k] * di[k]);
                d[k] = dn;
                x[k] = x[k] + dn;
            }
        }
    }
    // the tile's x, the residual into LDS for the restriction
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int fy = w * R + k;
        sR[fy][lane] = r[k];
        if (lane >= K && lane < K + TX && fy >= K && fy < K + TY) {
            const int J = J0 - K + fy;
            if (I < L.nx && J < L.ny) L.x[(J + 1) * L.pitch + I + MG_XOFF] = x[k];
        }
    }
    __syncthreads();
    auto restrict_cell = [&](int cx, int cy, int Q, uint8_t a, int dq) {
        float sum = 0.0f;
        for (int dj = 0; dj < 2; ++dj)
            for (int dx = 0; dx < 2; ++dx) {
                const int tx = 2 * cx + dx, ty = 2 * cy + dj;
                if (I0 + tx < L.nx && J0 + ty < L.ny) sum += sR[ty + K][tx + K];
            }
        const float bc = a ? sum : 0.0f;
        Lc.b[Q] = bc;
        if (dq >= 0) Lc.bd[dq] = bc;
    };
    if (cown) restrict_cell(ccx, ccy, QC, cact, cdq);
    if (NC > 64 * NW)
        for (int l = threadIdx.x + 64 * NW; l < NC; l += 64 * NW) {
            const int cy = l / (TX / 2), cx = l - cy * (TX / 2);
            const int ic = I0 / 2 + cx, jc = J0 / 2 + cy;
            if (ic < Lc.nx && jc < Lc.ny) {
                const int Q = (jc + 1) * Lc.pitch + ic + MG_XOFF;
                restrict_cell(cx, cy, Q, Lc.act[Q], Lc.didx ? Lc.didx[Q] : -1);
            }
        }
}

// Prolongation of the coarse correction (weight ch.cscale, active cells) and
// K Chebyshev steps of post-smoothing; the result goes to y.
template <int K, int NW, int R>
__global__ void __launch_bounds__(64 * NW) k_mg_post(MgLev L, MgLev Lc, MgCheb ch,
                                                   const MgScal* __restrict__ S) {
    constexpr int FH = NW * R, TX = MG_FW - 2 * K, TY = FH - 2 * K;
    static_assert(TY >= 2, "tile too short for the degree");
    __shared__ MgXch<NW> X;
    const int done = S->done;
    const int tiles_x = (L.nx + TX - 1) / TX;
    // each XCD sweeps a contiguous band of tiles, so the frame margins shared
    // with the neighbouring tiles are read through the same L2
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int I0 = (tile % tiles_x) * TX, J0 = (tile / tiles_x) * TY;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float r[R], y[R], d[R], cC[R], cE[R], cN[R], cW[R], cS[R], di[R];
    // x, the coarse correction, the mask and the couplings of every row in
    // one round of loads, at a safe cell outside the level (no branch), all
    // issued before any is used
    float lx[R], lyc[R], lc[R], le[R], ln[R], lb[R];
    int la[R];
    bool in[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int I = I0 - K + lane;
        const int J = J0 - K + w * R + k;
        const int q = mg_frame_cell(L, I, J), qs = mg_safe_cell(L, q);
        if (L.cyc) I = (I % L.nx + L.nx) % L.nx;
        const int qc = q >= 0 ? (J / 2 + 1) * Lc.pitch + I / 2 + MG_XOFF : mg_safe_cell(Lc, -1);
        in[k] = q >= 0;
        lx[k] = L.x[qs];
        lyc[k] = Lc.y[qc];
        lc[k] = L.C[qs];
        le[k] = L.E[qs];
        ln[k] = L.N[qs];
        lb[k] = L.b[qs];
        la[k] = L.act[qs];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const float c = in[k] ? lc[k] : 1.0f;
        y[k] = in[k] && la[k] ? lx[k] + ch.cscale * lyc[k] : 0.0f;
        cE[k] = in[k] ? le[k] : 0.0f;
        cN[k] = in[k] ? ln[k] : 0.0f;
        cC[k] = c;
        di[k] = 1.0f / c;
        r[k] = in[k] ? lb[k] : 0.0f;   // b
    }
    int buf = 0;
    mg_band_couplings(X, buf, cE, cN, cW, cS);   // (uses every loaded row: no load below the test)
    if (done) return;
    // step 0: r = b - S y', d0 = D^-1 r / theta, y = y' + d0
    {
        float below, above, sv[R];
        mg_xch(X, buf, y[0], y[R - 1], below, above);
        mg_band_apply(y, below, above, cC, cE, cW, cN, cS, sv);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            r[k] = r[k] - sv[k];
            d[k] = ch.c0 * (r[k] * di[k]);
            y[k] = y[k] + d[k];
        }
    }
#pragma unroll
    for (int s = 1; s < K; ++s) {
        float below, above, sv[R];
        mg_xch(X, buf, d[0], d[R - 1], below, above);
        mg_band_apply(d, below, above, cC, cE, cW, cN, cS, sv);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            r[k] = r[k] - sv[k];
            const float dn = ch.c1[s] * d[k] + ch.c2[s] * (r[k] * di[k]);
            d[k] = dn;
            y[k] = y[k] + dn;
        }
    }
    const int I = I0 - K + lane;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int fy = w * R + k;
        if (lane >= K && lane < K + TX && fy >= K && fy < K + TY) {
            const int J = J0 - K + fy;
            if (I < L.nx && J < L.ny) L.y[(J + 1) * L.pitch + I + MG_XOFF] = y[k];
        }
    }
}

// ---- outer flexible CG (FP64, solver grid layout) ---------------------------
// z = y0 / s_r at interior point (i, j) and its four neighbours, zero
// outside the interior and wrapped in x when cyclic.  Every load is inside
// the level layout (its halo row above and below, MG_XOFF columns to the
// left and at least 4 to the right), so all five are issued before the
// out-of-interior ones are replaced by zero.
__device__ __forceinline__ void mg_z5(const Geom& g, const MgLev& L0, int i, int j, double inv,
                                      double& zc, double& zs, double& zw, double& ze,
                                      double& zn) {
    const int P = L0.pitch, q = (j + 1) * P + i + MG_XOFF;
    int dw = -1, de = 1;
    if (g.cyclic) {
        if (i == 0) dw = g.nx - 1;
        if (i == g.nx - 1) de = 1 - g.nx;
    }
    const float yc = L0.y[q], ys = L0.y[q - P], yn = L0.y[q + P];
    const float yw = L0.y[q + dw], ye = L0.y[q + de];
    zc = (double)yc * inv;
    zs = j > 0 ? (double)ys * inv : 0.0;
    zn = j < g.ny - 1 ? (double)yn * inv : 0.0;
    zw = (g.cyclic || i > 0) ? (double)yw * inv : 0.0;
    ze = (g.cyclic || i < g.nx - 1) ? (double)ye * inv : 0.0;
}

__device__ __forceinline__ double mg_area(const double* dxu, const double* cosu,
                                          const double* dyu, int i, int j) {
    return dxu[i + 2] * cosu[j + 2] * dyu[j + 2];
}

// workgroups of the outer passes (grid-stride; pass A on pairs: per CU)
#ifndef PSI_MG_OUTER_BLOCKS
#define PSI_MG_OUTER_BLOCKS 2048
#endif
#ifndef PSI_MG_PASS_A_PER_CU
#define PSI_MG_PASS_A_PER_CU 2
#endif

// One partial per slot for this workgroup of NW waves (waves summed in order).
template <int NR, int NW = 4>
__device__ __forceinline__ void mg_block_partials(double (&v)[NR], double* part) {
    __shared__ double sh[NR * NW];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NR; ++k) sh[k * NW + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NR; ++k) {
            double a = sh[k * NW];
            for (int w = 1; w < NW; ++w) a += sh[k * NW + w];
            part[(size_t)k * gridDim.x + blockIdx.x] = a;
        }
}

// Pass A: with z = M r from the V-cycle, Az (kept for pass B) and the
// weighted products <z,Az>, <z,q>, <r,z> over active points.  Flexible CG
// also uses <Az,p> and <r,p>; the first equals <z,q> (W A is symmetric and
// q = A p) and the second vanishes by the previous line search, so p is not
// read here.
__global__ void __launch_bounds__(256) k_mg_fcg_a1(
    Geom g, MgLev L0, const double* __restrict__ E, const double* __restrict__ N,
    const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ dxu, const double* __restrict__ cosu,
    const double* __restrict__ dyu, const uint8_t* __restrict__ act,
    const double* __restrict__ q, const double* __restrict__ r, double* __restrict__ az_out,
    const double* __restrict__ zscale, const MgScal* __restrict__ S,
    double* __restrict__ part) {
    if (S->done) return;
    const double inv = 1.0 / *zscale;   // the scale of the V-cycle's input
    double acc[3] = {0, 0, 0};
    // row segments of 256 points, each XCD a contiguous range of rows
    const int segs = (g.nx + 255) / 256;
    int first, stride, end;
    range_schedule(g.ny * segs, first, stride, end);
    for (int it = first; it < end; it += stride) {
        const int j = it / segs, i = (it - j * segs) * 256 + threadIdx.x;
        if (i >= g.nx) continue;
        const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
        const int c = i + 2, rr = j + 2;
        // every load in bounds, issued before the active test
        double zc, zs, zw, ze, zn;
        mg_z5(g, L0, i, j, inv, zc, zs, zw, ze, zn);
        const double e = E[o], w = E[o - 1] * rx[c], n = N[o], so = N[o - g.pitch] * ry[rr];
        const double qv = q[o], rv = r[o], wa = mg_area(dxu, cosu, dyu, i, j);
        if (!act[o]) continue;
        const double cc = (((-e) - w) - n) - so;
        const double az = cc * zc + e * ze + w * zw + n * zn + so * zs;
        az_out[o] = az;
        acc[0] += wa * zc * az;
        acc[1] += wa * zc * qv;
        acc[2] += wa * rv * zc;
    }
    mg_block_partials<3>(acc, part);
}

// Fold of a pass's partials by one workgroup (fixed order).
template <int NR>
__device__ __forceinline__ void mg_fold(const double* __restrict__ part, int nblocks,
                                        double (&v)[NR]) {
    __shared__ double sh[NR * 16];
    // the first four partials of each thread loaded at once (past nblocks:
    // partial 0, dropped), then any further ones; the sum runs in block order
    double x[4][NR];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int b = threadIdx.x + q * blockDim.x;
#pragma unroll
        for (int k = 0; k < NR; ++k) x[q][k] = part[(size_t)k * nblocks + (b < nblocks ? b : 0)];
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        v[k] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (threadIdx.x + q * blockDim.x < nblocks) v[k] += x[q][k];
    }
    for (int b = threadIdx.x + 4 * blockDim.x; b < nblocks; b += blockDim.x)
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

// beta = -<z,q>/<p,q>_old, <p,q> = <z,Az> + 2 beta <z,q> + beta^2 <p,q>_old,
// alpha = <r,z> / <p,q>.
__global__ void __launch_bounds__(1024) k_mg_fold_a(const double* __restrict__ part, int nblocks,
                                                    MgScal* S) {
    if (S->done) return;
    double v[3];
    mg_fold<3>(part, nblocks, v);
    if (threadIdx.x == 0) {
        const double beta = S->pq_old != 0.0 ? -v[1] / S->pq_old : 0.0;
        const double pq = v[0] + 2.0 * beta * v[1] + beta * beta * S->pq_old;
        const double alpha = v[2] / pq;
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
// V-cycle's next input w r s_r on level 0; |r|^2 (active) and |w r|^2.
__global__ void __launch_bounds__(256) k_mg_fcg_b1(
    Geom g, MgLev L0, const double* __restrict__ dxu, const double* __restrict__ cosu,
    const double* __restrict__ dyu, const uint8_t* __restrict__ act, double* __restrict__ p,
    double* __restrict__ q, double* __restrict__ r, double* __restrict__ x,
    const double* __restrict__ az_in, const double* __restrict__ zscale,
    const MgScal* __restrict__ S, double* __restrict__ part) {
    if (S->done) return;
    const double inv = 1.0 / *zscale;
    const double beta = S->beta, alpha = S->alpha, sr = S->s_r;
    double acc[2] = {0, 0};
    // row segments of 256 points, each XCD a contiguous range of rows
    const int segs = (g.nx + 255) / 256;
    int first, stride, end;
    range_schedule(g.ny * segs, first, stride, end);
    for (int it = first; it < end; it += stride) {
        const int j = it / segs, i = (it - j * segs) * 256 + threadIdx.x;
        if (i >= g.nx) continue;
        const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
        const int ql = (j + 1) * L0.pitch + i + MG_XOFF;
        // every load in bounds, issued before the active test
        const float yv = L0.y[ql];
        const double pv = p[o], qv = q[o], xv = x[o], rv = r[o], azv = az_in[o];
        const double w = mg_area(dxu, cosu, dyu, i, j);
        if (!act[o]) {
            L0.b[ql] = 0.0f;
            continue;
        }
        const double z = (double)yv * inv;
        const double pn = z + beta * pv;
        const double qn = azv + beta * qv;
        p[o] = pn;
        q[o] = qn;
        x[o] = xv + alpha * pn;
        const double rn = rv - alpha * qn;
        r[o] = rn;
        L0.b[ql] = (float)(w * rn * sr);
        acc[0] += rn * rn;
        acc[1] += (w * rn) * (w * rn);
    }
    mg_block_partials<2>(acc, part);
}

// After pass B: the iteration count, convergence on the recursive residual,
// the scale of the next V-cycle input.
__global__ void __launch_bounds__(1024) k_mg_fold_b(const double* __restrict__ part, int nblocks,
                                                    int nact, MgScal* S, double* zscale) {
    const int done = S->done;   // tested once the fold's loads are in flight
    double v[2];
    mg_fold<2>(part, nblocks, v);
    if (done) return;
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


// ---- the outer passes on pairs of points -----------------------------------
// The same two passes with each thread taking the points (i, i+1), i even,
// of a row segment through 16-byte loads and stores (used when nx is even,
// so every row of the solver layout, pitch nx + 4, starts 16-byte aligned and
// no row ends in a single point).  Per point the arithmetic and its order are
// those of k_mg_fcg_a1 / k_mg_fcg_b1.  Pass A runs on MG_PASS_A_THREADS-thread
// workgroups, two per CU, so that it leaves few partials: pass B folds them
// itself instead of a one-workgroup fold kernel between the passes.
#ifndef PSI_MG_PASS_A_THREADS
#define PSI_MG_PASS_A_THREADS 1024
#endif
constexpr int MG_PASS_A_THREADS = PSI_MG_PASS_A_THREADS;
struct MgZ2 {
    double c0, c1;   // z at (i, j), (i+1, j)
    double w0, e1;   // z west of i, east of i+1
    double s0, s1, n0, n1;
};

// The V-cycle output around the pair (i, i+1) of row j, as stored (FP32).
struct MgY2 {
    float2 c, s, n;   // rows j, j-1, j+1
    float w, e;       // west of i, east of i+1
};

__device__ __forceinline__ MgY2 mg_y_pair(const Geom& g, const MgLev& L0, int i, int j) {
    const int P = L0.pitch, q = (j + 1) * P + i + MG_XOFF;   // even: 8-byte aligned pairs
    int dw = -1, de = 2;
    if (g.cyclic) {
        if (i == 0) dw = g.nx - 1;
        if (i + 1 == g.nx - 1) de = 2 - g.nx;
    }
    MgY2 y;
    y.c = *reinterpret_cast<const float2*>(L0.y + q);
    y.s = *reinterpret_cast<const float2*>(L0.y + q - P);
    y.n = *reinterpret_cast<const float2*>(L0.y + q + P);
    y.w = L0.y[q + dw];
    y.e = L0.y[q + de];
    return y;
}

// z = y / s_r with the values outside the interior replaced by zero.
__device__ __forceinline__ MgZ2 mg_z_pair(const Geom& g, const MgY2& y, int i, int j, double inv) {
    const bool hs = j > 0, hn = j < g.ny - 1;
    MgZ2 z;
    z.c0 = (double)y.c.x * inv;
    z.c1 = (double)y.c.y * inv;
    z.s0 = hs ? (double)y.s.x * inv : 0.0;
    z.s1 = hs ? (double)y.s.y * inv : 0.0;
    z.n0 = hn ? (double)y.n.x * inv : 0.0;
    z.n1 = hn ? (double)y.n.y * inv : 0.0;
    z.w0 = (g.cyclic || i > 0) ? (double)y.w * inv : 0.0;
    z.e1 = (g.cyclic || i + 2 < g.nx) ? (double)y.e * inv : 0.0;
    return z;
}

#ifndef PSI_MG_PASS_A_WAVES
#define PSI_MG_PASS_A_WAVES 8
#endif
__global__ void __launch_bounds__(MG_PASS_A_THREADS)
__attribute__((amdgpu_waves_per_eu(1, PSI_MG_PASS_A_WAVES))) k_mg_fcg_a(
    Geom g, MgLev L0, const double* __restrict__ E, const double* __restrict__ N,
    const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ dxu, const double* __restrict__ cosu,
    const double* __restrict__ dyu, const uint8_t* __restrict__ act,
    const double* __restrict__ q, const double* __restrict__ r, double* __restrict__ az_out,
    const double* __restrict__ zscale, const MgScal* __restrict__ S,
    double* __restrict__ part) {
    // tested once the first row segment's loads are in flight; every thread
    // of a finished solve returns before the partials' barrier
    const int done = S->done;
    const double inv = 1.0 / *zscale;   // the scale of the V-cycle's input
    double acc[3] = {0, 0, 0};
    // row segments of 2 MG_PASS_A_THREADS points, each XCD a contiguous range of rows
    const int segs = (g.nx + 2 * MG_PASS_A_THREADS - 1) / (2 * MG_PASS_A_THREADS);
    int first, stride, end;
    range_schedule(g.ny * segs, first, stride, end);
    for (int it = first; it < end; it += stride) {
        const int j = it / segs, i = (it - j * segs) * (2 * MG_PASS_A_THREADS) + 2 * threadIdx.x;
        if (i >= g.nx) continue;
        const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
        const int c = i + 2, rr = j + 2;
        // every load in bounds, issued before the active test (pinned there)
        MgY2 y = mg_y_pair(g, L0, i, j);
        double2 e2 = *reinterpret_cast<const double2*>(E + o);
        double2 n2 = *reinterpret_cast<const double2*>(N + o);
        double2 s2 = *reinterpret_cast<const double2*>(N + o - g.pitch);
        double2 q2 = *reinterpret_cast<const double2*>(q + o);
        double2 r2 = *reinterpret_cast<const double2*>(r + o);
        double em = E[o - 1], rx0 = rx[c], rx1 = rx[c + 1];
        const double ryv = ry[rr];
        double dx0 = dxu[c], dx1 = dxu[c + 1];
        const double cr = cosu[rr], dr = dyu[rr];
        const uint16_t a2 = *reinterpret_cast<const uint16_t*>(act + o);
        __builtin_amdgcn_sched_barrier(0);   // every load above, in flight together
        if (done) return;
        pin_loaded(e2), pin_loaded(n2), pin_loaded(s2), pin_loaded(q2), pin_loaded(r2);
        pin_loaded(em), pin_loaded(rx0), pin_loaded(rx1), pin_loaded(dx0), pin_loaded(dx1);
        pin_loaded(y.c), pin_loaded(y.s), pin_loaded(y.n), pin_loaded(y.w), pin_loaded(y.e);
        if (!a2) continue;
        const MgZ2 z = mg_z_pair(g, y, i, j, inv);
        const double wa0 = dx0 * cr * dr, wa1 = dx1 * cr * dr;
        double2 az2 = make_double2(0.0, 0.0);
        if (a2 & 0xff) {
            const double e = e2.x, w = em * rx0, n = n2.x, so = s2.x * ryv;
            const double cc = (((-e) - w) - n) - so;
            const double az = cc * z.c0 + e * z.c1 + w * z.w0 + n * z.n0 + so * z.s0;
            az2.x = az;
            acc[0] += wa0 * z.c0 * az;
            acc[1] += wa0 * z.c0 * q2.x;
            acc[2] += wa0 * r2.x * z.c0;
        }
        if (a2 >> 8) {
            const double e = e2.y, w = e2.x * rx1, n = n2.y, so = s2.y * ryv;
            const double cc = (((-e) - w) - n) - so;
            const double az = cc * z.c1 + e * z.e1 + w * z.c0 + n * z.n1 + so * z.s1;
            az2.y = az;
            acc[0] += wa1 * z.c1 * az;
            acc[1] += wa1 * z.c1 * q2.y;
            acc[2] += wa1 * r2.y * z.c1;
        }
        *reinterpret_cast<double2*>(az_out + o) = az2;
    }
    if (done) return;
    mg_block_partials<3, MG_PASS_A_THREADS / 64>(acc, part);
}

// Pass B's operands of the points (i, i+1) of a row.
struct MgB2 {
    float2 y;
    double2 p, q, x, r, az;
    double w0, w1;
    uint16_t a;
};

// Pass B's x and p are read and written by pass B alone, once per
// iteration: with non-temporal loads and stores they do not displace, in the
// 256 MB Infinity Cache, what the V-cycle and pass A read again within the
// iteration (the operator's FP64 couplings, q, r, the level vectors).  At
// 2048^2 the iteration went from 233.7 to 225.1 us (pass A 45.9 -> 42.2,
// finest post-smoothing 31.4 -> 28.5 us); x alone gave 230.0.
// PSI_MG_B_NT: 1 = x and p non-temporal, 2 = x alone, 0 = neither.
#ifndef PSI_MG_B_NT
#define PSI_MG_B_NT 1
#endif
__device__ __forceinline__ double2 mg_ld2(const double* a, bool nt) {
    if (nt) {
        const pin_d2 t = __builtin_nontemporal_load(reinterpret_cast<const pin_d2*>(a));
        return make_double2(t.x, t.y);
    }
    return *reinterpret_cast<const double2*>(a);
}
__device__ __forceinline__ void mg_st2(double* a, double2 v, bool nt) {
    if (nt) {
        const pin_d2 t = {v.x, v.y};
        __builtin_nontemporal_store(t, reinterpret_cast<pin_d2*>(a));
    } else {
        *reinterpret_cast<double2*>(a) = v;
    }
}

__device__ __forceinline__ void mg_b2_load(const Geom& g, const MgLev& L0,
                                           const double* __restrict__ dxu,
                                           const double* __restrict__ cosu,
                                           const double* __restrict__ dyu,
                                           const uint8_t* __restrict__ act, const double* p,
                                           const double* q, const double* r, const double* x,
                                           const double* __restrict__ az_in, int j, int i,
                                           MgB2& v) {
    const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
    const int ql = (j + 1) * L0.pitch + i + MG_XOFF;
    const int c = i + 2, rr = j + 2;
    v.y = *reinterpret_cast<const float2*>(L0.y + ql);
    v.p = mg_ld2(p + o, PSI_MG_B_NT == 1);
    v.q = *reinterpret_cast<const double2*>(q + o);
    v.x = mg_ld2(x + o, PSI_MG_B_NT >= 1);
    v.r = *reinterpret_cast<const double2*>(r + o);
    v.az = *reinterpret_cast<const double2*>(az_in + o);
    v.w0 = dxu[c] * cosu[rr] * dyu[rr];
    v.w1 = dxu[c + 1] * cosu[rr] * dyu[rr];
    v.a = *reinterpret_cast<const uint16_t*>(act + o);
}

// FOLD: the scalars come from pass A's partials (ga per slot), folded by
// every workgroup in the same fixed order, instead of from a separate fold
// kernel; workgroup 0 records them in S for the fold after pass B.  Else they
// are read from S.  (Holding the first row segment's operands in flight
// across the fold needs 72 VGPRs, 7 waves per SIMD, and spills at 8.)
template <bool FOLD>
__global__ void __launch_bounds__(256) k_mg_fcg_b(
    Geom g, MgLev L0, const double* __restrict__ dxu, const double* __restrict__ cosu,
    const double* __restrict__ dyu, const uint8_t* __restrict__ act, double* __restrict__ p,
    double* __restrict__ q, double* __restrict__ r, double* __restrict__ x,
    const double* __restrict__ az_in, const double* __restrict__ zscale, MgScal* S,
    const double* __restrict__ part_a, int ga, double* __restrict__ part) {
    __shared__ double sf[3 * 4];
    // tested after the fold (with FOLD: its loads need nothing from S) and
    // once the first row segment's loads are in flight; every thread of a
    // finished solve returns before the partials' barrier
    const int done = S->done;
    const double inv = 1.0 / *zscale;
    // row segments of 512 points, each XCD a contiguous range of rows
    const int segs = (g.nx + 511) / 512;
    int first, stride, end;
    range_schedule(g.ny * segs, first, stride, end);
    double beta, alpha, sr;
    if (FOLD) {
        // the first four partials per thread and slot loaded at once (past ga:
        // partial 0, dropped), any further ones after; block order throughout
        double v[3] = {0.0, 0.0, 0.0};
        double pa[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = threadIdx.x + 256 * q;
#pragma unroll
            for (int k = 0; k < 3; ++k) pa[q][k] = part_a[(size_t)k * ga + (b < ga ? b : 0)];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (threadIdx.x + 256 * q < ga)
#pragma unroll
                for (int k = 0; k < 3; ++k) v[k] += pa[q][k];
        for (int b = threadIdx.x + 1024; b < ga; b += 256)
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] += part_a[(size_t)k * ga + b];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double t = v[k];
            for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
            if (lane == 0) sf[k * 4 + wid] = t;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = ((sf[k * 4] + sf[k * 4 + 1]) + sf[k * 4 + 2]) + sf[k * 4 + 3];
        // beta = -<z,q>/<p,q>_old, <p,q> = <z,Az> + 2 beta <z,q> + beta^2 <p,q>_old,
        // alpha = <r,z> / <p,q>
        if (done) return;
        const double pq_old = S->pq_old;
        beta = pq_old != 0.0 ? -v[1] / pq_old : 0.0;
        const double pq = v[0] + 2.0 * beta * v[1] + beta * beta * pq_old;
        alpha = v[2] / pq;
        sr = S->s_pending;
        const bool bad = !(pq != 0.0) || alpha != alpha;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->beta = beta;
            S->pq = pq;
            S->alpha = alpha;
            S->s_r = sr;
            if (bad) {
                S->status = ST_NAN;
                S->done = 1;
            }
        }
        if (bad) return;
    } else {
        beta = S->beta;
        alpha = S->alpha;
        sr = S->s_r;
    }
    double acc[2] = {0, 0};
    for (int it = first; it < end; it += stride) {
        const int j = it / segs, i = (it - j * segs) * 512 + 2 * threadIdx.x;
        if (i < g.nx) {
            MgB2 cur;
            mg_b2_load(g, L0, dxu, cosu, dyu, act, p, q, r, x, az_in, j, i, cur);
            __builtin_amdgcn_sched_barrier(0);   // every load above, in flight together
            pin_loaded(cur.y), pin_loaded(cur.p), pin_loaded(cur.q), pin_loaded(cur.x);
            pin_loaded(cur.r), pin_loaded(cur.az), pin_loaded(cur.w0), pin_loaded(cur.w1);
            if (done) return;
            const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
            const int ql = (j + 1) * L0.pitch + i + MG_XOFF;
            float2 b2 = make_float2(0.0f, 0.0f);
            if (cur.a) {
                double2 pn = cur.p, qn = cur.q, xn = cur.x, rn = cur.r;
                if (cur.a & 0xff) {
                    const double z = (double)cur.y.x * inv;
                    pn.x = z + beta * cur.p.x;
                    qn.x = cur.az.x + beta * cur.q.x;
                    xn.x = cur.x.x + alpha * pn.x;
                    rn.x = cur.r.x - alpha * qn.x;
                    b2.x = (float)(cur.w0 * rn.x * sr);
                    acc[0] += rn.x * rn.x;
                    acc[1] += (cur.w0 * rn.x) * (cur.w0 * rn.x);
                }
                if (cur.a >> 8) {
                    const double z = (double)cur.y.y * inv;
                    pn.y = z + beta * cur.p.y;
                    qn.y = cur.az.y + beta * cur.q.y;
                    xn.y = cur.x.y + alpha * pn.y;
                    rn.y = cur.r.y - alpha * qn.y;
                    b2.y = (float)(cur.w1 * rn.y * sr);
                    acc[0] += rn.y * rn.y;
                    acc[1] += (cur.w1 * rn.y) * (cur.w1 * rn.y);
                }
                mg_st2(p + o, pn, PSI_MG_B_NT == 1);
                *reinterpret_cast<double2*>(q + o) = qn;
                mg_st2(x + o, xn, PSI_MG_B_NT >= 1);
                *reinterpret_cast<double2*>(r + o) = rn;
            }
            *reinterpret_cast<float2*>(L0.b + ql) = b2;
        }
    }
    if (done) return;
    mg_block_partials<2>(acc, part);
}

// The outer passes for a grid: on pairs of points when nx is even.
constexpr int MG_PASS_A_THREADS_DECL = 0;
// Pass B after a separate fold of pass A's partials (the scalars read from S).
template <class... A>
void mg_launch_pass_b(const Geom& g, unsigned nblk, hipStream_t s, const MgLev& L0,
                      const double* dxu, const double* cosu, const double* dyu,
                      const uint8_t* act, double* p, double* q, double* r, double* x,
                      const double* az, const double* zscale, MgScal* S, double* part) {
    if (g.nx & 1)
        hipLaunchKernelGGL(k_mg_fcg_b1, dim3(nblk), dim3(256), 0, s, g, L0, dxu, cosu, dyu, act, p,
                           q, r, x, az, zscale, (const MgScal*)S, part);
    else
        hipLaunchKernelGGL(k_mg_fcg_b<false>, dim3(nblk), dim3(256), 0, s, g, L0, dxu, cosu, dyu,
                           act, p, q, r, x, az, zscale, S, (const double*)nullptr, 0, part);
}

// Start of an outer leg, on the device, from the FP64 residual's norms res =
// {|r|^2, |b|^2, max|r|, count} (k_residual_fold): the state the passes
// test, the end tests of that residual (converged, NaN, iteration limit)
// and the V-cycle input scale s = 1/(w_mean rms(r)) in s_r, s_pending and
// zscale.  first: the reference norm bb is |b|^2 (|r|^2 when b = 0); later
// legs keep the first leg's.
__global__ void k_mg_init(const double* __restrict__ res, MgScal* S, double* zscale, int first,
                          int n, int max_it, double rtarget2, double w_mean) {
    if (threadIdx.x != 0) return;
    const double rr = res[0];
    MgScal v = {};
    v.bb = first ? (res[1] > 0.0 ? res[1] : rr) : S->bb;
    v.rr = rr;
    v.rtarget2 = rtarget2;
    v.n = n;
    v.max_it = max_it;
    v.status = ST_RUNNING;
    v.s_r = v.s_pending = 1.0 / (w_mean * sqrt(rr / (res[3] > 1.0 ? res[3] : 1.0)));
    if (rr <= rtarget2 * v.bb) {
        v.status = ST_CONVERGED;
        v.done = 1;
        v.confirmed = 1;
    } else if (rr != rr) {
        v.status = ST_NAN;
        v.done = 1;
    } else if (n >= max_it) {
        v.status = ST_MAXIT;
        v.done = 1;
    }
    *S = v;
    *zscale = v.s_r;
}

// After a leg's replays: when the recursive residual met the target, the
// FP64 true residual of x (res, k_residual_fold) confirms it or asks for a
// new leg; nothing while the leg runs or when it ended otherwise.
__global__ void k_mg_confirm(const double* __restrict__ res, MgScal* S) {
    if (threadIdx.x != 0) return;
    if (!S->done || S->status != ST_CONVERGED || S->confirmed != 0) return;
    S->confirmed = res[0] <= S->rtarget2 * S->bb ? 1 : -1;   // NaN: a new leg reports it
}

// Leg start from the residual r (after k_mg_init): p = q = 0 and the V-cycle
// input w r s on level 0 with s = *zscale; nothing when k_mg_init ended the
// solve.
__global__ void __launch_bounds__(256) k_mg_start(Geom g, MgLev L0, const double* __restrict__ dxu,
                                                  const double* __restrict__ cosu,
                                                  const double* __restrict__ dyu,
                                                  const uint8_t* __restrict__ act,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ p, double* __restrict__ q,
                                                  const double* __restrict__ zscale,
                                                  const MgScal* __restrict__ S) {
    if (S->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int i = t % g.nx, j = t / g.nx;
    const size_t o = (size_t)(j + 2) * g.pitch + (i + 2);
    const int ql = (j + 1) * L0.pitch + i + MG_XOFF;
    const double s = *zscale;
    p[o] = 0.0;
    q[o] = 0.0;
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
                cells.push_back((j + 1) * mg_pitch(L.nx) + i + MG_XOFF);   // level layout index
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
    float* Ainv = nullptr;            // ncoarse rows of dense_ld floats
    int dense_ld = 0;
    int* cells = nullptr;
    int nact0 = 0;
    double w_mean = 1.0;             // mean psi-cell area over active points
    double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *az = nullptr;
    MgScal* S = nullptr;
    MgScal* S_time = nullptr;        // scalars of psi_mg_kernel_times (never done)
    hipEvent_t t0 = nullptr, t1 = nullptr;
    MgScal* h_S = nullptr;           // pinned
    double* h_res = nullptr;         // pinned: a leg residual's norms
    double* zscale = nullptr;
    double* part = nullptr;
    int nblk = 0;                    // workgroups of the outer passes
    double* part_a = nullptr;        // pass A's partials when pass B folds them
    int nblk_a = 0;                  // ... and pass A's workgroups then
    hipStream_t gs = nullptr;        // stream the graph is replayed on
    hipEvent_t e_in = nullptr, e_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_K = 0, graph_Kc = 0;
    double graph_ratio = 0.0;
    // the same iteration with HIP events around its four heaviest kernels
    // (psi_opts.time_kernels): before/after pre level 0, before/after post
    // level 0, after pass A, before/after pass B
    hipEvent_t kev[7] = {};
    hipGraph_t graph_t = nullptr;
    hipGraphExec_t exec_t = nullptr;
    int graph_t_K = 0, graph_t_Kc = 0;
    double graph_t_ratio = 0.0;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    std::vector<int> level_nx, level_ny, level_nact;
    double cscale = MG_SCALE_PSI;    // coarse-correction weight of this operator
};

// Chebyshev scalars of K steps on [2/ratio, 2], and the coarse-correction
// weight.
inline MgCheb mg_cheb(int K, double ratio, double cscale) {
    MgCheb ch{};
    ch.cscale = (float)cscale;
    const double hi = 2.0, lo = hi / ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    ch.c0 = (float)(1.0 / theta);
    for (int k = 1; k < K && k < MG_MAXKC; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        ch.c1[k] = (float)(rn * rho);
        ch.c2[k] = (float)(2.0 * rn / delta);
        rho = rn;
    }
    return ch;
}

}  // namespace psi
