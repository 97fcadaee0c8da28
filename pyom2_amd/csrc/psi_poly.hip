// psi_poly.hip -- FP32 PCG sweeps with a Chebyshev polynomial preconditioner
// (included into psi_capi.hip after psi_kernels.hip).
//
// No counterpart in the reference, which preconditions congrad_streamfunction
// with Jacobi only (make_inv_sfc / inv_op_sfc, solve_streamfct.f90:364-422).
// Here the preconditioner is M = q(ZA) Z, Z the same Jacobi inverse (1/C on
// active points, 0 on land and island perimeters) and q the degree-K
// polynomial of the Chebyshev iteration for ZA e = Z r on the interval
// [a, b] (b = 2 bounds the spectrum of ZA: it is similar to a symmetric
// matrix, W A being symmetric, and Gershgorin puts it in [0, 2]).  M is
// linear and fixed, so plain CG applies.  PCG iterations drop about
// (K+1)-fold while the stencil applications per solve stay the same.  The
// sweep is bound by its fixed cost per launch rather than by bandwidth, so
// the K extra applications per iteration, done in registers on a frame whose
// halo grows to 2+K, cost far less than the launches they save.
//
// Per iteration the sweep keeps the single-reduction recurrence of the
// Jacobi sweep: p = z + beta p; r -= alpha A p; z = M r; x += alpha p;
// w = A z, s = A p; partials (r,z), (z,w), max|p|, |r|^2, (z,s)+(p,w), (p,s).
// Stored vectors (FP32): x, r, z, p (r and z are no longer proportional
// point by point, so both are kept).

namespace psi {

constexpr int PMAXK = 6;

// Chebyshev iteration scalars: d_0 = Zr / theta; d_j = c1[j-1] d_{j-1} +
// c2[j-1] res_{j-1}; res_j = res_{j-1} - ZA d_j; z = sum_j d_j.
struct ChebCoef {
    float inv_theta;
    float c1[PMAXK], c2[PMAXK];
};

enum : int { PM_SETUP = 0, PM_SWEEP = 1, PM_REFRESH = 2 };

// Frame of one tile: 64 columns (one per lane) x 4 * RPT rows, wave w of the
// workgroup holding rows [w RPT, (w+1) RPT) of every column in registers.
// The own points are the frame minus a margin R = K + 2.
template <int K, int RPT>
struct PolyShape {
    static constexpr int R = K + 2, FW = 64, FH = 4 * RPT;
    static constexpr int TXo = FW - 2 * R, TYo = FH - 2 * R;
};

// Rows per thread for degree K: the own tile is (64 - 2R) x (4 RPT - 2R).
// Taller frames re-read fewer margin cells but hold more registers: at
// degree 4 the single-grid kernel takes 164 VGPRs with 8 rows (3 waves per
// SIMD, 2% faster than 7 rows at 2048^2) and 176 with 9 (2 waves, 15% slower).
#ifndef PSI_POLY_RPT
#define PSI_POLY_RPT 0            // > 0: rows per thread for every degree
#endif
constexpr int poly_rows(int k) { return PSI_POLY_RPT > 0 ? PSI_POLY_RPT : (k <= 2 ? 6 : 8); }
template <int K>
struct PolyRows {
    static constexpr int value = poly_rows(K);
};
inline int PolyRowsHost(int k) { return poly_rows(k); }

// Neighbouring columns of the frame are neighbouring lanes, moved by DPP
// (wave_shr:1 = 0x138: lane l gets lane l-1; wave_shl:1 = 0x130: lane l gets
// lane l+1) rather than through the LDS crossbar.  Lanes 0 and 63 (frame
// edges, never inside a ring that is computed) keep their own value.
__device__ __forceinline__ float from_left(float v) {
    const int b = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float v) {
    const int b = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0x130, 0xf, 0xf, false));
}

// Rows above and below a wave's band, exchanged through LDS:
// xch[buf][vec][0 = first row | 1 = last row][wave][lane].
struct Exchange {
    float x[2][2][2][NW][64];
};

// The 5-point operator on a strip held in registers, summed S, W, C, E, N
// (the order of stencil5); up / dn are the rows above and below the strip.
// The west coupling is the west neighbour's east one times this column's
// metric factor rxv, formed here from cE rather than kept in registers.
template <int RPT>
__device__ __forceinline__ void apply_strip(const float (&v)[RPT], float up, float dn,
                                            const float (&cS)[RPT], float rxv,
                                            const float (&cE)[RPT], const float (&cN)[RPT],
                                            float (&out)[RPT]) {
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const float vu = i ? v[i - 1] : up, vd = (i < RPT - 1) ? v[i + 1] : dn;
        const float vl = from_left(v[i]), vr = from_right(v[i]);
        const float cW = from_left(cE[i]) * rxv;
        const float cc = (((-cE[i]) - cW) - cN[i]) - cS[i];
        float a = cS[i] * vu;
        a = a + cW * vl;
        a = a + cc * v[i];
        a = a + cE[i] * vr;
        a = a + cN[i] * vd;
        out[i] = a;
    }
}

// Publishes the first and last rows of up to two strips and returns the
// rows adjacent to this wave's band (one barrier; buffers alternate, so a
// buffer is rewritten only after the barrier that follows its last read).
template <int RPT, int NV>
__device__ __forceinline__ void exchange(Exchange& X, int& buf, const float (&a)[RPT],
                                         const float (&b)[RPT], float* up, float* dn) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    X.x[buf][0][0][w][lane] = a[0];
    X.x[buf][0][1][w][lane] = a[RPT - 1];
    if (NV > 1) {
        X.x[buf][1][0][w][lane] = b[0];
        X.x[buf][1][1][w][lane] = b[RPT - 1];
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        up[v] = w > 0 ? X.x[buf][v][1][w - 1][lane] : 0.f;
        dn[v] = w < NW - 1 ? X.x[buf][v][0][w + 1][lane] : 0.f;
    }
    buf ^= 1;
}

// Element idx of a plane whose base address is wave-uniform: the byte
// offset is formed in 32 bits, so the access takes a scalar base plus one
// 32-bit VGPR offset instead of a 64-bit address per lane (planes here stay
// far below 4 GiB).
template <typename T>
__device__ __forceinline__ T ld_at(const T* base, unsigned idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_at(T* base, unsigned idx, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T)) = v;
}

// Vectors of one polynomial-preconditioned launch over one grid.
struct PolyPtrs {
    const float *E, *N, *rx, *ry;
    const uint8_t* act;
    float* x;
    const float *z_in, *p_in, *r_in;
    float *z_out, *p_out, *r_out;
    const double* r64;
    double rscale;
};

// after n iterations z_n, r_n and p_{n-1} live in buffer n & 1; a sweep
// writes buffer (n+1) & 1, setup (n = 0) and refresh write buffer n & 1
template <int MODE>
__device__ __forceinline__ void poly_buffers(PolyPtrs& q, int n, float* zb0, float* zb1,
                                             float* pb0, float* pb1, float* rb0, float* rb1) {
    const int odd = n & 1;
    q.z_in = odd ? zb1 : zb0;
    q.p_in = odd ? pb1 : pb0;
    q.r_in = odd ? rb1 : rb0;
    q.z_out = (MODE == PM_SWEEP) ? (odd ? zb0 : zb1) : (odd ? zb1 : zb0);
    q.p_out = (MODE == PM_SWEEP) ? (odd ? pb0 : pb1) : (odd ? pb1 : pb0);
    q.r_out = (MODE == PM_SWEEP) ? (odd ? rb0 : rb1) : (odd ? rb1 : rb0);
}

// One tile of g (tiled (64 - 2R) x (4 RPT - 2R)).  MODE:
//  PM_SETUP   r = r64 * rscale; p = 0; x = 0; z = M r; partials for
//             cg_from_setup (slot 2 = max|z|, slot 4 unused: bb is fixed).
//  PM_SWEEP   one fused iteration with alpha, beta.
//  PM_REFRESH r = r64 * rscale, z = M r next to the kept p; partials for
//             cg_from_refresh.
template <int K, int MODE, int RPT>
__device__ __forceinline__ void poly_tile(const Geom& g, int tile, const PolyPtrs& q, float alpha,
                                          float beta, const ChebCoef& ch, Exchange& X, int& xbuf,
                                          double (&acc)[NRED]) {
    using PS = PolyShape<K, RPT>;
    constexpr int R = PS::R, FW = PS::FW, FH = PS::FH, TXo = PS::TXo, TYo = PS::TYo;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c0 = (tile % g.tiles_x) * TXo, r0 = (tile / g.tiles_x) * TYo;
    // frame row 0 / column 0 sit R - 2 rows / columns before the tile's
    // first own point (array row / column 2 + r0 / c0)
    const int fr0 = r0 + 2 - R, fc0 = c0 + 2 - R;
    const int col = map_col(fc0 + lane, g);
    const bool col_in = col >= 0 && col < g.nx + 4;
    const float rxv = col_in ? ld_at(q.rx, (unsigned)col) : 0.f;
    // ---- load the band: vectors, coefficients, own x -----------------------
    float P[RPT], Rr[RPT], cS[RPT], cE[RPT], cN[RPT], zi[RPT], xo[RPT];
    unsigned amask = 0u;
    float nprev;
    {
        const int r = fr0 + w * RPT - 1;
        nprev = (col_in && r >= 0 && r < g.ny + 4) ? ld_at(q.N, (unsigned)(r * g.pitch + col))
                                                   : 0.f;
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int gg = w * RPT + i, r = fr0 + gg;
        float pv = 0.f, rv = 0.f, e = 0.f, nn = 0.f, ryv = 0.f, xv = 0.f;
        uint8_t a = 0;
        if (col_in && r >= 0 && r < g.ny + 4) {
            const unsigned o = (unsigned)(r * g.pitch + col);
            if (MODE == PM_SWEEP) {
                pv = ld_at(q.z_in, o) + beta * ld_at(q.p_in, o);
                rv = ld_at(q.r_in, o);
            } else {
                if (MODE == PM_REFRESH) pv = ld_at(q.p_in, o);
                rv = (float)(ld_at(q.r64, o) * q.rscale);
            }
            e = ld_at(q.E, o);
            nn = ld_at(q.N, o);
            a = ld_at(q.act, o);
            ryv = ld_at(q.ry, (unsigned)r);
            if (MODE == PM_SWEEP && gg >= R && gg < FH - R && lane >= R && lane < FW - R)
                xv = ld_at((const float*)q.x, o);
        }
        P[i] = pv;
        Rr[i] = rv;
        cE[i] = e;
        cN[i] = nn;
        cS[i] = nprev * ryv;
        nprev = nn;
        xo[i] = xv;
        amask |= (a ? 1u : 0u) << i;
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const float cW = from_left(cE[i]) * rxv;
        const float cc = (((-cE[i]) - cW) - cN[i]) - cS[i];
        zi[i] = ((amask >> i) & 1u) ? 1.f / cc : 0.f;
    }
    float up[2], dn[2], tmp[RPT];
    if (MODE == PM_SWEEP) {
        // r_{k+1} = r_k - alpha A p_k (valid on radius K+1)
        exchange<RPT, 1>(X, xbuf, P, P, up, dn);
        apply_strip<RPT>(P, up[0], dn[0], cS, rxv, cE, cN, tmp);
#pragma unroll
        for (int i = 0; i < RPT; ++i)
            if ((amask >> i) & 1u) Rr[i] = Rr[i] - alpha * tmp[i];
    }
    // ---- z = q(ZA) Z r: Chebyshev iteration, valid on radius 1 ------------
    float D[RPT], S[RPT], Y[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const float v = zi[i] * Rr[i];
        S[i] = v;
        D[i] = v * ch.inv_theta;
        Y[i] = D[i];
    }
#pragma unroll
    for (int j = 1; j <= K; ++j) {
        exchange<RPT, 1>(X, xbuf, D, D, up, dn);
        apply_strip<RPT>(D, up[0], dn[0], cS, rxv, cE, cN, tmp);
        const float c1 = ch.c1[j - 1], c2 = ch.c2[j - 1];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            S[i] = S[i] - zi[i] * tmp[i];
            D[i] = c1 * D[i] + c2 * S[i];
            Y[i] = Y[i] + D[i];
        }
    }
    // ---- own points: w = A z, s = A p, outputs and partials ---------------
    float Wv[RPT], Sv[RPT];
    exchange<RPT, 2>(X, xbuf, Y, P, up, dn);
    apply_strip<RPT>(Y, up[0], dn[0], cS, rxv, cE, cN, Wv);
    apply_strip<RPT>(P, up[1], dn[1], cS, rxv, cE, cN, Sv);
    const int c = fc0 + lane;
    const bool own_col = lane >= R && lane < FW - R && c <= g.nx + 1;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int gg = w * RPT + i, r = fr0 + gg;
        if (own_col && gg >= R && gg < FH - R && r <= g.ny + 1) {
            const unsigned o = (unsigned)(r * g.pitch + c);
            const float zn = Y[i], pn = P[i], rn = Rr[i];
            if (MODE == PM_SWEEP) {
                st_at(q.x, o, xo[i] + alpha * pn);
                st_at(q.p_out, o, pn);
            }
            if (MODE == PM_SETUP) {
                st_at(q.x, o, 0.f);
                st_at(q.p_out, o, 0.f);
            }
            st_at(q.z_out, o, zn);
            st_at(q.r_out, o, rn);
            const double zd = (double)zn, pd = (double)pn, rd = (double)rn;
            const double wd = (double)Wv[i], sd = (double)Sv[i];
            if ((amask >> i) & 1u) {
                acc[0] += rd * zd;
                acc[3] += rd * rd;
            }
            acc[1] += zd * wd;
            acc[2] = fmax(acc[2], MODE == PM_SETUP ? fabs(zd) : fabs(pd));
            acc[4] += zd * sd + pd * wd;
            acc[5] += pd * sd;
        }
    }
}

// One launch of the polynomial-preconditioned FP32 PCG over the tiles of gp
// (the poly tiling of the grid); MODE as in poly_tile.  The sweep advances
// the state as k_cg_sweep does; the refresh handles it as k_cg_refresh.
template <int K, int MODE>
// Waves per SIMD the register allocation must allow.  Degree 4 takes 140
// VGPRs (3 waves); forcing 4 spills 12 and measured 15% slower at 2048^2.
#ifndef PSI_POLY_WAVES
#define PSI_POLY_WAVES 1
#endif
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PSI_POLY_WAVES))) void k_poly_pcg(
    Geom gp, ChebCoef ch, const float* __restrict__ E, const float* __restrict__ N,
    const float* __restrict__ rx, const float* __restrict__ ry, const uint8_t* __restrict__ act,
    float* __restrict__ x, float* zb0, float* zb1, float* pb0, float* pb1, float* rb0,
    float* rb1, const double* __restrict__ r64, double rscale,
    const double* __restrict__ true_res, double rr_thresh, int new_max_it,
    const CgState* __restrict__ st_in, CgState* __restrict__ st_out,
    const double* __restrict__ part_in, double* __restrict__ part_out) {
    __shared__ Exchange X;
    __shared__ double sRed[NRED * NW];
    __shared__ CgState sst;
    float alpha = 0.f, beta = 0.f;
    int n = 0;
    if (MODE == PM_SWEEP) {
        advance_state(st_in, st_out, part_in, PHASE_SWEEP, sst, sRed);
        if (sst.done) return;
        alpha = (float)sst.alpha;
        beta = (float)sst.beta;
        n = sst.n;
    } else if (MODE == PM_REFRESH) {
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
        n = st_in->n;
    }
    PolyPtrs q;
    q.E = E;
    q.N = N;
    q.rx = rx;
    q.ry = ry;
    q.act = act;
    q.x = x;
    q.r64 = r64;
    q.rscale = rscale;
    poly_buffers<MODE>(q, n, zb0, zb1, pb0, pb1, rb0, rb1);
    double acc[NRED] = {0, 0, 0, 0, 0, 0};
    int xbuf = 0;
    int t0, ts, te;
    tile_schedule(gp, t0, ts, te);
    for (int tile = t0; tile < te; tile += ts)
        poly_tile<K, MODE, PolyRows<K>::value>(gp, tile, q, alpha, beta, ch, X, xbuf, acc);
    block_reduce<NRED>(acc, 1u << 2, sRed);
    write_partials<NRED>(acc, part_out);
}

// Tiling of the polynomial kernels of degree K over the grid of g.
inline Geom poly_geom(const Geom& g, int K) {
    Geom gp = g;
    const int R = K + 2, rpt = PolyRowsHost(K);
    const int txo = 64 - 2 * R, tyo = 4 * rpt - 2 * R;
    gp.tiles_x = (g.nx + txo - 1) / txo;
    gp.tiles_y = (g.ny + tyo - 1) / tyo;
    gp.ntiles = gp.tiles_x * gp.tiles_y;
    return gp;
}

}  // namespace psi
