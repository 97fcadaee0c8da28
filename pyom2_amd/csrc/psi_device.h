// psi_device.h -- device-side building blocks shared by the solver kernels:
// tile geometry, XCD-aware tile mapping, the fused 5-point stencil on LDS
// frames, deterministic block reductions and the "last block" handoff.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psi {

// Output tile per workgroup and its LDS frame (2-point halo each side).
#ifndef PSI_TILE_Y
#define PSI_TILE_Y 16
#endif
constexpr int TX = 64;
constexpr int TY = PSI_TILE_Y;
constexpr int HX = TX + 4;
constexpr int HY = TY + 4;
constexpr int NT = 256;            // threads per workgroup (4 wave64)
constexpr int NW = NT / 64;
constexpr int NRED = 6;            // reduction slots per sweep

struct Geom {
    int nx, ny, pitch, cyclic;     // pitch = nx + 4 (elements)
    int tiles_x, tiles_y, ntiles;
};

// Solver state living in device memory; updated by the last workgroup of
// each sweep, read by every workgroup of the next one.
// Solver state in device memory, double-buffered by launch: launch L reads
// state (L-1) & 1 and the partials (L-1) & 1 of the previous launch, and
// workgroup 0 writes state L & 1.
struct CgState {
    double alpha, beta, gamma;     // next sweep's step, direction update, current (r, Z r)
    double betak_min, step1, est_error;
    double rr, rr0, bb, zmax0, pmax;
    double eps, rtarget2, bb_fixed;
    double rr_ref, red2;           // residual replacement: end the leg once
                                   // |r|^2 <= red2 * rr_ref (0: never)
    int n, status, done, max_it;
    int criterion, degenerate, phase, pad;
};

enum : int { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXIT = 2, ST_DIVERGED = 3, ST_NAN = 4, ST_LEG = 5 };
enum : int { PHASE_SETUP = 0, PHASE_SWEEP = 1, PHASE_REFRESH = 2 };
enum : int { CRIT_REFERENCE = 0, CRIT_RELRES = 1 };

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Give each XCD a contiguous range of tiles so neighbouring
// tiles, which share halo rows, hit the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int xcd = b & 7, k = b >> 3;
    const int q = n >> 3, rem = n & 7;
    return xcd * q + (xcd < rem ? xcd : rem) + k;
}

// Column of the stored array backing array column c of a frame (c may fall
// outside [0, nx+3]).  Cyclic in x: Fortran i = c-1 wraps into [1, nx].
__device__ __forceinline__ int map_col(int c, const Geom& g) {
    if (g.cyclic) {
        int i = c - 1 - 1;                 // 0-based interior index
        i %= g.nx;
        if (i < 0) i += g.nx;
        return i + 2;
    }
    return c;
}

__device__ __forceinline__ bool in_array(int r, int c, const Geom& g) {
    return r >= 0 && r < g.ny + 4 && c >= 0 && c < g.nx + 4;
}

// 5-point operator at LDS frame index `idx` (row stride HX).  Coefficients:
// E and N are stored; W(i,j) = E(i-1,j)*rx(i), S(i,j) = N(i,j-1)*ry(j);
// C = -E - W - N - S in the reference's order (solve_streamfct.f90:137-145).
// Terms are summed S, W, C, E, N: the order apply_op
// (solve_pressure.f90:264-273) visits the non-zero planes.
// S is the frame's row stride.
template <class T, int S = HX>
__device__ __forceinline__ T stencil5(const T* s, const T* E, const T* N, int idx, T rxv, T ryv,
                                      T& cc) {
    const T e = E[idx], w = E[idx - 1] * rxv, n = N[idx], so = N[idx - S] * ryv;
    cc = (((-e) - w) - n) - so;
    T r = so * s[idx - S];
    r = r + w * s[idx - 1];
    r = r + cc * s[idx];
    r = r + e * s[idx + 1];
    r = r + n * s[idx + S];
    return r;
}

// Reduce NR doubles across the workgroup; slot k is a max if bit k of
// maxmask is set, else a sum.  Fixed butterfly + fixed wave order: the
// result (valid in thread 0) is bit-reproducible run to run.
template <int NR>
__device__ __forceinline__ void block_reduce(double (&v)[NR], unsigned maxmask, double* sh) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_xor(x, off, 64);
            x = ((maxmask >> k) & 1u) ? fmax(x, y) : x + y;
        }
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) sh[k * NW + wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            double a = sh[k * NW];
            for (int w = 1; w < NW; ++w)
                a = ((maxmask >> k) & 1u) ? fmax(a, sh[k * NW + w]) : a + sh[k * NW + w];
            v[k] = a;
        }
    }
    __syncthreads();
}

// Publish this workgroup's partials (slot-major: part[k*nblocks + b]) and
// report whether it is the last workgroup of the grid to arrive.
template <int NR>
__device__ __forceinline__ bool publish_partials(const double (&v)[NR], double* part,
                                                 unsigned* counter, int b, int nblocks) {
    __shared__ int am_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) part[(size_t)k * nblocks + b] = v[k];
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        am_last = (prev == (unsigned)(nblocks - 1));
    }
    __syncthreads();
    if (am_last) __threadfence();
    return am_last != 0;
}

// Last workgroup: fold all partials in a fixed order.
template <int NR>
__device__ __forceinline__ void fold_partials(double (&v)[NR], unsigned maxmask, const double* part,
                                              int nblocks, double* sh) {
#pragma unroll
    for (int k = 0; k < NR; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += NT) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double x = __builtin_nontemporal_load(part + (size_t)k * nblocks + b);
            v[k] = ((maxmask >> k) & 1u) ? fmax(v[k], x) : v[k] + x;
        }
    }
    block_reduce<NR>(v, maxmask, sh);
}

}  // namespace psi
