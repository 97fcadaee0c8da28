// psi_noalloc.hip -- psi_congrad_noalloc: the plain FP64 Jacobi-PCG solve of
// psi_congrad (congrad_streamfunction, solve_streamfct.f90:153-327) without a
// context and without allocating anything, for callers that cannot call
// hipMalloc (code under graph capture elsewhere in the program, serving loops,
// harnesses that forbid it).  No counterpart in the reference.
//
// Every byte of device memory beyond the caller's inputs and outputs -- the
// stored operator, the Jacobi active mask, the CG vectors, the reduction
// partials and the solver state -- lives in one file-scope device array,
// `scratch`, of NA_SCRATCH_BYTES per GPU, held until the process ends.  Host
// code never passes `scratch` itself to a kernel or a copy: its device address
// comes from hipGetSymbolAddress, once per device (the first call on that
// device, with the device current), and is kept per device.  The solve runs
// the same kernels as psi_congrad's plain path (the context's driver,
// solve_plain, on a context whose pointers point into scratch), so results are
// identical.  Calls that overlap on one device share the array and race.
//
// Included into psi_capi.hip's translation unit (it uses that file's kernels
// and drivers).

#include <chrono>
#include <mutex>

namespace {

// Limits: an interior of at most 512 x 512 points (any shape with
// (nx+4)(ny+4) <= 516^2 and nx, ny <= 2048), at most 2^18 island perimeter
// points in total.
constexpr size_t NA_MAX_PLANE = 516 * 516;
constexpr int NA_MAX_SIDE = 2048;
constexpr int NA_MAX_POINTS = 1 << 18;
constexpr int NA_MAX_PART = 4096;     // reduction records (ntiles + 1024 fits)
constexpr int NA_MAX_DEV = 64;

constexpr size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }

// byte layout of scratch
constexpr size_t NA_PLANE64 = na_align(NA_MAX_PLANE * 8);
constexpr size_t NA_PLANE8 = na_align(NA_MAX_PLANE);
constexpr size_t NA_METRIC = na_align((size_t)(NA_MAX_SIDE + 4) * 8);
constexpr size_t NA_OFF_E = 0;
constexpr size_t NA_OFF_N = NA_OFF_E + NA_PLANE64;
constexpr size_t NA_OFF_C = NA_OFF_N + NA_PLANE64;
constexpr size_t NA_OFF_U = NA_OFF_C + NA_PLANE64;           // u64[0], u64[1]
constexpr size_t NA_OFF_P = NA_OFF_U + 2 * NA_PLANE64;       // p64[0], p64[1]
constexpr size_t NA_OFF_R = NA_OFF_P + 2 * NA_PLANE64;       // r64
constexpr size_t NA_OFF_ACT = NA_OFF_R + NA_PLANE64;
constexpr size_t NA_OFF_DIR = NA_OFF_ACT + NA_PLANE8;
constexpr size_t NA_OFF_MET = NA_OFF_DIR + NA_PLANE8;        // dxt dxu dyt dyu cost cosu rx ry
constexpr size_t NA_OFF_BI = NA_OFF_MET + 8 * NA_METRIC;
constexpr size_t NA_OFF_BJ = NA_OFF_BI + na_align((size_t)NA_MAX_POINTS * 4);
constexpr size_t NA_OFF_PART = NA_OFF_BJ + na_align((size_t)NA_MAX_POINTS * 4);
constexpr size_t NA_OFF_PART2 = NA_OFF_PART + na_align((size_t)NRED * NA_MAX_PART * 8);
constexpr size_t NA_OFF_ST = NA_OFF_PART2 + na_align((size_t)2 * NRED * NA_MAX_PART * 8);
constexpr size_t NA_OFF_RES = NA_OFF_ST + na_align(2 * sizeof(CgState));
constexpr size_t NA_SCRATCH_BYTES = NA_OFF_RES + na_align(8 * 8);
static_assert(NA_SCRATCH_BYTES <= (size_t)20 << 20, "scratch is documented as at most 20 MiB");

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

namespace {

struct NaDevice {
    unsigned char* ptr = nullptr;   // device address of this device's scratch
    int G = 0, ncu = 0;             // resident workgroups of k_cg_sweep<double>, CUs
    // host copies of the solver state and scalars (pageable: the polls copy
    // synchronously)
    CgState h_st[4];
    double h_res[8];
};
NaDevice g_na[NA_MAX_DEV];
std::mutex g_na_mu;

// Set-up of one device, with that device current: scratch's address and the
// occupancy of the sweep kernel.
int na_device(int dev, NaDevice** out) {
    if (dev < 0 || dev >= NA_MAX_DEV) return fail(nullptr, PSI_BAD_ARG, "device index out of range%s");
    std::lock_guard<std::mutex> lk(g_na_mu);
    NaDevice& d = g_na[dev];
    if (!d.ptr) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, PSI_HIP_ERROR, "hipGetSymbolAddress(scratch): %s",
                        hipGetErrorString(e));
        int ncu = 0, occ = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_cg_sweep<double>, NT, 0) !=
                hipSuccess)
            return fail(nullptr, PSI_HIP_ERROR, "device query failed%s");
        d.ncu = std::max(ncu, 1);
        d.G = std::max(1, d.ncu * std::max(occ, 1));
        d.ptr = (unsigned char*)p;
    }
    *out = &d;
    return PSI_OK;
}

__global__ void k_na_zero(uint4* __restrict__ a, size_t n16) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16;
         k += (size_t)gridDim.x * blockDim.x)
        a[k] = make_uint4(0u, 0u, 0u, 0u);
}

// 1-D metric ratios of the stored operator (psi_ctx_create forms the same
// quotients on the host): rx(i) = dxu(i-1)/dxu(i), ry(j) = dyu(j-1) cosu(j-1) /
// (dyu(j) cosu(j)), 0 at index 0.
__global__ void k_na_ratios(int nxh, int nyh, const double* __restrict__ dxu,
                            const double* __restrict__ dyu, const double* __restrict__ cosu,
                            double* __restrict__ rx, double* __restrict__ ry) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nxh) rx[k] = k == 0 ? 0.0 : dxu[k - 1] / dxu[k];
    if (k < nyh) ry[k] = k == 0 ? 0.0 : dyu[k - 1] * cosu[k - 1] / (dyu[k] * cosu[k]);
}

}  // namespace

extern "C" int psi_congrad_noalloc(const psi_grid* grid, const psi_fields* f,
                                   const psi_islands* isl, const psi_opts* o, const double* forc,
                                   double* sol, psi_info* info, void* stream) {
    const auto t_start = std::chrono::steady_clock::now();
    if (!grid || !f || !f->hur || !f->hvr || !o || !forc || !sol || !info)
        return fail(nullptr, PSI_BAD_ARG, "null argument%s");
    if (grid->nx < 2 || grid->ny < 2 || !grid->dxt || !grid->dxu || !grid->dyt || !grid->dyu ||
        !grid->cost || !grid->cosu)
        return fail(nullptr, PSI_BAD_ARG, "bad grid%s");
    if (grid->nx > NA_MAX_SIDE || grid->ny > NA_MAX_SIDE ||
        (size_t)(grid->nx + 4) * (size_t)(grid->ny + 4) > NA_MAX_PLANE)
        return fail(nullptr, PSI_BAD_ARG,
                    "psi_congrad_noalloc: grid larger than its static scratch (at most "
                    "(nx+4)(ny+4) <= 516^2 and nx, ny <= 2048)%s");
    if (o->mode != PSI_MODE_PLAIN)
        return fail(nullptr, PSI_BAD_ARG, "psi_congrad_noalloc: mode must be PSI_MODE_PLAIN%s");
    if (o->max_iterations < 1) return fail(nullptr, PSI_BAD_ARG, "max_iterations < 1%s");
    int np = 0;
    if (isl && isl->nisle > 0) {
        if (!isl->offsets || !isl->bi || !isl->bj)
            return fail(nullptr, PSI_BAD_ARG, "island lists incomplete%s");
        np = isl->offsets[isl->nisle];
        if (np < 0 || np > NA_MAX_POINTS)
            return fail(nullptr, PSI_BAD_ARG,
                        "psi_congrad_noalloc: more island perimeter points than its static "
                        "scratch holds (2^18)%s");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, PSI_HIP_ERROR, "hipGetDevice failed%s");
    NaDevice* d = nullptr;
    int rc = na_device(dev, &d);
    if (rc) return rc;

    // a context on the stack whose device pointers all point into scratch
    psi_ctx c;
    c.device = dev;
    c.stream = (hipStream_t)stream;
    Geom& g = c.g;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.pitch = grid->nx + 4;
    g.cyclic = grid->cyclic_x ? 1 : 0;
    g.tiles_x = (g.nx + TX - 1) / TX;
    g.tiles_y = (g.ny + TY - 1) / TY;
    g.ntiles = g.tiles_x * g.tiles_y;
    if (g.ntiles + 1024 > NA_MAX_PART) return fail(nullptr, PSI_BAD_ARG, "internal: tiles%s");
    c.plane = (size_t)(g.nx + 4) * (g.ny + 4);
    const int nxh = g.nx + 4, nyh = g.ny + 4;
    unsigned char* S = d->ptr;
    c.E = (double*)(S + NA_OFF_E);
    c.N = (double*)(S + NA_OFF_N);
    c.cC = (double*)(S + NA_OFF_C);
    c.u64[0] = (double*)(S + NA_OFF_U);
    c.u64[1] = (double*)(S + NA_OFF_U + NA_PLANE64);
    c.p64[0] = (double*)(S + NA_OFF_P);
    c.p64[1] = (double*)(S + NA_OFF_P + NA_PLANE64);
    c.r64 = (double*)(S + NA_OFF_R);
    c.act = S + NA_OFF_ACT;
    c.dirichlet = S + NA_OFF_DIR;
    double* met = (double*)(S + NA_OFF_MET);
    const size_t ms = NA_METRIC / 8;
    c.dxt = met;
    c.dxu = met + ms;
    c.dyt = met + 2 * ms;
    c.dyu = met + 3 * ms;
    c.cost = met + 4 * ms;
    c.cosu = met + 5 * ms;
    c.rx = met + 6 * ms;
    c.ry = met + 7 * ms;
    int32_t* bi = (int32_t*)(S + NA_OFF_BI);
    int32_t* bj = (int32_t*)(S + NA_OFF_BJ);
    c.part = (double*)(S + NA_OFF_PART);
    c.part2 = (double*)(S + NA_OFF_PART2);
    c.st = (CgState*)(S + NA_OFF_ST);
    c.result = (double*)(S + NA_OFF_RES);
    c.h_st = d->h_st;
    c.h_res = d->h_res;
    c.hur = f->hur;
    c.hvr = f->hvr;
    c.ncu = d->ncu;
    c.G = std::min(d->G, g.ntiles);
    if (c.G >= 8) c.G &= ~7;
    c.G32 = c.G;
    choose_march(&c);   // the same sweep kernel as psi_congrad's
    if (c.Gpart > NA_MAX_PART) return fail(nullptr, PSI_BAD_ARG, "internal: partials%s");
    memset(d->h_st, 0, sizeof d->h_st);

#define NA_HIP(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            g_create_err = std::string("HIP error: ") + hipGetErrorString(e_);            \
            return PSI_HIP_ERROR;                                                         \
        }                                                                                 \
    } while (0)
    // the last call's contents stay in scratch: clear everything this grid's
    // solve reads before writing it (the whole used part, 16 bytes at a time)
    const size_t used = NA_OFF_RES + na_align(8 * 8);
    hipLaunchKernelGGL(k_na_zero, dim3((unsigned)std::min<size_t>(2048, (used / 16 + 255) / 256)),
                       dim3(256), 0, c.stream, (uint4*)S, used / 16);
    NA_HIP(hipMemcpyAsync(c.dxt, grid->dxt, nxh * 8, hipMemcpyHostToDevice, c.stream));
    NA_HIP(hipMemcpyAsync(c.dxu, grid->dxu, nxh * 8, hipMemcpyHostToDevice, c.stream));
    NA_HIP(hipMemcpyAsync(c.dyt, grid->dyt, nyh * 8, hipMemcpyHostToDevice, c.stream));
    NA_HIP(hipMemcpyAsync(c.dyu, grid->dyu, nyh * 8, hipMemcpyHostToDevice, c.stream));
    NA_HIP(hipMemcpyAsync(c.cost, grid->cost, nyh * 8, hipMemcpyHostToDevice, c.stream));
    NA_HIP(hipMemcpyAsync(c.cosu, grid->cosu, nyh * 8, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_na_ratios, dim3(nblk(std::max(nxh, nyh))), dim3(256), 0, c.stream, nxh,
                       nyh, (const double*)c.dxu, (const double*)c.dyu, (const double*)c.cosu,
                       c.rx, c.ry);
    if (np > 0) {
        NA_HIP(hipMemcpyAsync(bi, isl->bi, (size_t)np * 4, hipMemcpyHostToDevice, c.stream));
        NA_HIP(hipMemcpyAsync(bj, isl->bj, (size_t)np * 4, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_mark_points, dim3(nblk(np)), dim3(256), 0, c.stream, g, np,
                           (const int*)bi, (const int*)bj, c.dirichlet);
    }
    hipLaunchKernelGGL(k_build_EN, dim3(nblk(c.plane)), dim3(256), 0, c.stream, g, c.hur, c.hvr,
                       c.dxt, c.dxu, c.dyt, c.dyu, c.cost, c.cosu, c.E, c.N);
    hipLaunchKernelGGL(k_build_active, dim3(nblk(c.plane)), dim3(256), 0, c.stream, g, c.E, c.N,
                       c.rx, c.ry, c.dirichlet, c.act, c.cC);
    if (g.cyclic)
        hipLaunchKernelGGL(k_setcyclic, dim3(nblk(g.ny + 4)), dim3(256), 0, c.stream, g, sol);
    NA_HIP(hipGetLastError());
    memset(info, 0, sizeof *info);
    rc = solve_plain(&c, o, forc, sol, info);
    if (rc == PSI_HIP_ERROR || rc == PSI_BAD_ARG) {
        g_create_err = c.err;
        return rc;
    }
    NA_HIP(hipStreamSynchronize(c.stream));
    info->elapsed_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
    info->sweep_ms = -1.0;   // no events: the sweeps are not timed separately
    return info->status;
#undef NA_HIP
}
