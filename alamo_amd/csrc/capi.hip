// C ABI, multigrid hierarchy and solve driver of libalamo_amd.so.
//
// The driver restates amrex::MLMG (AMReX 26.06, which the reference drives through
// Solver/Nonlocal/Linear.H:255-290) from its documented algorithm: F-cycle every
// iteration (setMaxFmgIter(1e8), Linear.H:261) or V-cycles, pre/post smoothing by
// Operator<Node>::Fsmooth, restriction of the correction residual, a
// Jacobi-preconditioned BiCGStab bottom solve and the max-norm stopping test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/alamo_amd.h"
#include "kernels_generic.h"
#include "kernels_iso.h"
#include "kernels_iso_rows.h"
#include "kernels_tail.h"
#include "krylov.h"
#include "amr.h"
#include "post.h"
#include "galerkin.h"
#include "calib.h"
#include "plotfile.h"

using namespace alamo;

static thread_local std::string g_err;
static std::atomic<uint64_t> g_model_version{0};
// largest grain table (21 values per grain) a MajorMinor workgroup stages in LDS: 48 KB (292 grains
// in FP64, 585 in FP32); larger tables use the coefficient planes
constexpr size_t MM_GRAIN_LDS_MAX = 48 * 1024;
// The LDS-tile kernel serves isotropic levels whose arrays reach 4 GiB (the row kernel's 32-bit
// byte offsets end there, ~812^3 FP64 nodes).  ALAMO_ISO_KERNEL=tiles routes every level to it, so
// that fallback stays parity-tested at sizes the oracle checks quickly.
static const bool g_iso_tiles = [] {
    const char* e = getenv("ALAMO_ISO_KERNEL");
    return e && strcmp(e, "tiles") == 0;
}();

#define FAIL(code, msg)            \
    do {                           \
        g_err = (msg);             \
        return (code);             \
    } while (0)

#define HIPCHK(expr)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            g_err = std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" +        \
                    std::to_string(__LINE__) + " (" #expr ")";                                            \
            return ALAMO_ERR_HIP;                                                                         \
        }                                                                                                 \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int r_ = (expr);           \
        if (r_ != ALAMO_OK) return r_; \
    } while (0)

// Device scratch of alamo_mg_fsmooth_noalloc: the ping-pong partner of its sweeps,
// ALAMO_NOALLOC_MAX_DOUBLES doubles (64 MiB) on every device the library runs on, from library load
// to process end.  Host code never passes `scratch` itself to a kernel (on the host the name is a
// stand-in, not a device address): scratch_ptr() looks up this device's copy once with
// hipGetSymbolAddress and every use goes through that pointer.
static __device__ __align__(256) double scratch[ALAMO_NOALLOC_MAX_DOUBLES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// the address of the current device's `scratch` (looked up on the first call on each device)
static int scratch_ptr(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) FAIL(ALAMO_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) FAIL(ALAMO_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return ALAMO_OK;
}

namespace {

struct LevelBuf {
    Geom g{};
    double* C = nullptr;   // FP64 stiffness, ncoef SoA planes
    double* D = nullptr;   // FP64 diagonal, 3 planes
    double *cor = nullptr, *res = nullptr, *rescor = nullptr, *tmp = nullptr, *hold = nullptr;
    float* Cf = nullptr;
    float* Df = nullptr;
    float *corf = nullptr, *resf = nullptr, *rescorf = nullptr, *tmpf = nullptr, *holdf = nullptr;
    double* dsm = nullptr;  // l1-Jacobi smoother diagonal (3 planes), allocated when that smoother is used
    float* dsmf = nullptr;
    uint16_t* dsmr = nullptr;  // BF16 reciprocal of dsm for the FP32 row-kernel sweeps (mixed precision)
    double* S = nullptr;    // Galerkin stencil of a coarse level (GAL_NPLANE planes), coarse_op GALERKIN
    float* Sf = nullptr;
    double* psi = nullptr;  // Elastic::m_psi_mf of the level (cells), when set
    float* psif = nullptr;
    uint16_t* gid = nullptr;  // per-grain model (finest level): grain ids and the grain table
    double* gtab = nullptr;
    float* gtabf = nullptr;
    int ngrains_cap = 0;
    double* mmd = nullptr;  // MajorMinor gradient contraction (18 planes), prepared with the level
    float* mmdf = nullptr;
};

}  // namespace

struct alamo_mg {
    alamo_mg_desc d{};
    BCTable bc{};
    int sym = 0, nc = 2;
    hipStream_t stream = nullptr;
    std::vector<LevelBuf> L;
    double* partial = nullptr;      // reduction partials
    double* red = nullptr;          // device reduction result (4 doubles)
    double* red_host = nullptr;     // pinned host mirror (KRY_MAXV + 4 doubles)
    double* bottom_work = nullptr;  // 8 x 3 x tot(bottom)
    int* bottom_iters = nullptr;
    double* dense_A = nullptr;      // dense bottom: assembled coarsest operator (column-major n x n)
    double* dense_inv = nullptr;    // unit columns during assembly, then the row-major inverse
    int64_t dense_n = 0;
    double* sol_res = nullptr;      // level-0 FP64 residual for the mixed-precision outer loop
    bool model_set = false, prepared = false;
    uint64_t model_version = 0;  // unique per SetModel call (the network's features follow it)
    int nblk_red = 1024;
    // flexible GMRES workspace (allocated on first use)
    double* kV = nullptr;     // (m+1) basis vectors
    double* kZ = nullptr;     // m preconditioned vectors
    int kry_m = 0;
    float* kVf = nullptr;     // FP32 basis and preconditioned vectors of the mixed-precision inner iterations
    float* kZf = nullptr;
    int kry_mf = 0;
    bool gal = false;         // levels >= 1 run the prepared Galerkin stencils
    bool psi_set = false;     // Elastic::m_psi_set
    bool uniform = false;     // Elastic::m_uniform
    double psi_small = 1e-8;  // Elastic::m_psi_small
    double* kcoef = nullptr;  // device copy of Gram-Schmidt / solution coefficients
    double* kpart = nullptr;  // multi-dot partials
    double* kdots = nullptr;  // multi-dot results (device)
    alamo_net* net = nullptr; // learned correction (non-owning)
    double* lq = nullptr;     // FP64 scratch of the learned stage
    double* rowtab = nullptr; // explicit level-0 rows, 243 doubles per node (transpose applies)
    std::vector<void*> allocs;
};

static Geom make_geom(const int* n, const double* dx) {
    Geom g{};
    for (int d = 0; d < 3; d++) {
        g.n[d] = n[d];
        g.N[d] = n[d] + 1;
        g.dx[d] = dx[d];
    }
    g.sy = g.N[0];
    g.sz = g.sy * g.N[1];
    g.tot = g.sz * g.N[2];
    return g;
}

static inline unsigned nblocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
static inline unsigned gs_blocks(int64_t n) { return (unsigned)std::min<int64_t>(4096, (n + 255) / 256); }

template <typename T>
static int dalloc(alamo_mg* h, T** p, int64_t n) {
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<int64_t>(1, n) * sizeof(T)));
    h->allocs.push_back(q);
    HIPCHK(hipMemset(q, 0, std::max<int64_t>(1, n) * sizeof(T)));
    HIPCHK(hipDeviceSynchronize());  // complete before work on any stream uses the buffer
    *p = (T*)q;
    return ALAMO_OK;
}

// ---------------------------------------------------------------------------
// Kernel dispatch
// ---------------------------------------------------------------------------
template <typename T, int SYM>
static void launch_node_sym(int mode, const Geom& g, const BCTable& bc, const T* C, const T* u, const T* b, T* out,
                            T omega, hipStream_t s) {
    dim3 grid(nblocks(g.tot)), blk(256);
    switch (mode) {
        case APPLY: k_node<T, SYM, APPLY><<<grid, blk, 0, s>>>(g, bc, C, u, b, out, omega); break;
        case RESID: k_node<T, SYM, RESID><<<grid, blk, 0, s>>>(g, bc, C, u, b, out, omega); break;
        case JACOBI: k_node<T, SYM, JACOBI><<<grid, blk, 0, s>>>(g, bc, C, u, b, out, omega); break;
        default: k_node<T, SYM, DIAG><<<grid, blk, 0, s>>>(g, bc, C, u, b, out, omega); break;
    }
}

template <typename T>
static void launch_node(int sym, int mode, const Geom& g, const BCTable& bc, const T* C, const T* u, const T* b,
                        T* out, T omega, hipStream_t s) {
    if (g.psi || g.uniform) {
        // the psi terms and the uniform-model rows live in the node-parallel kernels
        if (sym == ISO) launch_node_sym<T, ISO>(mode, g, bc, C, u, b, out, omega, s);
        else if (sym == MAJORMINOR) launch_node_sym<T, MAJORMINOR>(mode, g, bc, C, u, b, out, omega, s);
        else launch_node_sym<T, DIAGONAL>(mode, g, bc, C, u, b, out, omega, s);
        return;
    }
    if (sym == ISO) {
        // isotropic levels of >= 4 cells per direction take the row kernel (every index it forms
        // is clamped into the level); the row kernel addresses each array with 32-bit byte
        // offsets, so arrays of 4 GiB and more (beyond ~812^3 FP64 nodes) take the LDS-tiled kernel
        if (mode != DIAG) {
            const bool huge = (int64_t)sizeof(T) * g.tot >= ((int64_t)1 << 32);
            if ((huge || g_iso_tiles) && iso_tiled_ok(g)) {
                launch_iso_tiled<T>(mode, g, bc, C, u, b, out, omega, s);
                return;
            }
            if (!huge && !g_iso_tiles && iso_rows_ok(g)) {
                launch_iso_rows<T>(mode, g, bc, C, u, b, out, omega, s);
                return;
            }
        }
        launch_node_sym<T, ISO>(mode, g, bc, C, u, b, out, omega, s);
    } else if (sym == MAJORMINOR) {
        if (mode != DIAG && g.n[0] >= 2 && g.n[1] >= 2 && g.n[2] >= 2) {
            // interior rows slot-major (no spills), boundary rows by the face kernel
            const int64_t ni = (int64_t)(g.N[0] - 2) * (g.N[1] - 2) * (g.N[2] - 2);
            const int64_t nb = boundary_count(g);
            const unsigned gi = (unsigned)((ni + 255) / 256), gb = (unsigned)((nb + 255) / 256);
            // a per-grain model whose table fits the LDS budget reads it from there
            const size_t tabb = (size_t)21 * g.ngrains * sizeof(T);
            const bool grains = g.gid && tabb <= MM_GRAIN_LDS_MAX;
            // a prepared level reads its node-local gradient contraction instead of differencing C
            const bool dcon = std::is_same<T, double>::value ? g.mmd != nullptr : g.mmdf != nullptr;
            if (mode == APPLY) {
                if (grains) k_mm_interior_gid<T, APPLY><<<gi, 256, tabb, s>>>(g, u, b, out, omega);
                else if (dcon) k_mm_interior_d<T, APPLY><<<gi, 256, 0, s>>>(g, C, u, b, out, omega);
                else k_mm_interior<T, APPLY><<<gi, 256, 0, s>>>(g, C, u, b, out, omega);
                k_boundary<T, MAJORMINOR, APPLY><<<gb, 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
            } else if (mode == RESID) {
                if (grains) k_mm_interior_gid<T, RESID><<<gi, 256, tabb, s>>>(g, u, b, out, omega);
                else if (dcon) k_mm_interior_d<T, RESID><<<gi, 256, 0, s>>>(g, C, u, b, out, omega);
                else k_mm_interior<T, RESID><<<gi, 256, 0, s>>>(g, C, u, b, out, omega);
                k_boundary<T, MAJORMINOR, RESID><<<gb, 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
            } else {
                if (grains) k_mm_interior_gid<T, JACOBI><<<gi, 256, tabb, s>>>(g, u, b, out, omega);
                else if (dcon) k_mm_interior_d<T, JACOBI><<<gi, 256, 0, s>>>(g, C, u, b, out, omega);
                else k_mm_interior<T, JACOBI><<<gi, 256, 0, s>>>(g, C, u, b, out, omega);
                k_boundary<T, MAJORMINOR, JACOBI><<<gb, 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
            }
            return;
        }
        launch_node_sym<T, MAJORMINOR>(mode, g, bc, C, u, b, out, omega, s);
    } else {
        launch_node_sym<T, DIAGONAL>(mode, g, bc, C, u, b, out, omega, s);
    }
}

template <typename T>
static void launch_bottom(int sym, const Geom& g, const BCTable& bc, const T* C, const T* D, T* x, const T* b, T* w,
                          double rtol, int maxit, int* it, hipStream_t s) {
    if (sym == ISO) k_bottom_bicgstab<T, ISO><<<1, BOTTOM_BLOCK, 0, s>>>(g, bc, C, D, x, b, w, rtol, maxit, it);
    else if (sym == MAJORMINOR)
        k_bottom_bicgstab<T, MAJORMINOR><<<1, BOTTOM_BLOCK, 0, s>>>(g, bc, C, D, x, b, w, rtol, maxit, it);
    else k_bottom_bicgstab<T, DIAGONAL><<<1, BOTTOM_BLOCK, 0, s>>>(g, bc, C, D, x, b, w, rtol, maxit, it);
}

template <typename T>
static void vfill(T* x, int64_t n, T v, hipStream_t s) {
    k_fill<T><<<gs_blocks(n), 256, 0, s>>>(x, n, v);
}

// dst = src over n elements on the handle's stream: 16-byte loads and stores when both are
// 16-byte aligned (every level buffer and Krylov vector here is), one work-item per 16 bytes;
// the staging copies of the F-cycle preconditioner use it instead of the runtime's copy.
template <typename T>
__global__ void __launch_bounds__(256) k_copy16(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    constexpr int E = 16 / sizeof(T);
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = n / E;
    if (q < nv) ((uint4*)dst)[q] = ((const uint4*)src)[q];
    if (q < n - nv * E) dst[nv * E + q] = src[nv * E + q];  // the tail of fewer than E elements
}
template <typename T>
static void vcopy(T* dst, const T* src, int64_t n, hipStream_t s) {
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) != 0) {
        (void)hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s);
        return;
    }
    const int64_t nv = std::max<int64_t>(n / (16 / (int64_t)sizeof(T)), 1);
    k_copy16<T><<<(unsigned)((nv + 255) / 256), 256, 0, s>>>(src, dst, n);
}

template <typename T>
static int norms_dev(alamo_mg* h, const T* x, int64_t n, double* ninf, double* nl2) {
    const int nb = (int)std::min<int64_t>(h->nblk_red, (n + RED_BLOCK - 1) / RED_BLOCK);
    k_norm_partial<T><<<nb, RED_BLOCK, 0, h->stream>>>(x, n, h->partial);
    k_norm_final<<<1, RED_BLOCK, 0, h->stream>>>(h->partial, nb, h->red);
    HIPCHK(hipMemcpyAsync(h->red_host, h->red, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (ninf) *ninf = h->red_host[0];
    if (nl2) *nl2 = std::sqrt(h->red_host[1]);
    return ALAMO_OK;
}

// ---------------------------------------------------------------------------
// Multigrid cycles (templated on the cycle precision T)
// ---------------------------------------------------------------------------
template <typename T> struct LB;
template <> struct LB<double> {
    static double*& C(LevelBuf& l) { return l.C; }
    static double*& D(LevelBuf& l) { return l.D; }
    static double*& cor(LevelBuf& l) { return l.cor; }
    static double*& res(LevelBuf& l) { return l.res; }
    static double*& rescor(LevelBuf& l) { return l.rescor; }
    static double*& tmp(LevelBuf& l) { return l.tmp; }
    static double*& hold(LevelBuf& l) { return l.hold; }
    static double*& dsm(LevelBuf& l) { return l.dsm; }
    static double*& S(LevelBuf& l) { return l.S; }
};
template <> struct LB<float> {
    static float*& C(LevelBuf& l) { return l.Cf; }
    static float*& D(LevelBuf& l) { return l.Df; }
    static float*& cor(LevelBuf& l) { return l.corf; }
    static float*& res(LevelBuf& l) { return l.resf; }
    static float*& rescor(LevelBuf& l) { return l.rescorf; }
    static float*& tmp(LevelBuf& l) { return l.tmpf; }
    static float*& hold(LevelBuf& l) { return l.holdf; }
    static float*& dsm(LevelBuf& l) { return l.dsmf; }
    static float*& S(LevelBuf& l) { return l.Sf; }
};

// The operator of level l in the cycle precision: the level's rows (level 0 always; every level
// with coarse_op RESTRICT, the reference's re-discretisation), or its Galerkin stencil.
template <typename T>
static void level_op(alamo_mg* h, int l, int mode, const T* u, const T* b, T* out, T omega, hipStream_t s) {
    LevelBuf& lb = h->L[l];
    if (l > 0 && h->gal) launch_stencil<T>(mode, lb.g, LB<T>::S(lb), u, b, out, omega, s);
    else launch_node<T>(h->sym, mode, lb.g, h->bc, LB<T>::C(lb), u, b, out, omega, s);
}

// Diagonal and damping of a first sweep from a zero guess, x = w b / d: the prepared point
// diagonal with omega (the reference smoother), or the l1 smoother diagonal with weight 1.
template <typename T>
static const T* sweep_diag(alamo_mg* h, LevelBuf& lb) {
    return h->d.smoother == ALAMO_SMOOTHER_L1 ? LB<T>::dsm(lb) : LB<T>::D(lb);
}
static double sweep_omega(const alamo_mg* h) { return h->d.smoother == ALAMO_SMOOTHER_L1 ? 1.0 : h->d.omega; }

template <typename T>
static void sweeps(alamo_mg* h, int l, int count) {
    LevelBuf& lb = h->L[l];
    for (int s = 0; s < count; s++) {
        level_op<T>(h, l, JACOBI, LB<T>::cor(lb), LB<T>::res(lb), LB<T>::tmp(lb), T(h->d.omega), h->stream);
        std::swap(LB<T>::cor(lb), LB<T>::tmp(lb));
    }
}

// Jacobi sweeps per Fsmooth call on level l of a cycle (desc.sweeps_finest on level 0 when set)
static int sps(const alamo_mg* h, int l) {
    return (l == 0 && h->d.sweeps_finest > 0) ? h->d.sweeps_finest : h->d.sweeps_per_smooth;
}

template <typename T>
static void smooth(alamo_mg* h, int l, int ncalls) {
    sweeps<T>(h, l, ncalls * sps(h, l));
}

// First smoothing sweep from a zero correction: A 0 = 0, so x = omega b / D pointwise (the
// prepared point diagonal with omega, or the l1 smoother diagonal with 1; D = 0 leaves 0).  b_out, when given,
// receives b in the cycle precision, so the same pass stages the cycle's right-hand side.
template <typename Ti, typename T>
__global__ void k_jacobi_from_zero(const Ti* __restrict__ b_in, T* __restrict__ b_out, T* __restrict__ x,
                                   const T* __restrict__ D, int64_t n, T omega) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const T bq = (T)b_in[q];
        if (b_out) b_out[q] = bq;
        const T d = D[q];
        x[q] = d != T(0) ? omega * bq / d : T(0);
    }
}

// pre-smoothing of a V-cycle level from a zero correction; first_done: the first sweep has
// already been applied (by the caller's staging pass)
template <typename T>
static void presmooth_from_zero(alamo_mg* h, int l, bool first_done) {
    LevelBuf& lb = h->L[l];
    const int total = h->d.pre_smooth * sps(h, l);
    const int64_t n = 3 * lb.g.tot;
    if (total == 0) {
        vfill<T>(LB<T>::cor(lb), n, T(0), h->stream);
        return;
    }
    if (!first_done)
        k_jacobi_from_zero<T, T><<<gs_blocks(n), 256, 0, h->stream>>>(LB<T>::res(lb), nullptr, LB<T>::cor(lb),
                                                                      sweep_diag<T>(h, lb), n, T(sweep_omega(h)));
    sweeps<T>(h, l, total - 1);
}

template <typename T>
static void bottom(alamo_mg* h) {
    const int bot = (int)h->L.size() - 1;
    LevelBuf& lb = h->L[bot];
    const int64_t n = 3 * lb.g.tot;
    if (h->d.bottom_solver == ALAMO_BOTTOM_DENSE) {
        // the FP64 inverse applied to the cycle-precision vectors in place (FP64 products and sum)
        k_dense_matvec<T, T><<<nblocks(n, 4), 256, 0, h->stream>>>(h->dense_inv, LB<T>::res(lb), LB<T>::cor(lb), n);
        return;
    }
    // the other bottom solves run in FP64
    if constexpr (std::is_same<T, float>::value)
        k_convert<float, double><<<gs_blocks(n), 256, 0, h->stream>>>(lb.resf, lb.res, n);
    if (h->d.bottom_solver == ALAMO_BOTTOM_SMOOTHER) {
        vfill<double>(lb.cor, n, 0.0, h->stream);
        smooth<double>(h, bot, 8);
    } else {
        vfill<double>(lb.cor, n, 0.0, h->stream);
        launch_bottom<double>(h->sym, lb.g, h->bc, lb.C, lb.D, lb.cor, lb.res, h->bottom_work, h->d.bottom_tol_rel,
                              h->d.bottom_max_iter, h->bottom_iters, h->stream);
    }
    if constexpr (std::is_same<T, float>::value)
        k_convert<double, float><<<gs_blocks(n), 256, 0, h->stream>>>(lb.cor, lb.corf, n);
}

// The FP32 l1 cycle's last smoothed level runs as two single-workgroup launches around the dense
// bottom (k_cycle_tail) when it has at most TAIL_MAX nodes, all-Dirichlet isotropic rows and a
// prepared l1 diagonal; ALAMO_CYCLE_TAIL=0 (read at every call) runs it as separate launches.
template <typename T>
static bool cycle_tail_ok(const alamo_mg* h) {
    if constexpr (!std::is_same<T, float>::value) return false;
    else {
        // level 0 is never the tail level: the preconditioner entry points alias its buffers to the
        // Krylov vectors and count on one buffer swap per sweep there
        const int bot = (int)h->L.size() - 1;
        if (bot < 2 || h->d.bottom_solver != ALAMO_BOTTOM_DENSE || h->gal || h->sym != ISO || !all_dirichlet(h->bc))
            return false;
        const Geom& g = h->L[bot - 1].g;
        const Geom& gc = h->L[bot].g;
        if (g.psi || g.uniform || !g.dsmr || !g.dsmf || g.tot > TAIL_MAX || gc.tot > TAIL_MAXC || !h->dense_inv)
            return false;
        for (int d = 0; d < 3; d++)
            if (g.n[d] < 2 || g.n[d] != 2 * gc.n[d] || g.sy != g.N[0] || g.sz != (int64_t)g.N[0] * g.N[1])
                return false;
        if (gc.sy != gc.N[0] || gc.sz != (int64_t)gc.N[0] * gc.N[1]) return false;
        const char* e = getenv("ALAMO_CYCLE_TAIL");
        return !(e && e[0] == '0');
    }
}

// guess: the top level's cor already holds an initial guess (its pre-smoothing sweeps start from it)
template <typename T>
static void vcycle(alamo_mg* h, int top, bool first_done = false, bool guess = false) {
    const int bot = (int)h->L.size() - 1;
    const bool tail = top <= bot - 1 && cycle_tail_ok<T>(h);
    const int last = tail ? bot - 1 : bot;  // levels top .. last - 1 by separate launches
    bool done = first_done;
    for (int l = top; l < last; l++) {
        LevelBuf& lb = h->L[l];
        if (guess && l == top) sweeps<T>(h, l, h->d.pre_smooth * sps(h, l));
        else presmooth_from_zero<T>(h, l, done);
        level_op<T>(h, l, RESID, LB<T>::cor(lb), LB<T>::res(lb), LB<T>::rescor(lb), T(0), h->stream);
        LevelBuf& lc = h->L[l + 1];
        // the restriction also applies the next level's first pre-smoothing sweep (from zero)
        done = l + 1 < bot && h->d.pre_smooth * sps(h, l + 1) > 0;
        k_restrict<T><<<nblocks(lc.g.tot), 256, 0, h->stream>>>(lc.g, lb.g, 3, LB<T>::rescor(lb), LB<T>::res(lc),
                                                                 done ? LB<T>::cor(lc) : nullptr, sweep_diag<T>(h, lc),
                                                                 T(sweep_omega(h)));
    }
    if (tail) {
        if constexpr (std::is_same<T, float>::value) {
            const int l = bot - 1;
            LevelBuf& lb = h->L[l];
            const int total = h->d.pre_smooth * sps(h, l);
            int init = 0, npre = total;  // guess at the top level: sweeps from cor
            if (!(guess && l == top)) {
                if (total == 0) init = 2;
                else {
                    init = done ? 0 : 1;  // the first sweep from zero already written by the restriction, or here
                    npre = total - 1;
                }
            }
            LevelBuf& lc = h->L[bot];
            k_cycle_tail<0><<<1, TAIL_NT, 0, h->stream>>>(lb.g, lc.g, lb.Cf, lb.dsmf, lb.dsmr, lb.resf, lb.corf,
                                                          lc.resf, nullptr, init, npre);
            bottom<T>(h);
            k_cycle_tail<1><<<1, TAIL_NT, 0, h->stream>>>(lb.g, lc.g, lb.Cf, lb.dsmf, lb.dsmr, lb.resf, lb.corf,
                                                          nullptr, lc.corf, 0, h->d.post_smooth * sps(h, l));
        }
    } else {
        bottom<T>(h);
    }
    for (int l = last - 1; l >= top; l--) {
        LevelBuf& lb = h->L[l];
        LevelBuf& lc = h->L[l + 1];
        launch_interp<T>(lc.g, lb.g, LB<T>::cor(lc), LB<T>::cor(lb), h->stream);
        smooth<T>(h, l, h->d.post_smooth);
    }
}

template <typename T>
static void fcycle(alamo_mg* h) {
    const int bot = (int)h->L.size() - 1;
    for (int l = 1; l <= bot; l++) {
        LevelBuf& lf = h->L[l - 1];
        LevelBuf& lc = h->L[l];
        k_restrict<T><<<nblocks(lc.g.tot), 256, 0, h->stream>>>(lc.g, lf.g, 3, LB<T>::res(lf), LB<T>::res(lc));
    }
    bottom<T>(h);
    for (int l = bot - 1; l >= 0; l--) {
        LevelBuf& lb = h->L[l];
        LevelBuf& lc = h->L[l + 1];
        const int64_t n = 3 * lb.g.tot;
        vfill<T>(LB<T>::cor(lb), n, T(0), h->stream);
        launch_interp<T>(lc.g, lb.g, LB<T>::cor(lc), LB<T>::cor(lb), h->stream);
        // res = res - L(cor)
        level_op<T>(h, l, RESID, LB<T>::cor(lb), LB<T>::res(lb), LB<T>::rescor(lb), T(0), h->stream);
        std::swap(LB<T>::res(lb), LB<T>::rescor(lb));
        std::swap(LB<T>::hold(lb), LB<T>::cor(lb));
        vcycle<T>(h, l);
        k_add<T><<<gs_blocks(n), 256, 0, h->stream>>>(LB<T>::cor(lb), LB<T>::cor(lb), LB<T>::hold(lb), n);
    }
}

// z = F-cycle(v) from a zero guess for a preconditioner, v and z in the cycle precision: the steps
// of fcycle with each level's V-cycle run on the solution itself instead of on a correction --
// the level's sweeps start from u = P cor_{l+1} with the level's own right-hand side, where fcycle
// forms res - A hold, smooths a correction from zero and adds hold back.  The smoother is affine
// in u, so both are the same cycle up to rounding, without the residual, the first from-zero pass
// and the final addition on every level.  The first restriction reads v in place; on level 0 the
// sweeps read v in place and the solution lives in z (cor takes z's place when the level's sweep
// count is even, its partner's when odd, so the last sweep writes z).
template <typename T>
static void fcycle_into(alamo_mg* h, const T* v, T* z) {
    const int bot = (int)h->L.size() - 1;
    LevelBuf& l0 = h->L[0];
    LevelBuf& l1 = h->L[1];
    k_restrict<T><<<nblocks(l1.g.tot), 256, 0, h->stream>>>(l1.g, l0.g, 3, v, LB<T>::res(l1));
    for (int l = 2; l <= bot; l++) {
        LevelBuf& lf = h->L[l - 1];
        LevelBuf& lc = h->L[l];
        k_restrict<T><<<nblocks(lc.g.tot), 256, 0, h->stream>>>(lc.g, lf.g, 3, LB<T>::res(lf), LB<T>::res(lc));
    }
    bottom<T>(h);
    for (int l = bot - 1; l >= 1; l--) {
        LevelBuf& lb = h->L[l];
        LevelBuf& lc = h->L[l + 1];
        launch_interp<T, true>(lc.g, lb.g, LB<T>::cor(lc), LB<T>::cor(lb), h->stream);
        vcycle<T>(h, l, false, true);
    }
    T* own_cor = LB<T>::cor(l0);
    T* own_tmp = LB<T>::tmp(l0);
    T* own_res = LB<T>::res(l0);
    const int nsw = (h->d.pre_smooth + h->d.post_smooth) * sps(h, 0);
    if (nsw % 2) LB<T>::tmp(l0) = z;
    else LB<T>::cor(l0) = z;
    LB<T>::res(l0) = const_cast<T*>(v);
    launch_interp<T, true>(l1.g, l0.g, LB<T>::cor(l1), LB<T>::cor(l0), h->stream);
    vcycle<T>(h, 0, false, true);
    LB<T>::cor(l0) = own_cor;
    LB<T>::tmp(l0) = own_tmp;
    LB<T>::res(l0) = own_res;
}

// amrex::MLMG::oneIter: an F-cycle while iter < max_fmg_iter, then V-cycles (the reference sets
// max_fmg_iter to 1e8, Linear.H:261, so every iteration is an F-cycle)
template <typename T>
static void one_cycle(alamo_mg* h, int iter = 0) {
    if (h->d.cycle == ALAMO_CYCLE_F && iter < h->d.max_fmg_iter) fcycle<T>(h);
    else vcycle<T>(h, 0);
}

// ---------------------------------------------------------------------------
// Setup
// ---------------------------------------------------------------------------
// In-place Gauss-Jordan inverse with partial pivoting of the column-major n x n matrix A into the
// row-major inverse: the row interchanges are recorded and undone as column interchanges at the
// end (n^3 multiply-adds, half of eliminating into a separate identity), and each row's update
// runs over the pivot row's nonzeros when that row is at most half full (the identity rows of a
// Dirichlet boundary then cost one entry per row).  false when a pivot falls below 1e-11 max|A|
// (a singular coarsest operator, e.g. one that no displacement condition pins).
static bool dense_invert(const std::vector<double>& A, int64_t n, std::vector<double>& inv) {
    std::vector<double>& M = inv;
    M.assign(n * n, 0.0);
    double amax = 0.0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
            M[i * n + j] = A[j * n + i];
            amax = std::max(amax, std::fabs(M[i * n + j]));
        }
    std::vector<int64_t> piv(n), nz;
    nz.reserve(n);
    for (int64_t c = 0; c < n; c++) {
        int64_t p = c;
        for (int64_t r = c + 1; r < n; r++)
            if (std::fabs(M[r * n + c]) > std::fabs(M[p * n + c])) p = r;
        if (!(std::fabs(M[p * n + c]) > 1e-11 * amax)) return false;
        piv[c] = p;
        if (p != c)
            for (int64_t j = 0; j < n; j++) std::swap(M[p * n + j], M[c * n + j]);
        double* rc = &M[c * n];
        const double d = 1.0 / rc[c];
        rc[c] = 1.0;
        nz.clear();
        for (int64_t j = 0; j < n; j++) {
            rc[j] *= d;
            if (rc[j] != 0.0) nz.push_back(j);
        }
        const bool dense = 2 * (int64_t)nz.size() > n;
        for (int64_t r = 0; r < n; r++) {
            if (r == c) continue;
            double* rr = &M[r * n];
            const double f = rr[c];
            if (f == 0.0) continue;
            rr[c] = 0.0;
            if (dense) {
                for (int64_t j = 0; j < n; j++) rr[j] -= f * rc[j];
            } else {
                for (int64_t j : nz) rr[j] -= f * rc[j];
            }
        }
    }
    for (int64_t c = n - 1; c >= 0; c--)
        if (piv[c] != c)
            for (int64_t r = 0; r < n; r++) std::swap(M[r * n + c], M[r * n + piv[c]]);
    return true;
}

// solver.bottom_solver "dense": assemble the coarsest FP64 operator (with its psi and boundary
// rows) column by column on the device into A (host, column-major; synchronous), ...
static int dense_bottom_assemble(alamo_mg* h, std::vector<double>& A) {
    LevelBuf& lb = h->L.back();
    const int64_t n = 3 * lb.g.tot;
    if (n > ALAMO_DENSE_BOTTOM_MAX)
        FAIL(ALAMO_ERR_ARG, "bottom_solver dense: the coarsest level has " + std::to_string(n) + " unknowns, more than " +
                                std::to_string(ALAMO_DENSE_BOTTOM_MAX));
    if (h->dense_n != n) {
        TRY(dalloc(h, &h->dense_A, n * n));
        TRY(dalloc(h, &h->dense_inv, n * n));
        h->dense_n = n;
    }
    A.assign(n * n, 0.0);
    if (h->gal && h->L.size() > 1) {
        // the coarsest Galerkin stencil written out column-major
        const Geom& g = lb.g;
        std::vector<double> S((size_t)GAL_NPLANE * g.tot);
        HIPCHK(hipMemcpyAsync(S.data(), lb.S, S.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int k = 0; k <= g.n[2]; k++)
            for (int j = 0; j <= g.n[1]; j++)
                for (int i = 0; i <= g.n[0]; i++) {
                    const int64_t m = i + g.sy * j + g.sz * k;
                    for (int o = 0; o < 27; o++) {
                        const int ii = i + o % 3 - 1, jj = j + (o / 3) % 3 - 1, kk = k + o / 9 - 1;
                        if (ii < 0 || jj < 0 || kk < 0 || ii > g.n[0] || jj > g.n[1] || kk > g.n[2]) continue;
                        const int64_t mn = ii + g.sy * jj + g.sz * kk;
                        for (int q = 0; q < 3; q++)
                            for (int pp = 0; pp < 3; pp++)
                                A[(pp * g.tot + mn) * n + q * g.tot + m] = S[(size_t)gal_plane(o, q, pp) * g.tot + m];
                    }
                }
    } else {
        k_unit_columns<<<gs_blocks(n * n), 256, 0, h->stream>>>(h->dense_inv, n);
        const unsigned cols = (unsigned)n;
        if (h->sym == ISO)
            k_bottom_columns<ISO><<<cols, 256, 0, h->stream>>>(lb.g, h->bc, lb.C, h->dense_inv, h->dense_A);
        else if (h->sym == MAJORMINOR)
            k_bottom_columns<MAJORMINOR><<<cols, 256, 0, h->stream>>>(lb.g, h->bc, lb.C, h->dense_inv, h->dense_A);
        else k_bottom_columns<DIAGONAL><<<cols, 256, 0, h->stream>>>(lb.g, h->bc, lb.C, h->dense_inv, h->dense_A);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(A.data(), h->dense_A, n * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return ALAMO_OK;
}

// ... then invert it on the host and keep the inverse on the device.  *singular: the operator has
// no inverse (nothing on the device is changed then).
static int dense_bottom_finish(alamo_mg* h, const std::vector<double>& A, bool* singular) {
    *singular = false;
    const int64_t n = h->dense_n;
    std::vector<double> inv;
    if (!dense_invert(A, n, inv)) {
        *singular = true;
        return ALAMO_OK;
    }
    HIPCHK(hipMemcpyAsync(h->dense_inv, inv.data(), n * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return ALAMO_OK;
}

static int dense_bottom_factor(alamo_mg* h, bool* singular) {
    std::vector<double> A;
    TRY(dense_bottom_assemble(h, A));
    return dense_bottom_finish(h, A, singular);
}

static int prepare(alamo_mg* h) {
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "alamo_mg_prepare: SetModel has not been called");
    const int nl = (int)h->L.size();
    for (int l = 1; l < nl; l++) {
        LevelBuf& lf = h->L[l - 1];
        LevelBuf& lc = h->L[l];
        k_restrict<double><<<nblocks(lc.g.tot), 256, 0, h->stream>>>(lc.g, lf.g, h->nc, lf.C, lc.C);
    }
    if (h->psi_set) {
        // Elastic::averageDownCoeffs restricts psi to every level too (Elastic.cpp:839-907)
        for (int l = 0; l < nl; l++) {
            LevelBuf& lb = h->L[l];
            const int64_t nc = (int64_t)lb.g.n[0] * lb.g.n[1] * lb.g.n[2];
            if (l > 0)
                k_restrict_psi<<<nblocks(nc), 256, 0, h->stream>>>(lb.g, h->L[l - 1].g, h->L[l - 1].psi, lb.psi);
            if (h->d.precision == ALAMO_PREC_MIXED)
                k_convert<double, float><<<gs_blocks(nc), 256, 0, h->stream>>>(lb.psi, lb.psif, nc);
            lb.g.psi = lb.psi;
            lb.g.psif = lb.psif;
            lb.g.psi_small = h->psi_small;
        }
    }
    const bool l1 = h->d.smoother == ALAMO_SMOOTHER_L1;
    const bool mixed = h->d.precision == ALAMO_PREC_MIXED;
    if (l1 && h->psi_set) FAIL(ALAMO_ERR_ARG, "the l1 smoother has no psi terms; use smoother JACOBI with SetPsi");
    h->gal = false;
    if (h->d.coarse_op == ALAMO_COARSE_GALERKIN && nl > 1) {
        if (h->psi_set) FAIL(ALAMO_ERR_ARG, "coarse_op GALERKIN has no psi terms; use coarse_op RESTRICT with SetPsi");
        if (h->d.bottom_solver == ALAMO_BOTTOM_BICGSTAB)
            FAIL(ALAMO_ERR_ARG, "coarse_op GALERKIN needs bottom_solver dense or smoother");
        // A_l = R A_{l-1} P: level 1 from the level-0 rows, deeper levels from the stencil above
        for (int l = 1; l < nl; l++) {
            LevelBuf& lc = h->L[l];
            LevelBuf& lf = h->L[l - 1];
            if (!lc.S) TRY(dalloc(h, &lc.S, (int64_t)GAL_NPLANE * lc.g.tot));
            if (mixed && !lc.Sf) TRY(dalloc(h, &lc.Sf, (int64_t)GAL_NPLANE * lc.g.tot));
            const unsigned nb = nblocks(27 * lc.g.tot);
            if (l == 1) {
                if (h->sym == ISO) k_rap0<ISO><<<nb, 256, 0, h->stream>>>(lc.g, lf.g, h->bc, lf.C, lc.S);
                else if (h->sym == MAJORMINOR)
                    k_rap0<MAJORMINOR><<<nb, 256, 0, h->stream>>>(lc.g, lf.g, h->bc, lf.C, lc.S);
                else k_rap0<DIAGONAL><<<nb, 256, 0, h->stream>>>(lc.g, lf.g, h->bc, lf.C, lc.S);
            } else {
                k_rap<<<nb, 256, 0, h->stream>>>(lc.g, lf.g, lf.S, lc.S);
            }
            if (mixed)
                k_convert<double, float><<<gs_blocks(GAL_NPLANE * lc.g.tot), 256, 0, h->stream>>>(
                    lc.S, lc.Sf, (int64_t)GAL_NPLANE * lc.g.tot);
        }
        h->gal = true;
    }
    // the dense bottom's operator is assembled before the diagonal passes are queued, so that its
    // inversion on the host (the end of this function) overlaps them
    std::vector<double> Abot;
    const bool dense = h->d.bottom_solver == ALAMO_BOTTOM_DENSE;
    if (dense) TRY(dense_bottom_assemble(h, Abot));
    for (int l = 0; l < nl; l++) {
        LevelBuf& lb = h->L[l];
        const bool st = h->gal && l > 0;
        if (st) launch_stencil<double>(DIAG, lb.g, lb.S, nullptr, nullptr, lb.D, 0.0, h->stream);
        else launch_node<double>(h->sym, DIAG, lb.g, h->bc, lb.C, nullptr, nullptr, lb.D, 0.0, h->stream);
        lb.g.dsm = nullptr;
        lb.g.dsmf = nullptr;
        lb.g.dsmr = nullptr;
        if (l1) {
            // the smoother diagonal, from the level's coefficients and boundary table (or stencil)
            if (!lb.dsm) TRY(dalloc(h, &lb.dsm, 3 * lb.g.tot));
            const unsigned nb = nblocks(lb.g.tot);
            if (st) k_stencil_l1<double><<<nb, 256, 0, h->stream>>>(lb.g, lb.S, lb.dsm);
            else if (h->sym == ISO) k_l1_diag<double, ISO><<<nb, 256, 0, h->stream>>>(lb.g, h->bc, lb.C, lb.dsm);
            else if (h->sym == MAJORMINOR)
                k_l1_diag<double, MAJORMINOR><<<nb, 256, 0, h->stream>>>(lb.g, h->bc, lb.C, lb.dsm);
            else k_l1_diag<double, DIAGONAL><<<nb, 256, 0, h->stream>>>(lb.g, h->bc, lb.C, lb.dsm);
            lb.g.dsm = lb.dsm;
            if (h->d.precision == ALAMO_PREC_MIXED) {
                if (!lb.dsmf) TRY(dalloc(h, &lb.dsmf, 3 * lb.g.tot));
                k_convert<double, float><<<gs_blocks(3 * lb.g.tot), 256, 0, h->stream>>>(lb.dsm, lb.dsmf,
                                                                                         3 * lb.g.tot);
                lb.g.dsmf = lb.dsmf;
                if (!lb.dsmr) TRY(dalloc(h, &lb.dsmr, 3 * lb.g.tot));
                k_recip_bf16<<<gs_blocks(3 * lb.g.tot), 256, 0, h->stream>>>(lb.dsm, lb.dsmr, 3 * lb.g.tot);
                lb.g.dsmr = lb.dsmr;
            }
        }
        if (h->d.precision == ALAMO_PREC_MIXED) {
            k_convert<double, float><<<gs_blocks(h->nc * lb.g.tot), 256, 0, h->stream>>>(lb.C, lb.Cf,
                                                                                         h->nc * lb.g.tot);
            k_convert<double, float><<<gs_blocks(3 * lb.g.tot), 256, 0, h->stream>>>(lb.D, lb.Df, 3 * lb.g.tot);
        }
        // MajorMinor levels with coefficient planes: the node-local gradient contraction their
        // interior rows read
        lb.g.mmd = nullptr;
        lb.g.mmdf = nullptr;
        const int64_t ni = (int64_t)(lb.g.N[0] - 2) * (lb.g.N[1] - 2) * (lb.g.N[2] - 2);
        if (h->sym == MAJORMINOR && !st && ni > 0) {
            if (!lb.mmd) TRY(dalloc(h, &lb.mmd, 18 * lb.g.tot));
            k_mm_dcoef<<<nblocks(ni), 256, 0, h->stream>>>(lb.g, lb.C, lb.mmd);
            lb.g.mmd = lb.mmd;
            if (h->d.precision == ALAMO_PREC_MIXED) {
                if (!lb.mmdf) TRY(dalloc(h, &lb.mmdf, 18 * lb.g.tot));
                k_convert<double, float><<<gs_blocks(18 * lb.g.tot), 256, 0, h->stream>>>(lb.mmd, lb.mmdf,
                                                                                          18 * lb.g.tot);
                lb.g.mmdf = lb.mmdf;
            }
        }
    }
    HIPCHK(hipGetLastError());
    if (dense) {
        bool singular = false;
        TRY(dense_bottom_finish(h, Abot, &singular));
        // A coarsest level whose own boundary rows leave it singular (2 cells per direction with
        // traction-free sides: the one-sided traction rows of the two faces meet) ends the hierarchy
        // one level earlier, as long as that level still fits the dense bottom.
        while (singular && h->L.size() > 1 && 3 * h->L[h->L.size() - 2].g.tot <= ALAMO_DENSE_BOTTOM_MAX) {
            if (h->d.verbose)
                fprintf(stderr, "alamo: coarsest level (%d cells) singular; the hierarchy ends at %d cells\n",
                        h->L.back().g.n[0], h->L[h->L.size() - 2].g.n[0]);
            h->L.pop_back();
            // the new bottom level needs the FP64 cycle buffers of the FP64 bottom solve, which
            // mixed-precision handles allocate on level 0 and the original bottom only
            LevelBuf& nb = h->L.back();
            const int64_t v = 3 * nb.g.tot;
            if (!nb.cor) TRY(dalloc(h, &nb.cor, v));
            if (!nb.res) TRY(dalloc(h, &nb.res, v));
            if (!nb.tmp) TRY(dalloc(h, &nb.tmp, v));
            // and a BiCGStab workspace of its size, should the bottom solver be switched later
            if (nb.g.tot <= 17 * 17 * 17) TRY(dalloc(h, &h->bottom_work, 8 * v));
            TRY(dense_bottom_factor(h, &singular));
        }
        if (singular)
            FAIL(ALAMO_ERR_ARG, "bottom_solver dense: the coarsest-level operator is singular (no displacement "
                                "condition pins it); use bottom_solver bicgstab or smoother");
    }
    h->prepared = true;
    return ALAMO_OK;
}

static int check_level(alamo_mg* h, int amrlev, int mglev) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    if (amrlev != 0) FAIL(ALAMO_ERR_ARG, "only amrlev 0 is supported by this handle");
    if (mglev < 0 || mglev >= (int)h->L.size()) FAIL(ALAMO_ERR_ARG, "mglev out of range");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    if (!h->prepared) TRY(prepare(h));
    return ALAMO_OK;
}

// ---------------------------------------------------------------------------
// Solve
// ---------------------------------------------------------------------------
static int solve_impl(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                      alamo_solve_stats* st) {
    if (h->L.back().g.tot > 17 * 17 * 17)
        FAIL(ALAMO_ERR_ARG, "coarsest level has more than 17^3 nodes; use n_cell divisible by a larger power of 2");
    if (!h->prepared) TRY(prepare(h));
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t0 = std::chrono::steady_clock::now();
    LevelBuf& l0 = h->L[0];
    const int64_t n = 3 * l0.g.tot;
    const bool mixed = h->d.precision == ALAMO_PREC_MIXED;
    // The F-cycle swaps the level-0 residual buffer with its scratch partner, so the
    // outer residual location is looked up afresh every iteration.
    auto res0 = [&]() -> double* { return mixed ? h->sol_res : h->L[0].res; };
    launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, sol, rhs, res0(), 0.0, h->stream);
    double r0 = 0, bn = 0, rl2 = 0;
    TRY(norms_dev<double>(h, res0(), n, &r0, &rl2));
    TRY(norms_dev<double>(h, rhs, n, &bn, nullptr));
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * std::max(r0, bn));
    double rn = r0;
    int it = 0;
    if (h->d.verbose) fprintf(stderr, "alamo: iter 0 |r|inf %.6e target %.3e\n", r0, target);
    // solver.fixed_iter: exactly that many iterations, no early exit and no convergence failure
    const int fixed = h->d.fixed_iter > 0 ? h->d.fixed_iter : 0;
    while ((fixed ? it < fixed : (rn > target && it < h->d.max_iter)) && std::isfinite(rn)) {
        if (mixed) {
            k_convert<double, float><<<gs_blocks(n), 256, 0, h->stream>>>(res0(), h->L[0].resf, n);
            one_cycle<float>(h, it);
            k_axpy<float, double><<<gs_blocks(n), 256, 0, h->stream>>>(sol, h->L[0].corf, n, 1.0);
        } else {
            one_cycle<double>(h, it);
            k_axpy<double, double><<<gs_blocks(n), 256, 0, h->stream>>>(sol, h->L[0].cor, n, 1.0);
        }
        launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, sol, rhs, res0(), 0.0, h->stream);
        TRY(norms_dev<double>(h, res0(), n, &rn, &rl2));
        it++;
        if (h->d.verbose) fprintf(stderr, "alamo: iter %d |r|inf %.6e\n", it, rn);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t1 = std::chrono::steady_clock::now();
    const bool conv = rn <= target;
    if (st) {
        st->iterations = it;
        st->converged = conv ? 1 : 0;
        st->initial_res_inf = r0;
        st->final_res_inf = rn;
        st->final_res_l2 = rl2;
        st->rhs_inf = bn;
        st->target = target;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->fallback_events = 0;
    }
    if (!std::isfinite(rn)) FAIL(ALAMO_ERR_NAN, "alamo_linear_solve: residual is not finite (diverged)");
    if (!conv && !fixed) FAIL(ALAMO_ERR_NOCONV, "alamo_linear_solve: max_iter reached before the tolerance");
    return ALAMO_OK;
}

// ---------------------------------------------------------------------------
// Flexible GMRES(m) (FP64) with a multigrid-cycle preconditioner
// ---------------------------------------------------------------------------
static int multidot(alamo_mg* h, const double* w, const double* V, int64_t n, int nv, double* out);
#include "learned_host.h"

// z = V-cycle(b) from a zero guess, b given in precision Ti, with the level-0 correction in z
// itself: the correction and its ping-pong partner on level 0 swap once per sweep after the
// first, so z takes the place of the partner when that count (pre + post sweeps - 1) is odd and
// of the correction when it is even, and the cycle's last sweep writes z (no copy).  A V-cycle
// only reads its level-0 right-hand side, so when b is already in the cycle precision the level
// reads b in place; otherwise the first sweep's pass stages it.  The level's own buffers are
// restored afterwards.
template <typename Ti, typename T>
static void vcycle_into(alamo_mg* h, const Ti* b, T* z) {
    LevelBuf& l0 = h->L[0];
    const int64_t n = 3 * l0.g.tot;
    T* own_cor = LB<T>::cor(l0);
    T* own_tmp = LB<T>::tmp(l0);
    T* own_res = LB<T>::res(l0);
    const int swaps = (h->d.pre_smooth + h->d.post_smooth) * sps(h, 0) - 1;
    if (swaps % 2) LB<T>::tmp(l0) = z;
    else LB<T>::cor(l0) = z;
    T* stage = nullptr;
    if constexpr (std::is_same<Ti, T>::value) LB<T>::res(l0) = const_cast<T*>(b);
    else stage = own_res;
    k_jacobi_from_zero<Ti, T><<<gs_blocks(n), 256, 0, h->stream>>>(b, stage, LB<T>::cor(l0), sweep_diag<T>(h, l0),
                                                                   n, T(sweep_omega(h)));
    vcycle<T>(h, 0, true);
    LB<T>::cor(l0) = own_cor;
    LB<T>::tmp(l0) = own_tmp;
    LB<T>::res(l0) = own_res;
}

static int precond_apply(alamo_mg* h, int kind, const double* v, double* z) {
    LevelBuf& l0 = h->L[0];
    const int64_t n = 3 * l0.g.tot;
    if (kind == ALAMO_PRECOND_LEARNED) {
        // one multigrid cycle, then the network's correction of the cycle's remaining error
        TRY(precond_apply(h, ALAMO_PRECOND_MG, v, z));
        if (!h->lq) TRY(dalloc(h, &h->lq, 5 * n));
        return net_correct(h->net, h, v, z, h->lq);
    }
    if (kind == ALAMO_PRECOND_NONE) {
        HIPCHK(hipMemcpyAsync(z, v, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    } else if (h->d.precision == ALAMO_PREC_MIXED) {
        // a V-cycle with pre-smoothing stages its FP32 right-hand side in its first sweep's pass
        if (h->d.cycle == ALAMO_CYCLE_V && h->d.pre_smooth * h->d.sweeps_per_smooth > 0 && h->L.size() > 1) {
            k_jacobi_from_zero<double, float><<<gs_blocks(n), 256, 0, h->stream>>>(
                v, l0.resf, l0.corf, sweep_diag<float>(h, l0), n, (float)sweep_omega(h));
            vcycle<float>(h, 0, true);
        } else {
            k_convert<double, float><<<gs_blocks(n), 256, 0, h->stream>>>(v, l0.resf, n);
            one_cycle<float>(h);
        }
        k_convert<float, double><<<gs_blocks(n), 256, 0, h->stream>>>(l0.corf, z, n);
    } else {
        if (h->d.cycle == ALAMO_CYCLE_V && h->d.pre_smooth * h->d.sweeps_per_smooth > 0 && h->L.size() > 1) {
            vcycle_into<double, double>(h, v, z);
        } else if (h->d.cycle == ALAMO_CYCLE_F && h->d.max_fmg_iter > 0 && h->L.size() > 1) {
            fcycle_into<double>(h, v, z);
        } else {
            vcopy<double>(l0.res, v, n, h->stream);
            one_cycle<double>(h);
            vcopy<double>(z, l0.cor, n, h->stream);
        }
    }
    return ALAMO_OK;
}

// d[0..nv) = w . V_i  (device result copied to the pinned host buffer `out`)
// device scratch of the multi-dot / coefficient kernels, allocated on first use (the
// Krylov solver and the network stage both reach multidot)
static int krylov_scratch(alamo_mg* h) {
    if (h->kcoef) return ALAMO_OK;
    TRY(dalloc(h, &h->kcoef, KRY_MAXV));
    TRY(dalloc(h, &h->kpart, (int64_t)KRY_MAXV * KRY_DOT_BLOCKS));
    TRY(dalloc(h, &h->kdots, KRY_MAXV));
    return ALAMO_OK;
}

template <typename T>
static int multidot_t(alamo_mg* h, const T* w, const T* V, int64_t n, int nv, double* out, int64_t ldv = -1) {
    if (nv < 1 || nv > KRY_MAXV) FAIL(ALAMO_ERR_ARG, "multidot: vector count out of range");
    TRY(krylov_scratch(h));
    const int nb = (int)std::min<int64_t>(KRY_DOT_BLOCKS, (n + RED_BLOCK - 1) / RED_BLOCK);
    launch_multidot(w, V, n, ldv < 0 ? n : ldv, nv, nb, h->kpart, h->stream);
    k_sum_rows<<<nv, RED_BLOCK, 0, h->stream>>>(h->kpart, nb, h->kdots);
    HIPCHK(hipMemcpyAsync(h->red_host, h->kdots, nv * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < nv; i++) out[i] = h->red_host[i];
    return ALAMO_OK;
}
static int multidot(alamo_mg* h, const double* w, const double* V, int64_t n, int nv, double* out) {
    return multidot_t<double>(h, w, V, n, nv, out);
}

// w -= sum_{i < nv} c_i V_i with c already uploaded (upload_coef); *nrm = |w|_2 afterwards
template <typename T>
static int multiaxpy_norm(alamo_mg* h, T* w, const T* V, int64_t n, int nv, double* nrm, int64_t ldv = -1) {
    const int nb = (int)std::min<int64_t>(KRY_DOT_BLOCKS, (n + RED_BLOCK - 1) / RED_BLOCK);
    launch_multiaxpy_nrm<T>(w, V, n, ldv < 0 ? n : ldv, nv, h->kcoef, nb, h->kpart, h->stream);
    k_sum_rows<<<1, RED_BLOCK, 0, h->stream>>>(h->kpart, nb, h->kdots);
    HIPCHK(hipMemcpyAsync(h->red_host, h->kdots, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *nrm = std::sqrt(std::max(h->red_host[0], 0.0));
    return ALAMO_OK;
}

// second Gram-Schmidt pass when |w'| < eta |w| after the first.  With the multigrid cycle as
// preconditioner the first pass leaves |w'| / |w| = 0.3 .. 0.5 (A P v is close to v), so Kahan's
// eta = 1/sqrt 2 re-orthogonalised every step without changing the iterations or the final
// residual of any solve measured; eta = 0.1 keeps the second pass for severe cancellation only.
// ALAMO_KRY_REORTH overrides eta (0: never, above 1: always).
static const double g_reorth_eta = [] {
    const char* e = getenv("ALAMO_KRY_REORTH");
    return e ? atof(e) : 0.1;
}();
// Basis vectors kept unnormalised: vector j + 1 of the basis is stored as the Gram-Schmidt
// result w'' with its exact scale sc[j+1] = 1 / |w''| (the norm the axpy pass returns), so no
// pass over the vector normalises it.  The scales enter only host-side coefficients: the
// multi-dot values (h_i = sc_j sc_i w'.V_i), the axpy coefficients (sc_i^2 w'.V_i), the
// Hessenberg subdiagonal (sc_j |w''|) and the solution update (y_j sc_j), which needs the
// preconditioner to be homogeneous, M(s v) = s M(v): true of every multigrid cycle here
// (damped Jacobi, restriction, prolongation, the dense and smoother bottoms are linear;
// BiCGStab's iterates and relative stopping test scale with the right-hand side) and of
// none, not of the learned stage, whose steps normalise explicitly.  ALAMO_KRY_LAZY=0 always
// normalises.
static const bool g_kry_lazy = [] {
    const char* e = getenv("ALAMO_KRY_LAZY");
    return !(e && strcmp(e, "0") == 0);
}();

static int upload_coef(alamo_mg* h, const double* c, int nv) {
    TRY(krylov_scratch(h));
    for (int i = 0; i < nv; i++) h->red_host[i] = c[i];
    HIPCHK(hipMemcpyAsync(h->kcoef, h->red_host, nv * sizeof(double), hipMemcpyHostToDevice, h->stream));
    // the pinned staging buffer is reused by the next reduction: wait for the copy
    HIPCHK(hipStreamSynchronize(h->stream));
    return ALAMO_OK;
}

// One FP32 preconditioner application of the mixed-precision inner iterations: z = cycle(v), FP32
// in and out (the cycle's own FP32 level-0 buffers stage them); with the learned kind the network
// stage follows on FP64 copies of v and z (its residual, network and step are FP64 / BF16 as in the
// FP64 iteration).  Returns NET_NONFINITE (z holding the plain cycle's result) like precond_apply.
static int precond_f32(alamo_mg* h, int kind, const float* v, float* z) {
    LevelBuf& l0 = h->L[0];
    const int64_t n = 3 * l0.g.tot;
    if (h->d.cycle == ALAMO_CYCLE_V && h->d.pre_smooth * h->d.sweeps_per_smooth > 0 && h->L.size() > 1) {
        vcycle_into<float, float>(h, v, z);
    } else if (h->d.cycle == ALAMO_CYCLE_F && h->d.max_fmg_iter > 0 && h->L.size() > 1) {
        fcycle_into<float>(h, v, z);
    } else {
        vcopy<float>(l0.resf, v, n, h->stream);
        one_cycle<float>(h);
        vcopy<float>(z, l0.corf, n, h->stream);
    }
    if (kind != ALAMO_PRECOND_LEARNED) return ALAMO_OK;
    if (!h->lq) TRY(dalloc(h, &h->lq, 5 * n));
    double* v64 = h->lq + 3 * n;
    double* z64 = h->lq + 4 * n;
    k_convert<float, double><<<gs_blocks(n), 256, 0, h->stream>>>(v, v64, n);
    k_convert<float, double><<<gs_blocks(n), 256, 0, h->stream>>>(z, z64, n);
    const int rc = net_correct(h->net, h, v64, z64, h->lq);
    if (rc != ALAMO_OK) return rc;
    k_convert<double, float><<<gs_blocks(n), 256, 0, h->stream>>>(z64, z, n);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

// FGMRES(m) with the Krylov vectors in precision TB.  TB = double: the FP64 iteration, any
// preconditioner kind.  TB = float (mixed precision, cycle or learned preconditioner): the basis, the
// preconditioned vectors and A z are FP32 inside a restart cycle, a GMRES-based iterative
// refinement (Carson and Higham, SIAM J. Sci. Comput. 40 (2018)): every restart recomputes the
// FP64 true residual from the FP64 solution, and only that residual decides convergence.
template <typename TB>
static int fgmres_core(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                       const alamo_krylov_opts* o, alamo_solve_stats* st) {
    constexpr bool F32 = std::is_same<TB, float>::value;
    if (h->L.back().g.tot > 17 * 17 * 17 && o->precond != ALAMO_PRECOND_NONE)
        FAIL(ALAMO_ERR_ARG, "coarsest level has more than 17^3 nodes; use n_cell divisible by a larger power of 2");
    const int m = o->restart;
    if (m < 1 || m >= KRY_MAXV) FAIL(ALAMO_ERR_ARG, "restart must be in [1, 31]");
    if (!h->prepared) TRY(prepare(h));
    LevelBuf& l0 = h->L[0];
    const int64_t n = 3 * l0.g.tot;
    // the basis blocks' leading dimension: n rounded up to 64 elements, so every vector starts
    // 256-byte aligned and the Gram-Schmidt kernels take their 16-byte form
    const int64_t ld = (n + 63) / 64 * 64;
    TRY(krylov_scratch(h));
    TB *V, *Z;
    if constexpr (F32) {
        if (h->kry_mf < m) {
            TRY(dalloc(h, &h->kVf, (int64_t)(m + 1) * ld));
            TRY(dalloc(h, &h->kZf, (int64_t)m * ld));
            h->kry_mf = m;
        }
        V = h->kVf;
        Z = h->kZf;
    } else {
        if (h->kry_m < m) {
            TRY(dalloc(h, &h->kV, (int64_t)(m + 1) * ld));
            TRY(dalloc(h, &h->kZ, (int64_t)m * ld));
            h->kry_m = m;
        }
        V = h->kV;
        Z = h->kZ;
    }
    // FP64 true residual: its own buffer for FP32 inner vectors, else the first basis vector
    double* r = F32 ? h->sol_res : (double*)V;
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t0 = std::chrono::steady_clock::now();
    launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, sol, rhs, r, 0.0, h->stream);
    double r0 = 0, beta = 0, bn = 0;
    TRY(norms_dev<double>(h, r, n, &r0, &beta));
    TRY(norms_dev<double>(h, rhs, n, &bn, nullptr));
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * std::max(r0, bn));
    double rinf = r0;
    int it = 0, fallbacks = 0, reorths = 0;
    int kind = o->precond;
    double beta_prev = beta;
    double H[KRY_MAXV + 1][KRY_MAXV], cs[KRY_MAXV], sn[KRY_MAXV], gv[KRY_MAXV + 1], y[KRY_MAXV];
    double d1[KRY_MAXV], d2[KRY_MAXV], cf[KRY_MAXV];
    double sc[KRY_MAXV + 1];  // basis vector i = sc[i] x stored vector i
    if (o->verbose) fprintf(stderr, "alamo fgmres: iter 0 |r|inf %.6e |r|2 %.6e target %.3e\n", r0, beta, target);
    while (rinf > target && it < o->max_iter && std::isfinite(rinf) && beta > 0.0) {
        const double ratio = rinf / beta;
        k_scale<double, TB><<<gs_blocks(n), 256, 0, h->stream>>>(V, r, n, 1.0 / beta);
        sc[0] = 1.0;
        for (int i = 0; i <= m; i++) gv[i] = 0.0;
        gv[0] = beta;
        int jend = 0;
        for (int j = 0; j < m; j++) {
            TB* vj = V + (int64_t)j * ld;
            TB* zj = Z + (int64_t)j * ld;
            TB* w = V + (int64_t)(j + 1) * ld;
            int prc;
            if constexpr (F32) prc = precond_f32(h, kind, vj, zj);
            else prc = precond_apply(h, kind, vj, zj);
            if (prc == NET_NONFINITE) {
                // zj holds the plain cycle's result; the plain cycle takes over for good
                kind = ALAMO_PRECOND_MG;
                fallbacks++;
                if (o->verbose) fprintf(stderr, "alamo fgmres: learned stage not finite, falling back to the cycle\n");
            } else if (prc != ALAMO_OK) {
                return prc;
            }
            if constexpr (F32) launch_node<float>(h->sym, APPLY, l0.g, h->bc, l0.Cf, zj, nullptr, w, 0.f, h->stream);
            else launch_node<double>(h->sym, APPLY, l0.g, h->bc, l0.C, zj, nullptr, w, 0.0, h->stream);
            // classical Gram-Schmidt on the stored vectors (w' = A z'_j, z'_j = M(stored v_j)); w
            // is vector j + 1 of the basis block, so the same multi-dot pass returns |w'|^2 in
            // d1[j + 1]; the true Arnoldi quantities follow from the scales sc
            const bool lazy = g_kry_lazy && kind != ALAMO_PRECOND_LEARNED;
            const double sj = sc[j];
            TRY(multidot_t<TB>(h, w, V, n, j + 2, d1, ld));
            const double wn = std::sqrt(std::max(d1[j + 1], 0.0));
            for (int i = 0; i <= j; i++) {
                H[i][j] = sj * sc[i] * d1[i];
                cf[i] = sc[i] * sc[i] * d1[i];
            }
            TRY(upload_coef(h, cf, j + 1));
            double g = 0;  // |w''| of the stored vector
            TRY(multiaxpy_norm<TB>(h, w, V, n, j + 1, &g, ld));
            if (g < g_reorth_eta * wn) {
                // the first pass cancelled most of w: re-orthogonalise (CGS2)
                TRY(multidot_t<TB>(h, w, V, n, j + 1, d2, ld));
                for (int i = 0; i <= j; i++) {
                    H[i][j] += sj * sc[i] * d2[i];
                    cf[i] = sc[i] * sc[i] * d2[i];
                }
                TRY(upload_coef(h, cf, j + 1));
                TRY(multiaxpy_norm<TB>(h, w, V, n, j + 1, &g, ld));
                reorths++;
            }
            const double hn = sj * g;  // the true |w''|
            H[j + 1][j] = hn;
            if (lazy) {
                sc[j + 1] = g > 0.0 ? 1.0 / g : 1.0;
            } else {
                if (g > 0.0) k_scale<TB, TB><<<gs_blocks(n), 256, 0, h->stream>>>(w, w, n, 1.0 / g);
                sc[j + 1] = 1.0;
            }
            for (int i = 0; i < j; i++) {
                const double a = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
                H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
                H[i][j] = a;
            }
            const double den = std::hypot(H[j][j], H[j + 1][j]);
            cs[j] = den > 0 ? H[j][j] / den : 1.0;
            sn[j] = den > 0 ? H[j + 1][j] / den : 0.0;
            H[j][j] = den;
            H[j + 1][j] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            it++;
            jend = j + 1;
            if (o->verbose > 1) fprintf(stderr, "alamo fgmres: step %d est |r|2 %.6e\n", it, std::fabs(gv[j + 1]));
            // |g_{j+1}| estimates |r|_2; the stopping test is on |r|_inf, so scale by the inf/2 ratio
            // of the last true residual (the true residual at the restart decides convergence)
            if (std::fabs(gv[j + 1]) * ratio <= target || it >= o->max_iter || hn == 0.0 || !std::isfinite(hn))
                break;
        }
        for (int i = jend - 1; i >= 0; i--) {
            double s = gv[i];
            for (int k = i + 1; k < jend; k++) s -= H[i][k] * y[k];
            y[i] = H[i][i] != 0.0 ? s / H[i][i] : 0.0;
        }
        bool finite = true;
        for (int i = 0; i < jend; i++) finite = finite && std::isfinite(y[i]);
        if (!finite && kind == ALAMO_PRECOND_LEARNED) {
            // a non-finite learned stage leaves the iterate untouched; the plain cycle takes over
            kind = ALAMO_PRECOND_MG;
            fallbacks++;
            launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, sol, rhs, r, 0.0, h->stream);
            TRY(norms_dev<double>(h, r, n, &rinf, &beta));
            if (o->verbose) fprintf(stderr, "alamo fgmres: learned stage not finite, falling back to the cycle\n");
            beta_prev = beta;
            continue;
        }
        for (int i = 0; i < jend; i++) cf[i] = y[i] * sc[i];  // z_i = sc_i x stored z_i
        TRY(upload_coef(h, cf, jend));
        k_multiaxpy_add<TB><<<gs_blocks(n), 256, 0, h->stream>>>(sol, Z, n, ld, jend, h->kcoef);
        launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, sol, rhs, r, 0.0, h->stream);
        TRY(norms_dev<double>(h, r, n, &rinf, &beta));
        if (o->verbose)
            fprintf(stderr, "alamo fgmres: iter %d |r|inf %.6e |r|2 %.6e (%d second Gram-Schmidt passes)\n", it, rinf,
                    beta, reorths);
        // the plain cycle judges the learned stage: a restart that does not cut |r|_2 by the
        // stagnation ratio switches to the plain cycle for good
        if (kind == ALAMO_PRECOND_LEARNED && !(beta <= o->stagnation_ratio * beta_prev)) {
            kind = ALAMO_PRECOND_MG;
            fallbacks++;
            if (o->verbose) fprintf(stderr, "alamo fgmres: learned stage stagnated, falling back to the cycle\n");
        }
        beta_prev = beta;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t1 = std::chrono::steady_clock::now();
    const bool conv = rinf <= target;
    if (st) {
        st->iterations = it;
        st->converged = conv ? 1 : 0;
        st->initial_res_inf = r0;
        st->final_res_inf = rinf;
        st->final_res_l2 = beta;
        st->rhs_inf = bn;
        st->target = target;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->fallback_events = fallbacks;
    }
    if (!std::isfinite(rinf)) FAIL(ALAMO_ERR_NAN, "alamo_krylov_solve: residual is not finite");
    if (!conv) FAIL(ALAMO_ERR_NOCONV, "alamo_krylov_solve: max_iter reached before the tolerance");
    return ALAMO_OK;
}

// inner iterations in FP32 for mixed-precision handles with the cycle or learned preconditioner
// (ALAMO_KRY_F32=0 keeps them FP64, for comparisons)
static const bool g_kry_f32 = [] {
    const char* e = getenv("ALAMO_KRY_F32");
    return !(e && strcmp(e, "0") == 0);
}();

static int fgmres_impl(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                       const alamo_krylov_opts* o, alamo_solve_stats* st) {
    if (g_kry_f32 && h->d.precision == ALAMO_PREC_MIXED &&
        (o->precond == ALAMO_PRECOND_MG || o->precond == ALAMO_PRECOND_LEARNED))
        return fgmres_core<float>(h, sol, rhs, tol_rel, tol_abs, o, st);
    return fgmres_core<double>(h, sol, rhs, tol_rel, tol_abs, o, st);
}

// ===========================================================================
// C ABI
// ===========================================================================
template <int KIND>
static void launch_post(int sym, const Geom& g, const double* C, const double* u, double* out, int voigt,
                        const double* F0, hipStream_t s) {
    const unsigned nb = nblocks(g.tot);
    if (sym == ISO) k_post<ISO, KIND><<<nb, 256, 0, s>>>(g, C, u, out, voigt, F0);
    else if (sym == MAJORMINOR) k_post<MAJORMINOR, KIND><<<nb, 256, 0, s>>>(g, C, u, out, voigt, F0);
    else k_post<DIAGONAL, KIND><<<nb, 256, 0, s>>>(g, C, u, out, voigt, F0);
}

static int post_impl(alamo_mg* h, int amrlev, const double* u, double* out, int voigt, int kind, void* stream,
                     const double* F0 = nullptr) {
    if (!h || !u || !out) FAIL(ALAMO_ERR_ARG, "null argument");
    if (amrlev != 0) FAIL(ALAMO_ERR_ARG, "only amrlev 0 is supported by this handle");
    if (kind != POST_STRAIN && !h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    LevelBuf& l0 = h->L[0];
    const hipStream_t s = (hipStream_t)stream;
    if (kind == POST_STRAIN) launch_post<POST_STRAIN>(h->sym, l0.g, l0.C, u, out, voigt, nullptr, s);
    else if (kind == POST_STRESS) launch_post<POST_STRESS>(h->sym, l0.g, l0.C, u, out, voigt, nullptr, s);
    else if (kind == POST_ENERGY) launch_post<POST_ENERGY>(h->sym, l0.g, l0.C, u, out, 0, nullptr, s);
    else if (kind == POST_DW) launch_post<POST_DW>(h->sym, l0.g, l0.C, u, out, 0, F0, s);
    else launch_post<POST_W>(h->sym, l0.g, l0.C, u, out, 0, F0, s);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}
extern "C" {

const char* alamo_last_error(void) { return g_err.c_str(); }
int alamo_version(void) { return 1; }

int alamo_mg_desc_default(alamo_mg_desc* d) {
    if (!d) FAIL(ALAMO_ERR_ARG, "null desc");
    memset(d, 0, sizeof(*d));
    for (int f = 0; f < 26; f++)
        for (int c = 0; c < 3; c++) {
            d->bc_type[f][c] = ALAMO_BC_DISPLACEMENT;
            d->bc_val[f][c] = 0.0;
        }
    d->sym = ALAMO_SYM_ISOTROPIC;
    d->max_coarsening_level = -1;
    d->precision = ALAMO_PREC_FP64;
    d->cycle = ALAMO_CYCLE_F;
    d->pre_smooth = 2;
    d->post_smooth = 2;
    d->sweeps_per_smooth = 2;
    d->omega = 2.0 / 3.0;
    d->max_iter = 200;
    d->bottom_max_iter = 200;
    d->bottom_tol_rel = 1e-4;
    d->verbose = 0;
    d->smoother = ALAMO_SMOOTHER_JACOBI;
    d->coarse_op = ALAMO_COARSE_RESTRICT;
    d->fixed_iter = 0;
    d->bottom_solver = ALAMO_BOTTOM_BICGSTAB;
    d->max_fmg_iter = 100000000;
    d->sweeps_finest = 0;
    return ALAMO_OK;
}

int alamo_mg_create(const alamo_mg_desc* d, alamo_mg** out) {
    if (!d || !out) FAIL(ALAMO_ERR_ARG, "null argument");
    for (int q = 0; q < 3; q++) {
        if (d->n_cell[q] < 2) FAIL(ALAMO_ERR_ARG, "n_cell must be >= 2 in every direction");
        if (!(d->dx[q] > 0)) FAIL(ALAMO_ERR_ARG, "dx must be positive");
    }
    if (d->sym < 0 || d->sym > 2) FAIL(ALAMO_ERR_ARG, "unknown sym");
    for (int f = 0; f < 26; f++)
        for (int c = 0; c < 3; c++) {
            int t = d->bc_type[f][c];
            if (t < 0 || t > 4) FAIL(ALAMO_ERR_ARG, "unknown bc type");
            if (t == ALAMO_BC_PERIODIC) FAIL(ALAMO_ERR_ARG, "periodic boundaries are not supported by this handle");
        }
    if (d->sweeps_per_smooth < 1 || d->pre_smooth < 0 || d->post_smooth < 0) FAIL(ALAMO_ERR_ARG, "bad smoothing");
    if (d->fixed_iter < 0) FAIL(ALAMO_ERR_ARG, "fixed_iter must be >= 0");
    if (d->bottom_solver < ALAMO_BOTTOM_BICGSTAB || d->bottom_solver > ALAMO_BOTTOM_DENSE)
        FAIL(ALAMO_ERR_ARG, "unknown bottom_solver");
    if (d->sweeps_finest < 0) FAIL(ALAMO_ERR_ARG, "sweeps_finest must be >= 0");
    if (d->smoother != ALAMO_SMOOTHER_JACOBI && d->smoother != ALAMO_SMOOTHER_L1) FAIL(ALAMO_ERR_ARG, "unknown smoother");
    if (d->coarse_op != ALAMO_COARSE_RESTRICT && d->coarse_op != ALAMO_COARSE_GALERKIN)
        FAIL(ALAMO_ERR_ARG, "unknown coarse_op");
    double* scr = nullptr;  // the no-allocation entry point's scratch on this device
    TRY(scratch_ptr(&scr));
    alamo_mg* h = new alamo_mg();
    h->d = *d;
    h->sym = d->sym;
    h->nc = ncoef_of(d->sym);
    for (int f = 0; f < 26; f++)
        for (int c = 0; c < 3; c++) h->bc.t[f][c] = (int8_t)d->bc_type[f][c];
    // hierarchy: coarsen by 2 while every direction keeps >= 2 cells (AMReX mg_domain_min_width)
    int n[3] = {d->n_cell[0], d->n_cell[1], d->n_cell[2]};
    double dx[3] = {d->dx[0], d->dx[1], d->dx[2]};
    std::vector<Geom> gs;
    gs.push_back(make_geom(n, dx));
    while ((d->max_coarsening_level < 0 || (int)gs.size() <= d->max_coarsening_level) && n[0] % 2 == 0 &&
           n[1] % 2 == 0 && n[2] % 2 == 0 && n[0] / 2 >= 2 && n[1] / 2 >= 2 && n[2] / 2 >= 2) {
        for (int q = 0; q < 3; q++) { n[q] /= 2; dx[q] *= 2; }
        gs.push_back(make_geom(n, dx));
    }
    const bool mixed = d->precision == ALAMO_PREC_MIXED;
    h->L.resize(gs.size());
    int rc = ALAMO_OK;
    for (size_t l = 0; l < gs.size() && rc == ALAMO_OK; l++) {
        LevelBuf& lb = h->L[l];
        lb.g = gs[l];
        const int64_t v = 3 * lb.g.tot;
        rc |= dalloc(h, &lb.C, h->nc * lb.g.tot);
        rc |= dalloc(h, &lb.D, v);
        // FP64 cycle buffers: all levels for FP64 solves; the bottom level always (FP64 bottom solve)
        if (!mixed || l == gs.size() - 1 || l == 0) {
            rc |= dalloc(h, &lb.cor, v);
            rc |= dalloc(h, &lb.res, v);
            if (mixed && l == gs.size() - 1) rc |= dalloc(h, &lb.tmp, v);  // the smoother bottom's sweeps
            if (!mixed) {
                rc |= dalloc(h, &lb.rescor, v);
                rc |= dalloc(h, &lb.tmp, v);
                rc |= dalloc(h, &lb.hold, v);
            }
        }
        if (mixed) {
            rc |= dalloc(h, &lb.Cf, h->nc * lb.g.tot);
            rc |= dalloc(h, &lb.Df, v);
            rc |= dalloc(h, &lb.corf, v);
            rc |= dalloc(h, &lb.resf, v);
            rc |= dalloc(h, &lb.rescorf, v);
            rc |= dalloc(h, &lb.tmpf, v);
            rc |= dalloc(h, &lb.holdf, v);
        }
    }
    rc |= dalloc(h, &h->partial, 2 * h->nblk_red);
    rc |= dalloc(h, &h->red, 4);
    // BiCGStab bottom workspace: only when the coarsest level is small enough for a solve (<= 17^3)
    if (gs.back().tot <= 17 * 17 * 17) rc |= dalloc(h, &h->bottom_work, 8 * 3 * gs.back().tot);
    rc |= dalloc(h, &h->bottom_iters, 1);
    if (mixed) rc |= dalloc(h, &h->sol_res, 3 * gs[0].tot);
    if (rc == ALAMO_OK && hipHostMalloc((void**)&h->red_host, (KRY_MAXV + 4) * sizeof(double)) != hipSuccess) {
        g_err = "hipHostMalloc failed";
        rc = ALAMO_ERR_HIP;
    }
    if (rc != ALAMO_OK) {
        std::string e = g_err;
        alamo_mg_destroy(h);
        g_err = e;
        return ALAMO_ERR_HIP;
    }
    *out = h;
    return ALAMO_OK;
}

int alamo_mg_destroy(alamo_mg* h) {
    if (!h) return ALAMO_OK;
    (void)hipDeviceSynchronize();
    for (void* p : h->allocs) (void)hipFree(p);
    if (h->red_host) (void)hipHostFree(h->red_host);
    delete h;
    return ALAMO_OK;
}

int alamo_mg_set_stream(alamo_mg* h, void* stream) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    h->stream = (hipStream_t)stream;
    return ALAMO_OK;
}

int alamo_mg_num_levels(alamo_mg* h, int* nlev) {
    if (!h || !nlev) FAIL(ALAMO_ERR_ARG, "null argument");
    *nlev = (int)h->L.size();
    return ALAMO_OK;
}

int alamo_mg_level_ncell(alamo_mg* h, int mglev, int* ncell3) {
    if (!h || !ncell3) FAIL(ALAMO_ERR_ARG, "null argument");
    if (mglev < 0 || mglev >= (int)h->L.size()) FAIL(ALAMO_ERR_ARG, "mglev out of range");
    for (int d = 0; d < 3; d++) ncell3[d] = h->L[mglev].g.n[d];
    return ALAMO_OK;
}

int alamo_mg_update_opts(alamo_mg* h, const alamo_mg_desc* d) {
    if (!h || !d) FAIL(ALAMO_ERR_ARG, "null argument");
    const alamo_mg_desc& c = h->d;
    bool same = d->sym == c.sym && d->precision == c.precision && d->max_coarsening_level == c.max_coarsening_level;
    for (int q = 0; q < 3; q++) same = same && d->n_cell[q] == c.n_cell[q] && d->dx[q] == c.dx[q];
    for (int f = 0; f < 26; f++)
        for (int q = 0; q < 3; q++)
            same = same && d->bc_type[f][q] == c.bc_type[f][q] && d->bc_val[f][q] == c.bc_val[f][q];
    if (!same) FAIL(ALAMO_ERR_ARG, "alamo_mg_update_opts: geometry, boundary, symmetry or precision changed");
    if (d->sweeps_per_smooth < 1 || d->pre_smooth < 0 || d->post_smooth < 0) FAIL(ALAMO_ERR_ARG, "bad smoothing");
    if (d->cycle != ALAMO_CYCLE_V && d->cycle != ALAMO_CYCLE_F) FAIL(ALAMO_ERR_ARG, "unknown cycle");
    if (d->fixed_iter < 0) FAIL(ALAMO_ERR_ARG, "fixed_iter must be >= 0");
    if (d->bottom_solver < ALAMO_BOTTOM_BICGSTAB || d->bottom_solver > ALAMO_BOTTOM_DENSE)
        FAIL(ALAMO_ERR_ARG, "unknown bottom_solver");
    if (d->sweeps_finest < 0) FAIL(ALAMO_ERR_ARG, "sweeps_finest must be >= 0");
    if (d->smoother != ALAMO_SMOOTHER_JACOBI && d->smoother != ALAMO_SMOOTHER_L1) FAIL(ALAMO_ERR_ARG, "unknown smoother");
    if (d->coarse_op != ALAMO_COARSE_RESTRICT && d->coarse_op != ALAMO_COARSE_GALERKIN)
        FAIL(ALAMO_ERR_ARG, "unknown coarse_op");
    // the dense inverse, the smoother diagonals and the Galerkin stencils are built at prepare
    if (d->bottom_solver != h->d.bottom_solver || d->smoother != h->d.smoother || d->coarse_op != h->d.coarse_op)
        h->prepared = false;
    h->d.smoother = d->smoother;
    h->d.coarse_op = d->coarse_op;
    h->d.bottom_solver = d->bottom_solver;
    h->d.max_fmg_iter = d->max_fmg_iter;
    h->d.sweeps_finest = d->sweeps_finest;
    h->d.cycle = d->cycle;
    h->d.pre_smooth = d->pre_smooth;
    h->d.post_smooth = d->post_smooth;
    h->d.sweeps_per_smooth = d->sweeps_per_smooth;
    h->d.omega = d->omega;
    h->d.max_iter = d->max_iter;
    h->d.bottom_max_iter = d->bottom_max_iter;
    h->d.bottom_tol_rel = d->bottom_tol_rel;
    h->d.fixed_iter = d->fixed_iter;
    h->d.verbose = d->verbose;
    return ALAMO_OK;
}

int alamo_mg_level_layout(alamo_mg* h, int mglev, int64_t* pitch3) {
    if (!h || !pitch3) FAIL(ALAMO_ERR_ARG, "null argument");
    if (mglev < 0 || mglev >= (int)h->L.size()) FAIL(ALAMO_ERR_ARG, "mglev out of range");
    const Geom& g = h->L[mglev].g;
    pitch3[0] = g.sy;
    pitch3[1] = g.sz;
    pitch3[2] = g.tot;
    return ALAMO_OK;
}

static int set_model_impl(alamo_mg* h, int amrlev, const double* C, int layout, bool host) {
    if (!h || !C) FAIL(ALAMO_ERR_ARG, "null argument");
    if (amrlev != 0) FAIL(ALAMO_ERR_ARG, "only amrlev 0 is supported by this handle");
    if (layout != ALAMO_LAYOUT_AOS && layout != ALAMO_LAYOUT_SOA) FAIL(ALAMO_ERR_ARG, "unknown layout");
    LevelBuf& l0 = h->L[0];
    const int64_t n = h->nc * l0.g.tot;
    const double* src = C;
    double* staged = nullptr;
    const int64_t Vn = (int64_t)l0.g.N[0] * l0.g.N[1] * l0.g.N[2];
    if (host) {
        // host arrays are compact: stage them and scatter into the level's layout
        HIPCHK(hipMalloc(&staged, h->nc * Vn * sizeof(double)));
        HIPCHK(hipMemcpy(staged, C, h->nc * Vn * sizeof(double), hipMemcpyHostToDevice));
        k_compact_to_soa<double><<<gs_blocks(Vn), 256, 0, h->stream>>>(l0.g, staged, l0.C, h->nc,
                                                                       layout == ALAMO_LAYOUT_AOS);
    } else if (layout == ALAMO_LAYOUT_SOA) {
        HIPCHK(hipMemcpyAsync(l0.C, src, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    } else {
        k_aos_to_soa<double><<<gs_blocks(l0.g.tot), 256, 0, h->stream>>>(src, l0.C, l0.g.tot, h->nc);
    }
    HIPCHK(hipGetLastError());
    if (staged) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipFree(staged));
    }
    h->model_set = true;
    h->prepared = false;
    h->model_version = ++g_model_version;
    for (LevelBuf& lb : h->L) {  // the gradient contractions follow the model at the next preparation
        lb.g.mmd = nullptr;
        lb.g.mmdf = nullptr;
    }
    l0.g.gid = nullptr;  // a full model replaces a per-grain one
    l0.g.gtab = nullptr;
    l0.g.gtabf = nullptr;
    l0.g.ngrains = 0;
    return ALAMO_OK;
}

int alamo_mg_set_model(alamo_mg* h, int amrlev, const double* C, int layout) {
    return set_model_impl(h, amrlev, C, layout, false);
}
int alamo_mg_set_model_host(alamo_mg* h, int amrlev, const double* C, int layout) {
    return set_model_impl(h, amrlev, C, layout, true);
}

int alamo_mg_set_model_grains_host(alamo_mg* h, const uint16_t* gid, const double* table, int ngrains) {
    if (!h || !gid || !table || ngrains <= 0) FAIL(ALAMO_ERR_ARG, "bad argument");
    if (h->sym != ALAMO_SYM_MAJORMINOR) FAIL(ALAMO_ERR_ARG, "grain tables need sym MAJORMINOR");
    const Geom& g0 = h->L[0].g;
    const int64_t tot = (int64_t)g0.N[0] * g0.N[1] * g0.N[2];  // compact host arrays
    std::vector<double> soa((size_t)21 * tot);
    for (int64_t m = 0; m < tot; m++) {
        const int gi = gid[m];
        if (gi >= ngrains) FAIL(ALAMO_ERR_ARG, "grain id out of range");
        for (int q = 0; q < 21; q++) soa[(size_t)q * tot + m] = table[21 * gi + q];
    }
    // the full records serve the boundary rows and the coefficient restriction
    TRY(set_model_impl(h, 0, soa.data(), ALAMO_LAYOUT_SOA, true));
    // the interior rows of the finest level read the grain ids and the table
    LevelBuf& l0 = h->L[0];
    if (!l0.gid) TRY(dalloc(h, &l0.gid, l0.g.tot));
    if (l0.ngrains_cap < ngrains) {
        TRY(dalloc(h, &l0.gtab, (int64_t)21 * ngrains));
        TRY(dalloc(h, &l0.gtabf, (int64_t)21 * ngrains));
        l0.ngrains_cap = ngrains;
    }
    uint16_t* staged = nullptr;
    HIPCHK(hipMalloc(&staged, tot * sizeof(uint16_t)));
    HIPCHK(hipMemcpy(staged, gid, tot * sizeof(uint16_t), hipMemcpyHostToDevice));
    k_scatter_u16<<<gs_blocks(tot), 256, 0, h->stream>>>(l0.g, staged, l0.gid);
    HIPCHK(hipMemcpyAsync(l0.gtab, table, (size_t)21 * ngrains * sizeof(double), hipMemcpyHostToDevice, h->stream));
    k_convert<double, float><<<1, 256, 0, h->stream>>>(l0.gtab, l0.gtabf, (int64_t)21 * ngrains);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(staged));
    l0.g.gid = l0.gid;
    l0.g.gtab = l0.gtab;
    l0.g.gtabf = l0.gtabf;
    l0.g.ngrains = ngrains;
    return ALAMO_OK;
}

static int set_psi_impl(alamo_mg* h, int amrlev, const double* psi, double psi_small, bool host) {
    if (!h || !psi) FAIL(ALAMO_ERR_ARG, "null argument");
    if (amrlev != 0) FAIL(ALAMO_ERR_ARG, "only amrlev 0 is supported by this handle");
    if (!(psi_small >= 0.0 && psi_small < 1.0)) FAIL(ALAMO_ERR_ARG, "psi_small must be in [0, 1)");
    const bool mixed = h->d.precision == ALAMO_PREC_MIXED;
    for (LevelBuf& lb : h->L) {
        const int64_t nc = (int64_t)lb.g.n[0] * lb.g.n[1] * lb.g.n[2];
        if (!lb.psi) TRY(dalloc(h, &lb.psi, nc));
        if (mixed && !lb.psif) TRY(dalloc(h, &lb.psif, nc));
    }
    LevelBuf& l0 = h->L[0];
    const int64_t nc0 = (int64_t)l0.g.n[0] * l0.g.n[1] * l0.g.n[2];
    HIPCHK(hipMemcpyAsync(l0.psi, psi, nc0 * sizeof(double), host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                          h->stream));
    if (mixed) k_convert<double, float><<<gs_blocks(nc0), 256, 0, h->stream>>>(l0.psi, l0.psif, nc0);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->psi_set = true;
    h->psi_small = psi_small;
    // the finest level applies psi at once; the coarse levels get theirs at prepare
    l0.g.psi = l0.psi;
    l0.g.psif = l0.psif;
    l0.g.psi_small = psi_small;
    h->prepared = false;
    h->model_version = ++g_model_version;
    return ALAMO_OK;
}

int alamo_mg_set_psi(alamo_mg* h, int amrlev, const double* psi, double psi_small) {
    return set_psi_impl(h, amrlev, psi, psi_small, false);
}
int alamo_mg_set_psi_host(alamo_mg* h, int amrlev, const double* psi, double psi_small) {
    return set_psi_impl(h, amrlev, psi, psi_small, true);
}

int alamo_mg_trac_xhi(alamo_mg* h, const double* u, const double* F0, double* trac3) {
    if (!h || !u || !trac3) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    TRY(krylov_scratch(h));
    LevelBuf& l0 = h->L[0];
    const int64_t ncell = (int64_t)l0.g.n[1] * l0.g.n[2];
    const int nb = (int)std::min<int64_t>(KRY_DOT_BLOCKS, (ncell + RED_BLOCK - 1) / RED_BLOCK);
    if (h->sym == ISO) k_trac_xhi<ISO><<<nb, RED_BLOCK, 0, h->stream>>>(l0.g, l0.C, F0, u, h->kpart);
    else if (h->sym == MAJORMINOR) k_trac_xhi<MAJORMINOR><<<nb, RED_BLOCK, 0, h->stream>>>(l0.g, l0.C, F0, u, h->kpart);
    else k_trac_xhi<DIAGONAL><<<nb, RED_BLOCK, 0, h->stream>>>(l0.g, l0.C, F0, u, h->kpart);
    k_sum_rows<<<3, RED_BLOCK, 0, h->stream>>>(h->kpart, nb, h->kdots);
    HIPCHK(hipMemcpyAsync(h->red_host, h->kdots, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int a = 0; a < 3; a++) trac3[a] = h->red_host[a];
    return ALAMO_OK;
}

int alamo_mg_strain(alamo_mg* h, int amrlev, const double* u, double* eps, int voigt, void* stream) {
    return post_impl(h, amrlev, u, eps, voigt, POST_STRAIN, stream);
}
int alamo_mg_stress(alamo_mg* h, int amrlev, const double* u, double* sigma, int voigt, void* stream) {
    return post_impl(h, amrlev, u, sigma, voigt, POST_STRESS, stream);
}
int alamo_mg_energy(alamo_mg* h, int amrlev, const double* u, double* energy, void* stream) {
    return post_impl(h, amrlev, u, energy, 0, POST_ENERGY, stream);
}
int alamo_newton_dw(alamo_mg* h, int amrlev, const double* u, const double* F0, double* dw, void* stream) {
    return post_impl(h, amrlev, u, dw, 0, POST_DW, stream, F0);
}
int alamo_newton_w(alamo_mg* h, int amrlev, const double* u, const double* F0, double* w, void* stream) {
    return post_impl(h, amrlev, u, w, 0, POST_W, stream, F0);
}

int alamo_mg_set_uniform(alamo_mg* h, int uniform) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    h->uniform = uniform != 0;
    for (LevelBuf& lb : h->L) lb.g.uniform = h->uniform ? 1 : 0;
    return ALAMO_OK;
}

int alamo_mg_error0x(alamo_mg* h, int amrlev, int mglev, double* R0x, const double* x, void* stream) {
    TRY(check_level(h, amrlev, mglev));
    if (!R0x || !x) FAIL(ALAMO_ERR_ARG, "null field");
    LevelBuf& lb = h->L[mglev];
    const int64_t n = 3 * lb.g.tot;
    const hipStream_t s = (hipStream_t)stream;
    double* d0x = nullptr;
    HIPCHK(hipMallocAsync((void**)&d0x, n * sizeof(double), s));
    // D0x = x / D, R0x = A D0x, R0x = x - R0x
    HIPCHK(hipMemcpyAsync(d0x, x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    k_divide<<<gs_blocks(n), 256, 0, s>>>(d0x, lb.D, n);
    level_op<double>(h, mglev, RESID, d0x, x, R0x, 0.0, s);
    HIPCHK(hipFreeAsync(d0x, s));
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_blend_models(int64_t nnodes, int nmodels, int nrec, const double* eta, const double* table, double* out,
                       int layout, void* stream) {
    if (!eta || !table || !out || nnodes < 1 || nmodels < 1 || nrec < 1) FAIL(ALAMO_ERR_ARG, "bad argument");
    if (layout != ALAMO_LAYOUT_AOS && layout != ALAMO_LAYOUT_SOA) FAIL(ALAMO_ERR_ARG, "unknown layout");
    hipStream_t s = (hipStream_t)stream;
    double* dt = nullptr;
    const size_t tb = (size_t)nmodels * nrec * sizeof(double);
    HIPCHK(hipMallocAsync((void**)&dt, tb, s));
    HIPCHK(hipMemcpyAsync(dt, table, tb, hipMemcpyHostToDevice, s));
    k_blend<<<gs_blocks(nnodes), 256, 0, s>>>(nnodes, nmodels, nrec, eta, dt, out, layout == ALAMO_LAYOUT_AOS);
    HIPCHK(hipGetLastError());
    HIPCHK(hipFreeAsync(dt, s));
    // the host table may be reused by the caller as soon as this returns
    HIPCHK(hipStreamSynchronize(s));
    return ALAMO_OK;
}

int alamo_mg_prepare(alamo_mg* h) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    return prepare(h);
}

int alamo_mg_get_coeff(alamo_mg* h, int mglev, double* C_soa) {
    TRY(check_level(h, 0, mglev));
    LevelBuf& lb = h->L[mglev];
    HIPCHK(hipMemcpyAsync(C_soa, lb.C, h->nc * lb.g.tot * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return ALAMO_OK;
}

int alamo_mg_fapply(alamo_mg* h, int amrlev, int mglev, double* out, const double* in, void* stream) {
    TRY(check_level(h, amrlev, mglev));
    LevelBuf& lb = h->L[mglev];
    (void)lb;
    level_op<double>(h, mglev, APPLY, in, nullptr, out, 0.0, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_diagonal(alamo_mg* h, int amrlev, int mglev, double* diag, void* stream) {
    TRY(check_level(h, amrlev, mglev));
    LevelBuf& lb = h->L[mglev];
    (void)lb;
    level_op<double>(h, mglev, DIAG, nullptr, nullptr, diag, 0.0, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_normalize(alamo_mg* h, int amrlev, int mglev, double* x, void* stream) {
    TRY(check_level(h, amrlev, mglev));
    if (!h->prepared) FAIL(ALAMO_ERR_STATE, "Operator::Diagonal() must be called before using normalize");
    if (!x) FAIL(ALAMO_ERR_ARG, "null field");
    LevelBuf& lb = h->L[mglev];
    k_divide<<<gs_blocks(3 * lb.g.tot), 256, 0, (hipStream_t)stream>>>(x, lb.D, 3 * lb.g.tot);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_fsmooth(alamo_mg* h, int amrlev, int mglev, double* x, const double* b, void* stream) {
    TRY(check_level(h, amrlev, mglev));
    LevelBuf& lb = h->L[mglev];
    double* tmp = nullptr;
    HIPCHK(hipMallocAsync((void**)&tmp, 3 * lb.g.tot * sizeof(double), (hipStream_t)stream));
    double* a = x;
    double* c = tmp;
    for (int s = 0; s < h->d.sweeps_per_smooth; s++) {
        level_op<double>(h, mglev, JACOBI, a, b, c, h->d.omega, (hipStream_t)stream);
        std::swap(a, c);
    }
    if (a != x) HIPCHK(hipMemcpyAsync(x, a, 3 * lb.g.tot * sizeof(double), hipMemcpyDeviceToDevice,
                                      (hipStream_t)stream));
    HIPCHK(hipFreeAsync(tmp, (hipStream_t)stream));
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_stream_probe(const double* a, double* o, int64_t n, void* stream) {
    if (!a || !o) FAIL(ALAMO_ERR_ARG, "null argument");
    if (n < 1) FAIL(ALAMO_ERR_ARG, "n must be positive");
    k_stream_5r3w<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, a, o);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_fsmooth_noalloc(alamo_mg* h, int amrlev, int mglev, double* x, const double* b, void* stream) {
    if (!h || !x || !b) FAIL(ALAMO_ERR_ARG, "null argument");
    if (amrlev != 0) FAIL(ALAMO_ERR_ARG, "only amrlev 0 is supported by this handle");
    if (mglev < 0 || mglev >= (int)h->L.size()) FAIL(ALAMO_ERR_ARG, "mglev out of range");
    // preparing allocates (diagonals, stencils, dense bottom): the caller prepares beforehand
    if (!h->prepared)
        FAIL(ALAMO_ERR_STATE, "alamo_mg_fsmooth_noalloc: the handle is not prepared; call alamo_mg_prepare first");
    LevelBuf& lb = h->L[mglev];
    const int64_t n = 3 * lb.g.tot;
    if (n > (int64_t)ALAMO_NOALLOC_MAX_DOUBLES)
        FAIL(ALAMO_ERR_ARG, "alamo_mg_fsmooth_noalloc: level " + std::to_string(mglev) + " has " +
                                std::to_string(lb.g.tot) + " nodes; the scratch holds 3 x " +
                                std::to_string((int64_t)ALAMO_NOALLOC_MAX_DOUBLES / 3) + " values");
    double* tmp = nullptr;
    TRY(scratch_ptr(&tmp));
    const hipStream_t s = (hipStream_t)stream;
    // every sweep writes all 3 n entries of its output, so nothing in the scratch needs clearing
    double* a = x;
    double* c = tmp;
    for (int sw = 0; sw < h->d.sweeps_per_smooth; sw++) {
        level_op<double>(h, mglev, JACOBI, a, b, c, h->d.omega, s);
        std::swap(a, c);
    }
    if (a != x) k_convert<double, double><<<gs_blocks(n), 256, 0, s>>>(a, x, n);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_residual(alamo_mg* h, int amrlev, int mglev, double* r, const double* x, const double* b,
                      void* stream) {
    TRY(check_level(h, amrlev, mglev));
    LevelBuf& lb = h->L[mglev];
    (void)lb;
    level_op<double>(h, mglev, RESID, x, b, r, 0.0, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_restrict(alamo_mg* h, int amrlev, int cmglev, double* crse, const double* fine, void* stream) {
    TRY(check_level(h, amrlev, cmglev));
    if (cmglev < 1) FAIL(ALAMO_ERR_ARG, "cmglev must be >= 1");
    LevelBuf& lc = h->L[cmglev];
    LevelBuf& lf = h->L[cmglev - 1];
    k_restrict<double><<<nblocks(lc.g.tot), 256, 0, (hipStream_t)stream>>>(lc.g, lf.g, 3, fine, crse);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_interpolate(alamo_mg* h, int amrlev, int fmglev, double* fine, const double* crse, void* stream) {
    TRY(check_level(h, amrlev, fmglev));
    if (fmglev >= (int)h->L.size() - 1) FAIL(ALAMO_ERR_ARG, "fmglev must have a coarser level");
    LevelBuf& lc = h->L[fmglev + 1];
    LevelBuf& lf = h->L[fmglev];
    launch_interp<double>(lc.g, lf.g, crse, fine, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_norms(alamo_mg* h, int mglev, const double* x, double* norm_inf, double* norm_l2, void* stream) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    if (mglev < 0 || mglev >= (int)h->L.size()) FAIL(ALAMO_ERR_ARG, "mglev out of range");
    hipStream_t keep = h->stream;
    h->stream = (hipStream_t)stream;
    int rc = norms_dev<double>(h, x, 3 * h->L[mglev].g.tot, norm_inf, norm_l2);
    h->stream = keep;
    return rc;
}
int alamo_mg_norm_inf(alamo_mg* h, int mglev, const double* x, double* norm, void* stream) {
    return alamo_mg_norms(h, mglev, x, norm, nullptr, stream);
}
int alamo_mg_norm_l2(alamo_mg* h, int mglev, const double* x, double* norm, void* stream) {
    return alamo_mg_norms(h, mglev, x, nullptr, norm, stream);
}

int alamo_bc_init(alamo_mg* h, double* b, int homogeneous);
int alamo_newton_rhs(alamo_mg* h, const double* u, const double* b, const double* F0, double* rhs);

int alamo_krylov_opts_default(alamo_krylov_opts* o) {
    if (!o) FAIL(ALAMO_ERR_ARG, "null opts");
    o->restart = 16;
    o->precond = ALAMO_PRECOND_MG;
    o->max_iter = 400;
    o->stagnation_ratio = 0.9;
    o->verbose = 0;
    return ALAMO_OK;
}

int alamo_krylov_solve(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                       const alamo_krylov_opts* opts, alamo_solve_stats* stats) {
    if (!h || !sol || !rhs) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    alamo_krylov_opts o;
    if (opts) o = *opts;
    else alamo_krylov_opts_default(&o);
    if (o.precond != ALAMO_PRECOND_NONE && o.precond != ALAMO_PRECOND_MG && o.precond != ALAMO_PRECOND_LEARNED)
        FAIL(ALAMO_ERR_ARG, "unknown preconditioner");
    if (o.precond == ALAMO_PRECOND_LEARNED && !h->net) FAIL(ALAMO_ERR_STATE, "no network attached (alamo_mg_set_net)");
    return fgmres_impl(h, sol, rhs, tol_rel, tol_abs, &o, stats);
}

int alamo_learned_solve(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                        const alamo_krylov_opts* opts, alamo_solve_stats* stats) {
    alamo_krylov_opts o;
    if (opts) o = *opts;
    else alamo_krylov_opts_default(&o);
    o.precond = ALAMO_PRECOND_LEARNED;
    return alamo_krylov_solve(h, sol, rhs, tol_rel, tol_abs, &o, stats);
}

int alamo_mg_precond(alamo_mg* h, int kind, const double* r, double* z) {
    if (!h || !r || !z) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    if (kind != ALAMO_PRECOND_MG && kind != ALAMO_PRECOND_LEARNED) FAIL(ALAMO_ERR_ARG, "unknown preconditioner");
    if (kind == ALAMO_PRECOND_LEARNED && !h->net) FAIL(ALAMO_ERR_STATE, "no network attached (alamo_mg_set_net)");
    if (h->L.back().g.tot > 17 * 17 * 17)
        FAIL(ALAMO_ERR_ARG, "coarsest level has more than 17^3 nodes; use n_cell divisible by a larger power of 2");
    if (!h->prepared) TRY(prepare(h));
    const int rc = precond_apply(h, kind, r, z);
    HIPCHK(hipStreamSynchronize(h->stream));
    if (rc == NET_NONFINITE) FAIL(ALAMO_ERR_NAN, "network output is not finite (z holds the plain cycle's result)");
    return rc;
}

// (Re)build the explicit level-0 row table from the current coefficients and boundary table.
static int row_table(alamo_mg* h) {
    if (h->psi_set) FAIL(ALAMO_ERR_ARG, "the transposed operator has no psi terms");
    const Geom& g = h->L[0].g;
    if (!h->rowtab) TRY(dalloc(h, &h->rowtab, 243 * g.tot));
    const unsigned nb = nblocks(g.tot);
    if (h->sym == ISO) k_row_table<ISO><<<nb, 256, 0, h->stream>>>(g, h->bc, h->L[0].C, h->rowtab);
    else if (h->sym == MAJORMINOR) k_row_table<MAJORMINOR><<<nb, 256, 0, h->stream>>>(g, h->bc, h->L[0].C, h->rowtab);
    else k_row_table<DIAGONAL><<<nb, 256, 0, h->stream>>>(g, h->bc, h->L[0].C, h->rowtab);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_mg_apply_transpose(alamo_mg* h, const double* r, double* out) {
    if (!h || !r || !out) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    if (r == out) FAIL(ALAMO_ERR_ARG, "r and out must not alias");
    TRY(row_table(h));
    const Geom& g = h->L[0].g;
    k_row_apply_t<<<nblocks(g.tot), 256, 0, h->stream>>>(g, h->rowtab, r, out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return ALAMO_OK;
}

int alamo_linear_solve(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                       alamo_solve_stats* stats) {
    if (!h || !sol || !rhs) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    return solve_impl(h, sol, rhs, tol_rel, tol_abs, stats);
}

int alamo_newton_solve(alamo_mg* h, double* u, const double* b, const double* F0, int nriters, double nrtolerance,
                       double tol_rel, double tol_abs, alamo_solve_stats* stats, double* relnorm) {
    if (!h || !u || !b) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    if (nriters < 0) FAIL(ALAMO_ERR_ARG, "nriters must be >= 0");
    const int64_t n = 3 * h->L[0].g.tot;
    double *dsol = nullptr, *rhs = nullptr;
    HIPCHK(hipMallocAsync((void**)&dsol, n * sizeof(double), h->stream));
    HIPCHK(hipMallocAsync((void**)&rhs, n * sizeof(double), h->stream));
    vfill<double>(dsol, n, 0.0, h->stream);  // zeroed once (Newton.H:313)
    vfill<double>(rhs, n, 0.0, h->stream);
    int rc = ALAMO_OK;
    double rel = 0.0;
    for (int it = 0; it < nriters && rc == ALAMO_OK; it++) {
        rc = alamo_newton_rhs(h, u, b, F0, rhs);
        if (rc == ALAMO_OK) rc = solve_impl(h, dsol, rhs, tol_rel, tol_abs, stats);
        if (rc != ALAMO_OK) break;
        double cor = 0.0;
        rc = norms_dev<double>(h, dsol, n, &cor, nullptr);  // max |du|; solnorm stays 0 (Newton.H:330-346)
        if (rc != ALAMO_OK) break;
        k_axpy<double, double><<<gs_blocks(n), 256, 0, h->stream>>>(u, dsol, n, 1.0);
        if (cor < nrtolerance) {
            rel = cor;
            break;
        }
    }
    const std::string err = g_err;
    (void)hipFreeAsync(dsol, h->stream);
    (void)hipFreeAsync(rhs, h->stream);
    if (hipStreamSynchronize(h->stream) != hipSuccess && rc == ALAMO_OK) FAIL(ALAMO_ERR_HIP, "hipStreamSynchronize failed");
    g_err = err;
    if (relnorm) *relnorm = rel;
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Boundary values and the Newton residual-form right-hand side.
// ---------------------------------------------------------------------------
struct BCVals {
    double v[26][3];
};

__global__ void k_bc_init(Geom g, BCTable bc, BCVals vals, int homogeneous, double* __restrict__ b) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int f = classify(i, j, k, hi);
    if (f == INT) return;
    for (int c = 0; c < 3; c++) {
        const int t = bc.t[f][c];
        if (t == PERIODIC) continue;
        b[c * g.tot + m] = (homogeneous && t == DISPLACEMENT) ? 0.0 : vals.v[f][c];
    }
}

// DW = C : (grad u - F0) with GetStencil gradients (Newton.H:157-184)
template <int SYM>
__global__ void k_newton_dw(Geom g, const double* __restrict__ C, const double* __restrict__ F0,
                            const double* __restrict__ u, double* __restrict__ DW) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int idx[3] = {i, j, k};
    const int64_t st[3] = {1, g.sy, g.sz};
    double G[3][3], S[3][3], c[NCoef<SYM>::value];
    for (int p = 0; p < 3; p++)
        for (int d = 0; d < 3; d++) {
            G[p][d] = d1(u + p * g.tot, m, st[d], idx[d], hi[d], g.dx[d]);
            if (F0) G[p][d] -= F0[9 * m + 3 * p + d];
        }
    load_coef<double, SYM>(C, g.tot, m, c);
    c_dot<SYM>(c, G, S);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) DW[(3 * a + b) * g.tot + m] = S[a][b];
}

// rhs = b - div(DW) inside, b - BC(u, grad u, DW) on the boundary (Newton.H:199-231)
__global__ void k_newton_rhs(Geom g, BCTable bc, const double* __restrict__ DW, const double* __restrict__ u,
                             const double* __restrict__ b, double* __restrict__ rhs) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int idx[3] = {i, j, k};
    const int64_t st[3] = {1, g.sy, g.sz};
    double f[3];
    if (i == 0 || j == 0 || k == 0 || i == hi[0] || j == hi[1] || k == hi[2]) {
        double G[3][3], S[3][3], uu[3];
        for (int p = 0; p < 3; p++) {
            uu[p] = u[p * g.tot + m];
            for (int d = 0; d < 3; d++) G[p][d] = d1(u + p * g.tot, m, st[d], idx[d], hi[d], g.dx[d]);
        }
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) S[a][c] = DW[(3 * a + c) * g.tot + m];
        bc_row(bc, classify(i, j, k, hi), uu, G, S, f);
    } else {
        for (int a = 0; a < 3; a++) {
            double s = 0.0;
            for (int d = 0; d < 3; d++) {
                const double* P = DW + (3 * a + d) * g.tot;
                s += (P[m + st[d]] - P[m - st[d]]) / 2. / g.dx[d];
            }
            f[a] = s;
        }
        if (g.psi) {
            // divdw *= CellToNodeAverage(psi); divdw += dw grad psi (Newton.H:258-263, no psi_small);
            // the boundary rows above keep psiavg = 1 (the node lies on the box's edge, :247-252)
            const double pm = psi_node_mean(g.psi, g, i, j, k);
            double gp[3];
            psi_grad(g.psi, g, i, j, k, gp);
            for (int a = 0; a < 3; a++) {
                double s = 0.0;
                for (int d = 0; d < 3; d++) s += DW[(3 * a + d) * g.tot + m] * gp[d];
                f[a] = f[a] * pm + s;
            }
        }
    }
    for (int p = 0; p < 3; p++) rhs[p * g.tot + m] = b[p * g.tot + m] - f[p];
}

extern "C" int alamo_bc_init(alamo_mg* h, double* b, int homogeneous) {
    if (!h || !b) FAIL(ALAMO_ERR_ARG, "null argument");
    BCVals v;
    for (int f = 0; f < 26; f++)
        for (int c = 0; c < 3; c++) v.v[f][c] = h->d.bc_val[f][c];
    const Geom& g = h->L[0].g;
    k_bc_init<<<nblocks(g.tot), 256, 0, h->stream>>>(g, h->bc, v, homogeneous, b);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

extern "C" int alamo_newton_rhs(alamo_mg* h, const double* u, const double* b, const double* F0, double* rhs) {
    if (!h || !u || !b || !rhs) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    const Geom& g = h->L[0].g;
    double* DW = nullptr;
    HIPCHK(hipMallocAsync((void**)&DW, 9 * g.tot * sizeof(double), h->stream));
    if (h->sym == ISO) k_newton_dw<ISO><<<nblocks(g.tot), 256, 0, h->stream>>>(g, h->L[0].C, F0, u, DW);
    else if (h->sym == MAJORMINOR)
        k_newton_dw<MAJORMINOR><<<nblocks(g.tot), 256, 0, h->stream>>>(g, h->L[0].C, F0, u, DW);
    else k_newton_dw<DIAGONAL><<<nblocks(g.tot), 256, 0, h->stream>>>(g, h->L[0].C, F0, u, DW);
    k_newton_rhs<<<nblocks(g.tot), 256, 0, h->stream>>>(g, h->bc, DW, u, b, rhs);
    HIPCHK(hipFreeAsync(DW, h->stream));
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

// ===========================================================================
// Learned correction network: C ABI
// ===========================================================================
namespace {
__global__ void k_pack_target(Geom g, int64_t V, const double* __restrict__ e, double inv_s, float* __restrict__ t) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t m = node_offset(g, v);
    for (int c = 0; c < 3; c++) t[v * 3 + c] = (float)(e[c * g.tot + m] * inv_s);
}
// z (FP64 planes) = t[v][0..2]
__global__ void k_unpack3(Geom g, int64_t V, const float* __restrict__ t, double* __restrict__ z) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t m = node_offset(g, v);
    for (int c = 0; c < 3; c++) z[c * g.tot + m] = (double)t[v * 3 + c];
}
// dy[v][c] = gy_c(node v) for c < 3, 0 for the other channels
__global__ void k_pack_grad(Geom g, int64_t V, const double* __restrict__ gy, float* __restrict__ dy) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t m = node_offset(g, v);
#pragma unroll
    for (int c = 0; c < 16; c++) dy[v * 16 + c] = c < 3 ? (float)gy[c * g.tot + m] : 0.f;
}
// out = a x + b y
__global__ void k_lincomb(int64_t n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                          double* __restrict__ out) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = a * x[q] + b * y[q];
}
__global__ void k_f32_to_bf16_host_layout(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        b[q] = net::f2bf(a[q]);
}
}  // namespace

static const char NET_MAGIC[8] = {'A', 'L', 'A', 'M', 'O', 'N', 'E', 'T'};

extern "C" {

int alamo_net_create(alamo_net** net, uint64_t seed) {
    if (!net) FAIL(ALAMO_ERR_ARG, "null argument");
    return net_create(net, seed);
}

int alamo_net_destroy(alamo_net* n) {
    if (!n) return ALAMO_OK;
    (void)hipDeviceSynchronize();
    net_free_bufs(n->B);
    for (void* p : n->allocs) (void)hipFree(p);
    delete n;
    return ALAMO_OK;
}

int alamo_net_num_params(alamo_net* n, int64_t* np) {
    if (!n || !np) FAIL(ALAMO_ERR_ARG, "null argument");
    *np = n->nparam;
    return ALAMO_OK;
}

int alamo_net_get_params(alamo_net* n, float* host) {
    if (!n || !host) FAIL(ALAMO_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(n->s));
    HIPCHK(hipMemcpy(host, n->w, n->nparam * 4, hipMemcpyDeviceToHost));
    return ALAMO_OK;
}

int alamo_net_set_params(alamo_net* n, const float* host) {
    if (!n || !host) FAIL(ALAMO_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(n->s));
    HIPCHK(hipMemcpy(n->w, host, n->nparam * 4, hipMemcpyHostToDevice));
    n->wdirty = true;
    return ALAMO_OK;
}

int alamo_net_get_grads(alamo_net* n, float* host) {
    if (!n || !host) FAIL(ALAMO_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(n->s));
    HIPCHK(hipMemcpy(host, n->gw, n->nparam * 4, hipMemcpyDeviceToHost));
    return ALAMO_OK;
}

int alamo_net_zero_grads(alamo_net* n) {
    if (!n) FAIL(ALAMO_ERR_ARG, "null argument");
    HIPCHK(hipMemsetAsync(n->gw, 0, n->nparam * 4, n->s));
    return ALAMO_OK;
}

// Forward on host data: x [V][16] f32 (rounded to bf16), y [V][16] f32.
int alamo_net_forward_host(alamo_net* n, const int* nnodes, const float* x, float* y) {
    if (!n || !nnodes || !x || !y) FAIL(ALAMO_ERR_ARG, "null argument");
    TRY(net_bufs(n, nnodes));
    const int64_t V = n->B.g[0].V;
    float* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, V * 16 * 4));
    HIPCHK(hipMemcpy(tmp, x, V * 16 * 4, hipMemcpyHostToDevice));
    k_f32_to_bf16_host_layout<<<nb256(V * 16), 256, 0, n->s>>>(tmp, n->B.x, V * 16);
    int rc = net_forward(n);
    if (rc == ALAMO_OK) {
        HIPCHK(hipStreamSynchronize(n->s));
        HIPCHK(hipMemcpy(y, n->B.y, V * 16 * 4, hipMemcpyDeviceToHost));
    }
    (void)hipFree(tmp);
    return rc;
}

// One layer's weight gradient on host data (a check of the kernel alone): dW[27][cout][ci] +=
// sum_v dout[v][co] in[v + o][ci]; in [V][cin], dout [V][cout] bf16 bit patterns, channel-last.
int alamo_net_wgrad_host(alamo_net* n, const int* nnodes, int cin, int cout, const uint16_t* in, const uint16_t* dout,
                         float* dW) {
    if (!n || !nnodes || !in || !dout || !dW) FAIL(ALAMO_ERR_ARG, "null argument");
    const bool ok = (cin == 16 && (cout == 16 || cout == 32)) || (cin == 32 && (cout == 16 || cout == 64)) ||
                    (cin == 64 && cout == 32);
    if (!ok) FAIL(ALAMO_ERR_ARG, "channel pair not in the network (16-16, 16-32, 32-16, 32-64, 64-32)");
    TRY(net_bufs(n, nnodes));
    const net::Grid3 g = n->B.g[0];
    const int64_t nw = (int64_t)27 * cin * cout;
    uint16_t *din = nullptr, *ddo = nullptr;
    float* ddw = nullptr;
    HIPCHK(hipMalloc(&din, g.V * cin * 2));
    HIPCHK(hipMalloc(&ddo, g.V * cout * 2));
    HIPCHK(hipMalloc(&ddw, nw * 4));
    HIPCHK(hipMemcpy(din, in, g.V * cin * 2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ddo, dout, g.V * cout * 2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ddw, dW, nw * 4, hipMemcpyHostToDevice));
    wgrad_launch(n, g, cin, cout, din, ddo, ddw, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(n->s));
    HIPCHK(hipMemcpy(dW, ddw, nw * 4, hipMemcpyDeviceToHost));
    (void)hipFree(din);
    (void)hipFree(ddo);
    (void)hipFree(ddw);
    return ALAMO_OK;
}

// Forward + backward on host data; gradients accumulate into the network; dx (may be NULL)
// receives d(loss)/dx [V][16] for the given dy [V][16].
int alamo_net_backward_host(alamo_net* n, const int* nnodes, const float* x, const float* dy, float* dx) {
    if (!n || !nnodes || !x || !dy) FAIL(ALAMO_ERR_ARG, "null argument");
    TRY(net_bufs(n, nnodes));
    const int64_t V = n->B.g[0].V;
    float* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, V * 16 * 4));
    HIPCHK(hipMemcpy(tmp, x, V * 16 * 4, hipMemcpyHostToDevice));
    k_f32_to_bf16_host_layout<<<nb256(V * 16), 256, 0, n->s>>>(tmp, n->B.x, V * 16);
    TRY(net_forward(n));
    HIPCHK(hipMemcpy(n->B.dy, dy, V * 16 * 4, hipMemcpyHostToDevice));
    TRY(net_backward(n, dx != nullptr));
    HIPCHK(hipStreamSynchronize(n->s));
    if (dx) HIPCHK(hipMemcpy(dx, n->B.dx, V * 16 * 4, hipMemcpyDeviceToHost));
    (void)hipFree(tmp);
    return ALAMO_OK;
}

int alamo_net_adam_step(alamo_net* n, double lr) {
    if (!n) FAIL(ALAMO_ERR_ARG, "null argument");
    return net_adam(n, (float)lr);
}

int alamo_mg_set_net(alamo_mg* h, alamo_net* n) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    h->net = n;
    return ALAMO_OK;
}

int alamo_net_save(alamo_net* n, const char* path) {
    if (!n || !path) FAIL(ALAMO_ERR_ARG, "null argument");
    std::vector<float> w(n->nparam);
    TRY(alamo_net_get_params(n, w.data()));
    FILE* f = fopen(path, "wb");
    if (!f) FAIL(ALAMO_ERR_ARG, std::string("cannot open ") + path);
    const int32_t ver = 1, nl = (int32_t)n->L.size();
    fwrite(NET_MAGIC, 1, 8, f);
    fwrite(&ver, 4, 1, f);
    fwrite(&nl, 4, 1, f);
    for (auto& ly : n->L) {
        const int32_t d[5] = {ly.cin, ly.cout, ly.cin_real, ly.level, ly.relu};
        fwrite(d, 4, 5, f);
    }
    fwrite(&n->seed, 8, 1, f);
    fwrite(&n->train_seed, 8, 1, f);
    fwrite(&n->nparam, 8, 1, f);
    fwrite(w.data(), 4, w.size(), f);
    fclose(f);
    return ALAMO_OK;
}

int alamo_net_load(alamo_net* n, const char* path) {
    if (!n || !path) FAIL(ALAMO_ERR_ARG, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) FAIL(ALAMO_ERR_ARG, std::string("cannot open ") + path);
    char mg[8];
    int32_t ver = 0, nl = 0;
    int64_t np = 0;
    uint64_t s0 = 0, s1 = 0;
    bool ok = fread(mg, 1, 8, f) == 8 && memcmp(mg, NET_MAGIC, 8) == 0 && fread(&ver, 4, 1, f) == 1 && ver == 1 &&
              fread(&nl, 4, 1, f) == 1 && nl == (int32_t)n->L.size();
    for (int l = 0; ok && l < nl; l++) {
        int32_t d[5];
        ok = fread(d, 4, 5, f) == 5 && d[0] == n->L[l].cin && d[1] == n->L[l].cout;
    }
    ok = ok && fread(&s0, 8, 1, f) == 1 && fread(&s1, 8, 1, f) == 1 && fread(&np, 8, 1, f) == 1 && np == n->nparam;
    std::vector<float> w(ok ? np : 0);
    ok = ok && fread(w.data(), 4, np, f) == (size_t)np;
    fclose(f);
    if (!ok) FAIL(ALAMO_ERR_ARG, "not a network checkpoint of this architecture");
    n->seed = s0;
    n->train_seed = s1;
    return alamo_net_set_params(n, w.data());
}

// Generate (residual after one cycle, remaining error) pairs with the handle's operator and
// train the network on them with Adam.
int alamo_net_train(alamo_net* n, alamo_mg* h, const alamo_train_opts* o, alamo_train_stats* st) {
    if (!n || !h || !o) FAIL(ALAMO_ERR_ARG, "null argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    if (o->npairs < 1 || o->steps < 1) FAIL(ALAMO_ERR_ARG, "npairs and steps must be positive");
    if (!h->prepared) TRY(prepare(h));
    if (h->L.size() < 2) FAIL(ALAMO_ERR_ARG, "training needs at least two multigrid levels");
    if (o->loss != ALAMO_TRAIN_LOSS_ERROR && o->loss != ALAMO_TRAIN_LOSS_RESIDUAL) FAIL(ALAMO_ERR_ARG, "unknown loss");
    const bool resloss = o->loss == ALAMO_TRAIN_LOSS_RESIDUAL;
    TRY(net_level0(n, h));
    if (resloss) TRY(row_table(h));
    n->train_seed = o->seed;
    LevelBuf& l0 = h->L[0];
    // V network voxels; n3 elements of a level-0 vector
    const int64_t V = (int64_t)l0.g.N[0] * l0.g.N[1] * l0.g.N[2], n3 = 3 * l0.g.tot;
    const int nlev = (int)h->L.size();
    std::vector<void*> tmp;
    auto talloc = [&](void** p, size_t b) -> int { return net_alloc(tmp, p, b); };
    int rc = ALAMO_OK;
    double *r = nullptr, *z = nullptr, *q = nullptr, *e = nullptr, *qa = nullptr;
    std::vector<double*> lv(nlev, nullptr);
    uint16_t* X = nullptr;
    float* T = nullptr;
    rc |= talloc((void**)&r, n3 * 8);
    rc |= talloc((void**)&z, n3 * 8);
    rc |= talloc((void**)&q, n3 * 8);
    rc |= talloc((void**)&e, n3 * 8);
    if (resloss) rc |= talloc((void**)&qa, 2 * n3 * 8);  // q, then A y: one block for the two dots
    for (int l = 1; l < nlev; l++) rc |= talloc((void**)&lv[l], 3 * h->L[l].g.tot * 8);
    rc |= talloc((void**)&X, (size_t)o->npairs * V * 16 * 2);
    rc |= talloc((void**)&T, (size_t)o->npairs * V * 3 * 4);
    auto cleanup = [&]() { (void)hipStreamSynchronize(h->stream); for (void* p : tmp) (void)hipFree(p); };
    if (rc != ALAMO_OK) { cleanup(); return ALAMO_ERR_HIP; }
    auto t0 = std::chrono::steady_clock::now();
    alamo_krylov_opts ko;
    alamo_krylov_opts_default(&ko);
    ko.precond = ALAMO_PRECOND_MG;
    ko.max_iter = o->solve_max_iter;
    int solved = 0;
    for (int p = 0; p < o->npairs; p++) {
        // smooth random right-hand side: white noise on a coarse level, prolonged to level 0
        const int lc = 1 + (int)((o->seed + 7919ull * p) % (uint64_t)std::max(1, std::min(3, nlev - 1)));
        double* cur = lv[lc];
        k_noise<<<gs_blocks(3 * h->L[lc].g.tot), 256, 0, h->stream>>>(3 * h->L[lc].g.tot,
                                                                       o->seed * 1000003ull + p, cur);
        for (int l = lc - 1; l >= 0; l--) {
            double* fine = l == 0 ? r : lv[l];
            vfill<double>(fine, 3 * h->L[l].g.tot, 0.0, h->stream);
            launch_interp<double>(h->L[l + 1].g, h->L[l].g, cur, fine, h->stream);
            cur = fine;
        }
        if (resloss && o->solve_max_iter > 0) {
            // the residual loss trains on what the Krylov solve leaves: r becomes the residual after
            // k in [0, solve_max_iter] plain FGMRES iterations on it (its slowest modes)
            const int k = (int)((o->seed * 31ull + 7919ull * p) % (uint64_t)(o->solve_max_iter + 1));
            if (k > 0) {
                vfill<double>(e, n3, 0.0, h->stream);
                alamo_krylov_opts kk = ko;
                kk.max_iter = k;
                alamo_solve_stats ss{};
                const int src = fgmres_impl(h, e, r, 1e-30, 0.0, &kk, &ss);
                if (src != ALAMO_OK && src != ALAMO_ERR_NOCONV) { rc = src; break; }
                launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, e, r, q, 0.0, h->stream);
                std::swap(r, q);
            }
        }
        // z = V(r), q = r - A z, e = A^-1 q
        rc = precond_apply(h, ALAMO_PRECOND_MG, r, z);
        if (rc) break;
        launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, z, r, q, 0.0, h->stream);
        double ql2 = 0;
        rc = norms_dev<double>(h, q, n3, nullptr, &ql2);
        if (rc) break;
        const double s = std::max(ql2 / std::sqrt(3.0 * (double)V), 1e-300);
        net::k_pack_input<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, q, 1.0 / s, n->B.feat,
                                                           X + (int64_t)p * V * 16);
        if (resloss) {
            // the residual loss needs q only
            k_pack_target<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, q, 1.0 / s, T + (int64_t)p * V * 3);
            if (o->verbose) fprintf(stderr, "alamo train: pair %d generated\n", p);
            continue;
        }
        vfill<double>(e, n3, 0.0, h->stream);
        alamo_solve_stats ss{};
        const int src = fgmres_impl(h, e, q, o->solve_tol, 0.0, &ko, &ss);
        if (src != ALAMO_OK && src != ALAMO_ERR_NOCONV) { rc = src; break; }
        solved += ss.converged;
        double el2 = 0;
        rc = norms_dev<double>(h, e, n3, nullptr, &el2);
        if (rc) break;
        const double se = std::max(el2 / std::sqrt(3.0 * (double)V), 1e-300);
        k_pack_target<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, e, 1.0 / se, T + (int64_t)p * V * 3);
        if (o->verbose) fprintf(stderr, "alamo train: pair %d generated (solve %d its)\n", p, ss.iterations);
    }
    if (rc) { cleanup(); return rc; }
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t1 = std::chrono::steady_clock::now();
    double first = -1, last = -1;
    n->s = h->stream;
    for (int step = 0; step < o->steps && rc == ALAMO_OK; step++) {
        const int p = (int)(((uint64_t)step * 2654435761ull + o->seed) % (uint64_t)o->npairs);
        HIPCHK(hipMemcpyAsync(n->B.x, X + (int64_t)p * V * 16, V * 16 * 2, hipMemcpyDeviceToDevice, n->s));
        rc = net_forward(n);
        if (rc) break;
        const bool rep = step == 0 || step == o->steps - 1 || (o->verbose && step % 50 == 0);
        double lv_ = 0;
        if (resloss) {
            // L = 1 - c <q, A y> / <q, q> with c = <q, A y> / <A y, A y>;
            // dL/d(A y) = -(2 c / <q, q>) (q - c A y), and dL/dy = A^T dL/d(A y)
            double *qv = qa, *Ay = qa + n3, *yv = z, *gu = r, *gy = e;
            net::k_unpack_output<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, n->B.y, 1.0, yv, 0);
            k_unpack3<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, T + (int64_t)p * V * 3, qv);
            launch_node<double>(h->sym, APPLY, l0.g, h->bc, l0.C, yv, nullptr, Ay, 0.0, h->stream);
            double d[2], qq = 0;
            rc = multidot(h, Ay, qv, n3, 2, d);  // <A y, q>, <A y, A y>
            if (rc) break;
            rc = multidot(h, qv, qv, n3, 1, &qq);
            if (rc) break;
            if (!std::isfinite(d[0]) || !std::isfinite(d[1])) {
                g_err = "alamo_net_train: the network output is not finite";
                rc = ALAMO_ERR_NAN;
                break;
            }
            if (d[1] > 0.0 && qq > 0.0) {
                const double c = d[0] / d[1];
                lv_ = 1.0 - c * d[0] / qq;
                k_lincomb<<<gs_blocks(n3), 256, 0, h->stream>>>(n3, -2.0 * c / qq, qv, 2.0 * c * c / qq, Ay, gu);
                k_row_apply_t<<<nblocks(l0.g.tot), 256, 0, h->stream>>>(l0.g, h->rowtab, gu, gy);
                k_pack_grad<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, gy, n->B.dy);
            } else {
                lv_ = 1.0;
                HIPCHK(hipMemsetAsync(n->B.dy, 0, V * 16 * 4, n->s));
            }
        } else {
            HIPCHK(hipMemsetAsync(n->B.loss, 0, 4, n->s));
            net::k_mse<<<(unsigned)std::min<int64_t>(1024, (V + 255) / 256), 256, 0, n->s>>>(
                V, n->B.y, T + (int64_t)p * V * 3, n->B.dy, n->B.loss);
        }
        rc = net_backward(n, false);
        if (rc) break;
        if (rep) {
            if (!resloss) {
                float lf = 0;
                HIPCHK(hipMemcpyAsync(&lf, n->B.loss, 4, hipMemcpyDeviceToHost, n->s));
                HIPCHK(hipStreamSynchronize(n->s));
                lv_ = lf;
            }
            if (step == 0) first = lv_;
            last = lv_;
            if (o->verbose) fprintf(stderr, "alamo train: step %d loss %.6e\n", step, lv_);
        }
        rc = net_adam(n, (float)o->lr);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t2 = std::chrono::steady_clock::now();
    cleanup();
    if (st) {
        st->pairs = o->npairs;
        st->pairs_converged = solved;
        st->steps = o->steps;
        st->first_loss = first;
        st->last_loss = last;
        st->seconds_data = std::chrono::duration<double>(t1 - t0).count();
        st->seconds_train = std::chrono::duration<double>(t2 - t1).count();
    }
    return rc;
}

// Mean duration of each convolution layer's forward kernel at the handle's level-0 size,
// timed with HIP events on the handle's stream (reps launches per layer); ms[6].
int alamo_net_layer_times(alamo_net* n, alamo_mg* h, int reps, double* ms) {
    if (!n || !h || !ms || reps < 1) FAIL(ALAMO_ERR_ARG, "bad argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    TRY(net_level0(n, h));
    TRY(net_sync_weights(n));
    NetBufs& B = n->B;
    struct L { const net::Grid3* g; int cin, cout; const uint16_t* in; void* out; bool f32; int relu; };
    const L layers[6] = {{&B.g[0], 16, 16, B.x, B.a0, false, 1},  {&B.g[1], 16, 32, B.d1, B.a1, false, 1},
                         {&B.g[2], 32, 64, B.d2, B.a2, false, 1}, {&B.g[2], 64, 32, B.a2, B.b2, false, 0},
                         {&B.g[1], 32, 16, B.u1, B.b1, false, 1}, {&B.g[0], 16, 16, B.u0, B.y, true, 0}};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    for (int l = 0; l < 6; l++) {
        const L& y = layers[l];
        conv_launch(n, *y.g, y.cin, y.cout, y.in, n->wf[l], n->w + n->L[l].boff, y.out, y.f32, y.relu);  // warm
        HIPCHK(hipEventRecord(e0, n->s));
        for (int r = 0; r < reps; r++)
            conv_launch(n, *y.g, y.cin, y.cout, y.in, n->wf[l], n->w + n->L[l].boff, y.out, y.f32, y.relu);
        HIPCHK(hipEventRecord(e1, n->s));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        ms[l] = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ALAMO_OK;
}

// mean backward-kernel times of each layer (backward-data convolution, weight gradient with its
// reduction, bias gradient excluded), HIP events on the network's stream
int alamo_net_layer_times_bwd(alamo_net* n, alamo_mg* h, int reps, double* ms_data, double* ms_wgrad) {
    if (!n || !h || !ms_data || !ms_wgrad || reps < 1) FAIL(ALAMO_ERR_ARG, "bad argument");
    if (!h->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called");
    TRY(net_level0(n, h));
    TRY(net_sync_weights(n));
    NetBufs& B = n->B;
    // per layer: grid, cin, cout, forward input, a gradient-of-output buffer (bf16, cout channels)
    // and an output buffer of cin f32 channels for the backward-data convolution
    struct L { const net::Grid3* g; int cin, cout; const uint16_t* in; float* dx; };
    const L layers[6] = {{&B.g[0], 16, 16, B.x, B.dx},   {&B.g[1], 16, 32, B.d1, B.dd1},
                         {&B.g[2], 32, 64, B.d2, B.dd2}, {&B.g[2], 64, 32, B.a2, B.da2},
                         {&B.g[1], 32, 16, B.u1, B.du1}, {&B.g[0], 16, 16, B.u0, B.du0}};
    float* gscratch = nullptr;  // dW target of the timed weight gradients, discarded
    HIPCHK(hipMalloc(&gscratch, (size_t)27 * 64 * 32 * 4));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    for (int l = 0; l < 6; l++) {
        const L& y = layers[l];
        for (int pass = 0; pass < 2; pass++) {
            auto run = [&] {
                if (pass == 0) conv_launch(n, *y.g, y.cout, y.cin, B.gb, n->wb[l], nullptr, y.dx, true, 0);
                else wgrad_launch(n, *y.g, y.cin, y.cout, y.in, B.gb, gscratch, nullptr);
            };
            run();  // warm
            HIPCHK(hipEventRecord(e0, n->s));
            for (int r = 0; r < reps; r++) run();
            HIPCHK(hipEventRecord(e1, n->s));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            (pass == 0 ? ms_data : ms_wgrad)[l] = t / reps;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(gscratch);
    return ALAMO_OK;
}

// MFMA lane-layout probe (one 16x16x32 bf16 product of host matrices, A 16 x 32, B 32 x 16)
int alamo_net_mfma_probe32(const uint16_t* A, const uint16_t* B, float* D) {
    uint16_t *a = nullptr, *b = nullptr;
    float* d = nullptr;
    HIPCHK(hipMalloc(&a, 1024));
    HIPCHK(hipMalloc(&b, 1024));
    HIPCHK(hipMalloc(&d, 1024));
    HIPCHK(hipMemcpy(a, A, 1024, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b, B, 1024, hipMemcpyHostToDevice));
    net::k_mfma_probe32<<<1, 64>>>(a, b, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(D, d, 1024, hipMemcpyDeviceToHost));
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(d);
    return ALAMO_OK;
}

// MFMA lane-layout probe (one 16x16x16 bf16 product of host matrices), for the tests.
int alamo_net_mfma_probe(const uint16_t* A, const uint16_t* B, float* D) {
    uint16_t *a = nullptr, *b = nullptr;
    float* d = nullptr;
    HIPCHK(hipMalloc(&a, 512));
    HIPCHK(hipMalloc(&b, 512));
    HIPCHK(hipMalloc(&d, 1024));
    HIPCHK(hipMemcpy(a, A, 512, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b, B, 512, hipMemcpyHostToDevice));
    net::k_mfma_probe<<<1, 64>>>(a, b, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(D, d, 1024, hipMemcpyDeviceToHost));
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(d);
    return ALAMO_OK;
}

}  // extern "C"

// ===========================================================================
// Two-level AMR helpers: C ABI
// ===========================================================================
static int amr_dom(const int* nc, amr::Dom& d) {
    if (!nc) FAIL(ALAMO_ERR_ARG, "null ncell");
    for (int q = 0; q < 3; q++) {
        if (nc[q] < 1) FAIL(ALAMO_ERR_ARG, "coarse ncell must be positive");
        d.nc[q] = nc[q];
    }
    d.N0 = nc[0] + 1;
    d.N1 = nc[1] + 1;
    d.tot = d.N0 * d.N1 * (nc[2] + 1);
    return ALAMO_OK;
}

static int amr_check_box(const amr::Dom& d, const int* flo, const int* fhi) {
    if (!flo || !fhi) FAIL(ALAMO_ERR_ARG, "null fine box");
    for (int q = 0; q < 3; q++) {
        if (flo[q] % 2 != 0 || fhi[q] % 2 != 1 || flo[q] > fhi[q])
            FAIL(ALAMO_ERR_ARG, "fine boxes must be ratio-2 aligned cell boxes (lo even, hi odd, lo <= hi)");
        if (flo[q] < 0 || fhi[q] > 2 * d.nc[q] - 1) FAIL(ALAMO_ERR_ARG, "fine box outside the refined domain");
    }
    return ALAMO_OK;
}

static amr::FBox amr_fbox(const amr::Dom& d, const int* flo, const int* fhi) {
    amr::FBox fb;
    fb.vf = 1;
    for (int q = 0; q < 3; q++) {
        fb.vlo[q] = flo[q] / 2;
        fb.vhi[q] = (fhi[q] + 1) / 2;
        fb.blo[q] = std::max(fb.vlo[q] - 1, 0);
        fb.bhi[q] = std::min(fb.vhi[q] + 1, d.nc[q]);
        fb.flo[q] = flo[q];
        fb.gf[q] = fhi[q] - flo[q] + 4;
        fb.vf *= fb.gf[q];
    }
    return fb;
}

extern "C" {

int alamo_amr_fine_mask(const int* ncell_crse, int nbox, const int* fine_lo, const int* fine_hi, int* mask,
                        void* stream) {
    amr::Dom d;
    TRY(amr_dom(ncell_crse, d));
    if (!mask) FAIL(ALAMO_ERR_ARG, "null mask");
    if (nbox < 0 || nbox > amr::MAX_BOXES) FAIL(ALAMO_ERR_ARG, "nbox must be in [0, 64]");
    amr::Boxes b;
    b.n = nbox;
    for (int t = 0; t < nbox; t++) {
        TRY(amr_check_box(d, fine_lo + 3 * t, fine_hi + 3 * t));
        for (int q = 0; q < 3; q++) {
            b.lo[t][q] = fine_lo[3 * t + q] / 2;
            b.hi[t][q] = (fine_hi[3 * t + q] + 1) / 2 - 1;
        }
    }
    amr::k_fine_mask<<<nblocks(d.tot), 256, 0, (hipStream_t)stream>>>(d, b, mask);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_amr_reflux(const int* ncell_crse, const int* fine_lo, const int* fine_hi, const int* mask, double* crse_res,
                     const double* fine_res, void* stream) {
    amr::Dom d;
    TRY(amr_dom(ncell_crse, d));
    TRY(amr_check_box(d, fine_lo, fine_hi));
    if (!mask || !crse_res || !fine_res) FAIL(ALAMO_ERR_ARG, "null argument");
    const amr::FBox fb = amr_fbox(d, fine_lo, fine_hi);
    const int64_t nv = (int64_t)(fb.vhi[0] - fb.vlo[0] + 1) * (fb.vhi[1] - fb.vlo[1] + 1) * (fb.vhi[2] - fb.vlo[2] + 1);
    amr::k_reflux<<<nblocks(nv), 256, 0, (hipStream_t)stream>>>(d, fb, mask, crse_res, fine_res);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}
int alamo_mg_reflux(alamo_mg* h, const int* fine_lo, const int* fine_hi, const int* mask, double* crse_res,
                    const double* fine_res, void* stream) {
    if (!h) FAIL(ALAMO_ERR_ARG, "null handle");
    if (h->L[0].g.sy != h->L[0].g.N[0]) FAIL(ALAMO_ERR_ARG, "alamo_mg_reflux: the AMR helpers take compact fields");
    return alamo_amr_reflux(h->d.n_cell, fine_lo, fine_hi, mask, crse_res, fine_res, stream);
}

int alamo_amr_average_down(const int* ncell_crse, const int* fine_lo, const int* fine_hi, double* crse_sol,
                           const double* fine_sol, void* stream) {
    amr::Dom d;
    TRY(amr_dom(ncell_crse, d));
    TRY(amr_check_box(d, fine_lo, fine_hi));
    if (!crse_sol || !fine_sol) FAIL(ALAMO_ERR_ARG, "null argument");
    const amr::FBox fb = amr_fbox(d, fine_lo, fine_hi);
    const int64_t nv = (int64_t)(fb.vhi[0] - fb.vlo[0] + 1) * (fb.vhi[1] - fb.vlo[1] + 1) * (fb.vhi[2] - fb.vlo[2] + 1);
    amr::k_average_down<<<nblocks(nv), 256, 0, (hipStream_t)stream>>>(d, fb, crse_sol, fine_sol);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_amr_fixup_residual_mask(const int* ncell_crse, const int* mask, int* resmsk, void* stream) {
    amr::Dom d;
    TRY(amr_dom(ncell_crse, d));
    if (!mask || !resmsk) FAIL(ALAMO_ERR_ARG, "null argument");
    amr::k_fixup_resmask<<<nblocks(d.tot), 256, 0, (hipStream_t)stream>>>(d.tot, mask, resmsk);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Composite multi-level (AMR) solve.  Design: every AMR level is a full-domain alamo_mg handle at
// its own resolution (HBM is plentiful, and no ghost-exchange plan is needed: a level's region
// and the nodes around it live in one array); the region of level l >= 1 is the union of its
// boxes, given by a node mask, and its coverage by level l + 1 by a second mask (the nodal fine
// mask of amrex::MLNodeLinOp::buildMasks, read at Operator.cpp:236,553).
//
// The composite system (FP64):
//   * unknowns: level-0 nodes and the nodes strictly inside each fine region (region mask 2);
//     fine nodes on or outside the region boundary (mask 1, 0) follow the coarser level,
//     u_l = P u_{l-1} (the nodal prolongation of Operator.cpp:365-434, the reference's ghost-node
//     coarse/fine strategy, Linear.H:259); coarse nodes covered by level l + 1 (fine mask 1, 2)
//     hold the coincident fine values (averageDownSolutionRHS, Operator.cpp:435-447);
//   * level l rows (Elastic::Fapply with its BC rows) on level 0 and on each fine region's
//     interior; a coarse/fine interface node (fine mask 1) keeps its coarse row, which reads the
//     fine values of the covered nodes next to it;
//   * strictly covered coarse nodes (fine mask 2): the full-weighting restriction of level l + 1's
//     residual (Operator<Grid::Node>::reflux, Operator.cpp:526-633, whose valid nodes all take the
//     27-point interior stencil) -- the coarse-grid-correction equation of the fine level.
// Solver: composite V-cycles (amrex::MLMG's multi-level cycle): down from the finest level,
// masked smoothing of each level's correction and reflux of its residual, the level-0 multigrid
// V-cycle, up again with prolonged corrections and smoothing; stop on the max norm of the
// composite residual (Linear.H's tolerances).  Fine regions must lie inside the domain with one
// coarse cell of margin and nest properly in the next coarser level's region.
// ---------------------------------------------------------------------------
struct alamo_amr {
    int nl = 0;
    std::vector<alamo_mg*> H;
    std::vector<int*> region;  // level l >= 1 (level 0: null): 2 interior, 1 boundary, 0 outside
    std::vector<int*> fmask;   // level l < nl - 1: coverage by level l + 1 (2, 1, 0)
    std::vector<double*> e, r, c, rr, t;
    double* e0 = nullptr;
    std::vector<void*> allocs;
};

static int amr_alloc(alamo_amr* a, double** p, int64_t n) {
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<int64_t>(n, 1) * sizeof(double)));
    HIPCHK(hipMemset(q, 0, std::max<int64_t>(n, 1) * sizeof(double)));
    a->allocs.push_back(q);
    *p = (double*)q;
    return ALAMO_OK;
}

// e <- masked sweeps of level l's smoother on e with right-hand side c: only region-interior
// nodes change (the rest keeps its value: zero in the down sweep, the prolonged coarse
// correction in the up sweep)
static int amr_smooth(alamo_amr* a, int l, double* e, const double* c, int nsweeps) {
    alamo_mg* h = a->H[l];
    const Geom& g = h->L[0].g;
    for (int s = 0; s < nsweeps; s++) {
        level_op<double>(h, 0, JACOBI, e, c, a->t[l], h->d.omega, h->stream);
        amr::k_masked_copy<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->region[l], 2, 2, e, a->t[l]);
    }
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

// covered coarse nodes (fine mask 1, 2) take the coincident fine values, finest level first
static void amr_inject(alamo_amr* a, double* const* u) {
    for (int l = a->nl - 2; l >= 0; l--) {
        const Geom& g = a->H[l]->L[0].g;
        const Geom& gf = a->H[l + 1]->L[0].g;
        amr::k_inject_covered<<<nblocks(g.tot), 256, 0, a->H[l]->stream>>>(g.tot, g.N[0], g.N[1], gf.N[0], gf.N[1],
                                                                          gf.tot, a->fmask[l], u[l], u[l + 1]);
    }
}

// covered coarse nodes first take the fine values; then r_l = b_l - A_l u_l on level 0 and on each
// fine region's interior (zero elsewhere), the strictly covered nodes replaced by the restriction
// of r_{l+1} (reflux)
static int amr_residual(alamo_amr* a, double* const* u, const double* const* b) {
    amr_inject(a, u);
    for (int l = a->nl - 1; l >= 0; l--) {
        alamo_mg* h = a->H[l];
        const Geom& g = h->L[0].g;
        level_op<double>(h, 0, RESID, u[l], b[l], a->r[l], 0.0, h->stream);
        if (l > 0) amr::k_masked_zero<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->region[l], 2, 2, a->r[l]);
        if (l < a->nl - 1) {
            const Geom& gf = a->H[l + 1]->L[0].g;
            k_restrict<double><<<nblocks(g.tot), 256, 0, h->stream>>>(g, gf, 3, a->r[l + 1], a->t[l]);
            amr::k_masked_copy<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->fmask[l], 2, 2, a->r[l], a->t[l]);
        }
    }
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

// one composite V-cycle (correction scheme) on the residuals a->r; corrections added to u
static int amr_vcycle(alamo_amr* a, double* const* u) {
    const int L = a->nl - 1;
    alamo_mg* h0 = a->H[0];
    const int nsw = h0->d.pre_smooth * h0->d.sweeps_per_smooth;
    for (int l = L; l >= 0; l--) {
        alamo_mg* h = a->H[l];
        const Geom& g = h->L[0].g;
        HIPCHK(hipMemcpyAsync(a->c[l], a->r[l], 3 * g.tot * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        if (l < L) {
            // strictly covered nodes: the restricted residual of the finer correction equation;
            // coarse/fine nodes: their coarse rows less the effect of the finer correction injected
            // at the strictly covered nodes (rr[l] serves as scratch: it is written after smoothing)
            const Geom& gf = a->H[l + 1]->L[0].g;
            k_restrict<double><<<nblocks(g.tot), 256, 0, h->stream>>>(g, gf, 3, a->rr[l + 1], a->t[l]);
            amr::k_masked_copy<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->fmask[l], 2, 2, a->c[l], a->t[l]);
            amr::k_inject_strict<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, g.N[0], g.N[1], gf.N[0], gf.N[1], gf.tot,
                                                                       a->fmask[l], a->rr[l], a->e[l + 1]);
            level_op<double>(h, 0, RESID, a->rr[l], a->c[l], a->t[l], 0.0, h->stream);
            amr::k_masked_copy<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->fmask[l], 1, 1, a->c[l], a->t[l]);
        }
        if (l == 0) break;
        vfill<double>(a->e[l], 3 * g.tot, 0.0, h->stream);
        TRY(amr_smooth(a, l, a->e[l], a->c[l], nsw));
        level_op<double>(h, 0, RESID, a->e[l], a->c[l], a->rr[l], 0.0, h->stream);
        amr::k_masked_zero<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->region[l], 2, 2, a->rr[l]);
    }
    TRY(precond_apply(h0, ALAMO_PRECOND_MG, a->c[0], a->e[0]));
    for (int l = 1; l <= L; l++) {
        alamo_mg* h = a->H[l];
        const Geom& gc = a->H[l - 1]->L[0].g;
        const Geom& g = h->L[0].g;
        k_interp_add<double><<<interp_blocks(gc, g), 256, 0, h->stream>>>(gc, g, a->e[l - 1], a->e[l]);
        TRY(amr_smooth(a, l, a->e[l], a->c[l], h0->d.post_smooth * h0->d.sweeps_per_smooth));
    }
    for (int l = 0; l <= L; l++) {
        const int64_t n = 3 * a->H[l]->L[0].g.tot;
        k_axpy<double, double><<<gs_blocks(n), 256, 0, a->H[l]->stream>>>(u[l], a->e[l], n, 1.0);
    }
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

static int amr_norm(alamo_amr* a, double* const* v, double* ninf) {
    double m = 0.0;
    for (int l = 0; l < a->nl; l++) {
        double x = 0.0;
        TRY(norms_dev<double>(a->H[l], v[l], 3 * a->H[l]->L[0].g.tot, &x, nullptr));
        if (!(x <= m)) m = x;  // a NaN propagates
    }
    *ninf = m;
    return ALAMO_OK;
}

extern "C" {

int alamo_amr_create(const alamo_mg_desc* base, int nlevels, const int* nbox, const int* box_lo, const int* box_hi,
                     alamo_amr** out) {
    if (!base || !out) FAIL(ALAMO_ERR_ARG, "null argument");
    if (nlevels < 1 || nlevels > 6) FAIL(ALAMO_ERR_ARG, "nlevels must be in [1, 6]");
    if (nlevels > 1 && (!nbox || !box_lo || !box_hi)) FAIL(ALAMO_ERR_ARG, "null box list");
    // boxes per level (levels >= 1), inclusive cell boxes in the level's own index space
    std::vector<std::vector<std::array<int, 6>>> B(nlevels);
    int off = 0;
    for (int l = 1; l < nlevels; l++) {
        if (nbox[l - 1] < 1 || nbox[l - 1] > amr::MAX_BOXES) FAIL(ALAMO_ERR_ARG, "each fine level needs 1 to 64 boxes");
        const int f = 1 << l;
        for (int t = 0; t < nbox[l - 1]; t++, off++) {
            std::array<int, 6> bx{};
            for (int q = 0; q < 3; q++) {
                bx[q] = box_lo[3 * off + q];
                bx[3 + q] = box_hi[3 * off + q];
                if (bx[q] % 2 != 0 || bx[3 + q] % 2 != 1 || bx[q] > bx[3 + q])
                    FAIL(ALAMO_ERR_ARG, "fine boxes must cover whole coarse cells (lo even, hi odd, lo <= hi)");
                // one coarse cell of margin to the domain boundary
                if (bx[q] < 2 || bx[3 + q] > f * base->n_cell[q] - 3)
                    FAIL(ALAMO_ERR_ARG, "fine boxes must stay one coarse cell inside the domain");
            }
            B[l].push_back(bx);
        }
        if (l >= 2) {
            // proper nesting: every box, coarsened and grown by one cell, lies in the union of the
            // coarser level's boxes (so the coarse/fine interface nodes carry coarser-level rows)
            const int64_t n0 = (int64_t)base->n_cell[0] << (l - 1), n1 = (int64_t)base->n_cell[1] << (l - 1);
            const int64_t n2 = (int64_t)base->n_cell[2] << (l - 1);
            std::vector<uint8_t> cov((size_t)(n0 * n1 * n2), 0);
            for (const auto& cb : B[l - 1])
                for (int64_t k = cb[2]; k <= cb[5]; k++)
                    for (int64_t j = cb[1]; j <= cb[4]; j++)
                        for (int64_t i = cb[0]; i <= cb[3]; i++) cov[(size_t)(i + n0 * (j + n1 * k))] = 1;
            for (const auto& bx : B[l])
                for (int64_t k = bx[2] / 2 - 1; k <= (bx[5] + 1) / 2; k++)
                    for (int64_t j = bx[1] / 2 - 1; j <= (bx[4] + 1) / 2; j++)
                        for (int64_t i = bx[0] / 2 - 1; i <= (bx[3] + 1) / 2; i++)
                            if (!cov[(size_t)(i + n0 * (j + n1 * k))])
                                FAIL(ALAMO_ERR_ARG, "each box, coarsened and grown by one cell, must lie in the next "
                                                    "coarser level's boxes");
        }
    }
    alamo_amr* a = new alamo_amr();
    a->nl = nlevels;
    int rc = ALAMO_OK;
    for (int l = 0; l < nlevels && rc == ALAMO_OK; l++) {
        alamo_mg_desc d = *base;
        for (int q = 0; q < 3; q++) {
            d.n_cell[q] = base->n_cell[q] << l;
            d.dx[q] = base->dx[q] / (double)(1 << l);
        }
        if (l > 0) d.max_coarsening_level = 0;  // fine levels are smoothed, not coarsened
        alamo_mg* h = nullptr;
        rc = alamo_mg_create(&d, &h);
        if (rc != ALAMO_OK) break;
        a->H.push_back(h);
    }
    const std::string e0 = rc == ALAMO_OK ? std::string() : g_err;
    if (rc == ALAMO_OK) {
        a->region.assign(nlevels, nullptr);
        a->fmask.assign(nlevels, nullptr);
        a->e.assign(nlevels, nullptr);
        a->r = a->c = a->rr = a->t = a->e;
        for (int l = 0; l < nlevels && rc == ALAMO_OK; l++) {
            const Geom& g = a->H[l]->L[0].g;
            rc |= amr_alloc(a, &a->e[l], 3 * g.tot);
            rc |= amr_alloc(a, &a->r[l], 3 * g.tot);
            rc |= amr_alloc(a, &a->c[l], 3 * g.tot);
            rc |= amr_alloc(a, &a->rr[l], 3 * g.tot);
            rc |= amr_alloc(a, &a->t[l], 3 * g.tot);
            amr::Dom dom;
            dom.nc[0] = g.n[0]; dom.nc[1] = g.n[1]; dom.nc[2] = g.n[2];
            dom.N0 = g.N[0]; dom.N1 = g.N[1]; dom.tot = g.tot;
            // the level's own region, and its coverage by the next level
            for (int which = 0; which < 2 && rc == ALAMO_OK; which++) {
                const int src = which == 0 ? l : l + 1;
                if ((which == 0 && l == 0) || src >= nlevels) continue;
                double* m = nullptr;
                rc |= amr_alloc(a, &m, (g.tot + 1) / 2);
                amr::Boxes bx;
                bx.n = (int)B[src].size();
                for (int t = 0; t < bx.n; t++)
                    for (int q = 0; q < 3; q++) {
                        bx.lo[t][q] = which == 0 ? B[src][t][q] : B[src][t][q] / 2;
                        bx.hi[t][q] = which == 0 ? B[src][t][3 + q] : (B[src][t][3 + q] + 1) / 2 - 1;
                    }
                amr::k_fine_mask<<<nblocks(g.tot), 256>>>(dom, bx, (int*)m);
                if (which == 0) a->region[l] = (int*)m;
                else a->fmask[l] = (int*)m;
            }
        }
        if (rc == ALAMO_OK && hipDeviceSynchronize() != hipSuccess) rc = ALAMO_ERR_HIP;
    }
    if (rc != ALAMO_OK) {
        std::string e = e0.empty() ? g_err : e0;
        alamo_amr_destroy(a);
        g_err = e;
        return rc;
    }
    *out = a;
    return ALAMO_OK;
}

int alamo_amr_destroy(alamo_amr* a) {
    if (!a) return ALAMO_OK;
    (void)hipDeviceSynchronize();
    for (alamo_mg* h : a->H) alamo_mg_destroy(h);
    for (void* p : a->allocs) (void)hipFree(p);
    delete a;
    return ALAMO_OK;
}

int alamo_amr_level(alamo_amr* a, int amrlev, alamo_mg** h) {
    if (!a || !h) FAIL(ALAMO_ERR_ARG, "null argument");
    if (amrlev < 0 || amrlev >= a->nl) FAIL(ALAMO_ERR_ARG, "amrlev out of range");
    *h = a->H[amrlev];
    return ALAMO_OK;
}

int alamo_amr_masks(alamo_amr* a, int amrlev, int* region, int* fine_mask, void* stream) {
    if (!a) FAIL(ALAMO_ERR_ARG, "null handle");
    if (amrlev < 0 || amrlev >= a->nl) FAIL(ALAMO_ERR_ARG, "amrlev out of range");
    const Geom& g = a->H[amrlev]->L[0].g;
    const hipStream_t s = (hipStream_t)stream;
    if (region) {
        if (a->region[amrlev]) HIPCHK(hipMemcpyAsync(region, a->region[amrlev], g.tot * sizeof(int), hipMemcpyDeviceToDevice, s));
        else k_fill<int><<<gs_blocks(g.tot), 256, 0, s>>>(region, g.tot, 2);
    }
    if (fine_mask) {
        if (a->fmask[amrlev]) HIPCHK(hipMemcpyAsync(fine_mask, a->fmask[amrlev], g.tot * sizeof(int), hipMemcpyDeviceToDevice, s));
        else k_fill<int><<<gs_blocks(g.tot), 256, 0, s>>>(fine_mask, g.tot, 0);
    }
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

int alamo_amr_solve(alamo_amr* a, double* const* sol, const double* const* rhs, double tol_rel, double tol_abs,
                    alamo_solve_stats* st) {
    if (!a || !sol || !rhs) FAIL(ALAMO_ERR_ARG, "null argument");
    for (int l = 0; l < a->nl; l++) {
        if (!sol[l] || !rhs[l]) FAIL(ALAMO_ERR_ARG, "null level field");
        if (!a->H[l]->model_set) FAIL(ALAMO_ERR_STATE, "SetModel has not been called on every AMR level");
        if (!a->H[l]->prepared) TRY(prepare(a->H[l]));
        if (a->H[l]->gal && l > 0) FAIL(ALAMO_ERR_ARG, "coarse_op GALERKIN is for the level-0 hierarchy only");
    }
    alamo_mg* h0 = a->H[0];
    if (h0->L.back().g.tot > 17 * 17 * 17)
        FAIL(ALAMO_ERR_ARG, "coarsest level has more than 17^3 nodes; use n_cell divisible by a larger power of 2");
    HIPCHK(hipDeviceSynchronize());
    // every level's work goes on level 0's stream, in order (restored on return)
    std::vector<hipStream_t> keep(a->nl);
    for (int l = 0; l < a->nl; l++) {
        keep[l] = a->H[l]->stream;
        a->H[l]->stream = h0->stream;
    }
    struct Restore {
        alamo_amr* a;
        std::vector<hipStream_t>& k;
        ~Restore() { for (int l = 0; l < a->nl; l++) a->H[l]->stream = k[l]; }
    } restore{a, keep};
    auto t0 = std::chrono::steady_clock::now();
    // fine-level nodes outside the region interior take the prolonged coarse solution
    for (int l = 1; l < a->nl; l++) {
        alamo_mg* h = a->H[l];
        const Geom& g = h->L[0].g;
        const Geom& gc = a->H[l - 1]->L[0].g;
        vfill<double>(a->t[l], 3 * g.tot, 0.0, h->stream);
        k_interp_add<double><<<interp_blocks(gc, g), 256, 0, h->stream>>>(gc, g, sol[l - 1], a->t[l]);
        amr::k_masked_copy<<<nblocks(g.tot), 256, 0, h->stream>>>(g.tot, a->region[l], 0, 1, sol[l], a->t[l]);
    }
    // |b|: level 0 and the fine levels' region interiors
    for (int l = 0; l < a->nl; l++) {
        const Geom& g = a->H[l]->L[0].g;
        HIPCHK(hipMemcpyAsync(a->c[l], rhs[l], 3 * g.tot * sizeof(double), hipMemcpyDeviceToDevice, a->H[l]->stream));
        if (l > 0) amr::k_masked_zero<<<nblocks(g.tot), 256, 0, a->H[l]->stream>>>(g.tot, a->region[l], 2, 2, a->c[l]);
    }
    double bn = 0.0, r0 = 0.0;
    TRY(amr_norm(a, a->c.data(), &bn));
    TRY(amr_residual(a, sol, rhs));
    TRY(amr_norm(a, a->r.data(), &r0));
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * std::max(r0, bn));
    double rn = r0;
    int it = 0;
    while (rn > target && it < h0->d.max_iter && std::isfinite(rn)) {
        TRY(amr_vcycle(a, sol));
        TRY(amr_residual(a, sol, rhs));
        TRY(amr_norm(a, a->r.data(), &rn));
        it++;
        if (h0->d.verbose) fprintf(stderr, "alamo amr: iter %d |r|inf %.6e target %.3e\n", it, rn, target);
    }
    // averageDownSolutionRHS: covered coarse nodes take the fine values
    amr_inject(a, sol);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    auto t1 = std::chrono::steady_clock::now();
    const bool conv = rn <= target;
    if (st) {
        st->iterations = it;
        st->converged = conv ? 1 : 0;
        st->initial_res_inf = r0;
        st->final_res_inf = rn;
        st->final_res_l2 = 0.0;
        st->rhs_inf = bn;
        st->target = target;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->fallback_events = 0;
    }
    if (!std::isfinite(rn)) FAIL(ALAMO_ERR_NAN, "alamo_amr_solve: residual is not finite (diverged)");
    if (!conv) FAIL(ALAMO_ERR_NOCONV, "alamo_amr_solve: max_iter reached before the tolerance");
    return ALAMO_OK;
}

}  // extern "C"

// ===========================================================================
// AMReX plotfile of nodal fields (Integrator::WritePlotFile, Integrator.cpp:808-1054)
// ===========================================================================
extern "C" {

static int write_plotfile_impl(const char* dir, int nlevels, const int* n_cell0, const double* prob_lo,
                               const double* prob_hi, const int* nbox, const int* box_lo, const int* box_hi, int ncomp,
                               const char* const* names, const double* const* fields, double time,
                               const int* level_steps, bool device) {
    if (!dir || !n_cell0 || !prob_lo || !prob_hi || !nbox || !names || !fields || !level_steps)
        FAIL(ALAMO_ERR_ARG, "null argument");
    if (nlevels < 1 || nlevels > 16) FAIL(ALAMO_ERR_ARG, "nlevels must be in [1, 16]");
    if (ncomp < 1) FAIL(ALAMO_ERR_ARG, "ncomp must be positive");
    for (int d = 0; d < 3; d++) {
        if (n_cell0[d] < 1) FAIL(ALAMO_ERR_ARG, "n_cell0 must be positive");
        if (!(prob_hi[d] > prob_lo[d])) FAIL(ALAMO_ERR_ARG, "prob_hi must exceed prob_lo");
    }
    std::vector<plot::Level> L(nlevels);
    std::vector<std::vector<double>> host(nlevels);
    int64_t bo = 0;
    for (int l = 0; l < nlevels; l++) {
        plot::Level& lv = L[l];
        int64_t tot = 1;
        for (int d = 0; d < 3; d++) {
            lv.n_cell[d] = n_cell0[d] << l;
            lv.dx[d] = (prob_hi[d] - prob_lo[d]) / lv.n_cell[d];
            tot *= lv.n_cell[d] + 1;
        }
        if (nbox[l] < 0) FAIL(ALAMO_ERR_ARG, "nbox must be >= 0");
        if (nbox[l] == 0) {
            if (l > 0) FAIL(ALAMO_ERR_ARG, "levels >= 1 need at least one box");
            lv.boxes.push_back({{0, 0, 0}, {lv.n_cell[0] - 1, lv.n_cell[1] - 1, lv.n_cell[2] - 1}});
        }
        for (int b = 0; b < nbox[l]; b++, bo++) {
            if (!box_lo || !box_hi) FAIL(ALAMO_ERR_ARG, "null box arrays");
            plot::CellBox cb;
            for (int d = 0; d < 3; d++) {
                cb.lo[d] = box_lo[3 * bo + d];
                cb.hi[d] = box_hi[3 * bo + d];
                if (cb.lo[d] < 0 || cb.hi[d] < cb.lo[d] || cb.hi[d] >= lv.n_cell[d])
                    FAIL(ALAMO_ERR_ARG, "box " + std::to_string(b) + " of level " + std::to_string(l) +
                                            " is empty or leaves the level's domain");
            }
            lv.boxes.push_back(cb);
        }
        if (!fields[l]) FAIL(ALAMO_ERR_ARG, "null field of level " + std::to_string(l));
        if (device) {
            host[l].resize((size_t)(ncomp * tot));
            HIPCHK(hipMemcpy(host[l].data(), fields[l], host[l].size() * sizeof(double), hipMemcpyDeviceToHost));
            lv.data = host[l].data();
        } else {
            lv.data = fields[l];
        }
    }
    std::vector<std::string> nm;
    for (int c = 0; c < ncomp; c++) {
        if (!names[c] || !names[c][0]) FAIL(ALAMO_ERR_ARG, "empty component name");
        nm.push_back(names[c]);
    }
    std::vector<int> steps(level_steps, level_steps + nlevels);
    const std::string err = plot::write_node_plotfile(dir, L, ncomp, nm, prob_lo, prob_hi, time, steps);
    if (!err.empty()) FAIL(ALAMO_ERR_ARG, "alamo_write_plotfile: " + err);
    return ALAMO_OK;
}

int alamo_write_plotfile(const char* dir, int nlevels, const int* n_cell0, const double* prob_lo,
                         const double* prob_hi, const int* nbox, const int* box_lo, const int* box_hi, int ncomp,
                         const char* const* names, const double* const* fields, double time, const int* level_steps) {
    return write_plotfile_impl(dir, nlevels, n_cell0, prob_lo, prob_hi, nbox, box_lo, box_hi, ncomp, names, fields,
                               time, level_steps, true);
}

int alamo_write_plotfile_host(const char* dir, int nlevels, const int* n_cell0, const double* prob_lo,
                              const double* prob_hi, const int* nbox, const int* box_lo, const int* box_hi, int ncomp,
                              const char* const* names, const double* const* fields, double time,
                              const int* level_steps) {
    return write_plotfile_impl(dir, nlevels, n_cell0, prob_lo, prob_hi, nbox, box_lo, box_hi, ncomp, names, fields,
                               time, level_steps, false);
}

}  // extern "C"
