// Flexible GMRES(m) in FP64 around a preconditioner that may change from one application
// to the next (a multigrid cycle with a BiCGStab bottom, or the learned network), after
// Saad, "A flexible inner-outer preconditioned GMRES algorithm", SIAM J. Sci. Comput. 14 (1993).
//
// The operator is the nonsymmetric strong-form Operator::Elastic, so the outer Krylov method
// is GMRES rather than CG.  Orthogonalisation is classical Gram-Schmidt: one fused multi-dot
// kernel (which also returns |w|^2, w being the next vector of the strided basis) and one
// multi-axpy kernel that also returns |w'|^2, so a step reads the basis twice instead of once
// per vector as modified Gram-Schmidt would.  A second pass (CGS2) runs when the first one
// cancelled w by more than a set factor, |w'| < eta |w| (Kahan's criterion for eta = 1/sqrt 2;
// Giraud, Langou, Rozloznik, "The loss of orthogonality in the Gram-Schmidt orthogonalization
// process", Comput. Math. Appl. 50 (2005)).
#pragma once
#include "kernels_generic.h"

namespace alamo {

constexpr int KRY_MAXV = 32;         // largest restart length + 1
constexpr int KRY_DOT_BLOCKS = 4096;  // workgroups of the multi-dot pass (enough to saturate HBM)

// Vectors are FP64, or FP32 for the inner iterations of the mixed-precision solve; sums and
// coefficients are FP64 in both.
// partial[i * nb + block] = sum over the block's share of w . V_i, i < NV.  NV is a
// compile-time count so the NV accumulators stay in registers.
template <int NV, typename T>
__global__ void __launch_bounds__(RED_BLOCK) k_multidot(const T* __restrict__ w, const T* __restrict__ V, int64_t n,
                                                        int64_t ldv, double* __restrict__ partial) {
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const double wq = (double)w[q];
#pragma unroll
        for (int i = 0; i < NV; i++) acc[i] += wq * (double)__builtin_nontemporal_load(V + i * ldv + q);
    }
    __shared__ double sh[NV][RED_BLOCK / 64];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const double v = wave_sum(acc[i]);
        if (threadIdx.x % 64 == 0) sh[i][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) s += sh[threadIdx.x][q];
        partial[threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// Elements of T in one 16-byte access, and whether w, V and the leading dimension allow them.
template <typename T> constexpr int vec16() { return 16 / (int)sizeof(T); }
template <typename T>
inline bool aligned16(const T* a, const T* b, int64_t ld) {
    return ((((uintptr_t)a) | ((uintptr_t)b)) & 15) == 0 && ld % vec16<T>() == 0;
}
template <typename T>
__device__ __forceinline__ void unpack16(const uint4& u, double* out) {
    if constexpr (sizeof(T) == 4) {
        out[0] = (double)__uint_as_float(u.x);
        out[1] = (double)__uint_as_float(u.y);
        out[2] = (double)__uint_as_float(u.z);
        out[3] = (double)__uint_as_float(u.w);
    } else {
        out[0] = __hiloint2double((int)u.y, (int)u.x);
        out[1] = __hiloint2double((int)u.w, (int)u.z);
    }
}

// The same multi-dot with 16-byte loads (w, V 16-byte aligned, ldv a multiple of the vector
// width): each work-item reads E = 16 / sizeof(T) consecutive elements of w and of every V_i per
// step, so a wavefront access covers 1 KB instead of 256 or 512 B; the sums per element run in
// the same FP64 arithmetic.  The last n % E elements go to the first work-items of block 0.
template <int NV, typename T>
__global__ void __launch_bounds__(RED_BLOCK) k_multidot16(const T* __restrict__ w, const T* __restrict__ V, int64_t n,
                                                          int64_t ldv, double* __restrict__ partial) {
    constexpr int E = vec16<T>();
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
    const int64_t nv = n / E;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += (int64_t)gridDim.x * blockDim.x) {
        double wv[E], vv[E];
        unpack16<T>(*(const uint4*)(w + q * E), wv);
#pragma unroll
        for (int i = 0; i < NV; i++) {
            typedef unsigned u4v __attribute__((ext_vector_type(4)));
            const u4v x = __builtin_nontemporal_load((const u4v*)(V + i * ldv + q * E));
            unpack16<T>(make_uint4(x[0], x[1], x[2], x[3]), vv);
#pragma unroll
            for (int e = 0; e < E; e++) acc[i] += wv[e] * vv[e];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < n - nv * E) {
        const int64_t q = nv * E + threadIdx.x;
        const double wq = (double)w[q];
#pragma unroll
        for (int i = 0; i < NV; i++) acc[i] += wq * (double)V[i * ldv + q];
    }
    __shared__ double sh[NV][RED_BLOCK / 64];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const double v = wave_sum(acc[i]);
        if (threadIdx.x % 64 == 0) sh[i][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) s += sh[threadIdx.x][q];
        partial[threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// Launch k_multidot over nv <= KRY_MAXV vectors in register-sized groups (the 16-byte form,
// in groups of up to 16, when the vectors are aligned for it; else groups of up to 8).
template <typename T>
inline void launch_multidot(const T* w, const T* V, int64_t n, int64_t ldv, int nv, int nb, double* partial,
                            hipStream_t s) {
    const bool v16 = aligned16(w, V, ldv);
    // groups of up to 16 vectors in the 16-byte form (each pass re-reads w), 8 in the scalar form
    const int gmax = v16 ? 16 : 8;
    for (int i0 = 0; i0 < nv; i0 += gmax) {
        const int k = nv - i0 < gmax ? nv - i0 : gmax;
        const T* Vi = V + i0 * ldv;
        double* P = partial + (int64_t)i0 * nb;
#define ALAMO_MD(K)                                                                        \
    if (v16) k_multidot16<K, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P);                \
    else k_multidot<K, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P);                      \
    break;
        switch (k) {
            case 1: ALAMO_MD(1)
            case 2: ALAMO_MD(2)
            case 3: ALAMO_MD(3)
            case 4: ALAMO_MD(4)
            case 5: ALAMO_MD(5)
            case 6: ALAMO_MD(6)
            case 7: ALAMO_MD(7)
            case 8: ALAMO_MD(8)
            default:
                break;
        }
#undef ALAMO_MD
        // the 16-byte form only: 9 .. 16 vectors in one pass
#define ALAMO_MD16(K)                                                                      \
    k_multidot16<K, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P);                         \
    break;
        switch (k) {
            case 9: ALAMO_MD16(9)
            case 10: ALAMO_MD16(10)
            case 11: ALAMO_MD16(11)
            case 12: ALAMO_MD16(12)
            case 13: ALAMO_MD16(13)
            case 14: ALAMO_MD16(14)
            case 15: ALAMO_MD16(15)
            case 16: ALAMO_MD16(16)
            default:
                break;
        }
#undef ALAMO_MD16
    }
}

// out[i] = sum_b partial[i * nb + b]; one block per vector
__global__ void __launch_bounds__(RED_BLOCK) k_sum_rows(const double* __restrict__ partial, int nb,
                                                        double* __restrict__ out) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[blockIdx.x * nb + b];
    s = wave_sum(s);
    __shared__ double sh[RED_BLOCK / 64];
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) t += sh[q];
        out[blockIdx.x] = t;
    }
}

// w -= sum_{i < nv} c_i V_i and partial[block] = the block's share of |w|^2 afterwards (of the
// stored values)
template <typename T>
__global__ void __launch_bounds__(RED_BLOCK) k_multiaxpy_nrm(T* __restrict__ w, const T* __restrict__ V, int64_t n,
                                                             int64_t ldv, int nv, const double* __restrict__ c,
                                                             double* __restrict__ partial) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        double s = (double)w[q];
        for (int i = 0; i < nv; i++) s -= c[i] * (double)V[i * ldv + q];
        const T t = (T)s;
        w[q] = t;
        acc += (double)t * (double)t;
    }
    __shared__ double sh[RED_BLOCK / 64];
    const double v = wave_sum(acc);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) t += sh[q];
        partial[blockIdx.x] = t;
    }
}

// k_multiaxpy_nrm with 16-byte accesses (w, V aligned, ldv a multiple of the vector width); the
// coefficients sit in LDS, the per-element arithmetic is the same; block 0 takes the n % E tail.
template <typename T>
__global__ void __launch_bounds__(RED_BLOCK) k_multiaxpy_nrm16(T* __restrict__ w, const T* __restrict__ V, int64_t n,
                                                               int64_t ldv, int nv, const double* __restrict__ c,
                                                               double* __restrict__ partial) {
    constexpr int E = vec16<T>();
    __shared__ double cs[KRY_MAXV];
    if (threadIdx.x < nv) cs[threadIdx.x] = c[threadIdx.x];
    __syncthreads();
    double acc = 0.0;
    const int64_t nvec = n / E;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvec; q += (int64_t)gridDim.x * blockDim.x) {
        double sv[E], vv[E];
        unpack16<T>(*(const uint4*)(w + q * E), sv);
#pragma unroll 4
        for (int i = 0; i < nv; i++) {
            unpack16<T>(*(const uint4*)(V + i * ldv + q * E), vv);
#pragma unroll
            for (int e = 0; e < E; e++) sv[e] -= cs[i] * vv[e];
        }
        T t[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            t[e] = (T)sv[e];
            acc += (double)t[e] * (double)t[e];
        }
        uint4 o;
        memcpy(&o, t, 16);
        *(uint4*)(w + q * E) = o;
    }
    if (blockIdx.x == 0 && threadIdx.x < n - nvec * E) {
        const int64_t q = nvec * E + threadIdx.x;
        double s = (double)w[q];
        for (int i = 0; i < nv; i++) s -= cs[i] * (double)V[i * ldv + q];
        const T t = (T)s;
        w[q] = t;
        acc += (double)t * (double)t;
    }
    __shared__ double sh[RED_BLOCK / 64];
    const double v = wave_sum(acc);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) t += sh[q];
        partial[blockIdx.x] = t;
    }
}

template <typename T>
inline void launch_multiaxpy_nrm(T* w, const T* V, int64_t n, int64_t ldv, int nv, const double* c, int nb,
                                 double* partial, hipStream_t s) {
    if (aligned16<T>(w, V, ldv)) k_multiaxpy_nrm16<T><<<nb, RED_BLOCK, 0, s>>>(w, V, n, ldv, nv, c, partial);
    else k_multiaxpy_nrm<T><<<nb, RED_BLOCK, 0, s>>>(w, V, n, ldv, nv, c, partial);
}

// x += sum_{i < nv} c_i Z_i (x FP64)
template <typename T>
__global__ void __launch_bounds__(256) k_multiaxpy_add(double* __restrict__ x, const T* __restrict__ Z, int64_t n,
                                                       int64_t ldz, int nv, const double* __restrict__ c) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        double s = x[q];
        for (int i = 0; i < nv; i++) s += c[i] * (double)Z[i * ldz + q];
        x[q] = s;
    }
}

// y = a * x, converted to the precision of y
template <typename Ti, typename To>
__global__ void k_scale(To* __restrict__ y, const Ti* __restrict__ x, int64_t n, double a) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        y[q] = (To)(a * (double)x[q]);
}

}  // namespace alamo
