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

// Launch k_multidot over nv <= KRY_MAXV vectors in register-sized groups of at most 8.
template <typename T>
inline void launch_multidot(const T* w, const T* V, int64_t n, int64_t ldv, int nv, int nb, double* partial,
                            hipStream_t s) {
    for (int i0 = 0; i0 < nv; i0 += 8) {
        const int k = nv - i0 < 8 ? nv - i0 : 8;
        const T* Vi = V + i0 * ldv;
        double* P = partial + (int64_t)i0 * nb;
        switch (k) {
            case 1: k_multidot<1, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            case 2: k_multidot<2, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            case 3: k_multidot<3, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            case 4: k_multidot<4, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            case 5: k_multidot<5, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            case 6: k_multidot<6, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            case 7: k_multidot<7, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
            default: k_multidot<8, T><<<nb, RED_BLOCK, 0, s>>>(w, Vi, n, ldv, P); break;
        }
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
