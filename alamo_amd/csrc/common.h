// Shared device/host definitions for the nodal elastic operator on gfx950.
//
// Layout of every level (one box covering the non-periodic domain):
//   nodes N[d] = n[d] + 1, node (i,j,k) at flat index i + N0*(j + N1*k);
//   vector fields: 3 SoA planes of `tot` values (component slowest, as amrex::Array4);
//   stiffness:     ncoef SoA planes (Isotropic: lambda, mu; MajorMinor: the 21 slots of
//                  Set/Matrix4_MajorMinor.H:321-372; Diagonal: A[3][3] row-major).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace alamo {

enum Sym { ISO = 0, MAJORMINOR = 1, DIAGONAL = 2 };
// BC::Operator::Elastic::Elastic::Type (BC/Operator/Elastic/Elastic.H:26)
enum BCType { DISPLACEMENT = 0, TRACTION = 1, PERIODIC = 2, NEUMANN = 3, NONE = 4 };
// BC::Operator::Elastic::Elastic::Face (BC/Operator/Elastic/Elastic.H:35-43)
enum Face {
    XLO, YLO, ZLO, XHI, YHI, ZHI,
    YLO_ZLO, YLO_ZHI, YHI_ZLO, YHI_ZHI,
    ZLO_XLO, ZLO_XHI, ZHI_XLO, ZHI_XHI,
    XLO_YLO, XLO_YHI, XHI_YLO, XHI_YHI,
    XLO_YLO_ZLO, XLO_YLO_ZHI, XLO_YHI_ZLO, XLO_YHI_ZHI,
    XHI_YLO_ZLO, XHI_YLO_ZHI, XHI_YHI_ZLO, XHI_YHI_ZHI,
    INT
};

template <int SYM> struct NCoef;
template <> struct NCoef<ISO> { static constexpr int value = 2; };
template <> struct NCoef<MAJORMINOR> { static constexpr int value = 21; };
template <> struct NCoef<DIAGONAL> { static constexpr int value = 9; };

inline int ncoef_of(int sym) { return sym == ISO ? 2 : sym == MAJORMINOR ? 21 : 9; }

// Geometry of one multigrid level, passed by value to kernels.
struct Geom {
    int n[3];       // cells
    int N[3];       // nodes
    int64_t sy, sz; // strides of j and k
    int64_t tot;    // nodes per component plane
    double dx[3];
    // Elastic's psi on this level (cell field, n[0] n[1] n[2] values, i fastest), FP64 and FP32
    // copies; null when psi is not set
    const double* psi;
    const float* psif;
    double psi_small;
    // MajorMinor stiffness given per grain (the finest level of a SetModelGrains model): the node's
    // grain id and the grains' 21-slot records, FP64 and FP32; null otherwise
    const uint16_t* gid;
    const double* gtab;
    const float* gtabf;
    int ngrains;
    // Elastic::SetUniform(true): the interior rows omit grad(C) . grad(u)
    int uniform;
    // Smoother diagonal of the l1-Jacobi smoother (3 planes, FP64 and FP32 copies; null for the
    // reference's damped point Jacobi): the row's absolute sum sum_j |a_ij| with the sign of a_ii,
    // so a sweep is x += (b - A x) / dsm
    const double* dsm;
    const float* dsmf;
    // BF16 reciprocal of the l1 diagonal (3 planes of bfloat16 bit patterns) for the FP32 row-kernel
    // sweeps, which multiply by it: 6 bytes per node instead of 12 and no division; null when absent
    const uint16_t* dsmr;
    // MajorMinor levels after preparation: the node-local contraction of the stiffness gradient,
    // D_ikl = sum_j d_j C_ijkl (central differences; symmetric in k, l), 18 planes i * 6 + pair(k, l),
    // FP64 and FP32 copies; null when not prepared (the rows then difference C themselves)
    const double* mmd;
    const float* mmdf;
};

// Boundary types per region and component, packed for kernel arguments.
struct BCTable {
    int8_t t[26][3];
};

// ---------------------------------------------------------------------------
// Stiffness element access C_ijkl with compile-time indices.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int pair_of(int a, int b) {
    return (a > b) ? pair_of(b, a) : (a == 0 ? b : (a == 1 ? 2 + b : 5));
}
// MajorMinor slot: upper triangle of the 6x6 pair matrix, pairs ordered 00,01,02,11,12,22
// (the uid table of Set/Matrix4_MajorMinor.H:327-375).
__host__ __device__ constexpr int mm_slot(int i, int j, int k, int l) {
    return (pair_of(i, j) <= pair_of(k, l))
               ? pair_of(i, j) * 6 - pair_of(i, j) * (pair_of(i, j) - 1) / 2 + (pair_of(k, l) - pair_of(i, j))
               : pair_of(k, l) * 6 - pair_of(k, l) * (pair_of(k, l) - 1) / 2 + (pair_of(i, j) - pair_of(k, l));
}

template <int SYM, typename T>
__host__ __device__ __forceinline__ T C4(const T* c, int i, int j, int k, int l) {
    if constexpr (SYM == ISO) {
        // Set/Matrix4_Isotropic.H:52-59
        T r = T(0);
        if (i == k && j == l) r += c[1];
        if (i == l && j == k) r += c[1];
        if (i == j && k == l) r += c[0];
        return r;
    } else if constexpr (SYM == MAJORMINOR) {
        return c[mm_slot(i, j, k, l)];
    } else {
        // Set/Matrix4_Diagonal.H operator(): A[i][j] if i==k && j==l
        return (i == k && j == l) ? c[3 * i + j] : T(0);
    }
}

// sigma = C : A
template <int SYM, typename T>
__host__ __device__ __forceinline__ void c_dot(const T* c, const T A[3][3], T S[3][3]) {
    if constexpr (SYM == ISO) {
        // lambda tr(A) delta_ij + mu (A_ij + A_ji)   (Set/Matrix4_Isotropic.H:137-157)
        const T tr = A[0][0] + A[1][1] + A[2][2];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) S[i][j] = c[1] * (A[i][j] + A[j][i]) + (i == j ? c[0] * tr : T(0));
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                T s = T(0);
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int l = 0; l < 3; l++) s += C4<SYM>(c, i, j, k, l) * A[k][l];
                S[i][j] = s;
            }
    }
}

// ---------------------------------------------------------------------------
// Boundary regions.  Precedence of BC/Operator/Elastic/Constant.H:373-400;
// normals with the truncated constants of BC/Operator/Elastic/Elastic.H:79-80.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int classify(int i, int j, int k, const int* hi) {
    bool xl = i == 0, xh = i == hi[0], yl = j == 0, yh = j == hi[1], zl = k == 0, zh = k == hi[2];
    if (xl && yl && zl) return XLO_YLO_ZLO;
    if (xl && yl && zh) return XLO_YLO_ZHI;
    if (xl && yh && zl) return XLO_YHI_ZLO;
    if (xl && yh && zh) return XLO_YHI_ZHI;
    if (xh && yl && zl) return XHI_YLO_ZLO;
    if (xh && yl && zh) return XHI_YLO_ZHI;
    if (xh && yh && zl) return XHI_YHI_ZLO;
    if (xh && yh && zh) return XHI_YHI_ZHI;
    if (yl && zl) return YLO_ZLO;
    if (yl && zh) return YLO_ZHI;
    if (yh && zl) return YHI_ZLO;
    if (yh && zh) return YHI_ZHI;
    if (zl && xl) return ZLO_XLO;
    if (zl && xh) return ZLO_XHI;
    if (zh && xl) return ZHI_XLO;
    if (zh && xh) return ZHI_XHI;
    if (xl && yl) return XLO_YLO;
    if (xl && yh) return XLO_YHI;
    if (xh && yl) return XHI_YLO;
    if (xh && yh) return XHI_YHI;
    if (xl) return XLO;
    if (xh) return XHI;
    if (yl) return YLO;
    if (yh) return YHI;
    if (zl) return ZLO;
    if (zh) return ZHI;
    return INT;
}

__host__ __device__ __forceinline__ void region_normal(int f, double* n) {
    const double S3 = 0.57735026919, S2 = 0.70710678118;
    // sign pattern per axis: -1 lo, +1 hi, 0 none; magnitude by how many axes are set
    int sx = 0, sy = 0, sz = 0;
    switch (f) {
        case XLO: sx = -1; break;
        case YLO: sy = -1; break;
        case ZLO: sz = -1; break;
        case XHI: sx = 1; break;
        case YHI: sy = 1; break;
        case ZHI: sz = 1; break;
        case YLO_ZLO: sy = -1; sz = -1; break;
        case YLO_ZHI: sy = -1; sz = 1; break;
        case YHI_ZLO: sy = 1; sz = -1; break;
        case YHI_ZHI: sy = 1; sz = 1; break;
        case ZLO_XLO: sz = -1; sx = -1; break;
        case ZLO_XHI: sz = -1; sx = 1; break;
        case ZHI_XLO: sz = 1; sx = -1; break;
        case ZHI_XHI: sz = 1; sx = 1; break;
        case XLO_YLO: sx = -1; sy = -1; break;
        case XLO_YHI: sx = -1; sy = 1; break;
        case XHI_YLO: sx = 1; sy = -1; break;
        case XHI_YHI: sx = 1; sy = 1; break;
        case XLO_YLO_ZLO: sx = -1; sy = -1; sz = -1; break;
        case XLO_YLO_ZHI: sx = -1; sy = -1; sz = 1; break;
        case XLO_YHI_ZLO: sx = -1; sy = 1; sz = -1; break;
        case XLO_YHI_ZHI: sx = -1; sy = 1; sz = 1; break;
        case XHI_YLO_ZLO: sx = 1; sy = -1; sz = -1; break;
        case XHI_YLO_ZHI: sx = 1; sy = -1; sz = 1; break;
        case XHI_YHI_ZLO: sx = 1; sy = 1; sz = -1; break;
        case XHI_YHI_ZHI: sx = 1; sy = 1; sz = 1; break;
        default: break;
    }
    int cnt = (sx != 0) + (sy != 0) + (sz != 0);
    double m = cnt == 3 ? S3 : cnt == 2 ? S2 : 1.0;
    n[0] = sx * m;
    n[1] = sy * m;
    n[2] = sz * m;
}

}  // namespace alamo
