// Kernels "a" and "c": element matrices / vectors on simplices and the
// fixed-order (ascending cell) gather of element contributions into CSR rows.
#include "common.hpp"
#include "reduce.hpp"

#include <algorithm>

namespace fa {

// ---- cell geometry ---------------------------------------------------------
// Triangle: grad lambda as numpy_backend.py:586-598, |T| = det(p1-p0, p2-p1)/2
// (simplex_measure, numpy_backend.py:413-421).
// Tetrahedron: grad lambda = cross(v_jm, v_jk) / (6 vol) over localFace
// (numpy_backend.py:619-629, tetrahedron_mesh.py:23-24), |T| by the triple
// product of tetrahedron_mesh.py:165-175.
template <int TD>
struct Geom;

template <>
struct Geom<2> {
    double gl[3][2];
    double meas;
    __device__ void compute(const double p[3][2]) {
        double e[3][2];
        for (int d = 0; d < 2; ++d) {
            e[0][d] = p[2][d] - p[1][d];
            e[1][d] = p[0][d] - p[2][d];
            e[2][d] = p[1][d] - p[0][d];
        }
        const double nv = e[0][0] * e[1][1] - e[0][1] * e[1][0];
        for (int i = 0; i < 3; ++i) {
            gl[i][0] = -e[i][1] / nv;
            gl[i][1] = e[i][0] / nv;
        }
        const double a0 = p[1][0] - p[0][0], a1 = p[1][1] - p[0][1];
        const double b0 = p[2][0] - p[1][0], b1 = p[2][1] - p[1][1];
        meas = (a0 * b1 - a1 * b0) / 2.0;
    }
};

__device__ inline void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <>
struct Geom<3> {
    double gl[4][3];
    double meas;
    __device__ void compute(const double p[4][3]) {
        // simplex_measure: det of consecutive edge vectors / 6
        double a[3], b[3], c[3], t[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = p[1][d] - p[0][d];
            b[d] = p[2][d] - p[1][d];
            c[d] = p[3][d] - p[2][d];
        }
        cross3(b, c, t);
        const double vol6 = a[0] * t[0] + a[1] * t[1] + a[2] * t[2];  // 6 * simplex_measure
        const int lf[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
        for (int i = 0; i < 4; ++i) {
            double vjk[3], vjm[3], n[3];
            for (int d = 0; d < 3; ++d) {
                vjk[d] = p[lf[i][1]][d] - p[lf[i][0]][d];
                vjm[d] = p[lf[i][2]][d] - p[lf[i][0]][d];
            }
            cross3(vjm, vjk, n);
            for (int d = 0; d < 3; ++d) gl[i][d] = n[d] / vol6;
        }
        double v01[3], v02[3], v03[3], x[3];
        for (int d = 0; d < 3; ++d) {
            v01[d] = p[1][d] - p[0][d];
            v02[d] = p[2][d] - p[0][d];
            v03[d] = p[3][d] - p[0][d];
        }
        cross3(v01, v02, x);
        meas = (v03[0] * x[0] + v03[1] * x[1] + v03[2] * x[2]) / 6.0;
    }
};

template <int TD>
__device__ inline void load_cell(const double* __restrict__ node, const int32_t* __restrict__ cell, int64_t c,
                                 double p[TD + 1][TD]) {
#pragma unroll
    for (int v = 0; v < TD + 1; ++v) {
        const int64_t n = cell[c * (TD + 1) + v];
#pragma unroll
        for (int d = 0; d < TD; ++d) p[v][d] = node[n * TD + d];
    }
}

__device__ inline double coef_value(int kind, double scale, const double* __restrict__ coef, int64_t c, int q,
                                    int nq) {
    switch (kind) {
        case FA_COEF_CONST: return scale;
        case FA_COEF_CELL: return scale * coef[c];
        case FA_COEF_CELLQ: return scale * coef[c * nq + q];
        default: return 0.0;
    }
}

// ---- separable coefficients (FA_COEF_SEP) -----------------------------------
// sin(pi x) and cos(pi x): x = k/2 + r with k = rint(2x), |r| <= 1/4 (r = x - k/2
// is exact for |x| < 2^51), then the Taylor series of sin(pi r) and cos(pi r)
// at 0 (coefficients (-1)^n pi^(2n+1)/(2n+1)! and (-1)^n pi^(2n)/(2n)!, through
// r^17 and r^18: the remainder is below 1e-19 for |r| <= 1/4) by Horner's rule,
// and the quadrant k mod 4.  Few registers, unlike the general library sin.
__device__ inline void sincospi_quarter(double r, double& s, double& c) {
    const double z = r * r;
    double ps = 7.952054001475513e-07;
    ps = fma(ps, z, -2.1915353447830217e-05);
    ps = fma(ps, z, 0.00046630280576761255);
    ps = fma(ps, z, -0.0073704309457143504);
    ps = fma(ps, z, 0.08214588661112823);
    ps = fma(ps, z, -0.5992645293207921);
    ps = fma(ps, z, 2.5501640398773455);
    ps = fma(ps, z, -5.16771278004997);
    ps = fma(ps, z, 3.141592653589793);
    s = r * ps;
    double pc = -1.3878952462213771e-07;
    pc = fma(pc, z, 4.303069587032947e-06);
    pc = fma(pc, z, -0.0001046381049248457);
    pc = fma(pc, z, 0.0019295743094039231);
    pc = fma(pc, z, -0.02580689139001406);
    pc = fma(pc, z, 0.2353306303588932);
    pc = fma(pc, z, -1.3352627688545895);
    pc = fma(pc, z, 4.0587121264167685);
    pc = fma(pc, z, -4.934802200544679);
    c = fma(pc, z, 1.0);
}

// sin(pi x) for shift 0, cos(pi x) for shift 1
__device__ inline double sinpi_shift(double x, int shift) {
    const double k = rint(2.0 * x);
    const double r = fma(-0.5, k, x);
    double s, c;
    sincospi_quarter(r, s, c);
    const int q = (static_cast<int>(static_cast<long long>(k) & 3) + shift) & 3;
    const double v = (q & 1) ? c : s;
    return (q & 2) ? -v : v;
}

constexpr double INV_PI = 0.31830988618379067154;

// exp(x) = 2^k e^r: k = rint(x / ln 2), r = x - k ln 2 with ln 2 split in two
// doubles (Cody and Waite's reduction, the second part applied by FMA), so
// |r| <= ln(2)/2; e^r from its Taylor series through r^14 (remainder below
// 1e-19), by Horner's rule; then the exponent is scaled in by ldexp.
__device__ inline double exp_reduced(double x) {
    if (x > 709.78) return HUGE_VAL;
    if (x < -745.2) return 0.0;
    const double k = rint(x * 1.4426950408889634);
    double r = fma(-k, 0.6931471805599453, x);
    r = fma(-k, 2.3190468138462996e-17, r);
    double e = 1.1470745597729725e-11;
    e = fma(e, r, 1.6059043836821613e-10);
    e = fma(e, r, 2.08767569878681e-09);
    e = fma(e, r, 2.505210838544172e-08);
    e = fma(e, r, 2.755731922398589e-07);
    e = fma(e, r, 2.7557319223985893e-06);
    e = fma(e, r, 2.48015873015873e-05);
    e = fma(e, r, 0.0001984126984126984);
    e = fma(e, r, 0.001388888888888889);
    e = fma(e, r, 0.008333333333333333);
    e = fma(e, r, 0.041666666666666664);
    e = fma(e, r, 0.16666666666666666);
    e = fma(e, r, 0.5);
    e = fma(e, r, 1.0);
    e = fma(e, r, 1.0);
    return ldexp(e, static_cast<int>(k));
}

// One factor g(s) of a term; every lane runs the same program (the struct is
// read through uniform loads), so the switch does not diverge.
__device__ inline double sep_factor(int kind, double a, double b, double s) {
    switch (kind) {
        case FA_SEP_SIN: return sinpi_shift((a * s + b) * INV_PI, 0);
        case FA_SEP_COS: return sinpi_shift((a * s + b) * INV_PI, 1);
        case FA_SEP_EXP: return exp_reduced(a * s + b);
        case FA_SEP_AFFINE: return a * s + b;
        case FA_SEP_SINPI: return sinpi_shift(a * s + b, 0);
        case FA_SEP_COSPI: return sinpi_shift(a * s + b, 1);
        default: {  // FA_SEP_POW: s^a, a a whole number
            const int k = static_cast<int>(a);
            double r = 1.0;
            for (int i = 0; i < k; ++i) r *= s;
            return r;
        }
    }
}

// f(x) at quadrature point q of cell c with vertices p: x = sum_v bcs[q][v] p_v
// (the order of bc_to_point), then the sum of the terms in order
template <int TD>
__device__ inline double sep_value(const fa_sep_t* __restrict__ e, int q, const double p[TD + 1][TD], int64_t c) {
    double x[TD];
#pragma unroll
    for (int d = 0; d < TD; ++d) x[d] = 0.0;
#pragma unroll
    for (int v = 0; v < TD + 1; ++v) {
        const double w = e->bcs[q * (TD + 1) + v];
#pragma unroll
        for (int d = 0; d < TD; ++d) x[d] += w * p[v][d];
    }
    double acc = 0.0;
    int k = 0;
    const int nt = e->nterms;
    for (int t = 0; t < nt; ++t) {
        double v = e->coef[t];
        const double* fld = e->field[t];
        if (fld) v *= fld[c];
        const int kend = e->term_end[t];
        for (; k < kend; ++k) {
            const int d = e->dim[k];
            const double s = (TD == 2) ? (d == 0 ? x[0] : x[TD - 1]) : (d == 0 ? x[0] : (d == 1 ? x[1] : x[TD - 1]));
            v *= sep_factor(e->kind[k], e->a[k], e->b[k], s);
        }
        acc += v;
    }
    return acc;
}

// coefficient at quadrature point q of cell c: `ci` indexes the per-cell /
// per-point arrays (which may be a block's staged copy), `cg` is the global
// cell id (fields of a separable coefficient)
// SEP: the kernel was instantiated for a separable coefficient; kernels
// without one do not carry the transcendental code (it costs registers).
template <int TD, bool SEP>
__device__ inline double coef_at(int kind, double scale, const double* __restrict__ coef, int64_t ci, int64_t cg,
                                 int q, int nq, const double p[TD + 1][TD]) {
    if constexpr (SEP) {
        if (kind == FA_COEF_SEP) return scale * sep_value<TD>(reinterpret_cast<const fa_sep_t*>(coef), q, p, cg);
    }
    return coef_value(kind, scale, coef, ci, q, nq);
}

// One thread per (cell, row i); the row of the element matrix is kept in
// registers (LDOF is a compile-time constant).
template <int TD, int LDOF, class T, bool SEP>
__global__ void __launch_bounds__(256) k_element_matrices(const double* __restrict__ node, const int32_t* __restrict__ cell,
                                                         int64_t nc, int nq, const double* __restrict__ qw,
                                                         const double* __restrict__ phi, const double* __restrict__ gphi,
                                                         int dk, double ds, const double* __restrict__ dc, int mk,
                                                         double ms, const double* __restrict__ mc, T* __restrict__ elem) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nc * LDOF) return;
    const int64_t c = t / LDOF;
    const int i = static_cast<int>(t - c * LDOF);
    double p[TD + 1][TD];
    load_cell<TD>(node, cell, c, p);
    Geom<TD> g;
    g.compute(p);
    double acc[LDOF];
#pragma unroll
    for (int j = 0; j < LDOF; ++j) acc[j] = 0.0;
    for (int q = 0; q < nq; ++q) {
        const double wq = qw[q] * g.meas;
        if (dk) {
            const double kd = wq * coef_at<TD, SEP>(dk, ds, dc, c, c, q, nq, p);
            const double* R = gphi + (int64_t)q * LDOF * (TD + 1);
            double gi[TD];
#pragma unroll
            for (int d = 0; d < TD; ++d) {
                double s = 0.0;
#pragma unroll
                for (int a = 0; a < TD + 1; ++a) s += R[i * (TD + 1) + a] * g.gl[a][d];
                gi[d] = s;
            }
#pragma unroll
            for (int j = 0; j < LDOF; ++j) {
                double dot = 0.0;
#pragma unroll
                for (int d = 0; d < TD; ++d) {
                    double s = 0.0;
#pragma unroll
                    for (int a = 0; a < TD + 1; ++a) s += R[j * (TD + 1) + a] * g.gl[a][d];
                    dot += gi[d] * s;
                }
                acc[j] += kd * dot;
            }
        }
        if (mk) {
            const double km = wq * coef_at<TD, SEP>(mk, ms, mc, c, c, q, nq, p);
            const double* ph = phi + (int64_t)q * LDOF;
            const double pi = ph[i] * km;
#pragma unroll
            for (int j = 0; j < LDOF; ++j) acc[j] += pi * ph[j];
        }
    }
    T* out = elem + t * LDOF;
#pragma unroll
    for (int j = 0; j < LDOF; ++j) out[j] = static_cast<T>(acc[j]);
}

template <int TD, class T, bool SEP>
__global__ void __launch_bounds__(256) k_element_vectors(const double* __restrict__ node, const int32_t* __restrict__ cell,
                                                        int64_t nc, int ldof, int nq, const double* __restrict__ qw,
                                                        const double* __restrict__ phi, int sk, double ss,
                                                        const double* __restrict__ src, T* __restrict__ elem) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nc * ldof) return;
    const int64_t c = t / ldof;
    const int i = static_cast<int>(t - c * ldof);
    double p[TD + 1][TD];
    load_cell<TD>(node, cell, c, p);
    Geom<TD> g;
    g.compute(p);
    double acc = 0.0;
    for (int q = 0; q < nq; ++q) acc += qw[q] * g.meas * coef_at<TD, SEP>(sk, ss, src, c, c, q, nq, p) * phi[(int64_t)q * ldof + i];
    elem[t] = static_cast<T>(acc);
}

// ---- fixed-order gather ---------------------------------------------------
__device__ inline int lower_bound_lds(const int32_t* a, int n, int32_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// G lanes per global row (G >= ldof); each row's columns and accumulators
// live in LDS.  The incident cells of a row are visited in ascending order
// and lane j adds element entry (i, j) of the current cell, so every CSR entry
// is summed in ascending cell order -- the order of the reference's stable
// lexsort + sequential np.add.at.  Distinct columns of one cell never collide.
template <class T, int G>
__global__ void k_gather_matrix(int64_t gdof, int ldof, const int32_t* __restrict__ c2d,
                                const int32_t* __restrict__ ptr, const int32_t* __restrict__ adj,
                                const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices, int lmax,
                                const T* __restrict__ elem, int64_t cstride, T* __restrict__ values) {
    extern __shared__ char smem[];
    const int rpb = blockDim.x / G;
    const int sub = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    int32_t* cols = reinterpret_cast<int32_t*>(smem) + sub * lmax;
    T* acc = reinterpret_cast<T*>(smem + align_up(sizeof(int32_t) * rpb * lmax, 16)) + sub * lmax;
    for (int64_t r = (int64_t)blockIdx.x * rpb + sub; r < gdof; r += (int64_t)gridDim.x * rpb) {
        const int32_t r0 = indptr[r];
        const int L = indptr[r + 1] - r0;
        for (int t = lane; t < L; t += G) {
            cols[t] = indices[r0 + t];
            acc[t] = T(0);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int32_t k1 = ptr[r + 1];
        if (ldof <= G) {
            // the row's incidences G at a time in one load, handed out by
            // shuffles; the next incidence's column and value loads are issued
            // before the current one's search and add (same adds, same order)
            for (int32_t kb = ptr[r]; kb < k1; kb += G) {
                const int32_t mine = kb + lane < k1 ? adj[kb + lane] : 0;
                const int nk = min(G, k1 - kb);
                const bool act = lane < ldof;
                int32_t e = __shfl(mine, 0, G);
                int32_t c = e / ldof;
                int32_t col = act ? c2d[(int64_t)c * ldof + lane] : 0;
                T v = act ? elem[(int64_t)c * cstride + (e - c * ldof) * ldof + lane] : T(0);
                for (int u = 0; u < nk; ++u) {
                    int32_t col_n = 0;
                    T v_n = T(0);
                    if (u + 1 < nk) {
                        const int32_t en = __shfl(mine, u + 1, G);
                        const int32_t cn = en / ldof;
                        if (act) {
                            col_n = c2d[(int64_t)cn * ldof + lane];
                            v_n = elem[(int64_t)cn * cstride + (en - cn * ldof) * ldof + lane];
                        }
                    }
                    if (act) {
                        const int pos = lower_bound_lds(cols, L, col);
                        if (pos < L && cols[pos] == col) acc[pos] += v;
                    }
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    col = col_n;
                    v = v_n;
                }
            }
        } else {
            for (int32_t k = ptr[r]; k < k1; ++k) {
                const int32_t e = adj[k];
                const int32_t c = e / ldof;
                for (int j = lane; j < ldof; j += G) {
                    const int32_t col = c2d[(int64_t)c * ldof + j];
                    const T v = elem[(int64_t)c * cstride + (e - c * ldof) * ldof + j];
                    const int pos = lower_bound_lds(cols, L, col);
                    if (pos < L && cols[pos] == col) acc[pos] += v;
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
        for (int t = lane; t < L; t += G) values[r0 + t] = acc[t];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// 16 lanes per DOF, lane l adds incidences l, l+16, ...; the lane partials
// are combined by a fixed tree (deterministic)
template <class T>
__global__ void __launch_bounds__(256) k_gather_vector(int64_t gdof, const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ adj, const T* __restrict__ elem,
                                                      T* __restrict__ out) {
    constexpr int G = 16;
    const int lane = threadIdx.x % G;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; r < gdof;
         r += (int64_t)gridDim.x * blockDim.x / G) {
        T s = T(0);
        const int32_t k1 = ptr[r + 1];
        for (int32_t k = ptr[r] + lane; k < k1; k += G) s += elem[adj[k]];
        s = group_sum<G>(s);
        if (lane == 0) out[r] = s;
    }
}

// physical coordinates of barycentric points on every cell: one thread per
// cell reads its vertices once; the block's nq*gd outputs per cell are staged
// in LDS and written back as one contiguous, coalesced range
template <int GD, int NV>
__global__ void k_bc_to_point(int64_t nc, int nq, const double* __restrict__ node, const int32_t* __restrict__ cell,
                              const double* __restrict__ bcs, double* __restrict__ out) {
    extern __shared__ double stage[];
    const int per = nq * GD;
    for (int64_t c0 = (int64_t)blockIdx.x * blockDim.x; c0 < nc; c0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = c0 + threadIdx.x;
        if (c < nc) {
            double v[NV][GD];
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const int64_t nd = cell[c * NV + a];
#pragma unroll
                for (int d = 0; d < GD; ++d) v[a][d] = node[nd * GD + d];
            }
            double* o = stage + threadIdx.x * per;
            for (int q = 0; q < nq; ++q) {
#pragma unroll
                for (int d = 0; d < GD; ++d) {
                    double x = 0.0;
#pragma unroll
                    for (int a = 0; a < NV; ++a) x += bcs[q * NV + a] * v[a][d];
                    o[q * GD + d] = x;
                }
            }
        }
        __syncthreads();
        const int64_t rem = nc - c0;
        const int64_t cnt = (rem < (int64_t)blockDim.x ? rem : (int64_t)blockDim.x) * per;
        double* o = out + c0 * per;
        for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) o[t] = stage[t];
        __syncthreads();
    }
}

// ---- linear elasticity (vector Lagrange, isotropic) --------------------------
// A_ab[i][j] = |T| sum_q w_q d_a phi_i d_b phi_j
//            = |T| sum_{m,n} gl[m][a] gl[n][b] M[i][j][m][n],
// M[i][j][m][n] = sum_q w_q R[q][i][m] R[q][j][n] precontracted on the host
// (R = d phi / d lambda).  Element entry of (i, a), (j, b):
//   a == b: cA A_aa + cB (sum of the other diagonal blocks, in axis order)
//   a != b: cC A_ab + cD A_ba
// laid out as i*TD + a (dof_priority = 0) or a*LDOF + i (dof_priority = 1).
// A block takes el_cells() cells at a time: their geometry (gl, |T|) is
// computed once per cell into LDS, then one thread per (cell, i, j) forms the
// TD x TD block of the pair (i, j) and stores its TD rows of TD entries.
// Consecutive threads take consecutive j, so each store of a wavefront covers
// contiguous runs of a row (one thread per (cell, i) scattered its stores over
// 64 cache lines).  Same operations in the same order as the per-(cell, i)
// form: tm[m][b] = sum_n M[i][j][m][n] gl[n][b], A_ab = |T| sum_m gl[m][a] tm[m][b].
template <int TD, int LDOF>
constexpr int el_cells() {
    return (1024 / (LDOF * LDOF)) < 1 ? 1 : ((1024 / (LDOF * LDOF)) > 256 ? 256 : 1024 / (LDOF * LDOF));
}

template <int TD, int LDOF, class T>
__global__ void __launch_bounds__(256) k_element_elasticity(const double* __restrict__ node,
                                                           const int32_t* __restrict__ cell, int64_t nc,
                                                           const double* __restrict__ M, double cA, double cB,
                                                           double cC, double cD, int dof_priority,
                                                           T* __restrict__ elem) {
    constexpr int NV = TD + 1;
    constexpr int N = TD * LDOF;
    constexpr int CB = el_cells<TD, LDOF>();
    __shared__ double sgl[CB][NV][TD];
    __shared__ double smeas[CB];
    for (int64_t c0 = (int64_t)blockIdx.x * CB; c0 < nc; c0 += (int64_t)gridDim.x * CB) {
        const int nb = static_cast<int>(nc - c0 < CB ? nc - c0 : CB);
        if (threadIdx.x < nb) {
            double p[NV][TD];
            load_cell<TD>(node, cell, c0 + threadIdx.x, p);
            Geom<TD> g;
            g.compute(p);
#pragma unroll
            for (int m = 0; m < NV; ++m)
#pragma unroll
                for (int d = 0; d < TD; ++d) sgl[threadIdx.x][m][d] = g.gl[m][d];
            smeas[threadIdx.x] = g.meas;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nb * LDOF * LDOF; e += blockDim.x) {
            const int cl = e / (LDOF * LDOF);
            const int ij = e - cl * LDOF * LDOF;
            const int i = ij / LDOF, j = ij - i * LDOF;
            const double* Mij = M + (int64_t)ij * NV * NV;
            double mm[NV * NV];
#pragma unroll
            for (int q = 0; q < NV * NV; ++q) mm[q] = Mij[q];
            double gl[NV][TD];
#pragma unroll
            for (int m = 0; m < NV; ++m)
#pragma unroll
                for (int d = 0; d < TD; ++d) gl[m][d] = sgl[cl][m][d];
            const double meas = smeas[cl];
            double tm[NV][TD];
#pragma unroll
            for (int m = 0; m < NV; ++m)
#pragma unroll
                for (int bb = 0; bb < TD; ++bb) {
                    double s = 0.0;
#pragma unroll
                    for (int n = 0; n < NV; ++n) s += mm[m * NV + n] * gl[n][bb];
                    tm[m][bb] = s;
                }
            double A[TD][TD];
#pragma unroll
            for (int aa = 0; aa < TD; ++aa)
#pragma unroll
                for (int bb = 0; bb < TD; ++bb) {
                    double s = 0.0;
#pragma unroll
                    for (int m = 0; m < NV; ++m) s += gl[m][aa] * tm[m][bb];
                    A[aa][bb] = meas * s;
                }
            T* out = elem + (c0 + cl) * (int64_t)N * N;
#pragma unroll
            for (int aa = 0; aa < TD; ++aa) {
                const int row = dof_priority ? aa * LDOF + i : i * TD + aa;
#pragma unroll
                for (int bb = 0; bb < TD; ++bb) {
                    double v;
                    if (aa == bb) {
                        double others = 0.0;
#pragma unroll
                        for (int f = 0; f < TD; ++f)
                            if (f != aa) others += A[f][f];
                        v = cA * A[aa][aa] + cB * others;
                    } else {
                        v = cC * A[aa][bb] + cD * A[bb][aa];
                    }
                    const int col = dof_priority ? bb * LDOF + j : j * TD + bb;
                    out[(int64_t)row * N + col] = static_cast<T>(v);
                }
            }
        }
        __syncthreads();
    }
}

// vector load: out[c][I(i, a)] = |T| sum_q w_q phi_i(q) f[c][q][a], f (NC, NQ, ncomp)
template <int TD, class T>
__global__ void __launch_bounds__(256) k_element_vectors_tensor(const double* __restrict__ node,
                                                               const int32_t* __restrict__ cell, int64_t nc,
                                                               int ldof, int ncomp, int dof_priority, int nq,
                                                               const double* __restrict__ qw,
                                                               const double* __restrict__ phi,
                                                               const double* __restrict__ src,
                                                               T* __restrict__ elem) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nc * ldof) return;
    const int64_t c = t / ldof;
    const int i = static_cast<int>(t - c * ldof);
    double p[TD + 1][TD];
    load_cell<TD>(node, cell, c, p);
    Geom<TD> g;
    g.compute(p);
    T* out = elem + c * (int64_t)ldof * ncomp;
    for (int a = 0; a < ncomp; ++a) {
        double acc = 0.0;
        for (int q = 0; q < nq; ++q) acc += qw[q] * g.meas * src[(c * nq + q) * ncomp + a] * phi[(int64_t)q * ldof + i];
        out[dof_priority ? a * ldof + i : i * ncomp + a] = static_cast<T>(acc);
    }
}

// ---- fused P1 assembly: element rows computed where they are summed ---------
// P1 gradients are constant on a cell, so the element row i of cell c is
//   A_ij = kd |T| (gl_i . gl_j) + |T| sum_q w_q k_q phi_i(q) phi_j(q),
// kd = sum_q w_q k_q, gl_i = n_i / (6|T|) with n_i the face normals of
// Geom<3> (tetrahedron) or the rotated edges of Geom<2> (triangle).
template <int TD>
struct P1Cell {
    double n[TD + 1][TD];  // grad lambda times the scale `den` (den = 6|T| or 2|T|)
    double den, meas;
    __device__ void compute(const double p[TD + 1][TD]);
};

template <>
__device__ inline void P1Cell<3>::compute(const double p[4][3]) {
    double a[3], b[3], c[3], t[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = p[1][d] - p[0][d];
        b[d] = p[2][d] - p[1][d];
        c[d] = p[3][d] - p[2][d];
    }
    cross3(b, c, t);
    den = a[0] * t[0] + a[1] * t[1] + a[2] * t[2];
    const int lf[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double vjk[3], vjm[3];
        for (int d = 0; d < 3; ++d) {
            vjk[d] = p[lf[i][1]][d] - p[lf[i][0]][d];
            vjm[d] = p[lf[i][2]][d] - p[lf[i][0]][d];
        }
        cross3(vjm, vjk, n[i]);
    }
    double v01[3], v02[3], v03[3], x[3];
    for (int d = 0; d < 3; ++d) {
        v01[d] = p[1][d] - p[0][d];
        v02[d] = p[2][d] - p[0][d];
        v03[d] = p[3][d] - p[0][d];
    }
    cross3(v01, v02, x);
    meas = (v03[0] * x[0] + v03[1] * x[1] + v03[2] * x[2]) / 6.0;
}

template <>
__device__ inline void P1Cell<2>::compute(const double p[3][2]) {
    double e[3][2];
    for (int d = 0; d < 2; ++d) {
        e[0][d] = p[2][d] - p[1][d];
        e[1][d] = p[0][d] - p[2][d];
        e[2][d] = p[1][d] - p[0][d];
    }
    den = e[0][0] * e[1][1] - e[0][1] * e[1][0];
    for (int i = 0; i < 3; ++i) {
        n[i][0] = -e[i][1];
        n[i][1] = e[i][0];
    }
    const double a0 = p[1][0] - p[0][0], a1 = p[1][1] - p[0][1];
    const double b0 = p[2][0] - p[1][0], b1 = p[2][1] - p[1][1];
    meas = (a0 * b1 - a1 * b0) / 2.0;
}

// sum_q w_q k(x_q): the diffusion coefficient's share of a P1 element matrix
// (grad lambda is constant on the cell)
template <int TD, bool SEP>
__device__ inline double p1_kd(int nq, const double* __restrict__ qw, int dk, double ds, const double* __restrict__ dc,
                               int64_t c, int64_t cg, const double p[TD + 1][TD]) {
    double kd = 0.0;
    if (dk)
        for (int q = 0; q < nq; ++q) kd += qw[q] * coef_at<TD, SEP>(dk, ds, dc, c, cg, q, nq, p);
    return kd;
}

template <int TD, bool SEP>
__device__ inline void p1_row(const P1Cell<TD>& g, const double p[TD + 1][TD], int64_t cg, double kd, int i, int64_t c,
                              int nq, const double* __restrict__ qw,
                              const double* __restrict__ phi, int dk, double ds, const double* __restrict__ dc, int mk,
                              double ms, const double* __restrict__ mc, double out[TD + 1]) {
#pragma unroll
    for (int j = 0; j < TD + 1; ++j) out[j] = 0.0;
    if (dk) {
        const double s = kd * g.meas / (g.den * g.den);
        double ni[TD];  // row i selected without a runtime index into g.n (keeps it in registers)
#pragma unroll
        for (int d = 0; d < TD; ++d) ni[d] = g.n[0][d];
#pragma unroll
        for (int a = 1; a < TD + 1; ++a)
            if (i == a) {
#pragma unroll
                for (int d = 0; d < TD; ++d) ni[d] = g.n[a][d];
            }
#pragma unroll
        for (int j = 0; j < TD + 1; ++j) {
            double dot = 0.0;
#pragma unroll
            for (int d = 0; d < TD; ++d) dot += ni[d] * g.n[j][d];
            out[j] = s * dot;
        }
    }
    if (mk) {
        for (int q = 0; q < nq; ++q) {
            const double* ph = phi + q * (TD + 1);
            const double w = qw[q] * g.meas * coef_at<TD, SEP>(mk, ms, mc, c, cg, q, nq, p) * ph[i];
#pragma unroll
            for (int j = 0; j < TD + 1; ++j) out[j] += w * ph[j];
        }
    }
}

// G lanes per CSR row, K row positions owned per lane.  Incidences of row r
// (cell c, local i; ascending c) are taken G at a time: lane l computes the
// element row of its incidence and parks value j at tab[pos(col_j)][l]; the
// owner of each position then adds tab[pos][0..G) in lane order.  Every CSR
// entry is therefore the sequential sum of its cell contributions in
// ascending cell order, the order of the reference's stable lexsort followed
// by sequential np.add.at (coo_tensor.py:184-213).
template <int TD, int G, int K, class T, bool SEP>
__global__ void __launch_bounds__(256) k_assemble_p1(int64_t gdof, const double* __restrict__ node,
                                                    const int32_t* __restrict__ cell, int nq,
                                                    const double* __restrict__ qw, const double* __restrict__ phi,
                                                    int dk, double ds, const double* __restrict__ dc, int mk, double ms,
                                                    const double* __restrict__ mc, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ adj, const int32_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices, int lmax,
                                                    T* __restrict__ values) {
    constexpr int NV = TD + 1;
    constexpr int GS = G + 1;  // table row stride: odd, so owner lanes reading their rows hit distinct banks
    extern __shared__ char smem[];
    const int rpb = blockDim.x / G;
    const int sub = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    int32_t* cols = reinterpret_cast<int32_t*>(smem) + sub * lmax;
    double* tab = reinterpret_cast<double*>(smem + align_up(sizeof(int32_t) * rpb * lmax, 16)) + (int64_t)sub * lmax * GS;
    for (int64_t r = (int64_t)blockIdx.x * rpb + sub; r < gdof; r += (int64_t)gridDim.x * rpb) {
        const int32_t r0 = indptr[r];
        const int L = indptr[r + 1] - r0;
        for (int t = lane; t < L; t += G) cols[t] = indices[r0 + t];
        for (int t = 0; t < L; ++t) tab[t * GS + lane] = 0.0;
        double acc[K];
#pragma unroll
        for (int m = 0; m < K; ++m) acc[m] = 0.0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int32_t k1 = ptr[r + 1];
        for (int32_t kb = ptr[r]; kb < k1; kb += G) {
            const int32_t k = kb + lane;
            int pos[NV];
            if (k < k1) {
                const int32_t e = adj[k];
                const int64_t c = e / NV;
                const int i = e - static_cast<int>(c) * NV;
                double p[NV][TD];
                load_cell<TD>(node, cell, c, p);
                P1Cell<TD> g;
                g.compute(p);
                double v[NV];
                const double kd = p1_kd<TD, SEP>(nq, qw, dk, ds, dc, c, c, p);
                p1_row<TD, SEP>(g, p, c, kd, i, c, nq, qw, phi, dk, ds, dc, mk, ms, mc, v);
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    pos[j] = lower_bound_lds(cols, L, cell[c * NV + j]);
                    tab[pos[j] * GS + lane] = v[j];
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const int t = lane + m * G;
                if (t < L) {
#pragma unroll
                    for (int l = 0; l < G; ++l) acc[m] += tab[t * GS + l];
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (k < k1) {
#pragma unroll
                for (int j = 0; j < NV; ++j) tab[pos[j] * GS + lane] = 0.0;
            }
        }
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const int t = lane + m * G;
            if (t < L) values[r0 + t] = static_cast<T>(acc[m]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// load vector, P1: b_r = sum over incidences (c, i) of |T_c| sum_q w_q f_cq phi_i(q);
// G lanes per DOF, lane partials combined by a fixed tree
template <int TD, int G, class T, bool SEP>
__global__ void __launch_bounds__(256) k_assemble_p1_vec(int64_t gdof, const double* __restrict__ node,
                                                        const int32_t* __restrict__ cell, int nq,
                                                        const double* __restrict__ qw,
                                                        const double* __restrict__ phi, int sk, double ss,
                                                        const double* __restrict__ src,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ adj, T* __restrict__ out) {
    constexpr int NV = TD + 1;
    const int lane = threadIdx.x % G;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; r < gdof;
         r += (int64_t)gridDim.x * blockDim.x / G) {
        double s = 0.0;
        const int32_t k1 = ptr[r + 1];
        for (int32_t k = ptr[r] + lane; k < k1; k += G) {
            const int32_t e = adj[k];
            const int64_t c = e / NV;
            const int i = e - static_cast<int>(c) * NV;
            double p[NV][TD];
            load_cell<TD>(node, cell, c, p);
            P1Cell<TD> g;
            g.compute(p);
            double v = 0.0;
            for (int q = 0; q < nq; ++q) v += qw[q] * g.meas * coef_at<TD, SEP>(sk, ss, src, c, c, q, nq, p) * phi[q * NV + i];
            s += v;
        }
        s = group_sum<G>(s);
        if (lane == 0) out[r] = static_cast<T>(s);
    }
}

template <int TD, int G, int K, class T>
static int launch_p1(int64_t gdof, const double* node, const int32_t* cell, int nq, const double* qw,
                     const double* phi, int dk, double ds, const double* dc, int mk, double ms, const double* mc,
                     const int32_t* ptr, const int32_t* adj, const int32_t* indptr, const int32_t* indices, int lmax,
                     void* values, hipStream_t s) {
    const size_t per_row = sizeof(int32_t) * lmax + sizeof(double) * lmax * (G + 1);
    int rpb = 256 / G;
    while (rpb > 1 && rpb * per_row + 16 > 48 * 1024) rpb >>= 1;
    const size_t lds = align_up(sizeof(int32_t) * rpb * lmax, 16) + sizeof(double) * rpb * lmax * (G + 1);
    if (lds > 64 * 1024) {
        set_error("fa_assemble_p1_matrix: row length %d exceeds the LDS capacity", lmax);
        return FA_ECAPACITY;
    }
    const unsigned grid = grid_for(gdof, rpb, 1 << 20);
    if (dk == FA_COEF_SEP || mk == FA_COEF_SEP)
        hipLaunchKernelGGL((k_assemble_p1<TD, G, K, T, true>), dim3(grid), dim3(rpb * G), lds, s, gdof, node, cell, nq,
                           qw, phi, dk, ds, dc, mk, ms, mc, ptr, adj, indptr, indices, lmax, static_cast<T*>(values));
    else
        hipLaunchKernelGGL((k_assemble_p1<TD, G, K, T, false>), dim3(grid), dim3(rpb * G), lds, s, gdof, node, cell, nq,
                           qw, phi, dk, ds, dc, mk, ms, mc, ptr, adj, indptr, indices, lmax, static_cast<T*>(values));
    FA_LAUNCH("k_assemble_p1");
    return FA_OK;
}

template <int TD, int LDOF, class T>
static void launch_elem(const double* node, const int32_t* cell, int64_t nc, int nq, const double* qw,
                        const double* phi, const double* gphi, int dk, double ds, const double* dc, int mk, double ms,
                        const double* mc, void* elem, hipStream_t s) {
    const int64_t n = nc * LDOF;
    if (dk == FA_COEF_SEP || mk == FA_COEF_SEP)
        hipLaunchKernelGGL((k_element_matrices<TD, LDOF, T, true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           node, cell, nc, nq, qw, phi, gphi, dk, ds, dc, mk, ms, mc, static_cast<T*>(elem));
    else
        hipLaunchKernelGGL((k_element_matrices<TD, LDOF, T, false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           node, cell, nc, nq, qw, phi, gphi, dk, ds, dc, mk, ms, mc, static_cast<T*>(elem));
}

template <class T>
static int dispatch_elem(int td, int ldof, const double* node, const int32_t* cell, int64_t nc, int nq,
                         const double* qw, const double* phi, const double* gphi, int dk, double ds, const double* dc,
                         int mk, double ms, const double* mc, void* elem, hipStream_t s) {
#define FA_ELEM_CASE(TD_, L_)                                                                             \
    if (td == TD_ && ldof == L_) {                                                                        \
        launch_elem<TD_, L_, T>(node, cell, nc, nq, qw, phi, gphi, dk, ds, dc, mk, ms, mc, elem, s);      \
        FA_LAUNCH("k_element_matrices");                                                                  \
        return FA_OK;                                                                                     \
    }
    FA_ELEM_CASE(2, 3) FA_ELEM_CASE(2, 6) FA_ELEM_CASE(2, 10) FA_ELEM_CASE(2, 15)
    FA_ELEM_CASE(3, 4) FA_ELEM_CASE(3, 10) FA_ELEM_CASE(3, 20) FA_ELEM_CASE(3, 35)
#undef FA_ELEM_CASE
    set_error("fa_element_matrices: unsupported (td=%d, ldof=%d); p must be 1..4", td, ldof);
    return FA_EINVAL;
}

template <class T, int G>
static int launch_gather(int64_t gdof, int ldof, const int32_t* c2d, const int32_t* ptr, const int32_t* adj,
                         const int32_t* indptr, const int32_t* indices, int lmax, const void* elem, int64_t cstride,
                         void* values, hipStream_t s) {
    const size_t per_row = sizeof(int32_t) * lmax + sizeof(T) * lmax;
    int rpb = 256 / G;
    while (rpb > 1 && rpb * per_row + 16 > 48 * 1024) rpb >>= 1;
    const size_t lds = align_up(sizeof(int32_t) * rpb * lmax, 16) + sizeof(T) * rpb * lmax;
    if (lds > 64 * 1024) {
        set_error("fa_gather_matrix: row length %d exceeds the LDS capacity", lmax);
        return FA_ECAPACITY;
    }
    const unsigned grid = grid_for(gdof, rpb, 1 << 20);
    hipLaunchKernelGGL((k_gather_matrix<T, G>), dim3(grid), dim3(rpb * G), lds, s, gdof, ldof, c2d, ptr, adj, indptr,
                       indices, lmax, static_cast<const T*>(elem), cstride, static_cast<T*>(values));
    FA_LAUNCH("k_gather_matrix");
    return FA_OK;
}

template <class T>
static int dispatch_gather(int64_t gdof, int ldof, const int32_t* c2d, const int32_t* ptr, const int32_t* adj,
                           const int32_t* indptr, const int32_t* indices, int lmax, const void* elem, int64_t cs,
                           void* values, hipStream_t s) {
    if (ldof <= 4) return launch_gather<T, 4>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, cs, values, s);
    if (ldof <= 8) return launch_gather<T, 8>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, cs, values, s);
    if (ldof <= 16)
        return launch_gather<T, 16>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, cs, values, s);
    if (ldof <= 32)
        return launch_gather<T, 32>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, cs, values, s);
    return launch_gather<T, 64>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, cs, values, s);
}

}  // namespace fa

using namespace fa;

extern "C" int fa_element_matrices(int dtype, int td, int ldof, int nq, const double* node, const int32_t* cell,
                                   int64_t nc, const double* qw, const double* phi, const double* gphi_u,
                                   int diff_kind, double diff_scale, const double* diff_coef, int mass_kind,
                                   double mass_scale, const double* mass_coef, void* elem, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_element_matrices: bad dtype %d", dtype);
    FA_REQUIRE(nq > 0 && nc >= 0, "fa_element_matrices: bad sizes");
    FA_REQUIRE(!diff_kind || gphi_u, "fa_element_matrices: diffusion needs gphi_u");
    FA_REQUIRE(!mass_kind || phi, "fa_element_matrices: mass needs phi");
    FA_REQUIRE(!(diff_kind >= FA_COEF_CELL && !diff_coef) && !(mass_kind >= FA_COEF_CELL && !mass_coef),
               "fa_element_matrices: coefficient array missing");
    if (nc == 0) return FA_OK;
    if (dtype == FA_F64)
        return dispatch_elem<double>(td, ldof, node, cell, nc, nq, qw, phi, gphi_u, diff_kind, diff_scale, diff_coef,
                                     mass_kind, mass_scale, mass_coef, elem, S(stream));
    return dispatch_elem<float>(td, ldof, node, cell, nc, nq, qw, phi, gphi_u, diff_kind, diff_scale, diff_coef,
                                mass_kind, mass_scale, mass_coef, elem, S(stream));
}

extern "C" int fa_element_vectors(int dtype, int td, int ldof, int nq, const double* node, const int32_t* cell,
                                  int64_t nc, const double* qw, const double* phi, int src_kind, double src_scale,
                                  const double* src, void* elem, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_element_vectors: bad dtype %d", dtype);
    FA_REQUIRE(td == 2 || td == 3, "fa_element_vectors: td must be 2 or 3");
    FA_REQUIRE(src_kind < FA_COEF_CELL || src, "fa_element_vectors: source array missing");
    if (nc == 0) return FA_OK;
    const int64_t n = nc * ldof;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipStream_t s = S(stream);
#define FA_EV(TD_, T_, SEP_)                                                                                     \
    hipLaunchKernelGGL((k_element_vectors<TD_, T_, SEP_>), dim3(grid), dim3(256), 0, s, node, cell, nc, ldof, nq, qw, \
                       phi, src_kind, src_scale, src, static_cast<T_*>(elem))
#define FA_EV2(TD_, T_) \
    if (src_kind == FA_COEF_SEP) FA_EV(TD_, T_, true); else FA_EV(TD_, T_, false)
    if (dtype == FA_F64) {
        if (td == 2) { FA_EV2(2, double); } else { FA_EV2(3, double); }
    } else {
        if (td == 2) { FA_EV2(2, float); } else { FA_EV2(3, float); }
    }
#undef FA_EV2
#undef FA_EV
    FA_LAUNCH("k_element_vectors");
    return FA_OK;
}

extern "C" int fa_gather_matrix(int dtype, int64_t gdof, int32_t ldof, const int32_t* cell2dof,
                                const int32_t* adj_ptr, const int32_t* adj, const int32_t* indptr,
                                const int32_t* indices, int32_t max_row, const void* elem, int64_t elem_stride,
                                void* values, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_gather_matrix: bad dtype %d", dtype);
    FA_REQUIRE(gdof > 0 && ldof > 0 && max_row > 0, "fa_gather_matrix: bad sizes");
    const int64_t cs = elem_stride > 0 ? elem_stride : (int64_t)ldof * ldof;
    FA_REQUIRE(cs >= (int64_t)ldof * ldof, "fa_gather_matrix: element stride %lld < ldof^2", (long long)cs);
    if (dtype == FA_F64)
        return dispatch_gather<double>(gdof, ldof, cell2dof, adj_ptr, adj, indptr, indices, max_row, elem, cs, values,
                                       S(stream));
    return dispatch_gather<float>(gdof, ldof, cell2dof, adj_ptr, adj, indptr, indices, max_row, elem, cs, values,
                                  S(stream));
}

extern "C" int fa_gather_vector(int dtype, int64_t gdof, int32_t ldof, const int32_t* adj_ptr, const int32_t* adj,
                                const void* elem, void* out, fa_stream_t stream) {
    (void)ldof;
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_gather_vector: bad dtype %d", dtype);
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(gdof * 16, 256, 1 << 16);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_gather_vector<double>, dim3(grid), dim3(256), 0, s, gdof, adj_ptr, adj,
                           static_cast<const double*>(elem), static_cast<double*>(out));
    else
        hipLaunchKernelGGL(k_gather_vector<float>, dim3(grid), dim3(256), 0, s, gdof, adj_ptr, adj,
                           static_cast<const float*>(elem), static_cast<float*>(out));
    FA_LAUNCH("k_gather_vector");
    return FA_OK;
}

// v[c][i] = sum_j elem[c][i][j] x[c2d[c][j]]: one thread per (cell, row),
// the row read contiguously, the cell's x entries gathered (shared by the
// cell's rows, so they come from L1 after the first)
template <class T>
__global__ void __launch_bounds__(256) k_element_apply(int64_t nc, int ldof, const int32_t* __restrict__ c2d,
                                                      const T* __restrict__ elem, int64_t stride,
                                                      const T* __restrict__ x, T* __restrict__ v) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nc * ldof;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = t / ldof;
        const int i = static_cast<int>(t - c * ldof);
        const T* row = elem + c * stride + static_cast<int64_t>(i) * ldof;
        const int32_t* dofs = c2d + c * ldof;
        T s = T(0);
        for (int j = 0; j < ldof; ++j) s += row[j] * x[dofs[j]];
        v[t] = s;
    }
}

extern "C" int fa_element_apply(int dtype, int64_t nc, int32_t ldof, const int32_t* cell2dof, const void* elem,
                                int64_t elem_stride, const void* x, void* v, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_element_apply: bad dtype %d", dtype);
    FA_REQUIRE(ldof >= 1 && elem_stride >= (int64_t)ldof * ldof, "fa_element_apply: bad sizes");
    if (nc == 0) return FA_OK;
    const unsigned grid = grid_for(nc * ldof, 256, 1 << 20);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_element_apply<double>, dim3(grid), dim3(256), 0, S(stream), nc, ldof, cell2dof,
                           static_cast<const double*>(elem), elem_stride, static_cast<const double*>(x),
                           static_cast<double*>(v));
    else
        hipLaunchKernelGGL(k_element_apply<float>, dim3(grid), dim3(256), 0, S(stream), nc, ldof, cell2dof,
                           static_cast<const float*>(elem), elem_stride, static_cast<const float*>(x),
                           static_cast<float*>(v));
    FA_LAUNCH("k_element_apply");
    return FA_OK;
}

extern "C" int fa_bc_to_point(int gd, int nv, int64_t nc, int nq, const double* node, const int32_t* cell,
                              const double* bcs, double* out, fa_stream_t stream) {
    FA_REQUIRE(nq >= 1 && nc >= 0, "fa_bc_to_point: bad sizes");
    if (nc == 0) return FA_OK;
    hipStream_t s = S(stream);
    const size_t per = static_cast<size_t>(nq) * gd * sizeof(double);
    FA_REQUIRE(16 * per <= 64 * 1024, "fa_bc_to_point: %d points per cell exceed the staging capacity", nq);
    unsigned blk = 256;
    while (blk > 16 && blk * per > 48 * 1024) blk >>= 1;
    const unsigned grid = grid_for(nc, blk, 1 << 16);
#define FA_BP(GD, NV) \
    hipLaunchKernelGGL((k_bc_to_point<GD, NV>), dim3(grid), dim3(blk), blk * per, s, nc, nq, node, cell, bcs, out)
    if (gd == 2 && nv == 3) FA_BP(2, 3);
    else if (gd == 3 && nv == 3) FA_BP(3, 3);
    else if (gd == 3 && nv == 4) FA_BP(3, 4);
    else {
        set_error("fa_bc_to_point: unsupported (gd, nv) = (%d, %d)", gd, nv);
        return FA_EINVAL;
    }
#undef FA_BP
    FA_LAUNCH("k_bc_to_point");
    return FA_OK;
}

extern "C" int fa_assemble_p1_matrix(int dtype, int td, int64_t gdof, int nq, const double* node, const int32_t* cell,
                                     const double* qw, const double* phi, int diff_kind, double diff_scale,
                                     const double* diff_coef, int mass_kind, double mass_scale,
                                     const double* mass_coef, const int32_t* adj_ptr, const int32_t* adj,
                                     const int32_t* indptr, const int32_t* indices, int32_t max_row, void* values,
                                     fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_assemble_p1_matrix: bad dtype %d", dtype);
    FA_REQUIRE(td == 2 || td == 3, "fa_assemble_p1_matrix: td must be 2 or 3");
    FA_REQUIRE(gdof > 0 && nq > 0 && max_row > 0, "fa_assemble_p1_matrix: bad sizes");
    FA_REQUIRE(!mass_kind || phi, "fa_assemble_p1_matrix: mass needs phi");
    FA_REQUIRE(!(diff_kind >= FA_COEF_CELL && !diff_coef) && !(mass_kind >= FA_COEF_CELL && !mass_coef),
               "fa_assemble_p1_matrix: coefficient array missing");
    hipStream_t s = S(stream);
#define FA_P1(TD_, G_, K_)                                                                                     \
    return dtype == FA_F64                                                                                     \
               ? launch_p1<TD_, G_, K_, double>(gdof, node, cell, nq, qw, phi, diff_kind, diff_scale, diff_coef, \
                                                mass_kind, mass_scale, mass_coef, adj_ptr, adj, indptr, indices,   \
                                                max_row, values, s)                                              \
               : launch_p1<TD_, G_, K_, float>(gdof, node, cell, nq, qw, phi, diff_kind, diff_scale, diff_coef,  \
                                               mass_kind, mass_scale, mass_coef, adj_ptr, adj, indptr, indices,    \
                                               max_row, values, s)
    if (td == 3) {
        if (max_row <= 32) FA_P1(3, 16, 2);
        if (max_row <= 64) FA_P1(3, 16, 4);
        if (max_row <= 128) FA_P1(3, 32, 4);
    } else {
        if (max_row <= 16) FA_P1(2, 8, 2);
        if (max_row <= 32) FA_P1(2, 8, 4);
        if (max_row <= 128) FA_P1(2, 32, 4);
    }
#undef FA_P1
    set_error("fa_assemble_p1_matrix: rows of %d entries exceed the kernel capacity (128)", max_row);
    return FA_ECAPACITY;
}

extern "C" int fa_assemble_p1_vector(int dtype, int td, int64_t gdof, int nq, const double* node, const int32_t* cell,
                                     const double* qw, const double* phi, int src_kind, double src_scale,
                                     const double* src, const int32_t* adj_ptr, const int32_t* adj, void* out,
                                     fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_assemble_p1_vector: bad dtype %d", dtype);
    FA_REQUIRE(td == 2 || td == 3, "fa_assemble_p1_vector: td must be 2 or 3");
    FA_REQUIRE(gdof > 0 && nq > 0 && phi, "fa_assemble_p1_vector: bad sizes");
    FA_REQUIRE(src_kind < FA_COEF_CELL || src, "fa_assemble_p1_vector: source array missing");
    hipStream_t s = S(stream);
    constexpr int G = 8;
    const unsigned grid = grid_for(gdof * G, 256, 1 << 16);
#define FA_V1(TD_, T_, SEP_)                                                                                     \
    hipLaunchKernelGGL((k_assemble_p1_vec<TD_, G, T_, SEP_>), dim3(grid), dim3(256), 0, s, gdof, node, cell, nq, qw, \
                       phi, src_kind, src_scale, src, adj_ptr, adj, static_cast<T_*>(out))
#define FA_V(TD_, T_) \
    if (src_kind == FA_COEF_SEP) FA_V1(TD_, T_, true); else FA_V1(TD_, T_, false)
    if (td == 3) {
        if (dtype == FA_F64) FA_V(3, double); else FA_V(3, float);
    } else {
        if (dtype == FA_F64) FA_V(2, double); else FA_V(2, float);
    }
#undef FA_V
#undef FA_V1
    FA_LAUNCH("k_assemble_p1_vec");
    return FA_OK;
}

extern "C" int fa_element_elasticity(int dtype, int td, int ldof, const double* node, const int32_t* cell, int64_t nc,
                                     const double* M, double cA, double cB, double cC, double cD, int dof_priority,
                                     void* elem, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_element_elasticity: bad dtype %d", dtype);
    FA_REQUIRE(M && nc >= 0, "fa_element_elasticity: bad arguments");
    if (nc == 0) return FA_OK;
    hipStream_t s = S(stream);
#define FA_EL(TD_, L_)                                                                                         \
    if (td == TD_ && ldof == L_) {                                                                             \
        const unsigned grid = grid_for(nc, el_cells<TD_, L_>(), 1 << 20);                                        \
        if (dtype == FA_F64)                                                                                   \
            hipLaunchKernelGGL((k_element_elasticity<TD_, L_, double>), dim3(grid), dim3(256), 0, s, node, cell, \
                               nc, M, cA, cB, cC, cD, dof_priority, static_cast<double*>(elem));               \
        else                                                                                                   \
            hipLaunchKernelGGL((k_element_elasticity<TD_, L_, float>), dim3(grid), dim3(256), 0, s, node, cell,  \
                               nc, M, cA, cB, cC, cD, dof_priority, static_cast<float*>(elem));                \
        FA_LAUNCH("k_element_elasticity");                                                                     \
        return FA_OK;                                                                                          \
    }
    FA_EL(2, 3) FA_EL(2, 6) FA_EL(2, 10) FA_EL(2, 15) FA_EL(3, 4) FA_EL(3, 10) FA_EL(3, 20) FA_EL(3, 35)
#undef FA_EL
    set_error("fa_element_elasticity: unsupported (td=%d, ldof=%d); p must be 1..4", td, ldof);
    return FA_EINVAL;
}

extern "C" int fa_element_vectors_tensor(int dtype, int td, int ldof, int ncomp, int dof_priority, int nq,
                                         const double* node, const int32_t* cell, int64_t nc, const double* qw,
                                         const double* phi, const double* src, void* elem, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_element_vectors_tensor: bad dtype %d", dtype);
    FA_REQUIRE(td == 2 || td == 3, "fa_element_vectors_tensor: td must be 2 or 3");
    FA_REQUIRE(src && ncomp >= 1 && nq >= 1, "fa_element_vectors_tensor: bad arguments");
    if (nc == 0) return FA_OK;
    hipStream_t s = S(stream);
    const unsigned grid = (unsigned)((nc * ldof + 255) / 256);
#define FA_VT(TD_, T_)                                                                                          \
    hipLaunchKernelGGL((k_element_vectors_tensor<TD_, T_>), dim3(grid), dim3(256), 0, s, node, cell, nc, ldof,   \
                       ncomp, dof_priority, nq, qw, phi, src, static_cast<T_*>(elem))
    if (td == 3) {
        if (dtype == FA_F64) FA_VT(3, double); else FA_VT(3, float);
    } else {
        if (dtype == FA_F64) FA_VT(2, double); else FA_VT(2, float);
    }
#undef FA_VT
    FA_LAUNCH("k_element_vectors_tensor");
    return FA_OK;
}

// ---- incidence-ordered two-pass P1 assembly -----------------------------------
// The fused kernel above computes cell geometry once per incidence, behind
// three dependent loads (incidence -> cell -> vertex coordinates): it is bound
// by load latency, not by bytes.  Here pass 1 runs one thread per cell, reads
// the cell once, and writes element row i of cell c into the slot of
// incidence (row r = cell2dof[c][i], cell c).  A row's incidences are stored
// contiguously in ascending cell order (the adjacency), so pass 2 streams
// them: per CSR row, the element rows of its incidences (ldof doubles each)
// and their packed column positions (one byte per local DOF), summed through
// the same LDS table and in the same order as the fused kernel, which makes
// the values equal bit for bit.  The load vector is the same with one value
// per incidence and a plain ordered sum per row.
namespace fa {

// inc[c*ldof + i] = incidence slot of (c, i); pos[k] packs the CSR row
// positions of the cell's DOFs for incidence k (byte j: position of DOF j)
__global__ void __launch_bounds__(256) k_incidence_maps(int64_t gdof, int ldof, const int32_t* __restrict__ c2d,
                                                       const int32_t* __restrict__ adj_ptr,
                                                       const int32_t* __restrict__ adj,
                                                       const int32_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ indices, int32_t* __restrict__ inc,
                                                       uint32_t* __restrict__ pos, int32_t* __restrict__ bad) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < gdof; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r0 = indptr[r], L = indptr[r + 1] - indptr[r];
        for (int32_t k = adj_ptr[r]; k < adj_ptr[r + 1]; ++k) {
            const int32_t e = adj[k];
            inc[e] = k;
            const int64_t c = e / ldof;
            uint32_t packed = 0;
            for (int j = 0; j < ldof; ++j) {
                const int32_t col = c2d[c * ldof + j];
                int lo = 0, hi = L;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (indices[r0 + mid] < col) lo = mid + 1;
                    else hi = mid;
                }
                if (lo >= L || indices[r0 + lo] != col || lo > 255) atomicOr(bad, 1);
                packed |= static_cast<uint32_t>(lo & 255) << (8 * j);
            }
            pos[k] = packed;
        }
    }
}

// Pass 1 stages a block's 256 cells through LDS with coalesced loads (vertex
// ids, incidence slots, per-cell / per-point coefficients: contiguous per
// block) and reads vertex coordinates with one wide load per vertex (node
// rows padded to 4 doubles in 3D, 2 doubles are already 16-byte rows in 2D):
// the kernels are bound by the texture address unit, so fewer, wider memory
// instructions are what counts.  Values equal the fused kernel's bit for bit.
constexpr int P1_BLOCK = 256;

__device__ inline size_t coef_stage_len(int kind, int nq) {
    return kind == FA_COEF_CELL ? 1 : (kind == FA_COEF_CELLQ ? static_cast<size_t>(nq) : 0);
}

// stage cells [c0, c0 + P1_BLOCK) of `src` (len doubles per cell)
__device__ inline void stage_doubles(double* dst, const double* __restrict__ src, int64_t c0, int64_t nc, size_t len) {
    const int64_t n = (min<int64_t>(nc - c0, P1_BLOCK)) * static_cast<int64_t>(len);
    const double* s0 = src + c0 * static_cast<int64_t>(len);
    for (int64_t t = threadIdx.x; t < n; t += blockDim.x) dst[t] = s0[t];
}

__device__ inline void stage_ints(int32_t* dst, const int32_t* __restrict__ src, int64_t c0, int64_t nc, int len) {
    const int64_t n = (min<int64_t>(nc - c0, P1_BLOCK)) * len;
    const int32_t* s0 = src + c0 * len;
    for (int64_t t = threadIdx.x; t < n; t += blockDim.x) dst[t] = s0[t];
}

template <int TD>
__device__ inline void load_vertices_wide(const double* __restrict__ node, const int32_t* vid, double p[TD + 1][TD]) {
#pragma unroll
    for (int a = 0; a < TD + 1; ++a) {
        if (TD == 3) {
            const double4 q = reinterpret_cast<const double4*>(node)[vid[a]];
            p[a][0] = q.x;
            p[a][1] = q.y;
            p[a][TD - 1] = q.z;
        } else {
            const double2 q = reinterpret_cast<const double2*>(node)[vid[a]];
            p[a][0] = q.x;
            p[a][TD - 1] = q.y;
        }
    }
}

// pass 1, matrix: rows[inc[c*NV + i]][j] = element entry (i, j) of cell c,
// computed in FP64 and stored as R (FP32 rows for an FP32 matrix)
template <int TD, bool SEP, class R>
__global__ void __launch_bounds__(P1_BLOCK) k_p1_incidence_rows(int64_t nc, const double* __restrict__ nodew,
                                                               const int32_t* __restrict__ cell, int nq,
                                                               const double* __restrict__ qw,
                                                               const double* __restrict__ phi, int dk, double ds,
                                                               const double* __restrict__ dc, int mk, double ms,
                                                               const double* __restrict__ mc,
                                                               const int32_t* __restrict__ inc,
                                                               R* __restrict__ rows) {
    constexpr int NV = TD + 1;
    extern __shared__ double stage[];
    const size_t ld = coef_stage_len(dk, nq), lm = coef_stage_len(mk, nq);
    double* dS = stage;
    double* mS = dS + P1_BLOCK * ld;
    int32_t* vS = reinterpret_cast<int32_t*>(mS + P1_BLOCK * lm);
    int32_t* iS = vS + P1_BLOCK * NV;
    for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * P1_BLOCK; c0 < nc; c0 += static_cast<int64_t>(gridDim.x) * P1_BLOCK) {
        if (ld) stage_doubles(dS, dc, c0, nc, ld);
        if (lm) stage_doubles(mS, mc, c0, nc, lm);
        stage_ints(vS, cell, c0, nc, NV);
        stage_ints(iS, inc, c0, nc, NV);
        __syncthreads();
        const int t = threadIdx.x;
        if (c0 + t < nc) {
            double p[NV][TD];
            load_vertices_wide<TD>(nodew, vS + t * NV, p);
            P1Cell<TD> g;
            g.compute(p);
            // coefficients through the staged block: cell index t, row stride ld / lm
            // (a separable coefficient reads its fields at the global cell id;
            // without one the pointers stay LDS-only, so the loads are LDS reads)
            const double* dcl = ld ? dS : nullptr;
            const double* mcl = lm ? mS : nullptr;
            if constexpr (SEP) {
                if (dk == FA_COEF_SEP) dcl = dc;
                if (mk == FA_COEF_SEP) mcl = mc;
            }
            const double kd = p1_kd<TD, SEP>(nq, qw, dk, ds, dcl, t, c0 + t, p);
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double v[NV];
                p1_row<TD, SEP>(g, p, c0 + t, kd, i, t, nq, qw, phi, dk, ds, dcl, mk, ms, mcl, v);
                R* out = rows + static_cast<int64_t>(iS[t * NV + i]) * NV;
#pragma unroll
                for (int j = 0; j < NV; ++j) out[j] = static_cast<R>(v[j]);
            }
        }
        __syncthreads();
    }
}

// pass 1, load vector: vals[inc[c*NV + i]] = |T| sum_q w_q f_cq phi_i(q)
template <int TD, bool SEP>
__global__ void __launch_bounds__(P1_BLOCK) k_p1_incidence_vectors(int64_t nc, const double* __restrict__ nodew,
                                                                  const int32_t* __restrict__ cell, int nq,
                                                                  const double* __restrict__ qw,
                                                                  const double* __restrict__ phi, int sk, double ss,
                                                                  const double* __restrict__ src,
                                                                  const int32_t* __restrict__ inc,
                                                                  double* __restrict__ vals) {
    constexpr int NV = TD + 1;
    extern __shared__ double stage[];
    const size_t ls = coef_stage_len(sk, nq);
    double* sS = stage;
    int32_t* vS = reinterpret_cast<int32_t*>(sS + P1_BLOCK * ls);
    int32_t* iS = vS + P1_BLOCK * NV;
    for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * P1_BLOCK; c0 < nc; c0 += static_cast<int64_t>(gridDim.x) * P1_BLOCK) {
        if (ls) stage_doubles(sS, src, c0, nc, ls);
        stage_ints(vS, cell, c0, nc, NV);
        stage_ints(iS, inc, c0, nc, NV);
        __syncthreads();
        const int t = threadIdx.x;
        if (c0 + t < nc) {
            double p[NV][TD];
            load_vertices_wide<TD>(nodew, vS + t * NV, p);
            P1Cell<TD> g;
            g.compute(p);
            const double* sl = ls ? sS : nullptr;
            if constexpr (SEP) {
                if (sk == FA_COEF_SEP) sl = src;
            }
            // w_q |T| f(x_q) once per point, then its share of every basis
            // moment; each moment is summed over the points in order
            double v[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) v[i] = 0.0;
            for (int q = 0; q < nq; ++q) {
                const double fq = qw[q] * g.meas * coef_at<TD, SEP>(sk, ss, sl, t, c0 + t, q, nq, p);
#pragma unroll
                for (int i = 0; i < NV; ++i) v[i] += fq * phi[q * NV + i];
            }
#pragma unroll
            for (int i = 0; i < NV; ++i) vals[iS[t * NV + i]] = v[i];
        }
        __syncthreads();
    }
}

// pass 2, matrix: the fused kernel's LDS-table sum with the element rows and
// positions read from the incidence slots instead of computed.  R: the type
// of the stored rows, of the LDS table and of the sums (FP64 rows give the
// fused kernel's values bit for bit; FP32 rows an FP32 matrix)
template <int NV, int G, int K, class T, class R>
__global__ void __launch_bounds__(256) k_sum_incidence_rows(int64_t gdof, const int32_t* __restrict__ adj_ptr,
                                                           const int32_t* __restrict__ indptr,
                                                           const uint32_t* __restrict__ pos,
                                                           const R* __restrict__ rows, int lmax,
                                                           T* __restrict__ values) {
    constexpr int GS = G + 1;
    extern __shared__ double tabs_[];
    R* tabs = reinterpret_cast<R*>(tabs_);
    const int rpb = blockDim.x / G;
    const int sub = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    R* tab = tabs + (int64_t)sub * lmax * GS;
    (void)rpb;
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / G) + sub; r < gdof; r += (int64_t)gridDim.x * (blockDim.x / G)) {
        const int32_t r0 = indptr[r];
        const int L = indptr[r + 1] - r0;
        for (int t = 0; t < L; ++t) tab[t * GS + lane] = R(0);
        R acc[K];
#pragma unroll
        for (int m = 0; m < K; ++m) acc[m] = R(0);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // the next chunk's position word and element row are loaded before
        // the current chunk is summed (two chunks of loads in flight per wave)
        const int32_t kbeg = adj_ptr[r], k1 = adj_ptr[r + 1];
        uint32_t pk_next = 0;
        R v_next[NV];
        if (kbeg + lane < k1) {
            pk_next = pos[kbeg + lane];
#pragma unroll
            for (int j = 0; j < NV; ++j) v_next[j] = rows[static_cast<int64_t>(kbeg + lane) * NV + j];
        }
        for (int32_t kb = kbeg; kb < k1; kb += G) {
            const int32_t k = kb + lane;
            const uint32_t pk = pk_next;
            R v[NV];
#pragma unroll
            for (int j = 0; j < NV; ++j) v[j] = v_next[j];
            if (k + G < k1) {
                pk_next = pos[k + G];
#pragma unroll
                for (int j = 0; j < NV; ++j) v_next[j] = rows[static_cast<int64_t>(k + G) * NV + j];
            }
            int p[NV];
            if (k < k1) {
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    p[j] = static_cast<int>((pk >> (8 * j)) & 255u);
                    tab[p[j] * GS + lane] = v[j];
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const int t = lane + m * G;
                if (t < L) {
#pragma unroll
                    for (int l = 0; l < G; ++l) acc[m] += tab[t * GS + l];
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (k < k1) {
#pragma unroll
                for (int j = 0; j < NV; ++j) tab[p[j] * GS + lane] = R(0);
            }
        }
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const int t = lane + m * G;
            if (t < L) values[r0 + t] = static_cast<T>(acc[m]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// pass 2, load vector: out[r] = sum of vals[k] over the row's incidences, in order
template <class T>
__global__ void __launch_bounds__(256) k_sum_incidence_values(int64_t gdof, const int32_t* __restrict__ adj_ptr,
                                                             const double* __restrict__ vals, T* __restrict__ out) {
    constexpr int B = 8;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < gdof; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k0 = adj_ptr[r], k1 = adj_ptr[r + 1];
        double acc = 0.0;
        for (int32_t kb = k0; kb < k1; kb += B) {
            double v[B];
#pragma unroll
            for (int u = 0; u < B; ++u) v[u] = vals[min(kb + u, k1 - 1)];
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (kb + u < k1) acc += v[u];
        }
        out[r] = static_cast<T>(acc);
    }
}

template <int NV, int G, int K, class T, class R>
static int launch_sum_rows(int64_t gdof, const int32_t* adj_ptr, const int32_t* indptr, const uint32_t* pos,
                           const R* rows, int lmax, void* values, hipStream_t s) {
    const size_t per_row = sizeof(R) * lmax * (G + 1);
    int rpb = 256 / G;
    while (rpb > 1 && rpb * per_row > 48 * 1024) rpb >>= 1;
    const size_t lds = per_row * rpb;
    if (lds > 64 * 1024) {
        set_error("fa_sum_incidence_rows: row length %d exceeds the LDS capacity", lmax);
        return FA_ECAPACITY;
    }
    const unsigned grid = grid_for(gdof, rpb, 1 << 20);
    timer_begin(FA_TIMER_P1_SUM_ROWS, s);
    hipLaunchKernelGGL((k_sum_incidence_rows<NV, G, K, T, R>), dim3(grid), dim3(rpb * G), lds, s, gdof, adj_ptr,
                       indptr, pos, rows, lmax, static_cast<T*>(values));
    timer_end(FA_TIMER_P1_SUM_ROWS, s);
    FA_LAUNCH("k_sum_incidence_rows");
    return FA_OK;
}

}  // namespace fa

extern "C" int fa_incidence_maps(int64_t gdof, int32_t ldof, const int32_t* cell2dof, const int32_t* adj_ptr,
                                 const int32_t* adj, const int32_t* indptr, const int32_t* indices, int32_t* inc,
                                 uint32_t* pos, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(gdof > 0 && ldof >= 1 && ldof <= 4 && inc && pos, "fa_incidence_maps: bad arguments (ldof <= 4)");
    FA_REQUIRE(work && work_bytes >= sizeof(int32_t), "fa_incidence_maps: workspace too small");
    hipStream_t s = S(stream);
    int32_t* bad = static_cast<int32_t*>(work);
    FA_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_incidence_maps, dim3(grid_for(gdof, 256, 1 << 16)), dim3(256), 0, s, gdof, ldof, cell2dof,
                       adj_ptr, adj, indptr, indices, inc, pos, bad);
    FA_LAUNCH("k_incidence_maps");
    int32_t h = 0;
    FA_HIP(hipMemcpyAsync(&h, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (h) {
        set_error("fa_incidence_maps: a row is longer than 256 entries or misses a DOF of an incident cell");
        return FA_ECAPACITY;
    }
    return FA_OK;
}

static size_t p1_stage_bytes(int td, int nq, int k1, int k2) {
    auto len = [nq](int kind) -> size_t {
        return kind == FA_COEF_CELL ? 1 : (kind == FA_COEF_CELLQ ? static_cast<size_t>(nq) : 0);
    };
    return sizeof(double) * P1_BLOCK * (len(k1) + len(k2)) + 2 * sizeof(int32_t) * P1_BLOCK * (td + 1);
}

extern "C" int fa_p1_incidence_rows(int dtype_rows, int td, int64_t nc, int nq, const double* node_wide,
                                    const int32_t* cell, const double* qw, const double* phi, int diff_kind,
                                    double diff_scale, const double* diff_coef, int mass_kind, double mass_scale,
                                    const double* mass_coef, const int32_t* inc, void* rows, fa_stream_t stream) {
    FA_REQUIRE(dtype_rows == FA_F64 || dtype_rows == FA_F32, "fa_p1_incidence_rows: bad dtype %d", dtype_rows);
    FA_REQUIRE(td == 2 || td == 3, "fa_p1_incidence_rows: td must be 2 or 3");
    FA_REQUIRE(nq > 0 && inc && rows && node_wide, "fa_p1_incidence_rows: bad arguments");
    FA_REQUIRE(!mass_kind || phi, "fa_p1_incidence_rows: mass needs phi");
    FA_REQUIRE(!(diff_kind >= FA_COEF_CELL && !diff_coef) && !(mass_kind >= FA_COEF_CELL && !mass_coef),
               "fa_p1_incidence_rows: coefficient array missing");
    if (nc == 0) return FA_OK;
    const size_t lds = p1_stage_bytes(td, nq, diff_kind, mass_kind);
    if (lds > 64 * 1024) {
        set_error("fa_p1_incidence_rows: %d quadrature points per cell exceed the staging capacity", nq);
        return FA_ECAPACITY;
    }
    const unsigned grid = grid_for(nc, P1_BLOCK, 1 << 20);
    timer_begin(FA_TIMER_P1_ROWS, S(stream));
#define FA_R1(TD_, SEP_, R_)                                                                                        \
    hipLaunchKernelGGL((k_p1_incidence_rows<TD_, SEP_, R_>), dim3(grid), dim3(P1_BLOCK), lds, S(stream), nc,         \
                       node_wide, cell, nq, qw, phi, diff_kind, diff_scale, diff_coef, mass_kind, mass_scale, mass_coef, \
                       inc, static_cast<R_*>(rows))
#define FA_R(TD_, SEP_) \
    if (dtype_rows == FA_F64) FA_R1(TD_, SEP_, double); else FA_R1(TD_, SEP_, float)
    const bool sep = diff_kind == FA_COEF_SEP || mass_kind == FA_COEF_SEP;
    if (td == 3) {
        if (sep) FA_R(3, true); else FA_R(3, false);
    } else {
        if (sep) FA_R(2, true); else FA_R(2, false);
    }
#undef FA_R
#undef FA_R1
    timer_end(FA_TIMER_P1_ROWS, S(stream));
    FA_LAUNCH("k_p1_incidence_rows");
    return FA_OK;
}

extern "C" int fa_sum_incidence_rows(int dtype, int dtype_rows, int td, int64_t gdof, const int32_t* adj_ptr,
                                     const int32_t* indptr, const uint32_t* pos, const void* rows, int32_t max_row,
                                     void* values, fa_stream_t stream) {
    FA_REQUIRE(dtype_rows == FA_F64 || (dtype_rows == FA_F32 && dtype == FA_F32),
               "fa_sum_incidence_rows: rows of dtype %d for values of dtype %d", dtype_rows, dtype);
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_sum_incidence_rows: bad dtype %d", dtype);
    FA_REQUIRE(td == 2 || td == 3, "fa_sum_incidence_rows: td must be 2 or 3");
    FA_REQUIRE(gdof > 0 && max_row > 0 && max_row <= 256, "fa_sum_incidence_rows: bad sizes");
    hipStream_t s = S(stream);
#define FA_SR(NV_, G_, K_)                                                                                       \
    do {                                                                                                         \
        const double* r64 = static_cast<const double*>(rows);                                                   \
        const float* r32 = static_cast<const float*>(rows);                                                     \
        if (dtype_rows == FA_F32)                                                                                \
            return launch_sum_rows<NV_, G_, K_, float, float>(gdof, adj_ptr, indptr, pos, r32, max_row, values, s); \
        return dtype == FA_F64                                                                                   \
                   ? launch_sum_rows<NV_, G_, K_, double, double>(gdof, adj_ptr, indptr, pos, r64, max_row, values, s) \
                   : launch_sum_rows<NV_, G_, K_, float, double>(gdof, adj_ptr, indptr, pos, r64, max_row, values, s); \
    } while (0)
    // lanes per row G: the sum runs over the incidences in order for any G, so
    // G only sets the chunking and the values do not depend on it.  Built-in
    // choice: 8 for rows of up to 32 entries (3D P1: 15 entries, 24
    // incidences), else enough lanes for K <= 4 positions each; the autotuner
    // may set another (FA_TUNE_SUM_ROWS_LANES).  K = positions owned per lane.
    const int gdef = max_row <= 32 ? 8 : (max_row <= 64 ? 16 : 32);
    const int G = tuned(FA_TUNE_SUM_ROWS_LANES, fa_tune_key(FA_TUNE_SUM_ROWS_LANES, td, max_row), gdef);
    if (td == 3) {
        if (G == 4 && max_row <= 32) FA_SR(4, 4, 8);
        if (G == 16 && max_row <= 32) FA_SR(4, 16, 2);
        if (G == 32 && max_row <= 32) FA_SR(4, 32, 1);
        if (G <= 8 && max_row <= 32) FA_SR(4, 8, 4);
        if (G <= 16 && max_row <= 64) FA_SR(4, 16, 4);
        if (max_row <= 128) FA_SR(4, 32, 4);
        FA_SR(4, 32, 8);
    } else {
        if (G == 4 && max_row <= 16) FA_SR(3, 4, 4);
        if (G == 16 && max_row <= 16) FA_SR(3, 16, 1);
        if (G <= 8 && max_row <= 16) FA_SR(3, 8, 2);
        if (G <= 8 && max_row <= 32) FA_SR(3, 8, 4);
        if (max_row <= 128) FA_SR(3, 32, 4);
        FA_SR(3, 32, 8);
    }
#undef FA_SR
}

extern "C" int fa_p1_incidence_vectors(int td, int64_t nc, int nq, const double* node_wide, const int32_t* cell,
                                       const double* qw, const double* phi, int src_kind, double src_scale,
                                       const double* src, const int32_t* inc, double* vals, fa_stream_t stream) {
    FA_REQUIRE(td == 2 || td == 3, "fa_p1_incidence_vectors: td must be 2 or 3");
    FA_REQUIRE(nq > 0 && phi && inc && vals && node_wide, "fa_p1_incidence_vectors: bad arguments");
    FA_REQUIRE(src_kind < FA_COEF_CELL || src, "fa_p1_incidence_vectors: source array missing");
    if (nc == 0) return FA_OK;
    const size_t lds = p1_stage_bytes(td, nq, src_kind, FA_COEF_NONE);
    if (lds > 64 * 1024) {
        set_error("fa_p1_incidence_vectors: %d quadrature points per cell exceed the staging capacity", nq);
        return FA_ECAPACITY;
    }
    const unsigned grid = grid_for(nc, P1_BLOCK, 1 << 20);
    timer_begin(FA_TIMER_P1_VECTORS, S(stream));
#define FA_PV(TD_, SEP_)                                                                                              \
    hipLaunchKernelGGL((k_p1_incidence_vectors<TD_, SEP_>), dim3(grid), dim3(P1_BLOCK), lds, S(stream), nc, node_wide, \
                       cell, nq, qw, phi, src_kind, src_scale, src, inc, vals)
    if (td == 3) {
        if (src_kind == FA_COEF_SEP) FA_PV(3, true); else FA_PV(3, false);
    } else {
        if (src_kind == FA_COEF_SEP) FA_PV(2, true); else FA_PV(2, false);
    }
#undef FA_PV
    timer_end(FA_TIMER_P1_VECTORS, S(stream));
    FA_LAUNCH("k_p1_incidence_vectors");
    return FA_OK;
}

extern "C" int fa_sum_incidence_values(int dtype, int64_t gdof, const int32_t* adj_ptr, const double* vals, void* out,
                                       fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_sum_incidence_values: bad dtype %d", dtype);
    FA_REQUIRE(gdof > 0 && adj_ptr && vals && out, "fa_sum_incidence_values: bad arguments");
    const unsigned grid = grid_for(gdof, 256, 1 << 20);
    timer_begin(FA_TIMER_P1_SUM_VALUES, S(stream));
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_sum_incidence_values<double>, dim3(grid), dim3(256), 0, S(stream), gdof, adj_ptr, vals,
                           static_cast<double*>(out));
    else
        hipLaunchKernelGGL(k_sum_incidence_values<float>, dim3(grid), dim3(256), 0, S(stream), gdof, adj_ptr, vals,
                           static_cast<float*>(out));
    timer_end(FA_TIMER_P1_SUM_VALUES, S(stream));
    FA_LAUNCH("k_sum_incidence_values");
    return FA_OK;
}

// ---- P1 matrix on a from_box tetrahedral grid -------------------------------
// TetrahedronMesh.from_box (tetrahedron_mesh.py:1017-1086) numbers node
// (ix, iy, iz) as (ix (ny+1) + iy) (nz+1) + iz, cube (cx, cy, cz) as
// (cx ny + cy) nz + cz, and cell 6 cube + t, the Kuhn tetrahedron t over the
// cube corners BOX_KUHN[t] (corner k at offset BOX_CORNER[k]).  So the cells
// around a node, in ascending order, are known from its grid position: the
// cubes (ix - a, iy - b, iz - c) in the order (a, b, c) = (1,1,1), (1,1,0),
// ..., (0,0,0) (lexicographic cube index), and in each cube the tetrahedra
// that hold the node's corner, ascending t.  One thread per CSR row walks
// them in that order, computes each element row with the code of the
// two-pass kernels (P1Cell, p1_kd, p1_row on the cell's own vertex order) and
// adds it into 15 register accumulators, one per neighbour offset: every
// entry is the sequential sum of its cell contributions in ascending cell
// order, the order of the reference's lexsort + np.add.at.  The values equal
// those of the two-pass kernels to a few units in the last place (the
// compiler contracts multiply-adds in the inlined element geometry
// differently here) and are bitwise reproducible.  No incidence list, cell
// array or element-row buffer is read or written: the thread's loads are
// vertex coordinates (one 32-byte row each) and coefficients, all addressed
// from the grid position, so they issue back to back.
//
// The neighbour offsets of a node are the 14 Kuhn edge directions
// +-(1,0,0), +-(0,1,0), +-(0,0,1), +-(1,1,0), +-(1,0,1), +-(0,1,1),
// +-(1,1,1) and the node itself; a neighbour inside the grid is always in the
// CSR row (some cube holds the pair as a Kuhn edge), and ascending column
// order is lexicographic offset order, so the entry's position in the row is
// the number of present neighbours before it.
// offsets of corner k as bit masks over k (corners 1, 2, 5, 6 have x + 1;
// 2, 3, 6, 7 have y + 1; 4-7 have z + 1): branch-free for a runtime k
constexpr int box_cdx(int k) { return (0x66 >> k) & 1; }
constexpr int box_cdy(int k) { return (0xCC >> k) & 1; }
constexpr int box_cdz(int k) { return (k >> 2) & 1; }
constexpr int box_corner_of(int a, int b, int c) {
    return c == 0 ? (a == 0 ? (b == 0 ? 0 : 3) : (b == 0 ? 1 : 2)) : (a == 0 ? (b == 0 ? 4 : 7) : (b == 0 ? 5 : 6));
}
constexpr int box_kuhn(int t, int v) {
    constexpr int K[6][4] = {{0, 1, 2, 6}, {0, 5, 1, 6}, {0, 4, 5, 6}, {0, 7, 4, 6}, {0, 3, 7, 6}, {0, 2, 3, 6}};
    return K[t][v];
}
// local index of corner k in tetrahedron t, or -1
constexpr int box_local(int t, int k) {
    return box_kuhn(t, 0) == k ? 0 : box_kuhn(t, 1) == k ? 1 : box_kuhn(t, 2) == k ? 2 : box_kuhn(t, 3) == k ? 3 : -1;
}
// accumulator slot of offset (dx, dy, dz): rank among the 15 offsets in
// lexicographic order, -1 for an offset that is no Kuhn edge
constexpr int box_slot(int dx, int dy, int dz) {
    constexpr int S[27] = {0, 1, -1, 2, 3, -1, -1, -1, -1, 4, 5, -1, 6, 7, 8, -1, 9, 10,
                           -1, -1, -1, -1, 11, 12, -1, 13, 14};
    return S[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
}
constexpr int BOX_SLOTS = 15;

// Kuhn tetrahedra as paths: the corners of tetrahedron t, ordered by their
// number of +1 offsets, run from corner 0 to corner 6 one axis at a time.
// box_pos: path position of local vertex v; box_axis: the axis of path step
// p (1..3); box_sign: orientation of the local vertex order (the sign of
// det(o1 - o0, o2 - o0, o3 - o0) on the unit cube)
constexpr int box_pos(int t, int v) {
    return box_cdx(box_kuhn(t, v)) + box_cdy(box_kuhn(t, v)) + box_cdz(box_kuhn(t, v));
}
constexpr int box_corner_at(int t, int p) {
    return box_pos(t, 0) == p ? box_kuhn(t, 0)
           : box_pos(t, 1) == p ? box_kuhn(t, 1)
           : box_pos(t, 2) == p ? box_kuhn(t, 2)
                                : box_kuhn(t, 3);
}
constexpr int box_axis(int t, int p) {
    return box_cdx(box_corner_at(t, p)) != box_cdx(box_corner_at(t, p - 1))   ? 0
           : box_cdy(box_corner_at(t, p)) != box_cdy(box_corner_at(t, p - 1)) ? 1
                                                                               : 2;
}
constexpr int box_off(int t, int v, int d) {
    return (d == 0 ? box_cdx(box_kuhn(t, v)) : d == 1 ? box_cdy(box_kuhn(t, v)) : box_cdz(box_kuhn(t, v))) -
           (d == 0 ? box_cdx(box_kuhn(t, 0)) : d == 1 ? box_cdy(box_kuhn(t, 0)) : box_cdz(box_kuhn(t, 0)));
}
constexpr int box_sign(int t) {
    return box_off(t, 1, 0) * (box_off(t, 2, 1) * box_off(t, 3, 2) - box_off(t, 2, 2) * box_off(t, 3, 1)) -
           box_off(t, 1, 1) * (box_off(t, 2, 0) * box_off(t, 3, 2) - box_off(t, 2, 2) * box_off(t, 3, 0)) +
           box_off(t, 1, 2) * (box_off(t, 2, 0) * box_off(t, 3, 1) - box_off(t, 2, 1) * box_off(t, 3, 0));
}

// the contributions of the tetrahedra of cube (ix - A, iy - B, iz - C) that
// hold the node, in ascending cell order
template <int A, int B, int C, bool SEP, bool TENSOR>
__device__ inline void box_cube(int ix, int iy, int iz, int nx, int ny, int nz, const double* __restrict__ nodew,
                                const double* __restrict__ axes, int nq, const double* __restrict__ qw,
                                const double* __restrict__ phi, int dk, double ds, const double* __restrict__ dc,
                                int mk, double ms, const double* __restrict__ mc, double acc[BOX_SLOTS]) {
    const int cx = ix - A, cy = iy - B, cz = iz - C;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= nx || cy >= ny || cz >= nz) return;
    constexpr int KAP = box_corner_of(A, B, C);
    const int64_t cube = (static_cast<int64_t>(cx) * ny + cy) * nz + cz;
    double P[8][3];
    if constexpr (TENSOR) {  // corner coordinates from the grid lines (node (ix, iy, iz) at (X, Y, Z)[.])
        const double x0 = axes[cx], x1 = axes[cx + 1];
        const double y0 = axes[nx + 1 + cy], y1 = axes[nx + 2 + cy];
        const double z0 = axes[nx + ny + 2 + cz], z1 = axes[nx + ny + 3 + cz];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            P[k][0] = box_cdx(k) ? x1 : x0;
            P[k][1] = box_cdy(k) ? y1 : y0;
            P[k][2] = box_cdz(k) ? z1 : z0;
        }
    } else {
        const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1;
        const int64_t n0 = static_cast<int64_t>(cx) * nyz + static_cast<int64_t>(cy) * nz1 + cz;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double4 q =
                reinterpret_cast<const double4*>(nodew)[n0 + box_cdx(k) * nyz + box_cdy(k) * nz1 + box_cdz(k)];
            P[k][0] = q.x;
            P[k][1] = q.y;
            P[k][2] = q.z;
        }
    }
    // an axis-aligned box (every corner on the box of corners 0 and 6): the
    // closed form of a Kuhn tetrahedron's geometry.  Along its path with
    // steps of lengths h_1, h_2, h_3 (u_k = 1 / h_k^2) the barycentric
    // gradients are -e_1/h_1, e_1/h_1 - e_2/h_2, e_2/h_2 - e_3/h_3, e_3/h_3,
    // so grad(lambda_i).grad(lambda_j) is u_1, u_1 + u_2, u_2 + u_3, u_3 on
    // the diagonal (path positions 0..3), -u_{p+1} between positions p and
    // p + 1 and 0 otherwise, and |T| = sign * h_x h_y h_z / 6: the values of
    // P1Cell's formula to rounding, with no cross products per visit
    int ok = 1;
    if constexpr (!TENSOR) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            ok &= static_cast<int>(P[k][0] == (box_cdx(k) ? P[6][0] : P[0][0])) &
                  static_cast<int>(P[k][1] == (box_cdy(k) ? P[6][1] : P[0][1])) &
                  static_cast<int>(P[k][2] == (box_cdz(k) ? P[6][2] : P[0][2]));
    }
    const double hx = P[6][0] - P[0][0], hy = P[6][1] - P[0][1], hz = P[6][2] - P[0][2];
    const double ux = 1.0 / (hx * hx), uy = 1.0 / (hy * hy), uz = 1.0 / (hz * hz);
    const double vol6 = hx * hy * hz / 6.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int i = box_local(t, KAP);
        if (i < 0) continue;
        double p[4][3];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int d = 0; d < 3; ++d) p[v][d] = P[box_kuhn(t, v)][d];
        const int64_t c = 6 * cube + t;
        const double kd = p1_kd<3, SEP>(nq, qw, dk, ds, dc, c, c, p);
        double v[4];
        if (ok) {
            const double meas = box_sign(t) > 0 ? vol6 : -vol6;
            const double u[4] = {0.0, box_axis(t, 1) == 0 ? ux : box_axis(t, 1) == 1 ? uy : uz,
                                 box_axis(t, 2) == 0 ? ux : box_axis(t, 2) == 1 ? uy : uz,
                                 box_axis(t, 3) == 0 ? ux : box_axis(t, 3) == 1 ? uy : uz};
            const int pi = box_pos(t, i);
            const double km = kd * meas;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int pj = box_pos(t, j);
                double gij = 0.0;
                if (pi == pj) gij = pi == 0 ? u[1] : pi == 3 ? u[3] : u[pi] + u[pi + 1];
                else if (pj == pi + 1) gij = -u[pj];
                else if (pi == pj + 1) gij = -u[pi];
                v[j] = dk ? km * gij : 0.0;
            }
            if (mk) {
                for (int q = 0; q < nq; ++q) {
                    const double* ph = phi + q * 4;
                    const double w = qw[q] * meas * coef_at<3, SEP>(mk, ms, mc, c, c, q, nq, p) * ph[i];
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] += w * ph[j];
                }
            }
        } else if constexpr (!TENSOR) {
            P1Cell<3> g;
            g.compute(p);
            p1_row<3, SEP>(g, p, c, kd, i, c, nq, qw, phi, dk, ds, dc, mk, ms, mc, v);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // no contraction into the element value's last multiply: each
            // value is rounded before it is summed, as the two-pass kernels
            // round it when they store it and the reference when it forms
            // its element matrices
#pragma clang fp contract(off)
            const int k = box_kuhn(t, j);
            acc[box_slot(box_cdx(k) - A, box_cdy(k) - B, box_cdz(k) - C)] += v[j];
        }
    }
}

// TENSOR: the caller has checked that node (ix, iy, iz) sits at (X[ix],
// Y[iy], Z[iz]) with axes = X, Y, Z (fa_box_axes), so every cube is an
// axis-aligned box whose spacings come from the three short axis arrays and
// no node row is read
template <class T, bool SEP, bool TENSOR>
__global__ void __launch_bounds__(256) k_p1_box_rows(int nx, int ny, int nz, const double* __restrict__ nodew,
                                                    const double* __restrict__ axes, int nq,
                                                    const double* __restrict__ qw, const double* __restrict__ phi,
                                                    int dk, double ds, const double* __restrict__ dc, int mk,
                                                    double ms, const double* __restrict__ mc,
                                                    const int32_t* __restrict__ indptr, T* __restrict__ values) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1, nn = static_cast<int64_t>(nx + 1) * nyz;
    // XCD-aware order: each XCD's L2 holds one contiguous range of rows, so
    // the coordinate rows shared by neighbouring rows are read into one L2
    for (int64_t r = xcd_block() * (int64_t)blockDim.x + threadIdx.x; r < nn; r += (int64_t)gridDim.x * blockDim.x) {
        const int ix = static_cast<int>(r / nyz);
        const int64_t rem = r - ix * nyz;
        const int iy = static_cast<int>(rem / nz1);
        const int iz = static_cast<int>(rem - iy * nz1);
        double acc[BOX_SLOTS];
#pragma unroll
        for (int s = 0; s < BOX_SLOTS; ++s) acc[s] = 0.0;
        box_cube<1, 1, 1, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<1, 1, 0, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<1, 0, 1, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<1, 0, 0, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<0, 1, 1, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<0, 1, 0, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<0, 0, 1, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        box_cube<0, 0, 0, SEP, TENSOR>(ix, iy, iz, nx, ny, nz, nodew, axes, nq, qw, phi, dk, ds, dc, mk, ms, mc, acc);
        const int32_t r0 = indptr[r], len = indptr[r + 1] - r0;
        int pos = 0;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    const int s = box_slot(dx, dy, dz);
                    if (s < 0) continue;
                    const bool in = ix + dx >= 0 && ix + dx <= nx && iy + dy >= 0 && iy + dy <= ny && iz + dz >= 0 &&
                                    iz + dz <= nz;
                    if (in) {
                        if (pos < len) values[r0 + pos] = static_cast<T>(acc[s]);
                        ++pos;
                    }
                }
    }
}

// Load vector on the same grid, two passes without an incidence list:
// pass 1, one thread per cube, vals[4c + i] = |T| sum_q w_q f(x_q) phi_i(q)
// for its 6 cells (the arithmetic of k_p1_incidence_vectors), a 32-byte
// store per cell; pass 2, one thread per node, the sum of vals[4c + i] over
// the node's cells in ascending order, the addresses computed from the grid
// position (the order of the two-pass kernels' incidence list, so the values
// are equal bit for bit).
//
// A separable source on a tensor grid (node (ix, iy, iz) at (X[ix], Y[iy],
// Z[iz])): coordinate d of quadrature point q of tetrahedron t in a cube whose
// index along d is i is sum_v bcs[q][v] A_d[i + o_d(v)], o_d(v) the offset
// along d of the tetrahedron's vertex v, so it depends on (i, t, q) only, and
// so does every factor g_k of dimension d.  k_box_sep_table evaluates each
// factor once per (i, t, q), the same sum in the same order as the per-cell
// evaluation, and pass 1 multiplies table entries instead of evaluating
// sin/cos/exp per cell (12 transcendental factors a cell for the usual
// 3-factor source at 4 points).  Pass 1 uses the table for a cell only when
// its four vertices lie exactly on the grid lines the table was built from;
// any other cell evaluates the source directly.

// axes[0 .. nx] = X, axes[nx+1 .. nx+ny+1] = Y, then Z, from the nodes on the
// three grid axes; tab[((k * nmax + i) * 6 + t) * nq + q] = g_k at coordinate
// dim[k] of point q of tetrahedron t in the cubes of index i along dim[k]
__global__ void __launch_bounds__(256) k_box_sep_table(int nx, int ny, int nz, const double* __restrict__ nodew,
                                                      const fa_sep_t* __restrict__ e, int nq, int nmax,
                                                      double* __restrict__ axes, double* __restrict__ tab) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1;
    const double4* __restrict__ node = reinterpret_cast<const double4*>(nodew);
    const int64_t naxes = static_cast<int64_t>(nx) + ny + nz + 3;
    const int64_t ntab = static_cast<int64_t>(e->nfactors) * nmax * 6 * nq;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < naxes + ntab;
         j += (int64_t)gridDim.x * blockDim.x) {
        if (j < naxes) {
            if (j <= nx) axes[j] = node[j * nyz].x;
            else if (j <= nx + 1 + ny) axes[j] = node[(j - nx - 1) * nz1].y;
            else axes[j] = node[j - nx - ny - 2].z;
            continue;
        }
        const int64_t idx = j - naxes;
        const int q = static_cast<int>(idx % nq);
        int64_t rest = idx / nq;
        const int t = static_cast<int>(rest % 6);
        rest /= 6;
        const int i = static_cast<int>(rest % nmax);
        const int k = static_cast<int>(rest / nmax);
        const int d = e->dim[k];
        const int n = d == 0 ? nx : (d == 1 ? ny : nz);
        if (i >= n) continue;
        double lo, hi;
        if (d == 0) {
            lo = node[i * nyz].x;
            hi = node[(i + 1) * nyz].x;
        } else if (d == 1) {
            lo = node[i * nz1].y;
            hi = node[(i + 1) * nz1].y;
        } else {
            lo = node[i].z;
            hi = node[i + 1].z;
        }
        double x = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int kv = 0;
#pragma unroll
            for (int tt = 0; tt < 6; ++tt)
                if (t == tt) kv = box_kuhn(tt, v);
            const int o = d == 0 ? box_cdx(kv) : (d == 1 ? box_cdy(kv) : box_cdz(kv));
            x += e->bcs[q * 4 + v] * (o ? hi : lo);
        }
        tab[idx] = sep_factor(e->kind[k], e->a[k], e->b[k], x);
    }
}

// sep_value at the NQ points of each of the 6 tetrahedra of cube (cx, cy,
// cz) (cells c0 .. c0 + 5), the factors read from the table of
// k_box_sep_table: per factor one block load (6 x NQ entries, contiguous),
// per term the 6 cells' field values, so a cube waits on a few independent
// batches of loads rather than a chain of one load per point and factor; per
// point the operations and their order are those of sep_value
template <int NQ>
__device__ inline void sep_table_cube(const fa_sep_t* __restrict__ e, const double* __restrict__ tab, int nmax,
                                      int cx, int cy, int cz, int64_t c0, double f[6][NQ]) {
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
        for (int q = 0; q < NQ; ++q) f[t][q] = 0.0;
    int k = 0;
    const int nt = e->nterms;
    for (int m = 0; m < nt; ++m) {
        double v[6][NQ];
        const double cm = e->coef[m];
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
            for (int q = 0; q < NQ; ++q) v[t][q] = cm;
        const double* fld = e->field[m];
        if (fld) {
            double fc[6];
#pragma unroll
            for (int t = 0; t < 6; ++t) fc[t] = fld[c0 + t];
#pragma unroll
            for (int t = 0; t < 6; ++t)
#pragma unroll
                for (int q = 0; q < NQ; ++q) v[t][q] *= fc[t];
        }
        const int kend = e->term_end[m];
        for (; k < kend; ++k) {
            const int d = e->dim[k];
            const int i = d == 0 ? cx : (d == 1 ? cy : cz);
            const double* row = tab + (static_cast<int64_t>(k) * nmax + i) * 6 * NQ;
            double r[6][NQ];
            if constexpr (NQ == 4) {  // rows of 32 bytes, 32-byte aligned
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    const double4 w = reinterpret_cast<const double4*>(row)[t];
                    r[t][0] = w.x;
                    r[t][1] = w.y;
                    r[t][2] = w.z;
                    r[t][3] = w.w;
                }
            } else {
#pragma unroll
                for (int t = 0; t < 6; ++t)
#pragma unroll
                    for (int q = 0; q < NQ; ++q) r[t][q] = row[t * NQ + q];
            }
#pragma unroll
            for (int t = 0; t < 6; ++t)
#pragma unroll
                for (int q = 0; q < NQ; ++q) v[t][q] *= r[t][q];
        }
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
            for (int q = 0; q < NQ; ++q) f[t][q] += v[t][q];
    }
}

// sep_table_cube for one term of at most 3 factors at 4 points, the usual
// source (a coefficient, perhaps a cell field, and one factor per axis): the
// field values and the factors' table blocks loaded in one batch, then the
// products of sep_value in its order (coefficient, field, factors in turn)
__device__ inline void sep_table_cube_short(const fa_sep_t* __restrict__ e, const double* __restrict__ tab, int nmax,
                                            int cx, int cy, int cz, int64_t c0, double f[6][4]) {
    const int nf = e->term_end[0];
    const double cm = e->coef[0];
    const double* fld = e->field[0];
    double fc[6];
    const double* rows[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int kk = k < nf ? k : 0;
        const int d = e->dim[kk];
        const int i = d == 0 ? cx : (d == 1 ? cy : cz);
        rows[k] = tab + (static_cast<int64_t>(kk) * nmax + i) * 24;
    }
    double4 r[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int t = 0; t < 6; ++t) r[k][t] = reinterpret_cast<const double4*>(rows[k])[t];
#pragma unroll
    for (int t = 0; t < 6; ++t) fc[t] = fld ? fld[c0 + t] : 1.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        double v[4] = {cm, cm, cm, cm};
        if (fld) {
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] *= fc[t];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k >= nf) break;
            v[0] *= r[k][t].x;
            v[1] *= r[k][t].y;
            v[2] *= r[k][t].z;
            v[3] *= r[k][t].w;
        }
        // a single term: acc = 0 + v, as sep_value adds it
#pragma unroll
        for (int q = 0; q < 4; ++q) f[t][q] = 0.0 + v[q];
    }
}

// element vectors of the 6 cells of a cube from the source values at NQ
// points: v_i = sum_q (w_q |T| s f_q) phi_i(q), the arithmetic of the per-cell
// kernels
template <int NQ>
__device__ inline void box_cube_cell_vectors(const double f[6][NQ], const double meas[6], double ss,
                                             const double* __restrict__ qw, const double* __restrict__ phi,
                                             double4* __restrict__ out) {
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double fq = qw[q] * meas[t] * (ss * f[t][q]);
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] += fq * phi[q * 4 + i];
        }
        out[t] = double4{v[0], v[1], v[2], v[3]};
    }
}

// the same for one point of a rule with more than 4 points
__device__ inline double sep_table_value(const fa_sep_t* __restrict__ e, const double* __restrict__ tab, int nmax,
                                         int nq, int t, int q, int cx, int cy, int cz, int64_t c) {
    double acc = 0.0;
    int k = 0;
    const int nt = e->nterms;
    for (int m = 0; m < nt; ++m) {
        double v = e->coef[m];
        const double* fld = e->field[m];
        if (fld) v *= fld[c];
        const int kend = e->term_end[m];
        for (; k < kend; ++k) {
            const int d = e->dim[k];
            const int i = d == 0 ? cx : (d == 1 ? cy : cz);
            v *= tab[((static_cast<int64_t>(k) * nmax + i) * 6 + t) * nq + q];
        }
        acc += v;
    }
    return acc;
}

// Pass 1, one thread per cube: its 8 corners loaded once, the 6 cells'
// measures, and for a separable source on the grid lines the 6 x NQ source
// values from the table; the 6 element vectors are stored at the cells'
// positions (cell 6 cube + t).  Other sources, rules and cubes off the grid
// lines take the per-cell arithmetic cell by cell.
template <bool SEP>
__global__ void __launch_bounds__(256) k_p1_box_cell_vectors(int nx, int ny, int nz, const double* __restrict__ nodew,
                                                            int nq, const double* __restrict__ qw,
                                                            const double* __restrict__ phi, int sk, double ss,
                                                            const double* __restrict__ src,
                                                            const double* __restrict__ axes,
                                                            const double* __restrict__ tab, int nmax,
                                                            double* __restrict__ vals) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1;
    const int64_t ncube = static_cast<int64_t>(nx) * ny * nz;
    const double4* __restrict__ node = reinterpret_cast<const double4*>(nodew);
    for (int64_t cube = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cube < ncube;
         cube += (int64_t)gridDim.x * blockDim.x) {
        const int cx = static_cast<int>(cube / (static_cast<int64_t>(ny) * nz));
        const int64_t rem = cube - static_cast<int64_t>(cx) * ny * nz;
        const int cy = static_cast<int>(rem / nz);
        const int cz = static_cast<int>(rem - static_cast<int64_t>(cy) * nz);
        const int64_t n0 = static_cast<int64_t>(cx) * nyz + static_cast<int64_t>(cy) * nz1 + cz;
        const int64_t c0 = 6 * cube;
        double4* out = reinterpret_cast<double4*>(vals) + c0;
        if constexpr (SEP) {
            double P[8][3];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double4 w = node[n0 + box_cdx(k) * nyz + box_cdy(k) * nz1 + box_cdz(k)];
                P[k][0] = w.x;
                P[k][1] = w.y;
                P[k][2] = w.z;
            }
            // all six grid-line coordinates loaded at once, compared without branches
            const double x0 = axes[cx], x1 = axes[cx + 1];
            const double y0 = axes[nx + 1 + cy], y1 = axes[nx + 2 + cy];
            const double z0 = axes[nx + ny + 2 + cz], z1 = axes[nx + ny + 3 + cz];
            int ok = 1;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                ok &= static_cast<int>(P[k][0] == (box_cdx(k) ? x1 : x0)) &
                      static_cast<int>(P[k][1] == (box_cdy(k) ? y1 : y0)) &
                      static_cast<int>(P[k][2] == (box_cdz(k) ? z1 : z0));
            if (ok && (nq == 4 || nq == 1)) {
                double meas[6];
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    double p[4][3];
#pragma unroll
                    for (int v = 0; v < 4; ++v)
#pragma unroll
                        for (int d = 0; d < 3; ++d) p[v][d] = P[box_kuhn(t, v)][d];
                    P1Cell<3> g;
                    g.compute(p);
                    meas[t] = g.meas;
                }
                const fa_sep_t* e = reinterpret_cast<const fa_sep_t*>(src);
                if (nq == 4) {  // the usual rule (q = 2)
                    double f[6][4];
                    if (e->nterms == 1 && e->term_end[0] <= 3)
                        sep_table_cube_short(e, tab, nmax, cx, cy, cz, c0, f);
                    else
                        sep_table_cube<4>(e, tab, nmax, cx, cy, cz, c0, f);
                    box_cube_cell_vectors<4>(f, meas, ss, qw, phi, out);
                } else {
                    double f[6][1];
                    sep_table_cube<1>(e, tab, nmax, cx, cy, cz, c0, f);
                    box_cube_cell_vectors<1>(f, meas, ss, qw, phi, out);
                }
                continue;
            }
            // a cube on the grid lines with another rule: table entries point by point
            if (ok) {
#pragma unroll 1
                for (int t = 0; t < 6; ++t) {
                    double p[4][3];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        int k = 0;
#pragma unroll
                        for (int tt = 0; tt < 6; ++tt)
                            if (t == tt) k = box_kuhn(tt, v);
                        const double4 w = node[n0 + box_cdx(k) * nyz + box_cdy(k) * nz1 + box_cdz(k)];
                        p[v][0] = w.x;
                        p[v][1] = w.y;
                        p[v][2] = w.z;
                    }
                    P1Cell<3> g;
                    g.compute(p);
                    double v[4] = {0.0, 0.0, 0.0, 0.0};
                    for (int q = 0; q < nq; ++q) {
                        const double f = ss * sep_table_value(reinterpret_cast<const fa_sep_t*>(src), tab, nmax, nq,
                                                              t, q, cx, cy, cz, c0 + t);
                        const double fq = qw[q] * g.meas * f;
#pragma unroll
                        for (int i = 0; i < 4; ++i) v[i] += fq * phi[q * 4 + i];
                    }
                    out[t] = double4{v[0], v[1], v[2], v[3]};
                }
                continue;
            }
        }
        // any other source, or a cube off the grid lines: cell by cell
#pragma unroll 1
        for (int t = 0; t < 6; ++t) {
            double p[4][3];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // corner of vertex v of tetrahedron t, t a runtime index: the
                // table is small enough to select without memory
                int k = 0;
#pragma unroll
                for (int tt = 0; tt < 6; ++tt)
                    if (t == tt) k = box_kuhn(tt, v);
                const double4 w = node[n0 + box_cdx(k) * nyz + box_cdy(k) * nz1 + box_cdz(k)];
                p[v][0] = w.x;
                p[v][1] = w.y;
                p[v][2] = w.z;
            }
            P1Cell<3> g;
            g.compute(p);
            double v[4] = {0.0, 0.0, 0.0, 0.0};
            for (int q = 0; q < nq; ++q) {
                const double fq = qw[q] * g.meas * coef_at<3, SEP>(sk, ss, src, c0 + t, c0 + t, q, nq, p);
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] += fq * phi[q * 4 + i];
            }
            out[t] = double4{v[0], v[1], v[2], v[3]};
        }
    }
}

// the values of cube (ix - A, iy - B, iz - C) at the node: loads only (cell
// 0 for a cube outside the grid, masked when summed)
template <int A, int B, int C>
__device__ inline void box_cube_load(int ix, int iy, int iz, int nx, int ny, int nz, const double* __restrict__ vals,
                                     double v[6], bool& valid) {
    const int cx = ix - A, cy = iy - B, cz = iz - C;
    valid = cx >= 0 && cy >= 0 && cz >= 0 && cx < nx && cy < ny && cz < nz;
    constexpr int KAP = box_corner_of(A, B, C);
    const int64_t cube = valid ? (static_cast<int64_t>(cx) * ny + cy) * nz + cz : 0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int i = box_local(t, KAP);
        v[t] = i < 0 ? 0.0 : vals[(6 * cube + t) * 4 + i];
    }
}

// the loaded values held in registers from here on (placed after all the
// loads of a row, so they are in flight together; the compiler otherwise
// sinks each load to its use)
template <int A, int B, int C>
__device__ inline void box_cube_pin(const double v[6]) {
    constexpr int KAP = box_corner_of(A, B, C);
#pragma unroll
    for (int t = 0; t < 6; ++t)
        if (box_local(t, KAP) >= 0) asm volatile("" : : "v"(v[t]));
}

template <int A, int B, int C>
__device__ inline void box_cube_add(const double v[6], bool valid, double& acc) {
    constexpr int KAP = box_corner_of(A, B, C);
    if (!valid) return;
#pragma unroll
    for (int t = 0; t < 6; ++t)
        if (box_local(t, KAP) >= 0) acc += v[t];
}

template <class T>
__global__ void __launch_bounds__(256) k_p1_box_sum_vectors(int nx, int ny, int nz, const double* __restrict__ vals,
                                                           T* __restrict__ out) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1, nn = static_cast<int64_t>(nx + 1) * nyz;
    for (int64_t r = xcd_block() * (int64_t)blockDim.x + threadIdx.x; r < nn; r += (int64_t)gridDim.x * blockDim.x) {
        const int ix = static_cast<int>(r / nyz);
        const int64_t rem = r - ix * nyz;
        const int iy = static_cast<int>(rem / nz1);
        const int iz = static_cast<int>(rem - iy * nz1);
        double v[8][6];
        bool ok[8];
        box_cube_load<1, 1, 1>(ix, iy, iz, nx, ny, nz, vals, v[0], ok[0]);
        box_cube_load<1, 1, 0>(ix, iy, iz, nx, ny, nz, vals, v[1], ok[1]);
        box_cube_load<1, 0, 1>(ix, iy, iz, nx, ny, nz, vals, v[2], ok[2]);
        box_cube_load<1, 0, 0>(ix, iy, iz, nx, ny, nz, vals, v[3], ok[3]);
        box_cube_load<0, 1, 1>(ix, iy, iz, nx, ny, nz, vals, v[4], ok[4]);
        box_cube_load<0, 1, 0>(ix, iy, iz, nx, ny, nz, vals, v[5], ok[5]);
        box_cube_load<0, 0, 1>(ix, iy, iz, nx, ny, nz, vals, v[6], ok[6]);
        box_cube_load<0, 0, 0>(ix, iy, iz, nx, ny, nz, vals, v[7], ok[7]);
        box_cube_pin<1, 1, 1>(v[0]);
        box_cube_pin<1, 1, 0>(v[1]);
        box_cube_pin<1, 0, 1>(v[2]);
        box_cube_pin<1, 0, 0>(v[3]);
        box_cube_pin<0, 1, 1>(v[4]);
        box_cube_pin<0, 1, 0>(v[5]);
        box_cube_pin<0, 0, 1>(v[6]);
        box_cube_pin<0, 0, 0>(v[7]);
        double acc = 0.0;  // the cells in ascending order
        box_cube_add<1, 1, 1>(v[0], ok[0], acc);
        box_cube_add<1, 1, 0>(v[1], ok[1], acc);
        box_cube_add<1, 0, 1>(v[2], ok[2], acc);
        box_cube_add<1, 0, 0>(v[3], ok[3], acc);
        box_cube_add<0, 1, 1>(v[4], ok[4], acc);
        box_cube_add<0, 1, 0>(v[5], ok[5], acc);
        box_cube_add<0, 0, 1>(v[6], ok[6], acc);
        box_cube_add<0, 0, 0>(v[7], ok[7], acc);
        out[r] = static_cast<T>(acc);
    }
}

// nonzeros of the P1 pattern of an nx x ny x nz Kuhn grid: the nodes plus
// both directions of every edge
static int64_t box_nnz(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1);
    const int o[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    int64_t edges = 0;
    for (const auto& e : o) edges += (nx + 1 - e[0]) * (ny + 1 - e[1]) * (nz + 1 - e[2]);
    return nn + 2 * edges;
}

// ---- matrix-free P1 operator -------------------------------------------------
// y_r = sum over the incidences (c, i) of DOF r of sum_j A^c_ij x_{cell[c][j]},
// where A^c is the element matrix whose rows k_assemble_p1 sums into CSR
// entries (same row formula, computed where it is used and never stored).
// G lanes per DOF: lane l takes the incidences l, l + G, ... of the row in
// order and the lane partials are combined by a fixed butterfly, so y is the
// same run to run.  RES: y_r = b_r - (A x)_r, set to 0 on rows with bd_r != 0,
// and every workgroup stores the sum of its y_r^2 at part[blockIdx.x] (the
// grid depends on gdof only, so the norm's summation order is fixed).  Y: the
// type y is stored in (FP32 for the right-hand side of an FP32 inner solve);
// the sums and the norm are formed in FP64.
namespace fa {

template <int TD, int G, bool SEP, bool RES, class Y>
__global__ void __launch_bounds__(256) k_p1_apply(int64_t gdof, const double* __restrict__ nodew,
                                                 const int32_t* __restrict__ cell, int nq,
                                                 const double* __restrict__ qw, const double* __restrict__ phi,
                                                 int dk, double ds, const double* __restrict__ dc, int mk, double ms,
                                                 const double* __restrict__ mc, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ adj, const double* __restrict__ x,
                                                 const double* __restrict__ b, const uint8_t* __restrict__ bd,
                                                 Y* __restrict__ y, double* __restrict__ part) {
    constexpr int NV = TD + 1;
    const int lane = threadIdx.x % G;
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / G);
    double nrm = 0.0;
    for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; r < gdof; r += stride) {
        double s = 0.0;
        const int32_t k1 = ptr[r + 1];
        for (int32_t k = ptr[r] + lane; k < k1; k += G) {
            const int32_t e = adj[k];
            const int64_t c = e / NV;
            const int i = e - static_cast<int>(c) * NV;
            int32_t vid[NV];
#pragma unroll
            for (int a = 0; a < NV; ++a) vid[a] = cell[c * NV + a];
            double xv[NV];
#pragma unroll
            for (int a = 0; a < NV; ++a) xv[a] = x[vid[a]];
            double p[NV][TD];
            load_vertices_wide<TD>(nodew, vid, p);
            P1Cell<TD> g;
            g.compute(p);
            const double kd = p1_kd<TD, SEP>(nq, qw, dk, ds, dc, c, c, p);
            double v[NV];
            p1_row<TD, SEP>(g, p, c, kd, i, c, nq, qw, phi, dk, ds, dc, mk, ms, mc, v);
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < NV; ++j) t += v[j] * xv[j];
            s += t;
        }
        s = group_sum<G>(s);
        if (RES) {
            s = bd[r] ? 0.0 : b[r] - s;
            if (lane == 0) nrm += s * s;
        }
        if (lane == 0) y[r] = static_cast<Y>(s);
    }
    if (RES) store_partial(block_sum(nrm), part);
}

__global__ void k_p1_apply_finish(const double* __restrict__ part, int n, double* __restrict__ out) {
    const double s = finish_sum(part, n);
    if (threadIdx.x == 0) *out = s;
}

}  // namespace fa

extern "C" size_t fa_p1_apply_work_bytes(void) { return align_up(RED_GRID * sizeof(double)) + 256; }

extern "C" int fa_p1_apply(int dtype_y, int td, int64_t gdof, int nq, const double* node_wide, const int32_t* cell,
                           const double* qw, const double* phi, int diff_kind, double diff_scale,
                           const double* diff_coef, int mass_kind, double mass_scale, const double* mass_coef,
                           const int32_t* adj_ptr, const int32_t* adj, const double* x, const double* b,
                           const uint8_t* is_bd, void* y, double* norm2, void* work, size_t work_bytes,
                           fa_stream_t stream) {
    FA_REQUIRE(dtype_y == FA_F64 || dtype_y == FA_F32, "fa_p1_apply: bad dtype_y %d", dtype_y);
    FA_REQUIRE(td == 2 || td == 3, "fa_p1_apply: td must be 2 or 3");
    FA_REQUIRE(gdof > 0 && nq > 0 && node_wide && cell && adj_ptr && adj && x && y, "fa_p1_apply: bad arguments");
    FA_REQUIRE(!mass_kind || phi, "fa_p1_apply: mass needs phi");
    FA_REQUIRE(!(diff_kind >= FA_COEF_CELL && !diff_coef) && !(mass_kind >= FA_COEF_CELL && !mass_coef),
               "fa_p1_apply: coefficient array missing");
    const bool res = b != nullptr;
    FA_REQUIRE(!res || (is_bd && norm2), "fa_p1_apply: the residual form needs is_bd and norm2");
    Arena ar(work, work_bytes);
    double* part = res ? ar.take<double>(RED_GRID) : nullptr;
    if (res && !part) {
        set_error("fa_p1_apply: workspace too small");
        return FA_ENOMEM;
    }
    hipStream_t s = S(stream);
    constexpr int G = 8;
    // rows per workgroup 256 / G; the residual form keeps the grid within the partials' capacity
    const int64_t rows_per_block = 256 / G;
    const unsigned grid = res ? static_cast<unsigned>(std::min<int64_t>((gdof + rows_per_block - 1) / rows_per_block,
                                                                        RED_GRID))
                              : grid_for(gdof, static_cast<int>(rows_per_block), 1 << 20);
    const bool sep = diff_kind == FA_COEF_SEP || mass_kind == FA_COEF_SEP;
    timer_begin(FA_TIMER_P1_APPLY, s);
#define FA_AP1(TD_, SEP_, RES_, Y_)                                                                                 \
    hipLaunchKernelGGL((k_p1_apply<TD_, G, SEP_, RES_, Y_>), dim3(grid), dim3(256), 0, s, gdof, node_wide, cell, nq,   \
                       qw, phi, diff_kind, diff_scale, diff_coef, mass_kind, mass_scale, mass_coef, adj_ptr, adj, x, b, \
                       is_bd, static_cast<Y_*>(y), part)
#define FA_AP(TD_, SEP_, RES_)                          \
    if (dtype_y == FA_F64) FA_AP1(TD_, SEP_, RES_, double); \
    else FA_AP1(TD_, SEP_, RES_, float)
#define FA_AP2(TD_)                                   \
    if (sep) {                                        \
        if (res) FA_AP(TD_, true, true);              \
        else FA_AP(TD_, true, false);                 \
    } else {                                          \
        if (res) FA_AP(TD_, false, true);             \
        else FA_AP(TD_, false, false);                \
    }
    if (td == 3) {
        FA_AP2(3)
    } else {
        FA_AP2(2)
    }
#undef FA_AP2
#undef FA_AP
#undef FA_AP1
    timer_end(FA_TIMER_P1_APPLY, s);
    FA_LAUNCH("k_p1_apply");
    if (res) {
        hipLaunchKernelGGL(k_p1_apply_finish, dim3(1), dim3(RED_BLOCK), 0, s, part, static_cast<int>(grid), norm2);
        FA_LAUNCH("k_p1_apply_finish");
    }
    return FA_OK;
}

// Tensor grid, coefficients of the kinds none / constant / per cell: every
// load of a row issued up front (the nine grid-line coordinates around the
// node, the coefficient of each of its 24 cells, the row pointer), then the
// closed-form rows of box_cube in the same cell order and arithmetic (so the
// same values bit for bit as k_p1_box_rows<T, false, true>).  One template
// instance per incident cube keeps each cube's coefficients in registers.
struct BoxCoefs {
    double kd[6], km[6];
    bool valid;
};

// the coefficient of a cell as coef_at gives it for the kinds none, constant
// and per cell, the load issued for every kind (from `any`, a valid array,
// when the kind has no per-cell array) so no load waits behind a branch
__device__ inline double box_cell_coef(int kind, double scale, const double* __restrict__ coef,
                                       const double* __restrict__ any, int64_t c) {
    const bool cell = kind == FA_COEF_CELL;
    const double v = (cell ? coef : any)[cell ? c : 0];
    return cell ? scale * v : (kind ? scale : 0.0);
}

template <int A, int B, int C, bool DCELL, bool MASS>
__device__ inline BoxCoefs box_load(int ix, int iy, int iz, int nx, int ny, int nz, int dk, double ds,
                                    const double* __restrict__ dc, int mk, double ms, const double* __restrict__ mc,
                                    const double* __restrict__ any) {
    BoxCoefs o;
    const int cx = ix - A, cy = iy - B, cz = iz - C;
    o.valid = cx >= 0 && cy >= 0 && cz >= 0 && cx < nx && cy < ny && cz < nz;
    const int64_t cube = o.valid ? (static_cast<int64_t>(cx) * ny + cy) * nz + cz : 0;
    constexpr int KAP = box_corner_of(A, B, C);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        o.kd[t] = 0.0;
        o.km[t] = 0.0;
        if (box_local(t, KAP) < 0) continue;
        // DCELL: a per-cell diffusion coefficient, loaded with no branch
        o.kd[t] = DCELL ? ds * dc[6 * cube + t] : (dk ? ds : 0.0);
        if (MASS) o.km[t] = box_cell_coef(mk, ms, mc, any, 6 * cube + t);
    }
    return o;
}

// the loaded coefficients held in registers from here on: the compiler
// otherwise sinks each load to its use, one round trip per cell
template <int A, int B, int C, bool MASS>
__device__ inline void box_pin(const BoxCoefs& o) {
    constexpr int KAP = box_corner_of(A, B, C);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        if (box_local(t, KAP) < 0) continue;
        asm volatile("" : : "v"(o.kd[t]));
        if (MASS) asm volatile("" : : "v"(o.km[t]));
    }
}

template <int A, int B, int C>
__device__ inline void box_rows_tensor(const BoxCoefs& cf, double xm, double xc, double xp, double ym, double yc,
                                       double yp, double zm, double zc, double zp, int nq,
                                       const double* __restrict__ qw, const double* __restrict__ phi, int dk, int mk,
                                       double acc[BOX_SLOTS]) {
    if (!cf.valid) return;
    const double x0 = A ? xm : xc, x1 = A ? xc : xp;
    const double y0 = B ? ym : yc, y1 = B ? yc : yp;
    const double z0 = C ? zm : zc, z1 = C ? zc : zp;
    const double hx = x1 - x0, hy = y1 - y0, hz = z1 - z0;
    const double ux = 1.0 / (hx * hx), uy = 1.0 / (hy * hy), uz = 1.0 / (hz * hz);
    const double vol6 = hx * hy * hz / 6.0;
    constexpr int KAP = box_corner_of(A, B, C);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int i = box_local(t, KAP);
        if (i < 0) continue;
        double kd = 0.0;
        if (dk)
            for (int q = 0; q < nq; ++q) kd += qw[q] * cf.kd[t];
        const double meas = box_sign(t) > 0 ? vol6 : -vol6;
        const double u[4] = {0.0, box_axis(t, 1) == 0 ? ux : box_axis(t, 1) == 1 ? uy : uz,
                             box_axis(t, 2) == 0 ? ux : box_axis(t, 2) == 1 ? uy : uz,
                             box_axis(t, 3) == 0 ? ux : box_axis(t, 3) == 1 ? uy : uz};
        const int pi = box_pos(t, i);
        const double km = kd * meas;
        double v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int pj = box_pos(t, j);
            double gij = 0.0;
            if (pi == pj) gij = pi == 0 ? u[1] : pi == 3 ? u[3] : u[pi] + u[pi + 1];
            else if (pj == pi + 1) gij = -u[pj];
            else if (pi == pj + 1) gij = -u[pi];
            v[j] = dk ? km * gij : 0.0;
        }
        if (mk) {
            for (int q = 0; q < nq; ++q) {
                const double* ph = phi + q * 4;
                const double w = qw[q] * meas * cf.km[t] * ph[i];
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] += w * ph[j];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma clang fp contract(off)
            const int k = box_kuhn(t, j);
            acc[box_slot(box_cdx(k) - A, box_cdy(k) - B, box_cdz(k) - C)] += v[j];
        }
    }
}

template <class T, bool DCELL, bool MASS>
__global__ void __launch_bounds__(256) k_p1_box_rows_tensor(int nx, int ny, int nz, const double* __restrict__ axes,
                                                           int nq, const double* __restrict__ qw,
                                                           const double* __restrict__ phi, int dk, double ds,
                                                           const double* __restrict__ dc, int mk, double ms,
                                                           const double* __restrict__ mc,
                                                           const int32_t* __restrict__ indptr,
                                                           T* __restrict__ values) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1, nn = static_cast<int64_t>(nx + 1) * nyz;
    for (int64_t r = xcd_block() * (int64_t)blockDim.x + threadIdx.x; r < nn; r += (int64_t)gridDim.x * blockDim.x) {
        const int ix = static_cast<int>(r / nyz);
        const int64_t rem = r - ix * nyz;
        const int iy = static_cast<int>(rem / nz1);
        const int iz = static_cast<int>(rem - iy * nz1);
        // grid lines (indices clamped into range), the row pointer and the
        // cells' coefficients: all issued before any is used
        const double xm = axes[max(ix - 1, 0)], xc = axes[ix], xp = axes[min(ix + 1, nx)];
        const double ym = axes[nx + 1 + max(iy - 1, 0)], yc = axes[nx + 1 + iy], yp = axes[nx + 1 + min(iy + 1, ny)];
        const double zm = axes[nx + ny + 2 + max(iz - 1, 0)], zc = axes[nx + ny + 2 + iz],
                     zp = axes[nx + ny + 2 + min(iz + 1, nz)];
        const int32_t r0 = indptr[r], len = indptr[r + 1] - r0;
#define FA_BL(A_, B_, C_) box_load<A_, B_, C_, DCELL, MASS>(ix, iy, iz, nx, ny, nz, dk, ds, dc, mk, ms, mc, qw)
        const BoxCoefs c0 = FA_BL(1, 1, 1), c1 = FA_BL(1, 1, 0), c2 = FA_BL(1, 0, 1), c3 = FA_BL(1, 0, 0);
        const BoxCoefs c4 = FA_BL(0, 1, 1), c5 = FA_BL(0, 1, 0), c6 = FA_BL(0, 0, 1), c7 = FA_BL(0, 0, 0);
#undef FA_BL
        box_pin<1, 1, 1, MASS>(c0);
        box_pin<1, 1, 0, MASS>(c1);
        box_pin<1, 0, 1, MASS>(c2);
        box_pin<1, 0, 0, MASS>(c3);
        box_pin<0, 1, 1, MASS>(c4);
        box_pin<0, 1, 0, MASS>(c5);
        box_pin<0, 0, 1, MASS>(c6);
        box_pin<0, 0, 0, MASS>(c7);
        double acc[BOX_SLOTS];
#pragma unroll
        for (int s = 0; s < BOX_SLOTS; ++s) acc[s] = 0.0;
#define FA_BR(A_, B_, C_, CF_) \
    box_rows_tensor<A_, B_, C_>(CF_, xm, xc, xp, ym, yc, yp, zm, zc, zp, nq, qw, phi, dk, MASS ? mk : 0, acc)
        FA_BR(1, 1, 1, c0);
        FA_BR(1, 1, 0, c1);
        FA_BR(1, 0, 1, c2);
        FA_BR(1, 0, 0, c3);
        FA_BR(0, 1, 1, c4);
        FA_BR(0, 1, 0, c5);
        FA_BR(0, 0, 1, c6);
        FA_BR(0, 0, 0, c7);
#undef FA_BR
        int pos = 0;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    const int s = box_slot(dx, dy, dz);
                    if (s < 0) continue;
                    const bool in = ix + dx >= 0 && ix + dx <= nx && iy + dy >= 0 && iy + dy <= ny && iz + dz >= 0 &&
                                    iz + dz <= nz;
                    if (in) {
                        if (pos < len) values[r0 + pos] = static_cast<T>(acc[s]);
                        ++pos;
                    }
                }
    }
}

// axes[0 .. nx] = X (node (i, 0, 0)), then Y (node (0, j, 0)), then Z (node
// (0, 0, k)); flag[0] = 1 when some node is not at (X[ix], Y[iy], Z[iz])
__global__ void __launch_bounds__(256) k_box_axes(int nx, int ny, int nz, const double* __restrict__ nodew,
                                                 double* __restrict__ axes, int32_t* __restrict__ flag) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1;
    const double4* __restrict__ node = reinterpret_cast<const double4*>(nodew);
    const int64_t naxes = static_cast<int64_t>(nx) + ny + nz + 3;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < naxes; j += (int64_t)gridDim.x * blockDim.x) {
        if (j <= nx) axes[j] = node[j * nyz].x;
        else if (j <= nx + 1 + ny) axes[j] = node[(j - nx - 1) * nz1].y;
        else axes[j] = node[j - nx - ny - 2].z;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 0;
}

__global__ void __launch_bounds__(256) k_box_axes_check(int nx, int ny, int nz, const double* __restrict__ nodew,
                                                       const double* __restrict__ axes, int32_t* __restrict__ flag) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1, nn = static_cast<int64_t>(nx + 1) * nyz;
    const double4* __restrict__ node = reinterpret_cast<const double4*>(nodew);
    bool bad = false;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nn; r += (int64_t)gridDim.x * blockDim.x) {
        const int ix = static_cast<int>(r / nyz);
        const int64_t rem = r - ix * nyz;
        const int iy = static_cast<int>(rem / nz1);
        const int iz = static_cast<int>(rem - iy * nz1);
        const double4 q = node[r];
        bad = bad || !(q.x == axes[ix] && q.y == axes[nx + 1 + iy] && q.z == axes[nx + ny + 2 + iz]);
    }
    if (bad) flag[0] = 1;  // any writer writes 1
}

// One factor at a given point, computed as the torch evaluation of the same
// expression computes it: the argument a s + b rounded after the product and
// after the sum, then the device library's sin, cos or exp of that argument
// (the element kernels' sinpi form instead divides by pi first, which is
// exact where the argument is a multiple of pi, e.g. 0 at x = 1 for sin(pi x),
// where FP64 sin of the rounded argument gives 1.2e-16)
__device__ inline double point_factor(int kind, double a, double b, double s) {
    const double t = __dadd_rn(__dmul_rn(a, s), b);
    switch (kind) {
        case FA_SEP_SIN: return ::sin(t);
        case FA_SEP_COS: return ::cos(t);
        case FA_SEP_EXP: return ::exp(t);
        case FA_SEP_AFFINE: return t;
        default: return sep_factor(kind, a, b, s);
    }
}

// f at n points of dimension gd (row-major): the terms of sep_value in the
// same order, the coordinate dim[k] of the point itself in place of a
// quadrature point (fields are not read: a field-free f)
__global__ void __launch_bounds__(256) k_sep_points(const fa_sep_t* __restrict__ e, int64_t n, int gd,
                                                   const double* __restrict__ pts, double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double* x = pts + i * gd;
        double acc = 0.0;
        int k = 0;
        const int nt = e->nterms;
        for (int t = 0; t < nt; ++t) {
            double v = e->coef[t];
            const int kend = e->term_end[t];
            for (; k < kend; ++k) v *= point_factor(e->kind[k], e->a[k], e->b[k], x[e->dim[k]]);
            acc += v;
        }
        out[i] = acc;
    }
}

extern "C" int fa_sep_points(const void* sep, int64_t n, int gd, const double* points, double* out,
                             fa_stream_t stream) {
    FA_REQUIRE(sep && n >= 0 && gd >= 1 && gd <= 3 && (n == 0 || (points && out)), "fa_sep_points: bad arguments");
    if (n == 0) return FA_OK;
    hipLaunchKernelGGL(k_sep_points, dim3(grid_for(n, 256, 1 << 14)), dim3(256), 0, S(stream),
                       static_cast<const fa_sep_t*>(sep), n, gd, points, out);
    FA_LAUNCH("k_sep_points");
    return FA_OK;
}

extern "C" size_t fa_box_axes_work_bytes(void) { return 256; }

extern "C" int fa_box_axes(int nx, int ny, int nz, const double* node_wide, double* axes, int32_t* on_grid_host,
                           void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && node_wide && axes && on_grid_host, "fa_box_axes: bad arguments");
    FA_REQUIRE(work && work_bytes >= fa_box_axes_work_bytes(), "fa_box_axes: workspace too small");
    const int64_t nn = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
    FA_REQUIRE(nn < INT32_MAX, "fa_box_axes: grid too large for int32 indices");
    hipStream_t s = S(stream);
    int32_t* flag = static_cast<int32_t*>(work);
    const int64_t naxes = static_cast<int64_t>(nx) + ny + nz + 3;
    hipLaunchKernelGGL(k_box_axes, dim3(grid_for(naxes, 256, 64)), dim3(256), 0, s, nx, ny, nz, node_wide, axes, flag);
    FA_LAUNCH("k_box_axes");
    hipLaunchKernelGGL(k_box_axes_check, dim3(grid_for(nn, 256, 1 << 14)), dim3(256), 0, s, nx, ny, nz, node_wide,
                       axes, flag);
    FA_LAUNCH("k_box_axes_check");
    int32_t h = 1;
    FA_HIP(hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    *on_grid_host = h == 0;
    return FA_OK;
}

extern "C" int fa_p1_box_matrix(int dtype, int nx, int ny, int nz, const double* node_wide, const double* axes,
                                int nq, const double* qw, const double* phi, int diff_kind, double diff_scale,
                                const double* diff_coef, int mass_kind, double mass_scale, const double* mass_coef,
                                const int32_t* indptr, int64_t nnz, void* values, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_p1_box_matrix: bad dtype %d", dtype);
    FA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nq > 0 && node_wide && qw && indptr && values,
               "fa_p1_box_matrix: bad arguments");
    FA_REQUIRE(!mass_kind || phi, "fa_p1_box_matrix: mass needs phi");
    FA_REQUIRE(!(diff_kind >= FA_COEF_CELL && !diff_coef) && !(mass_kind >= FA_COEF_CELL && !mass_coef),
               "fa_p1_box_matrix: coefficient array missing");
    const int64_t nn = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
    FA_REQUIRE(nn < INT32_MAX && 6 * static_cast<int64_t>(nx) * ny * nz < INT32_MAX,
               "fa_p1_box_matrix: grid too large for int32 indices");
    FA_REQUIRE(nnz == box_nnz(nx, ny, nz),
               "fa_p1_box_matrix: %lld nonzeros, not the P1 pattern of a %d x %d x %d tetrahedral box (%lld)",
               (long long)nnz, nx, ny, nz, (long long)box_nnz(nx, ny, nz));
    const unsigned grid = xcd_grid(grid_for(nn, 256, 1 << 20));
    const bool sep = diff_kind == FA_COEF_SEP || mass_kind == FA_COEF_SEP;
    timer_begin(FA_TIMER_P1_BOX, S(stream));
#define FA_BOX(T_, SEP_, TEN_)                                                                                      \
    hipLaunchKernelGGL((k_p1_box_rows<T_, SEP_, TEN_>), dim3(grid), dim3(256), 0, S(stream), nx, ny, nz, node_wide,     \
                       axes, nq, qw, phi, diff_kind, diff_scale, diff_coef, mass_kind, mass_scale, mass_coef, indptr, \
                       static_cast<T_*>(values))
#define FA_BOX_T(T_, SEP_) \
    if (axes) FA_BOX(T_, SEP_, true); else FA_BOX(T_, SEP_, false)
    const auto simple = [](int kind) { return kind == FA_COEF_NONE || kind == FA_COEF_CONST || kind == FA_COEF_CELL; };
    if (axes && simple(diff_kind) && simple(mass_kind)) {
#define FA_BT(T_, D_, M_)                                                                                           \
    hipLaunchKernelGGL((k_p1_box_rows_tensor<T_, D_, M_>), dim3(grid), dim3(256), 0, S(stream), nx, ny, nz, axes, nq, \
                       qw, phi, diff_kind, diff_scale, diff_coef, mass_kind, mass_scale, mass_coef, indptr,           \
                       static_cast<T_*>(values))
#define FA_BT2(T_, M_) \
    if (diff_kind == FA_COEF_CELL) FA_BT(T_, true, M_); else FA_BT(T_, false, M_)
        if (dtype == FA_F64) {
            if (mass_kind) { FA_BT2(double, true); } else { FA_BT2(double, false); }
        } else {
            if (mass_kind) { FA_BT2(float, true); } else { FA_BT2(float, false); }
        }
#undef FA_BT2
#undef FA_BT
    } else if (dtype == FA_F64) {
        if (sep) { FA_BOX_T(double, true); } else { FA_BOX_T(double, false); }
    } else {
        if (sep) { FA_BOX_T(float, true); } else { FA_BOX_T(float, false); }
    }
#undef FA_BOX_T
#undef FA_BOX
    timer_end(FA_TIMER_P1_BOX, S(stream));
    FA_LAUNCH("k_p1_box_rows");
    return FA_OK;
}

// work: the cell-order element vectors (NC x 4 doubles), then the axis
// coordinates and the separable-source table of k_box_sep_table
static size_t box_vector_axes_offset(int nx, int ny, int nz) {
    return sizeof(double) * 4 * 6 * static_cast<size_t>(nx) * ny * nz;
}
static size_t box_vector_table_offset(int nx, int ny, int nz) {
    const size_t naxes = static_cast<size_t>(nx) + ny + nz + 3;
    return box_vector_axes_offset(nx, ny, nz) + ((naxes * sizeof(double) + 255) & ~static_cast<size_t>(255));
}

extern "C" size_t fa_p1_box_vector_work_bytes(int nx, int ny, int nz, int nq) {
    const size_t nmax = static_cast<size_t>(std::max(nx, std::max(ny, nz)));
    return box_vector_table_offset(nx, ny, nz) +
           sizeof(double) * FA_SEP_MAX_FACTORS * nmax * 6 * static_cast<size_t>(std::max(nq, 0));
}

extern "C" int fa_p1_box_vector(int dtype, int nx, int ny, int nz, const double* node_wide, int nq, const double* qw,
                                const double* phi, int src_kind, double src_scale, const double* src, void* out,
                                void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_p1_box_vector: bad dtype %d", dtype);
    FA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nq > 0 && node_wide && qw && phi && out && work,
               "fa_p1_box_vector: bad arguments");
    FA_REQUIRE(!(src_kind >= FA_COEF_CELL && !src), "fa_p1_box_vector: source array missing");
    const int64_t nn = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t nc = 6 * static_cast<int64_t>(nx) * ny * nz;
    FA_REQUIRE(nn < INT32_MAX && nc < INT32_MAX, "fa_p1_box_vector: grid too large for int32 indices");
    const size_t need = fa_p1_box_vector_work_bytes(nx, ny, nz, nq);
    FA_REQUIRE(work_bytes >= need && (reinterpret_cast<uintptr_t>(work) & 31) == 0,
               "fa_p1_box_vector: workspace of %zu bytes (32-byte aligned) too small, %zu needed", work_bytes, need);
    double* vals = static_cast<double*>(work);
    double* axes = reinterpret_cast<double*>(static_cast<char*>(work) + box_vector_axes_offset(nx, ny, nz));
    double* tab = reinterpret_cast<double*>(static_cast<char*>(work) + box_vector_table_offset(nx, ny, nz));
    const int nmax = std::max(nx, std::max(ny, nz));
    hipStream_t s = S(stream);
    const bool sep = src_kind == FA_COEF_SEP;
    timer_begin(FA_TIMER_P1_VECTORS, s);
    if (sep) {  // the table's size follows the struct's factor count, read on the device
        const int64_t most = static_cast<int64_t>(nx) + ny + nz + 3 +
                             static_cast<int64_t>(FA_SEP_MAX_FACTORS) * nmax * 6 * nq;
        hipLaunchKernelGGL(k_box_sep_table, dim3(grid_for(most, 256, 1 << 14)), dim3(256), 0, s, nx, ny, nz,
                           node_wide, reinterpret_cast<const fa_sep_t*>(src), nq, nmax, axes, tab);
        FA_LAUNCH("k_box_sep_table");
    }
    const unsigned g1 = grid_for(nc / 6, 256, 1 << 20);
    if (sep)
        hipLaunchKernelGGL(k_p1_box_cell_vectors<true>, dim3(g1), dim3(256), 0, s, nx, ny, nz, node_wide, nq, qw, phi,
                           src_kind, src_scale, src, axes, tab, nmax, vals);
    else
        hipLaunchKernelGGL(k_p1_box_cell_vectors<false>, dim3(g1), dim3(256), 0, s, nx, ny, nz, node_wide, nq, qw,
                           phi, src_kind, src_scale, src, axes, tab, nmax, vals);
    timer_end(FA_TIMER_P1_VECTORS, s);
    FA_LAUNCH("k_p1_box_cell_vectors");
    const unsigned g2 = xcd_grid(grid_for(nn, 256, 1 << 20));
    timer_begin(FA_TIMER_P1_SUM_VALUES, s);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_p1_box_sum_vectors<double>, dim3(g2), dim3(256), 0, s, nx, ny, nz, vals,
                           static_cast<double*>(out));
    else
        hipLaunchKernelGGL(k_p1_box_sum_vectors<float>, dim3(g2), dim3(256), 0, s, nx, ny, nz, vals,
                           static_cast<float*>(out));
    timer_end(FA_TIMER_P1_SUM_VALUES, s);
    FA_LAUNCH("k_p1_box_sum_vectors");
    return FA_OK;
}
