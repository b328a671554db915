// Kernels "a" and "c": element matrices / vectors on simplices and the
// fixed-order (ascending cell) gather of element contributions into CSR rows.
#include "common.hpp"

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

// One thread per (cell, row i); the row of the element matrix is kept in
// registers (LDOF is a compile-time constant).
template <int TD, int LDOF, class T>
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
            const double kd = wq * coef_value(dk, ds, dc, c, q, nq);
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
            const double km = wq * coef_value(mk, ms, mc, c, q, nq);
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

template <int TD, class T>
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
    for (int q = 0; q < nq; ++q) acc += qw[q] * g.meas * coef_value(sk, ss, src, c, q, nq) * phi[(int64_t)q * ldof + i];
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
                                const T* __restrict__ elem, T* __restrict__ values) {
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
        for (int32_t k = ptr[r]; k < k1; ++k) {
            const int32_t e = adj[k];
            const int32_t c = e / ldof;
            for (int j = lane; j < ldof; j += G) {
                const int32_t col = c2d[(int64_t)c * ldof + j];
                const T v = elem[(int64_t)e * ldof + j];
                const int pos = lower_bound_lds(cols, L, col);
                if (pos < L && cols[pos] == col) acc[pos] += v;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
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
// cell reads its vertices once and writes its nq*gd contiguous outputs
template <int GD, int NV>
__global__ void __launch_bounds__(256) k_bc_to_point(int64_t nc, int nq, const double* __restrict__ node,
                                                    const int32_t* __restrict__ cell, const double* __restrict__ bcs,
                                                    double* __restrict__ out) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        double v[NV][GD];
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const int64_t nd = cell[c * NV + a];
#pragma unroll
            for (int d = 0; d < GD; ++d) v[a][d] = node[nd * GD + d];
        }
        double* o = out + c * nq * GD;
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
}

template <int TD, int LDOF, class T>
static void launch_elem(const double* node, const int32_t* cell, int64_t nc, int nq, const double* qw,
                        const double* phi, const double* gphi, int dk, double ds, const double* dc, int mk, double ms,
                        const double* mc, void* elem, hipStream_t s) {
    const int64_t n = nc * LDOF;
    hipLaunchKernelGGL((k_element_matrices<TD, LDOF, T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, node,
                       cell, nc, nq, qw, phi, gphi, dk, ds, dc, mk, ms, mc, static_cast<T*>(elem));
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
                         const int32_t* indptr, const int32_t* indices, int lmax, const void* elem, void* values,
                         hipStream_t s) {
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
                       indices, lmax, static_cast<const T*>(elem), static_cast<T*>(values));
    FA_LAUNCH("k_gather_matrix");
    return FA_OK;
}

template <class T>
static int dispatch_gather(int64_t gdof, int ldof, const int32_t* c2d, const int32_t* ptr, const int32_t* adj,
                           const int32_t* indptr, const int32_t* indices, int lmax, const void* elem, void* values,
                           hipStream_t s) {
    if (ldof <= 4) return launch_gather<T, 4>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, values, s);
    if (ldof <= 8) return launch_gather<T, 8>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, values, s);
    if (ldof <= 16) return launch_gather<T, 16>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, values, s);
    if (ldof <= 32) return launch_gather<T, 32>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, values, s);
    return launch_gather<T, 64>(gdof, ldof, c2d, ptr, adj, indptr, indices, lmax, elem, values, s);
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
    if (dtype == FA_F64) {
        if (td == 2)
            hipLaunchKernelGGL((k_element_vectors<2, double>), dim3(grid), dim3(256), 0, s, node, cell, nc, ldof, nq,
                               qw, phi, src_kind, src_scale, src, static_cast<double*>(elem));
        else
            hipLaunchKernelGGL((k_element_vectors<3, double>), dim3(grid), dim3(256), 0, s, node, cell, nc, ldof, nq,
                               qw, phi, src_kind, src_scale, src, static_cast<double*>(elem));
    } else {
        if (td == 2)
            hipLaunchKernelGGL((k_element_vectors<2, float>), dim3(grid), dim3(256), 0, s, node, cell, nc, ldof, nq,
                               qw, phi, src_kind, src_scale, src, static_cast<float*>(elem));
        else
            hipLaunchKernelGGL((k_element_vectors<3, float>), dim3(grid), dim3(256), 0, s, node, cell, nc, ldof, nq,
                               qw, phi, src_kind, src_scale, src, static_cast<float*>(elem));
    }
    FA_LAUNCH("k_element_vectors");
    return FA_OK;
}

extern "C" int fa_gather_matrix(int dtype, int64_t gdof, int32_t ldof, const int32_t* cell2dof,
                                const int32_t* adj_ptr, const int32_t* adj, const int32_t* indptr,
                                const int32_t* indices, int32_t max_row, const void* elem, void* values,
                                fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_gather_matrix: bad dtype %d", dtype);
    FA_REQUIRE(gdof > 0 && ldof > 0 && max_row > 0, "fa_gather_matrix: bad sizes");
    if (dtype == FA_F64)
        return dispatch_gather<double>(gdof, ldof, cell2dof, adj_ptr, adj, indptr, indices, max_row, elem, values,
                                       S(stream));
    return dispatch_gather<float>(gdof, ldof, cell2dof, adj_ptr, adj, indptr, indices, max_row, elem, values,
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

extern "C" int fa_bc_to_point(int gd, int nv, int64_t nc, int nq, const double* node, const int32_t* cell,
                              const double* bcs, double* out, fa_stream_t stream) {
    FA_REQUIRE(nq >= 1 && nc >= 0, "fa_bc_to_point: bad sizes");
    if (nc == 0) return FA_OK;
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(nc, 256, 1 << 16);
#define FA_BP(GD, NV) \
    hipLaunchKernelGGL((k_bc_to_point<GD, NV>), dim3(grid), dim3(256), 0, s, nc, nq, node, cell, bcs, out)
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
