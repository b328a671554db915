// High-order element matrices (diffusion + mass on tetrahedra) as an MFMA
// contraction.  With the reduced gradients R'[q][i][m] = R[q][i][m] - R[q][i][0]
// (m = 1..3; sum_m grad lambda_m = 0) and the per-cell metric
// G_mn = grad lambda_m . grad lambda_n (6 distinct entries s),
//   A_ij(c) = sum_{s<6} (|T| G_s(c)) sum_q a_c[q] K_s[q][ij]
//           + |T| sum_q m_c[q] K_6[q][ij],
// a_c[q] = w_q k(c, q), m_c[q] = w_q kappa(c, q),
// K_s = R'_im R'_jn (+ R'_in R'_jm off the diagonal), K_6 = phi_i phi_j.
// Each inner sum is a GEMM [cells x q] . [q x ij] on v_mfma_f32_32x32x16_bf16
// (BF16 inputs, FP32 accumulation) or v_mfma_f32_32x32x2f32 (FP32); the sum
// over s with the per-cell metric is done on the FP32 accumulators.
//
// Work split: blockIdx.y picks a slice of TPG column tiles (32 ij each) whose
// reference tensors are staged in LDS once; the block's 4 wavefronts take
// pairs of 32-cell tiles of the block's cell range in turn.
#include "common.hpp"

namespace fa {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int MF_CELLS_PER_BLOCK = 512;  // 16 tiles of 32 cells, 4 per wavefront
constexpr int MF_NS = 7;                 // 6 metric components + mass

// per-cell A operand: out[c][q] = w_q coef(c, q) (zero beyond nq), in the MFMA input type
template <class E>
__global__ void __launch_bounds__(256) k_mfma_coef(int64_t nc, int nq, int kp, const double* __restrict__ qw,
                                                  int kind, double scale, const double* __restrict__ coef,
                                                  E* __restrict__ out) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nc * kp; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = t / kp;
        const int q = static_cast<int>(t - c * kp);
        double v = 0.0;
        if (q < nq) {
            switch (kind) {
                case FA_COEF_CONST: v = qw[q] * scale; break;
                case FA_COEF_CELL: v = qw[q] * scale * coef[c]; break;
                case FA_COEF_CELLQ: v = qw[q] * scale * coef[c * nq + q]; break;
                default: v = 0.0;
            }
        }
        out[t] = static_cast<E>(v);
    }
}

// |T| G_s for the 6 metric entries (11, 22, 33, 12, 13, 23) and |T| (mass)
__device__ inline void mf_metric(const double* __restrict__ node, const int32_t* __restrict__ cell, int64_t c,
                                 float g[MF_NS]) {
    double p[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t n = cell[c * 4 + v];
#pragma unroll
        for (int d = 0; d < 3; ++d) p[v][d] = node[n * 3 + d];
    }
    double a[3], b[3], cc[3], t[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = p[1][d] - p[0][d];
        b[d] = p[2][d] - p[1][d];
        cc[d] = p[3][d] - p[2][d];
    }
    t[0] = b[1] * cc[2] - b[2] * cc[1];
    t[1] = b[2] * cc[0] - b[0] * cc[2];
    t[2] = b[0] * cc[1] - b[1] * cc[0];
    const double vol6 = a[0] * t[0] + a[1] * t[1] + a[2] * t[2];
    const int lf[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    double gl[4][3];
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        double vjk[3], vjm[3];
        for (int d = 0; d < 3; ++d) {
            vjk[d] = p[lf[i][1]][d] - p[lf[i][0]][d];
            vjm[d] = p[lf[i][2]][d] - p[lf[i][0]][d];
        }
        gl[i][0] = (vjm[1] * vjk[2] - vjm[2] * vjk[1]) / vol6;
        gl[i][1] = (vjm[2] * vjk[0] - vjm[0] * vjk[2]) / vol6;
        gl[i][2] = (vjm[0] * vjk[1] - vjm[1] * vjk[0]) / vol6;
    }
    const double meas = vol6 / 6.0;
    const int mm[6][2] = {{1, 1}, {2, 2}, {3, 3}, {1, 2}, {1, 3}, {2, 3}};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int m = mm[s][0], n = mm[s][1];
        g[s] = static_cast<float>(meas * (gl[m][0] * gl[n][0] + gl[m][1] * gl[n][1] + gl[m][2] * gl[n][2]));
    }
    g[6] = static_cast<float>(meas);
}

// per-cell scaled metric, 8 floats per cell: |T| G_s (s < 6), |T|, 0
__global__ void __launch_bounds__(256) k_mfma_metric(int64_t nc, const double* __restrict__ node,
                                                    const int32_t* __restrict__ cell, float* __restrict__ out) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        float g[MF_NS];
        mf_metric(node, cell, c, g);
        float4* o = reinterpret_cast<float4*>(out + c * 8);
        o[0] = make_float4(g[0], g[1], g[2], g[3]);
        o[1] = make_float4(g[4], g[5], g[6], 0.f);
    }
}

// BF16: E = bf16x8 fragments, K = 16 per instruction.  FP32: E = float, K = 2.
template <int KS, bool BF16>
struct MfTraits;

template <int KS>
struct MfTraits<KS, true> {
    typedef __bf16 elem;
    typedef bf16x8 frag;
    static constexpr int KSTEP = 16, PER_LANE = 8;
    __device__ static floatx16 mma(const frag& a, const frag& b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static frag load(const elem* p) { return *reinterpret_cast<const frag*>(p); }
};

template <int KS>
struct MfTraits<KS, false> {
    typedef float elem;
    typedef float frag;
    static constexpr int KSTEP = 2, PER_LANE = 1;
    __device__ static floatx16 mma(const frag& a, const frag& b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static frag load(const elem* p) { return *p; }
};

// Btab: [ngroups][MF_NS][TPG][KS][64 lanes][PER_LANE]; lane l of step ks holds
// K_s[KSTEP*ks + PER_LANE*(l/32) + t][32*tile + l%32], t < PER_LANE.
// A operands ad / am: (nc or 1, KS*KSTEP) rows; row stride 0 broadcasts one row.
template <int KS, bool BF16, int TPG>
__global__ void __launch_bounds__(256) k_element_mfma(int64_t nc, int ldof, const double* __restrict__ node,
                                                     const int32_t* __restrict__ cell,
                                                     const typename MfTraits<KS, BF16>::elem* __restrict__ ad,
                                                     int64_t ad_stride,
                                                     const typename MfTraits<KS, BF16>::elem* __restrict__ am,
                                                     int64_t am_stride,
                                                     const typename MfTraits<KS, BF16>::elem* __restrict__ btab,
                                                     const float* __restrict__ metric, int64_t ld,
                                                     float* __restrict__ elem) {
    typedef MfTraits<KS, BF16> Tr;
    typedef typename Tr::elem E;
    typedef typename Tr::frag F;
    constexpr int KP = KS * Tr::KSTEP;
    constexpr int SLICE = MF_NS * TPG * KS * WAVE * Tr::PER_LANE;  // elements per column slice
    extern __shared__ char smem[];
    E* bl = reinterpret_cast<E*>(smem);
    float* gl = reinterpret_cast<float*>(smem + align_up(sizeof(E) * SLICE, 16));  // [4 waves][2][8][32]
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int group = blockIdx.y;
    {   // stage the group's reference tensors (16-byte copies)
        const int4* src = reinterpret_cast<const int4*>(btab + (int64_t)group * SLICE);
        int4* dst = reinterpret_cast<int4*>(bl);
        const int n16 = static_cast<int>(sizeof(E) * SLICE / 16);
        for (int t = threadIdx.x; t < n16; t += blockDim.x) dst[t] = src[t];
    }
    __syncthreads();
    const int nn = ldof * ldof;
    float* gw = gl + wave * 2 * 8 * 32;  // [2 tiles][8 metric rows][32 cells]
    const int64_t cbase = (int64_t)blockIdx.x * MF_CELLS_PER_BLOCK;
    // each wavefront takes two 32-cell tiles at a time: every B fragment read
    // from LDS feeds two MFMA chains
    for (int pt = wave; pt < MF_CELLS_PER_BLOCK / 64; pt += 4) {
        const int64_t c0 = cbase + pt * 64;
        if (c0 >= nc) break;
        F fa[2][KS], fm[2][KS];
#pragma unroll
        for (int h2 = 0; h2 < 2; ++h2) {
            const int64_t cl = c0 + 32 * h2 + (lane & 31);
            const bool valid = cl < nc;
            {   // metric of the tile's 32 cells, transposed to [s][cell]: two lanes per cell
                const int cc = lane & 31, h = lane >> 5;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (valid) v = reinterpret_cast<const float4*>(metric + cl * 8)[h];
                float* g = gw + h2 * 8 * 32;
                g[(4 * h + 0) * 32 + cc] = v.x;
                g[(4 * h + 1) * 32 + cc] = v.y;
                g[(4 * h + 2) * 32 + cc] = v.z;
                g[(4 * h + 3) * 32 + cc] = v.w;
            }
            const int64_t crow = valid ? cl : c0;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int k = ks * Tr::KSTEP + Tr::PER_LANE * (lane / 32);
                fa[h2][ks] = Tr::load(ad + crow * ad_stride + k);
                fm[h2][ks] = Tr::load(am + crow * am_stride + k);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const float* grow = gw + 4 * (lane / 32);  // row 8*(r/4) + (r&3) of this lane's half
#pragma unroll 1
        for (int tile = 0; tile < TPG; ++tile) {
            const int col = (group * TPG + tile) * 32 + (lane & 31);
            floatx16 tot0 = {}, tot1 = {};
#pragma unroll 1
            for (int s = 0; s < MF_NS; ++s) {  // one pair of accumulator chains live at a time
                floatx16 acc0 = {}, acc1 = {};
                const E* bs = bl + ((s * TPG + tile) * KS) * WAVE * Tr::PER_LANE + lane * Tr::PER_LANE;
                const bool mass = s == MF_NS - 1;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const F b = Tr::load(bs + ks * WAVE * Tr::PER_LANE);
                    acc0 = Tr::mma(mass ? fm[0][ks] : fa[0][ks], b, acc0);
                    acc1 = Tr::mma(mass ? fm[1][ks] : fa[1][ks], b, acc1);
                }
                const float4* g0 = reinterpret_cast<const float4*>(grow + s * 32);
                const float4* g1 = reinterpret_cast<const float4*>(grow + 8 * 32 + s * 32);
#pragma unroll
                for (int q4 = 0; q4 < 4; ++q4) {  // rows 8*q4 + (0..3) of this lane's half
                    const float4 a4 = g0[2 * q4], b4 = g1[2 * q4];
                    tot0[4 * q4 + 0] += a4.x * acc0[4 * q4 + 0];
                    tot0[4 * q4 + 1] += a4.y * acc0[4 * q4 + 1];
                    tot0[4 * q4 + 2] += a4.z * acc0[4 * q4 + 2];
                    tot0[4 * q4 + 3] += a4.w * acc0[4 * q4 + 3];
                    tot1[4 * q4 + 0] += b4.x * acc1[4 * q4 + 0];
                    tot1[4 * q4 + 1] += b4.y * acc1[4 * q4 + 1];
                    tot1[4 * q4 + 2] += b4.z * acc1[4 * q4 + 2];
                    tot1[4 * q4 + 3] += b4.w * acc1[4 * q4 + 3];
                }
            }
            if (col < nn) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = 8 * (r / 4) + 4 * (lane / 32) + (r & 3);
                    const int64_t ca = c0 + row, cb = c0 + 32 + row;
                    if (ca < nc) elem[ca * ld + col] = tot0[r];
                    if (cb < nc) elem[cb * ld + col] = tot1[r];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

template <int KS, bool BF16, int TPG>
static int launch_mfma(int64_t nc, int ldof, const double* node, const int32_t* cell, const void* ad,
                       int64_t ad_stride, const void* am, int64_t am_stride, const void* btab, int ngroups,
                       const float* metric, int64_t ld, float* elem, hipStream_t s) {
    typedef typename MfTraits<KS, BF16>::elem E;
    constexpr int SLICE = MF_NS * TPG * KS * WAVE * MfTraits<KS, BF16>::PER_LANE;
    const size_t lds = align_up(sizeof(E) * SLICE, 16) + sizeof(float) * 4 * 2 * 8 * 32;
    if (lds > 64 * 1024) {
        set_error("fa_element_matrices_mfma: LDS slice of %zu bytes exceeds 64 KiB", lds);
        return FA_ECAPACITY;
    }
    const dim3 grid((unsigned)((nc + MF_CELLS_PER_BLOCK - 1) / MF_CELLS_PER_BLOCK), (unsigned)ngroups);
    hipLaunchKernelGGL((k_element_mfma<KS, BF16, TPG>), grid, dim3(256), lds, s, nc, ldof, node, cell,
                       static_cast<const E*>(ad), ad_stride, static_cast<const E*>(am), am_stride,
                       static_cast<const E*>(btab), metric, ld, elem);
    FA_LAUNCH("k_element_mfma");
    return FA_OK;
}

}  // namespace fa

using namespace fa;

// Largest column-tile count per group whose LDS slice fits 64 KiB.
extern "C" int fa_mfma_tiles_per_group(int precision, int ks) {
    const int per = MF_NS * ks * WAVE * (precision == 0 ? 8 * 2 : 4);  // bytes per tile
    const int room = 64 * 1024 - 4 * 2 * 8 * 32 * 4 - 16;
    int t = room / per;
    if (t >= 4) return 4;
    if (t >= 2) return 2;
    return t >= 1 ? 1 : 0;
}

extern "C" size_t fa_mfma_work_bytes(int precision, int64_t nc, int ks) {
    const size_t e = precision == 0 ? 2 : 4;
    const size_t kp = (size_t)ks * (precision == 0 ? 16 : 2);
    return 2 * align_up(nc * kp * e) + align_up(nc * 8 * sizeof(float)) + 1024;
}

extern "C" int fa_element_matrices_mfma(int precision, int ldof, int nq, int ks, int tpg, int ngroups,
                                        const double* node, const int32_t* cell, int64_t nc, const double* qw,
                                        const void* btab, int diff_kind, double diff_scale, const double* diff_coef,
                                        int mass_kind, double mass_scale, const double* mass_coef, float* elem,
                                        int64_t elem_stride, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(precision == 0 || precision == 1, "fa_element_matrices_mfma: precision must be 0 (BF16) or 1 (FP32)");
    FA_REQUIRE(nc > 0 && ldof > 0 && nq > 0 && btab && elem, "fa_element_matrices_mfma: bad arguments");
    const int kstep = precision == 0 ? 16 : 2;
    FA_REQUIRE(ks * kstep >= nq, "fa_element_matrices_mfma: ks * %d < nq", kstep);
    FA_REQUIRE((int64_t)ngroups * tpg * 32 >= (int64_t)ldof * ldof, "fa_element_matrices_mfma: too few column tiles");
    const int64_t ld = elem_stride > 0 ? elem_stride : (int64_t)ldof * ldof;
    FA_REQUIRE(ld >= (int64_t)ldof * ldof, "fa_element_matrices_mfma: element stride below ldof^2");
    FA_REQUIRE(!(diff_kind >= FA_COEF_CELL && !diff_coef) && !(mass_kind >= FA_COEF_CELL && !mass_coef),
               "fa_element_matrices_mfma: coefficient array missing");
    hipStream_t s = S(stream);
    const int kp = ks * kstep;
    const size_t e = precision == 0 ? 2 : 4;
    Arena ar(work, work_bytes);
    const bool dcell = diff_kind >= FA_COEF_CELL, mcell = mass_kind >= FA_COEF_CELL;
    char* ad = ar.take<char>((dcell ? nc : 1) * kp * e);
    char* am = ar.take<char>((mcell ? nc : 1) * kp * e);
    float* metric = ar.take<float>(nc * 8);
    if (!ad || !am || !metric) {
        set_error("fa_element_matrices_mfma: workspace too small");
        return FA_ENOMEM;
    }
    const int64_t nd = dcell ? nc : 1, nm = mcell ? nc : 1;
#define FA_COEF(E_, out, n_, kind, scale, coef)                                                                  \
    hipLaunchKernelGGL(k_mfma_coef<E_>, dim3(grid_for((n_) * kp, 256, 1 << 16)), dim3(256), 0, s, (int64_t)(n_), \
                       nq, kp, qw, kind, scale, coef, reinterpret_cast<E_*>(out))
    if (precision == 0) {
        FA_COEF(__bf16, ad, nd, diff_kind, diff_scale, diff_coef);
        FA_COEF(__bf16, am, nm, mass_kind, mass_scale, mass_coef);
    } else {
        FA_COEF(float, ad, nd, diff_kind, diff_scale, diff_coef);
        FA_COEF(float, am, nm, mass_kind, mass_scale, mass_coef);
    }
#undef FA_COEF
    hipLaunchKernelGGL(k_mfma_metric, dim3(grid_for(nc, 256, 1 << 16)), dim3(256), 0, s, nc, node, cell, metric);
    FA_LAUNCH("k_mfma_coef");
    const int64_t das = dcell ? kp : 0, mas = mcell ? kp : 0;
#define FA_MF(KS_, BF_, TPG_) \
    return launch_mfma<KS_, BF_, TPG_>(nc, ldof, node, cell, ad, das, am, mas, btab, ngroups, metric, ld, elem, s)
#define FA_MF_TPG(KS_, BF_)                 \
    do {                                    \
        if (tpg == 4) FA_MF(KS_, BF_, 4);  \
        if (tpg == 2) FA_MF(KS_, BF_, 2);  \
        if (tpg == 1) FA_MF(KS_, BF_, 1);  \
    } while (0)
    if (precision == 0) {
        switch (ks) {
            case 1: FA_MF_TPG(1, true); break;
            case 2: FA_MF_TPG(2, true); break;
            case 3: FA_MF_TPG(3, true); break;
            case 4: FA_MF_TPG(4, true); break;
            case 6: FA_MF_TPG(6, true); break;
            default: break;
        }
    } else {
        switch (ks) {
            case 2: FA_MF_TPG(2, false); break;
            case 5: FA_MF_TPG(5, false); break;
            case 10: FA_MF_TPG(10, false); break;
            case 18: FA_MF_TPG(18, false); break;
            case 28: FA_MF_TPG(28, false); break;
            default: break;
        }
    }
#undef FA_MF_TPG
#undef FA_MF
    set_error("fa_element_matrices_mfma: no kernel for precision %d, ks %d, tiles per group %d", precision, ks, tpg);
    return FA_EINVAL;
}
