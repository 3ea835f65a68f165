// fno.hip -- the learned coarse propagator of the parareal driver: a spectral
// neural operator (Fourier neural operator) on the square grid of the 2-D
// path, trained and evaluated by this file's own kernels.
//
// There is no counterpart in the reference; the operator maps the physical
// fields [eta, phi, depth] (normalised) to the increment of [eta, phi] over
// one coarse time step.  Architecture: a pointwise lift in_ch -> width, L
// Fourier layers v <- act(W v + b + K v) with K the spectral convolution over
// the retained modes (kx in [0, modes), ky in [0, modes) and [n-modes, n)),
// act = GELU except after the last layer, and a two-layer pointwise
// projection width -> hidden (GELU) -> out_ch.
//
// Every matrix product (the lift and projections, the pointwise W, the four
// truncated DFTs as products with DFT matrices, and all their backward
// products) runs through one hand-written batched GEMM on the CDNA4 matrix
// cores: operands converted to BF16 on their way to LDS,
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation, FP32 results.  The
// per-mode complex channel mixing (a few MFLOP) and the elementwise work run
// in FP32 on the vector units.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/hawassi_hip.h"

namespace hw {
void set_error(const char* fmt, ...);
}

namespace hw {
namespace fno {

#define FNO_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::hw::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                            __FILE__, __LINE__);                                        \
            return HW_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)
#define FNO_CHECK() FNO_TRY(hipGetLastError())

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ------------------------------------------------------------------ GEMM --
// C[b] = alpha op(A)[b] op(B)[b] + beta C[b] with general element strides:
// A(m, k) = A[m a_rs + k a_cs + b a_bs], B(k, n) = B[k b_rs + n b_cs + b b_bs],
// C(m, n) = C[m c_rs + n c_cs + b c_bs].  A transposed operand is a stride
// swap.  With splits > 1 the K range is cut in `splits` chunks (grid.z =
// batch * splits) and each chunk writes alpha * (its partial product) to the
// workspace W[s][b][m][n]; k_splitk_reduce then forms C.
struct GemmArgs {
    int M, N, K, batch, splits, kchunk;
    const float* A; long long a_rs, a_cs, a_bs;
    const float* B; long long b_rs, b_cs, b_bs;
    float* C; long long c_rs, c_cs, c_bs;
    float alpha, beta;
    int to_ws;  // write alpha * partial products to W[s][b][m][n] (split-K / batch sums)
    float* W;   // workspace (splits * batch * M * N floats)
};

constexpr int KT = 32;          // K per LDS stage (two MFMA of K = 16)
constexpr int LDK = KT + 8;     // LDS row length in bf16 (padding against bank conflicts)

// BM x BN block tile, 4 waves in a (BM/32) x (BN/32) arrangement, each wave
// one 32 x 32 output tile.  MFMA lane layout (v_mfma_f32_32x32x16_bf16):
// lane l holds A[l % 32][8 (l/32) + j] and B[8 (l/32) + j][l % 32], j < 8, and
// result r < 16 is D[8 (r/4) + 4 (l/32) + r % 4][l % 32].
//
// X3 (split precision): every operand is split as v = hi + lo with hi =
// bf16(v), lo = bf16(v - hi), and the tile product is hi*hi + hi*lo + lo*hi,
// three MFMA per K step; operands then carry 16 mantissa bits and the product
// is good to about 1e-5 relative, where plain BF16 is good to about 4e-3.
template <int BM, int BN, bool X3>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    static_assert((BM / 32) * (BN / 32) == 4, "four waves");
    constexpr int NP = X3 ? 2 : 1;   // hi (and lo) planes
    __shared__ __attribute__((aligned(16))) __bf16 As[NP][BM * LDK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NP][BN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / (BN / 32), wn = wave % (BN / 32);
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int bz = blockIdx.z;
    const int b = bz % g.batch, s = bz / g.batch;
    const int k_beg = s * g.kchunk;
    const int k_end = min(g.K, k_beg + g.kchunk);
    const float* A = g.A + (long long)b * g.a_bs;
    const float* B = g.B + (long long)b * g.b_bs;
    const bool a_kfast = g.a_cs == 1, b_kfast = g.b_rs == 1;
    f32x16 acc = {};
    for (int k0 = k_beg; k0 < k_end; k0 += KT) {
        // A tile BM x KT -> As[m][k]
        for (int idx = tid; idx < BM * KT; idx += 256) {
            int mm, kk;
            if (a_kfast) { mm = idx / KT; kk = idx % KT; }
            else { mm = idx % BM; kk = idx / BM; }
            const int m = m0 + mm, k = k0 + kk;
            float v = (m < g.M && k < k_end) ? A[(long long)m * g.a_rs + (long long)k * g.a_cs] : 0.f;
            const __bf16 hi = (__bf16)v;
            As[0][mm * LDK + kk] = hi;
            if constexpr (X3) As[NP - 1][mm * LDK + kk] = (__bf16)(v - (float)hi);
        }
        // B tile KT x BN -> Bs[n][k]
        for (int idx = tid; idx < BN * KT; idx += 256) {
            int nn, kk;
            if (b_kfast) { nn = idx / KT; kk = idx % KT; }
            else { nn = idx % BN; kk = idx / BN; }
            const int n = n0 + nn, k = k0 + kk;
            float v = (n < g.N && k < k_end) ? B[(long long)k * g.b_rs + (long long)n * g.b_cs] : 0.f;
            const __bf16 hi = (__bf16)v;
            Bs[0][nn * LDK + kk] = hi;
            if constexpr (X3) Bs[NP - 1][nn * LDK + kk] = (__bf16)(v - (float)hi);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KT; kk += 16) {
            const int ao = (wm * 32 + (lane & 31)) * LDK + kk + 8 * (lane >> 5);
            const int bo = (wn * 32 + (lane & 31)) * LDK + kk + 8 * (lane >> 5);
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(&As[0][ao]);
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&Bs[0][bo]);
            if constexpr (X3) {
                const bf16x8 al = *reinterpret_cast<const bf16x8*>(&As[NP - 1][ao]);
                const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&Bs[NP - 1][bo]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bv, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bl, acc, 0, 0, 0);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m >= g.M || n >= g.N) continue;
        if (g.to_ws) {
            g.W[(((long long)s * g.batch + b) * g.M + m) * g.N + n] = g.alpha * acc[r];
        } else {
            float* c = g.C + (long long)b * g.c_bs + (long long)m * g.c_rs + (long long)n * g.c_cs;
            *c = g.alpha * acc[r] + (g.beta != 0.f ? g.beta * *c : 0.f);
        }
    }
}

// C(m, n) = sum over splits (and over the batch when sum_batch) of W + beta C
__global__ void k_splitk_reduce(const float* W, int splits, int batch, int M, int N, float* C, long long c_rs,
                                long long c_cs, long long c_bs, float beta, int sum_batch) {
    const long long MN = (long long)M * N;
    const int nb = sum_batch ? 1 : batch;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < MN * nb; i += stride) {
        const int b = (int)(i / MN);
        const long long mn = i % MN;
        const int m = (int)(mn / N), n = (int)(mn % N);
        float acc = 0.f;
        for (int s = 0; s < splits; ++s)
            for (int bb = (sum_batch ? 0 : b); bb < (sum_batch ? batch : b + 1); ++bb)
                acc += W[((long long)s * batch + bb) * MN + mn];
        float* c = C + (long long)b * c_bs + (long long)m * c_rs + (long long)n * c_cs;
        *c = acc + (beta != 0.f ? beta * *c : 0.f);
    }
}

// ------------------------------------------------------- elementwise ------
__device__ inline float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.3989422804014327f * expf(-0.5f * x * x);
}

// pre[b][c][p] += bias[c]; act = gelu(pre) (or pre when gelu_on == 0)
__global__ void k_bias_act(float* pre, float* act, const float* bias, int C, long long P, long long total, int gelu_on) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int c = (int)((i / P) % C);
        const float v = pre[i] + bias[c];
        pre[i] = v;
        act[i] = gelu_on ? gelu(v) : v;
    }
}

// d <- d * gelu'(pre)
__global__ void k_gelu_bwd(float* d, const float* pre, long long total) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) d[i] *= gelu_d(pre[i]);
}

// per-(channel, chunk) partial sums of x[b][c][p]: part[c][chunk]
__global__ void k_chan_partial(const float* x, int batch, int C, long long P, int chunks, float* part) {
    const int c = blockIdx.y, ch = blockIdx.x;
    const long long per = (P + chunks - 1) / chunks;
    const long long p0 = ch * per, p1 = p0 + per < P ? p0 + per : P;
    float acc = 0.f;
    for (int b = 0; b < batch; ++b)
        for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) acc += x[((long long)b * C + c) * P + p];
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[c * chunks + ch] = red[0];
}
__global__ void k_chan_final(const float* part, int C, int chunks, float* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float acc = 0.f;
    for (int i = 0; i < chunks; ++i) acc += part[c * chunks + i];
    out[c] = acc;
}

// squared-error partials and d(loss)/d(out) for loss = mean((out - t)^2)
__global__ void k_mse(const float* out, const float* t, long long total, float* dout, float* part) {
    __shared__ float red[256];
    float acc = 0.f;
    const float sc = 2.f / (float)total;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const float d = out[i] - t[i];
        acc += d * d;
        if (dout) dout[i] = sc * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba) with bias correction
__global__ void k_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// gather dataset samples: dst[j] = src[idx[j]], each sample `len` floats
__global__ void k_gather(const float* src, const int* idx, int nb, long long len, float* dst) {
    const long long total = (long long)nb * len;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int j = (int)(i / len);
        dst[i] = src[(long long)idx[j] * len + i % len];
    }
}

// ----------------------------------------------- per-mode channel mixing --
// Spectra zh[b][c][part][m] (part 0 real, 1 imaginary; Mm modes per part),
// weights R[part][i][o][m].  Forward: y[b][o][m] = sum_i z[b][i][m] R[i][o][m].
__global__ void k_mix_fwd(const float* z, const float* R, float* y, int batch, int C, int Mm) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)batch * C * Mm) return;
    const int m = (int)(t % Mm), o = (int)((t / Mm) % C), b = (int)(t / ((long long)Mm * C));
    const long long RC = (long long)C * C * Mm;
    float ar = 0.f, ai = 0.f;
    for (int i = 0; i < C; ++i) {
        const float zr = z[(((long long)b * C + i) * 2 + 0) * Mm + m], zi = z[(((long long)b * C + i) * 2 + 1) * Mm + m];
        const float rr = R[((long long)i * C + o) * Mm + m], ri = R[RC + ((long long)i * C + o) * Mm + m];
        ar += zr * rr - zi * ri;
        ai += zr * ri + zi * rr;
    }
    y[(((long long)b * C + o) * 2 + 0) * Mm + m] = ar;
    y[(((long long)b * C + o) * 2 + 1) * Mm + m] = ai;
}
// dz[b][i][m] = sum_o dy[b][o][m] conj(R[i][o][m])
__global__ void k_mix_bwd_data(const float* dy, const float* R, float* dz, int batch, int C, int Mm) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)batch * C * Mm) return;
    const int m = (int)(t % Mm), i = (int)((t / Mm) % C), b = (int)(t / ((long long)Mm * C));
    const long long RC = (long long)C * C * Mm;
    float ar = 0.f, ai = 0.f;
    for (int o = 0; o < C; ++o) {
        const float gr = dy[(((long long)b * C + o) * 2 + 0) * Mm + m], gi = dy[(((long long)b * C + o) * 2 + 1) * Mm + m];
        const float rr = R[((long long)i * C + o) * Mm + m], ri = R[RC + ((long long)i * C + o) * Mm + m];
        ar += gr * rr + gi * ri;
        ai += gi * rr - gr * ri;
    }
    dz[(((long long)b * C + i) * 2 + 0) * Mm + m] = ar;
    dz[(((long long)b * C + i) * 2 + 1) * Mm + m] = ai;
}
// dR[i][o][m] = sum_b conj(z[b][i][m]) dy[b][o][m]
__global__ void k_mix_bwd_w(const float* z, const float* dy, float* dR, int batch, int C, int Mm) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)C * C * Mm) return;
    const int m = (int)(t % Mm), o = (int)((t / Mm) % C), i = (int)(t / ((long long)Mm * C));
    const long long RC = (long long)C * C * Mm;
    float ar = 0.f, ai = 0.f;
    for (int b = 0; b < batch; ++b) {
        const float zr = z[(((long long)b * C + i) * 2 + 0) * Mm + m], zi = z[(((long long)b * C + i) * 2 + 1) * Mm + m];
        const float gr = dy[(((long long)b * C + o) * 2 + 0) * Mm + m], gi = dy[(((long long)b * C + o) * 2 + 1) * Mm + m];
        ar += zr * gr + zi * gi;
        ai += zr * gi - zi * gr;
    }
    dR[((long long)i * C + o) * Mm + m] = ar;
    dR[RC + ((long long)i * C + o) * Mm + m] = ai;
}

// ---------------------------------------------------------------- host --
static unsigned grid_for(long long n, int per = 256, long long cap = 65536) {
    long long g = (n + per - 1) / per;
    return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

struct Net {
    hw_fno_config cfg{};
    int n = 0, C = 0, L = 0, Hd = 0, mx = 0, KY = 0, Mm = 0, cin = 0, cout = 0, maxb = 0;
    long long P = 0;            // n * n
    hipStream_t stream = nullptr;
    // parameter offsets (floats) in the flat parameter array
    long long o_pw = 0, o_pb = 0, o_q1w = 0, o_q1b = 0, o_q2w = 0, o_q2b = 0, nparam = 0;
    std::vector<long long> o_w, o_b, o_r;
    float *prm = nullptr, *grd = nullptr, *am = nullptr, *av = nullptr;
    long long adam_t = 0;
    bool split_precision = false;   // products in split BF16 (HW_FNO_SPLIT) rather than BF16
    double gemm_flops = 0.0;        // 2 M N K per product launched (one MFMA product each, split or not)
    double gemm_bytes = 0.0;        // FP32 operands read once, result written once (read too when beta != 0)
    // DFT matrices
    float *Fx = nullptr, *A1 = nullptr, *A2 = nullptr, *A3 = nullptr, *A4 = nullptr, *Hx = nullptr;
    // activations (max batch)
    float *xin = nullptr, *tgt = nullptr, *out = nullptr, *dout = nullptr;
    std::vector<float*> V, Pre, Zh;     // V[0..L] layer inputs (V[L] = last output), Pre[l], Zh[l]
    float *Zx = nullptr, *Yh = nullptr, *Yx = nullptr, *Hpre = nullptr, *Hact = nullptr;
    float *dV = nullptr, *dV2 = nullptr, *dH = nullptr, *dZh = nullptr, *dYh = nullptr;
    float *ws = nullptr, *part = nullptr;
    size_t ws_floats = 0;
    // dataset
    float *dsx = nullptr, *dsy = nullptr;
    int ds_count = 0;
    int* d_idx = nullptr;
    std::vector<void*> allocs;

    ~Net() {
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int dalloc(float** p, size_t count) {
        FNO_TRY(hipMalloc((void**)p, sizeof(float) * (count ? count : 1)));
        allocs.push_back(*p);
        return HW_OK;
    }

    // ---------------------------------------------------------------- gemm
    template <bool X3>
    void launch_gemm(const GemmArgs& g, int M, int N, unsigned zs) {
        if (M <= 32) {
            hipLaunchKernelGGL((k_gemm<32, 128, X3>), dim3((N + 127) / 128, (M + 31) / 32, zs), dim3(256), 0, stream, g);
        } else if (N <= 32) {
            hipLaunchKernelGGL((k_gemm<128, 32, X3>), dim3((N + 31) / 32, (M + 127) / 128, zs), dim3(256), 0, stream, g);
        } else {
            hipLaunchKernelGGL((k_gemm<64, 64, X3>), dim3((N + 63) / 64, (M + 63) / 64, zs), dim3(256), 0, stream, g);
        }
    }

    int gemm(int M, int N, int K, int batch, const float* A, long long a_rs, long long a_cs, long long a_bs,
             const float* B, long long b_rs, long long b_cs, long long b_bs, float* C, long long c_rs, long long c_cs,
             long long c_bs, float alpha, float beta, int splits = 1, bool sum_batch = false) {
        GemmArgs g{};
        gemm_flops += 2.0 * M * N * (double)K * batch;
        gemm_bytes += 4.0 * batch * ((double)M * K + (double)K * N + (double)M * N * (beta != 0.f ? 2 : 1));
        g.M = M; g.N = N; g.K = K; g.batch = batch;
        g.A = A; g.a_rs = a_rs; g.a_cs = a_cs; g.a_bs = a_bs;
        g.B = B; g.b_rs = b_rs; g.b_cs = b_cs; g.b_bs = b_bs;
        g.C = C; g.c_rs = c_rs; g.c_cs = c_cs; g.c_bs = c_bs;
        g.alpha = alpha; g.beta = beta;
        if (splits < 1) splits = 1;
        g.kchunk = ((K + splits - 1) / splits + KT - 1) / KT * KT;
        splits = (K + g.kchunk - 1) / g.kchunk;
        g.splits = splits;
        g.to_ws = splits > 1 || sum_batch;
        if (g.to_ws) {
            if ((size_t)splits * batch * M * N > ws_floats) {
                hw::set_error("fno: split-K workspace too small (%d x %d x %d x %d)", splits, batch, M, N);
                return HW_ERR_ARG;
            }
            g.W = ws;
        }
        const unsigned zs = (unsigned)(batch * splits);
        if (split_precision) launch_gemm<true>(g, M, N, zs);
        else launch_gemm<false>(g, M, N, zs);
        FNO_CHECK();
        if (g.to_ws) {
            const long long outs = (long long)M * N * (sum_batch ? 1 : batch);
            hipLaunchKernelGGL(k_splitk_reduce, dim3(grid_for(outs)), dim3(256), 0, stream, ws, splits, batch, M, N, C,
                               c_rs, c_cs, c_bs, beta, sum_batch ? 1 : 0);
            FNO_CHECK();
        }
        return HW_OK;
    }

    // ------------------------------------------------------------- set-up
    int init(const hw_fno_config* c) {
        cfg = *c;
        n = c->n; C = c->width; L = c->layers; Hd = c->hidden; mx = c->modes; KY = 2 * c->modes; cin = c->in_ch;
        cout = c->out_ch; maxb = c->max_batch;
        if (n < 8 || C < 1 || L < 1 || Hd < 1 || mx < 1 || 2 * mx > n || cin < 1 || cout < 1 || maxb < 1) {
            hw::set_error("fno: bad configuration (n %d width %d layers %d hidden %d modes %d)", n, C, L, Hd, mx);
            return HW_ERR_ARG;
        }
        Mm = KY * mx;
        P = (long long)n * n;
        FNO_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        // parameters
        long long o = 0;
        o_pw = o; o += (long long)C * cin;
        o_pb = o; o += C;
        o_w.resize(L); o_b.resize(L); o_r.resize(L);
        for (int l = 0; l < L; ++l) {
            o_w[l] = o; o += (long long)C * C;
            o_b[l] = o; o += C;
            o_r[l] = o; o += 2LL * C * C * Mm;
        }
        o_q1w = o; o += (long long)Hd * C;
        o_q1b = o; o += Hd;
        o_q2w = o; o += (long long)cout * Hd;
        o_q2b = o; o += cout;
        nparam = o;
        int rc;
        if ((rc = dalloc(&prm, nparam)) || (rc = dalloc(&grd, nparam)) || (rc = dalloc(&am, nparam)) ||
            (rc = dalloc(&av, nparam)))
            return rc;
        FNO_TRY(hipMemsetAsync(am, 0, sizeof(float) * nparam, stream));
        FNO_TRY(hipMemsetAsync(av, 0, sizeof(float) * nparam, stream));
        // seeded initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for
        // the pointwise maps and biases, uniform(0, 1/(C C)) for the spectral
        // weights (real and imaginary parts)
        std::vector<float> hp(nparam);
        uint64_t st = c->seed ? c->seed : 0x9E3779B97F4A7C15ull;
        auto rnd = [&st]() {
            st += 0x9E3779B97F4A7C15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](long long off, long long cnt, double lo, double hi) {
            for (long long i = 0; i < cnt; ++i) hp[off + i] = (float)(lo + (hi - lo) * rnd());
        };
        fill(o_pw, (long long)C * cin, -1.0 / std::sqrt((double)cin), 1.0 / std::sqrt((double)cin));
        fill(o_pb, C, -1.0 / std::sqrt((double)cin), 1.0 / std::sqrt((double)cin));
        for (int l = 0; l < L; ++l) {
            fill(o_w[l], (long long)C * C, -1.0 / std::sqrt((double)C), 1.0 / std::sqrt((double)C));
            fill(o_b[l], C, -1.0 / std::sqrt((double)C), 1.0 / std::sqrt((double)C));
            fill(o_r[l], 2LL * C * C * Mm, 0.0, 1.0 / ((double)C * C));
        }
        fill(o_q1w, (long long)Hd * C, -1.0 / std::sqrt((double)C), 1.0 / std::sqrt((double)C));
        fill(o_q1b, Hd, -1.0 / std::sqrt((double)C), 1.0 / std::sqrt((double)C));
        fill(o_q2w, (long long)cout * Hd, -1.0 / std::sqrt((double)Hd), 1.0 / std::sqrt((double)Hd));
        fill(o_q2b, cout, -1.0 / std::sqrt((double)Hd), 1.0 / std::sqrt((double)Hd));
        FNO_TRY(hipMemcpyAsync(prm, hp.data(), sizeof(float) * nparam, hipMemcpyHostToDevice, stream));
        // DFT matrices (FP64 angles, rounded once)
        std::vector<int> kys(KY);
        for (int i = 0; i < mx; ++i) { kys[i] = i; kys[mx + i] = n - mx + i; }
        std::vector<float> hFx((size_t)n * 2 * mx), hA1((size_t)2 * KY * n), hA2((size_t)2 * KY * n),
            hA3((size_t)n * 2 * KY), hA4((size_t)n * 2 * KY), hHx((size_t)2 * mx * n);
        const double tp = 2.0 * M_PI / n, inv = 1.0 / ((double)n * n);
        for (int x = 0; x < n; ++x)
            for (int k = 0; k < mx; ++k) {
                const double a = tp * (double)((long)k * x % n);
                hFx[(size_t)x * 2 * mx + k] = (float)std::cos(a);
                hFx[(size_t)x * 2 * mx + mx + k] = (float)(-std::sin(a));
                const double w = (k == 0 ? 1.0 : 2.0) * inv;
                hHx[(size_t)k * n + x] = (float)(w * std::cos(a));
                hHx[(size_t)(mx + k) * n + x] = (float)(-w * std::sin(a));
            }
        for (int i = 0; i < KY; ++i)
            for (int y = 0; y < n; ++y) {
                const double a = tp * (double)((long)kys[i] * y % n);
                const double fr = std::cos(a), fi = -std::sin(a);     // forward exp(-i a)
                hA1[(size_t)i * n + y] = (float)fr;                   // [Fr; Fi] rows
                hA1[(size_t)(KY + i) * n + y] = (float)fi;
                hA2[(size_t)i * n + y] = (float)(-fi);                // [-Fi; Fr] rows
                hA2[(size_t)(KY + i) * n + y] = (float)fr;
                const double gr = std::cos(a), gi = std::sin(a);      // inverse exp(+i a)
                hA3[(size_t)y * 2 * KY + i] = (float)gr;              // [Gr, -Gi]
                hA3[(size_t)y * 2 * KY + KY + i] = (float)(-gi);
                hA4[(size_t)y * 2 * KY + i] = (float)gi;              // [Gi, Gr]
                hA4[(size_t)y * 2 * KY + KY + i] = (float)gr;
            }
        if ((rc = dalloc(&Fx, hFx.size())) || (rc = dalloc(&A1, hA1.size())) || (rc = dalloc(&A2, hA2.size())) ||
            (rc = dalloc(&A3, hA3.size())) || (rc = dalloc(&A4, hA4.size())) || (rc = dalloc(&Hx, hHx.size())))
            return rc;
        FNO_TRY(hipMemcpyAsync(Fx, hFx.data(), sizeof(float) * hFx.size(), hipMemcpyHostToDevice, stream));
        FNO_TRY(hipMemcpyAsync(A1, hA1.data(), sizeof(float) * hA1.size(), hipMemcpyHostToDevice, stream));
        FNO_TRY(hipMemcpyAsync(A2, hA2.data(), sizeof(float) * hA2.size(), hipMemcpyHostToDevice, stream));
        FNO_TRY(hipMemcpyAsync(A3, hA3.data(), sizeof(float) * hA3.size(), hipMemcpyHostToDevice, stream));
        FNO_TRY(hipMemcpyAsync(A4, hA4.data(), sizeof(float) * hA4.size(), hipMemcpyHostToDevice, stream));
        FNO_TRY(hipMemcpyAsync(Hx, hHx.data(), sizeof(float) * hHx.size(), hipMemcpyHostToDevice, stream));
        // activations
        const size_t act = (size_t)maxb * C * P, spec = (size_t)maxb * C * 2 * Mm;
        V.resize(L + 1); Pre.resize(L); Zh.resize(L);
        for (int l = 0; l <= L; ++l)
            if ((rc = dalloc(&V[l], act))) return rc;
        for (int l = 0; l < L; ++l)
            if ((rc = dalloc(&Pre[l], act)) || (rc = dalloc(&Zh[l], spec))) return rc;
        ws_floats = (size_t)maxb * 64 * (size_t)(Hd > C ? Hd : C) * (Hd > C ? Hd : C);
        if (ws_floats < (size_t)64 * C * C * 4) ws_floats = (size_t)64 * C * C * 4;
        if ((rc = dalloc(&xin, (size_t)maxb * cin * P)) || (rc = dalloc(&tgt, (size_t)maxb * cout * P)) ||
            (rc = dalloc(&out, (size_t)maxb * cout * P)) || (rc = dalloc(&dout, (size_t)maxb * cout * P)) ||
            (rc = dalloc(&Zx, (size_t)maxb * C * n * 2 * mx)) || (rc = dalloc(&Yh, spec)) ||
            (rc = dalloc(&Yx, (size_t)maxb * C * n * 2 * mx)) || (rc = dalloc(&Hpre, (size_t)maxb * Hd * P)) ||
            (rc = dalloc(&Hact, (size_t)maxb * Hd * P)) || (rc = dalloc(&dV, act)) || (rc = dalloc(&dV2, act)) ||
            (rc = dalloc(&dH, (size_t)maxb * Hd * P)) || (rc = dalloc(&dZh, spec)) || (rc = dalloc(&dYh, spec)) ||
            (rc = dalloc(&ws, ws_floats)) || (rc = dalloc(&part, 65536)))
            return rc;
        FNO_TRY(hipMalloc((void**)&d_idx, sizeof(int) * (size_t)maxb));
        allocs.push_back(d_idx);
        FNO_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    // ------------------------------------------------------------ forward
    // spectral convolution of layer l: Zh[l] from V[l], result into dst (b x C x P)
    int spectral(int l, int b, float* dst) {
        int rc;
        const float* v = V[l];
        // x transform: Zx[b C n][2 mx] = V[b C n][n] Fx[n][2 mx]
        if ((rc = gemm(b * C * n, 2 * mx, n, 1, v, n, 1, 0, Fx, 2 * mx, 1, 0, Zx, 2 * mx, 1, 0, 1.f, 0.f))) return rc;
        // y transform per (b, c): Zh[2 KY][mx] = A1 Zx_re + A2 Zx_im
        const long long zb = (long long)n * 2 * mx, hb = 2LL * Mm;
        if ((rc = gemm(2 * KY, mx, n, b * C, A1, n, 1, 0, Zx, 2 * mx, 1, zb, Zh[l], mx, 1, hb, 1.f, 0.f))) return rc;
        if ((rc = gemm(2 * KY, mx, n, b * C, A2, n, 1, 0, Zx + mx, 2 * mx, 1, zb, Zh[l], mx, 1, hb, 1.f, 1.f))) return rc;
        // mode mixing
        hipLaunchKernelGGL(k_mix_fwd, dim3(grid_for((long long)b * C * Mm)), dim3(256), 0, stream, Zh[l],
                           prm + o_r[l], Yh, b, C, Mm);
        FNO_CHECK();
        // inverse y: Yx[n][2 mx] (re | im) = A3 Yh, A4 Yh
        if ((rc = gemm(n, mx, 2 * KY, b * C, A3, 2 * KY, 1, 0, Yh, mx, 1, hb, Yx, 2 * mx, 1, zb, 1.f, 0.f))) return rc;
        if ((rc = gemm(n, mx, 2 * KY, b * C, A4, 2 * KY, 1, 0, Yh, mx, 1, hb, Yx + mx, 2 * mx, 1, zb, 1.f, 0.f))) return rc;
        // inverse x: dst[b C n][n] = Yx Hx
        return gemm(b * C * n, n, 2 * mx, 1, Yx, 2 * mx, 1, 0, Hx, n, 1, 0, dst, n, 1, 0, 1.f, 0.f);
    }

    int forward(int b) {
        int rc;
        const long long act = (long long)b * C * P;
        // lift: V0[b][C][P] = Pw[C][cin] x[b][cin][P] + pb
        if ((rc = gemm(C, (int)P, cin, b, prm + o_pw, cin, 1, 0, xin, P, 1, (long long)cin * P, V[0], P, 1,
                       (long long)C * P, 1.f, 0.f)))
            return rc;
        hipLaunchKernelGGL(k_bias_act, dim3(grid_for(act)), dim3(256), 0, stream, V[0], V[0], prm + o_pb, C, P, act, 0);
        FNO_CHECK();
        for (int l = 0; l < L; ++l) {
            if ((rc = spectral(l, b, Pre[l]))) return rc;
            // Pre += W V
            if ((rc = gemm(C, (int)P, C, b, prm + o_w[l], C, 1, 0, V[l], P, 1, (long long)C * P, Pre[l], P, 1,
                           (long long)C * P, 1.f, 1.f)))
                return rc;
            hipLaunchKernelGGL(k_bias_act, dim3(grid_for(act)), dim3(256), 0, stream, Pre[l], V[l + 1], prm + o_b[l],
                               C, P, act, l < L - 1 ? 1 : 0);
            FNO_CHECK();
        }
        // projection: Hpre = Q1 V[L] + q1b, Hact = gelu; out = Q2 Hact + q2b
        const long long hact = (long long)b * Hd * P, oact = (long long)b * cout * P;
        if ((rc = gemm(Hd, (int)P, C, b, prm + o_q1w, C, 1, 0, V[L], P, 1, (long long)C * P, Hpre, P, 1,
                       (long long)Hd * P, 1.f, 0.f)))
            return rc;
        hipLaunchKernelGGL(k_bias_act, dim3(grid_for(hact)), dim3(256), 0, stream, Hpre, Hact, prm + o_q1b, Hd, P,
                           hact, 1);
        FNO_CHECK();
        if ((rc = gemm(cout, (int)P, Hd, b, prm + o_q2w, Hd, 1, 0, Hact, P, 1, (long long)Hd * P, out, P, 1,
                       (long long)cout * P, 1.f, 0.f)))
            return rc;
        hipLaunchKernelGGL(k_bias_act, dim3(grid_for(oact)), dim3(256), 0, stream, out, out, prm + o_q2b, cout, P,
                           oact, 0);
        FNO_CHECK();
        return HW_OK;
    }

    // bias gradient: g[c] = sum over batch and points of d[b][c][P]
    int chan_sum(const float* d, int b, int ch, float* g) {
        const int chunks = 64;
        hipLaunchKernelGGL(k_chan_partial, dim3(chunks, ch), dim3(256), 0, stream, d, b, ch, P, chunks, part);
        FNO_CHECK();
        hipLaunchKernelGGL(k_chan_final, dim3((ch + 255) / 256), dim3(256), 0, stream, part, ch, chunks, g);
        FNO_CHECK();
        return HW_OK;
    }

    // weight gradient of y = Wt x (x: [b][K_in][P], dy: [b][M_out][P]): dW[M_out][K_in]
    int wgrad(const float* dy, int mo, const float* x, int ki, int b, float* dW) {
        const int splits = (int)((P + 4095) / 4096);
        return gemm(mo, ki, (int)P, b, dy, P, 1, (long long)mo * P, x, 1, P, (long long)ki * P, dW, ki, 1, 0, 1.f,
                    0.f, splits, true);
    }

    // ----------------------------------------------------------- backward
    // from dout (d loss / d out, b x cout x P) to grd
    int backward(int b) {
        int rc;
        const long long act = (long long)b * C * P, hact = (long long)b * Hd * P;
        // out = Q2 Hact + q2b
        if ((rc = chan_sum(dout, b, cout, grd + o_q2b))) return rc;
        if ((rc = wgrad(dout, cout, Hact, Hd, b, grd + o_q2w))) return rc;
        if ((rc = gemm(Hd, (int)P, cout, b, prm + o_q2w, 1, Hd, 0, dout, P, 1, (long long)cout * P, dH, P, 1,
                       (long long)Hd * P, 1.f, 0.f)))
            return rc;
        hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(hact)), dim3(256), 0, stream, dH, Hpre, hact);
        FNO_CHECK();
        if ((rc = chan_sum(dH, b, Hd, grd + o_q1b))) return rc;
        if ((rc = wgrad(dH, Hd, V[L], C, b, grd + o_q1w))) return rc;
        // dV (d loss / d V[L]) = Q1^T dH
        if ((rc = gemm(C, (int)P, Hd, b, prm + o_q1w, 1, C, 0, dH, P, 1, (long long)Hd * P, dV, P, 1,
                       (long long)C * P, 1.f, 0.f)))
            return rc;
        const long long zb = (long long)n * 2 * mx, hb = 2LL * Mm;
        for (int l = L - 1; l >= 0; --l) {
            // dV holds d/dV[l+1]; through the activation to d/dPre[l] (in place)
            if (l < L - 1) {
                hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(act)), dim3(256), 0, stream, dV, Pre[l], act);
                FNO_CHECK();
            }
            float* dP = dV;
            if ((rc = chan_sum(dP, b, C, grd + o_b[l]))) return rc;
            if ((rc = wgrad(dP, C, V[l], C, b, grd + o_w[l]))) return rc;
            // W path: dV2 = W^T dP
            if ((rc = gemm(C, (int)P, C, b, prm + o_w[l], 1, C, 0, dP, P, 1, (long long)C * P, dV2, P, 1,
                           (long long)C * P, 1.f, 0.f)))
                return rc;
            // spectral path: dYx = dP Hx^T
            if ((rc = gemm(b * C * n, 2 * mx, n, 1, dP, n, 1, 0, Hx, 1, n, 0, Yx, 2 * mx, 1, 0, 1.f, 0.f))) return rc;
            // dYh = A3^T dYx_re + A4^T dYx_im
            if ((rc = gemm(2 * KY, mx, n, b * C, A3, 1, 2 * KY, 0, Yx, 2 * mx, 1, zb, dYh, mx, 1, hb, 1.f, 0.f))) return rc;
            if ((rc = gemm(2 * KY, mx, n, b * C, A4, 1, 2 * KY, 0, Yx + mx, 2 * mx, 1, zb, dYh, mx, 1, hb, 1.f, 1.f)))
                return rc;
            hipLaunchKernelGGL(k_mix_bwd_w, dim3(grid_for((long long)C * C * Mm)), dim3(256), 0, stream, Zh[l], dYh,
                               grd + o_r[l], b, C, Mm);
            FNO_CHECK();
            hipLaunchKernelGGL(k_mix_bwd_data, dim3(grid_for((long long)b * C * Mm)), dim3(256), 0, stream, dYh,
                               prm + o_r[l], dZh, b, C, Mm);
            FNO_CHECK();
            // dZx_re = A1^T dZh, dZx_im = A2^T dZh
            if ((rc = gemm(n, mx, 2 * KY, b * C, A1, 1, n, 0, dZh, mx, 1, hb, Zx, 2 * mx, 1, zb, 1.f, 0.f))) return rc;
            if ((rc = gemm(n, mx, 2 * KY, b * C, A2, 1, n, 0, dZh, mx, 1, hb, Zx + mx, 2 * mx, 1, zb, 1.f, 0.f))) return rc;
            // dV2 += dZx Fx^T
            if ((rc = gemm(b * C * n, n, 2 * mx, 1, Zx, 2 * mx, 1, 0, Fx, 1, 2 * mx, 0, dV2, n, 1, 0, 1.f, 1.f)))
                return rc;
            std::swap(dV, dV2);
        }
        // lift: V0 = Pw x + pb (no activation)
        if ((rc = chan_sum(dV, b, C, grd + o_pb))) return rc;
        return wgrad(dV, C, xin, cin, b, grd + o_pw);
    }

    int loss(int b, double* lossv, bool with_grad) {
        const long long tot = (long long)b * cout * P;
        const int blocks = (int)grid_for(tot, 256, 1024);
        hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, stream, out, tgt, tot, with_grad ? dout : nullptr, part);
        FNO_CHECK();
        std::vector<float> h(blocks);
        FNO_TRY(hipMemcpyAsync(h.data(), part, sizeof(float) * blocks, hipMemcpyDeviceToHost, stream));
        FNO_TRY(hipStreamSynchronize(stream));
        double s = 0.0;
        for (float v : h) s += v;
        *lossv = s / (double)tot;
        return HW_OK;
    }

    int adam(double lr) {
        adam_t += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 - std::pow((double)b1, (double)adam_t));
        const float c2 = (float)(1.0 - std::pow((double)b2, (double)adam_t));
        hipLaunchKernelGGL(k_adam, dim3(grid_for(nparam)), dim3(256), 0, stream, prm, grd, am, av, nparam, (float)lr,
                           b1, b2, 1e-8f, c1, c2);
        FNO_CHECK();
        return HW_OK;
    }
};

}  // namespace fno
}  // namespace hw

using hw::fno::Net;

extern "C" {

int hw_fno_create(const hw_fno_config* cfg, void** ctx) {
    if (!cfg || !ctx) { hw::set_error("null argument"); return HW_ERR_ARG; }
    *ctx = nullptr;
    Net* net = new Net();
    int rc = net->init(cfg);
    if (rc) { delete net; return rc; }
    *ctx = net;
    return HW_OK;
}

void hw_fno_destroy(void* ctx) { delete static_cast<Net*>(ctx); }

long long hw_fno_param_count(void* ctx) { return ctx ? static_cast<Net*>(ctx)->nparam : -1; }

int hw_fno_get_params(void* ctx, float* out) {
    Net* s = static_cast<Net*>(ctx);
    if (!s || !out) { hw::set_error("null argument"); return HW_ERR_ARG; }
    FNO_TRY(hipMemcpyAsync(out, s->prm, sizeof(float) * s->nparam, hipMemcpyDeviceToHost, s->stream));
    FNO_TRY(hipStreamSynchronize(s->stream));
    return HW_OK;
}

int hw_fno_set_params(void* ctx, const float* in) {
    Net* s = static_cast<Net*>(ctx);
    if (!s || !in) { hw::set_error("null argument"); return HW_ERR_ARG; }
    FNO_TRY(hipMemcpyAsync(s->prm, in, sizeof(float) * s->nparam, hipMemcpyHostToDevice, s->stream));
    FNO_TRY(hipStreamSynchronize(s->stream));
    s->adam_t = 0;
    FNO_TRY(hipMemsetAsync(s->am, 0, sizeof(float) * s->nparam, s->stream));
    FNO_TRY(hipMemsetAsync(s->av, 0, sizeof(float) * s->nparam, s->stream));
    return HW_OK;
}

int hw_fno_forward(void* ctx, int batch, const float* x, float* y) {
    Net* s = static_cast<Net*>(ctx);
    if (!s || !x || !y) { hw::set_error("null argument"); return HW_ERR_ARG; }
    if (batch < 1 || batch > s->maxb) { hw::set_error("fno: batch %d outside 1..%d", batch, s->maxb); return HW_ERR_ARG; }
    FNO_TRY(hipMemcpyAsync(s->xin, x, sizeof(float) * batch * s->cin * s->P, hipMemcpyHostToDevice, s->stream));
    int rc = s->forward(batch);
    if (rc) return rc;
    FNO_TRY(hipMemcpyAsync(y, s->out, sizeof(float) * batch * s->cout * s->P, hipMemcpyDeviceToHost, s->stream));
    FNO_TRY(hipStreamSynchronize(s->stream));
    return HW_OK;
}

int hw_fno_loss_grad(void* ctx, int batch, const float* x, const float* target, double* loss, float* grad) {
    Net* s = static_cast<Net*>(ctx);
    if (!s || !x || !target || !loss) { hw::set_error("null argument"); return HW_ERR_ARG; }
    if (batch < 1 || batch > s->maxb) { hw::set_error("fno: batch %d outside 1..%d", batch, s->maxb); return HW_ERR_ARG; }
    FNO_TRY(hipMemcpyAsync(s->xin, x, sizeof(float) * batch * s->cin * s->P, hipMemcpyHostToDevice, s->stream));
    FNO_TRY(hipMemcpyAsync(s->tgt, target, sizeof(float) * batch * s->cout * s->P, hipMemcpyHostToDevice, s->stream));
    int rc = s->forward(batch);
    if (rc || (rc = s->loss(batch, loss, true)) || (rc = s->backward(batch))) return rc;
    if (grad) {
        FNO_TRY(hipMemcpyAsync(grad, s->grd, sizeof(float) * s->nparam, hipMemcpyDeviceToHost, s->stream));
    }
    FNO_TRY(hipStreamSynchronize(s->stream));
    return HW_OK;
}

int hw_fno_set_dataset(void* ctx, int count, const float* x, const float* y) {
    Net* s = static_cast<Net*>(ctx);
    if (!s || !x || !y || count < 1) { hw::set_error("null argument or empty dataset"); return HW_ERR_ARG; }
    if (s->dsx) { (void)hipFree(s->dsx); (void)hipFree(s->dsy); s->dsx = s->dsy = nullptr; }
    const size_t nx = (size_t)count * s->cin * s->P, ny = (size_t)count * s->cout * s->P;
    FNO_TRY(hipMalloc((void**)&s->dsx, sizeof(float) * nx));
    FNO_TRY(hipMalloc((void**)&s->dsy, sizeof(float) * ny));
    FNO_TRY(hipMemcpyAsync(s->dsx, x, sizeof(float) * nx, hipMemcpyHostToDevice, s->stream));
    FNO_TRY(hipMemcpyAsync(s->dsy, y, sizeof(float) * ny, hipMemcpyHostToDevice, s->stream));
    FNO_TRY(hipStreamSynchronize(s->stream));
    s->ds_count = count;
    return HW_OK;
}

int hw_fno_train_step(void* ctx, int batch, const int* idx, double lr, double* loss) {
    Net* s = static_cast<Net*>(ctx);
    if (!s || !idx || !loss) { hw::set_error("null argument"); return HW_ERR_ARG; }
    if (!s->dsx) { hw::set_error("fno: no dataset (hw_fno_set_dataset)"); return HW_ERR_ARG; }
    if (batch < 1 || batch > s->maxb) { hw::set_error("fno: batch %d outside 1..%d", batch, s->maxb); return HW_ERR_ARG; }
    for (int i = 0; i < batch; ++i)
        if (idx[i] < 0 || idx[i] >= s->ds_count) { hw::set_error("fno: sample index %d out of range", idx[i]); return HW_ERR_ARG; }
    FNO_TRY(hipMemcpyAsync(s->d_idx, idx, sizeof(int) * batch, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(hw::fno::k_gather, dim3(hw::fno::grid_for((long long)batch * s->cin * s->P)), dim3(256), 0,
                       s->stream, s->dsx, s->d_idx, batch, (long long)s->cin * s->P, s->xin);
    hipLaunchKernelGGL(hw::fno::k_gather, dim3(hw::fno::grid_for((long long)batch * s->cout * s->P)), dim3(256), 0,
                       s->stream, s->dsy, s->d_idx, batch, (long long)s->cout * s->P, s->tgt);
    FNO_CHECK();
    int rc = s->forward(batch);
    if (rc || (rc = s->loss(batch, loss, true)) || (rc = s->backward(batch)) || (rc = s->adam(lr))) return rc;
    return HW_OK;
}

int hw_fno_set_precision(void* ctx, int precision) {
    if (!ctx || (precision != HW_FNO_BF16 && precision != HW_FNO_SPLIT)) {
        hw::set_error("hw_fno_set_precision: bad argument");
        return HW_ERR_ARG;
    }
    static_cast<Net*>(ctx)->split_precision = precision == HW_FNO_SPLIT;
    return HW_OK;
}

double hw_fno_gemm_flops(void* ctx) { return ctx ? static_cast<Net*>(ctx)->gemm_flops : -1.0; }
double hw_fno_gemm_bytes(void* ctx) { return ctx ? static_cast<Net*>(ctx)->gemm_bytes : -1.0; }

int hw_fno_gemm_check(int M, int N, int K, const float* A, const float* B, float* C, int precision) {
    // one plain product C[M][N] = A[M][K] B[K][N] through the MFMA GEMM
    // (row-major host arrays), for the kernel's own tests
    if (!A || !B || !C || M < 1 || N < 1 || K < 1 || (precision != HW_FNO_BF16 && precision != HW_FNO_SPLIT)) {
        hw::set_error("bad argument");
        return HW_ERR_ARG;
    }
    Net net;
    net.split_precision = precision == HW_FNO_SPLIT;
    FNO_TRY(hipStreamCreateWithFlags(&net.stream, hipStreamNonBlocking));
    float *dA, *dB, *dC;
    int rc;
    if ((rc = net.dalloc(&dA, (size_t)M * K)) || (rc = net.dalloc(&dB, (size_t)K * N)) ||
        (rc = net.dalloc(&dC, (size_t)M * N)) || (rc = net.dalloc(&net.ws, (size_t)8 * M * N)))
        return rc;
    net.ws_floats = (size_t)8 * M * N;
    FNO_TRY(hipMemcpyAsync(dA, A, sizeof(float) * M * K, hipMemcpyHostToDevice, net.stream));
    FNO_TRY(hipMemcpyAsync(dB, B, sizeof(float) * K * N, hipMemcpyHostToDevice, net.stream));
    if ((rc = net.gemm(M, N, K, 1, dA, K, 1, 0, dB, N, 1, 0, dC, N, 1, 0, 1.f, 0.f, K > 4096 ? 4 : 1))) return rc;
    FNO_TRY(hipMemcpyAsync(C, dC, sizeof(float) * M * N, hipMemcpyDeviceToHost, net.stream));
    FNO_TRY(hipStreamSynchronize(net.stream));
    return HW_OK;
}

}  // extern "C"
