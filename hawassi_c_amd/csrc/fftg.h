// fftg.h -- full complex 2-D transforms of any grid whose sides factor into
// 2, 3 and 5, for the full-grid path (solver2dg.h).
//
// The fused half-spectrum pipeline (rhs2d_kernels.h) is built around
// register-resident radix-2^k transforms of compile-time lengths.  The
// reference's grids are not all powers of two (300x250, 600x300, 2000x2), and
// FFTW plans any length (2d/fun_odesolver.c:2885-2955), so this file gives the
// full-grid path transforms of runtime lengths: an LDS-resident Stockham
// autosort with radix 8, 4, 2, 5 and 3 passes, two LDS buffers (read one,
// write the other, one barrier per pass), twiddles from a table of
// exp(-2 pi i m / L) built on the host in FP64.
//   rows:    a block owns B whole rows of [ny][nx] (coalesced loads/stores);
//   columns: a block owns CW adjacent columns, loaded row by row (CW
//            consecutive elements per row), transformed in LDS, stored back.
// Conventions are FFTW's: forward exp(-i..), backward exp(+i..), both
// unnormalised.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>

#include "hw_fft.h"

namespace hw {

// ---------------------------------------------------------------- plan ----
struct PlanG {
    int L = 0;           // length
    int np = 0;          // passes
    int radix[24] = {};
    int ns[24] = {};     // product of the radices before pass p
};

// radix-8/4/2 passes first, then 5 and 3; returns false when L has another
// prime factor
inline bool plan_g(int L, PlanG& p) {
    p = PlanG{};
    p.L = L;
    if (L < 1) return false;
    int m = L, ns = 1;
    auto push = [&](int r) { p.radix[p.np] = r; p.ns[p.np] = ns; ns *= r; m /= r; p.np++; };
    while (m % 8 == 0 && p.np < 24) push(8);
    while (m % 4 == 0 && p.np < 24) push(4);
    while (m % 2 == 0 && p.np < 24) push(2);
    while (m % 5 == 0 && p.np < 24) push(5);
    while (m % 3 == 0 && p.np < 24) push(3);
    return m == 1;
}

// exp(-2 pi i m / L), m = 0..L-1, rounded once from FP64
template <typename C> std::vector<C> twiddles_g(int L) {
    std::vector<C> t(L);
    for (int m = 0; m < L; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)L;
        t[m].x = (decltype(t[m].x))std::cos(a);
        t[m].y = (decltype(t[m].y))std::sin(a);
    }
    return t;
}

// (the 3- and 5-point DFTs dft3g, dft5g are in hw_fft.h)

// one radix-R butterfly of pass (R, NS) of B sequences of length L stored at
// src[b * LS + n]; writes dst in Stockham order
template <int R, int DIR, typename C>
__device__ __forceinline__ void bfly_g(const C* __restrict__ src, C* __restrict__ dst, int LS, int L, int NS, int bf,
                                       const C* __restrict__ tw) {
    const int M = L / R;
    const int b = bf / M, j = bf - b * M, k = j % NS;
    const C* s = src + (size_t)b * LS;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * M];
    if (NS > 1) {
        const int step = L / (NS * R);
#pragma unroll
        for (int r = 1; r < R; ++r) {
            C w = tw[r * k * step];
            if (DIR > 0) w = cconj(w);
            v[r] = cmul(v[r], w);
        }
    }
    if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v);
    else if constexpr (R == 2) dft2<DIR>(v);
    else if constexpr (R == 5) dft5g<DIR>(v);
    else dft3g<DIR>(v);
    C* d = dst + (size_t)b * LS + (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * NS] = v[r];
}

// B sequences in bufA (stride LS); returns the buffer holding the result.
// Every thread of the block calls it; it starts and ends with a barrier.
template <int DIR, typename C>
__device__ C* lds_fft_g(C* bufA, C* bufB, int LS, int B, const PlanG& P, const C* __restrict__ tw) {
    C* src = bufA;
    C* dst = bufB;
    __syncthreads();
    for (int p = 0; p < P.np; ++p) {
        const int R = P.radix[p], NS = P.ns[p];
        const int nb = B * (P.L / R);
        for (int bf = threadIdx.x; bf < nb; bf += blockDim.x) {
            switch (R) {
                case 8: bfly_g<8, DIR>(src, dst, LS, P.L, NS, bf, tw); break;
                case 4: bfly_g<4, DIR>(src, dst, LS, P.L, NS, bf, tw); break;
                case 2: bfly_g<2, DIR>(src, dst, LS, P.L, NS, bf, tw); break;
                case 5: bfly_g<5, DIR>(src, dst, LS, P.L, NS, bf, tw); break;
                default: bfly_g<3, DIR>(src, dst, LS, P.L, NS, bf, tw); break;
            }
        }
        __syncthreads();
        C* t = src; src = dst; dst = t;
    }
    return src;
}

// ------------------------------------------------------------- kernels ----
// B rows per block; dynamic LDS 2 B nx elements
template <typename C, int DIR>
__global__ void __launch_bounds__(256) k_fftg_rows(C* __restrict__ a, int nx, int ny, int B, PlanG P,
                                                   const C* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* bufA = reinterpret_cast<C*>(lds_raw);
    C* bufB = bufA + (size_t)B * nx;
    const int r0 = blockIdx.x * B;
    const int nr = min(B, ny - r0);
    const size_t n = (size_t)nr * nx;
    C* g = a + (size_t)r0 * nx;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) bufA[i] = g[i];
    C* res = lds_fft_g<DIR>(bufA, bufB, nx, nr, P, tw);
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) g[i] = res[i];
}

// CW columns per block; dynamic LDS 2 CW ny elements, column c at c * ny
template <typename C, int DIR>
__global__ void __launch_bounds__(256) k_fftg_cols(C* __restrict__ a, int nx, int ny, int CW, PlanG P,
                                                   const C* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* bufA = reinterpret_cast<C*>(lds_raw);
    C* bufB = bufA + (size_t)CW * ny;
    const int c0 = blockIdx.x * CW;
    const int nc = min(CW, nx - c0);
    const int n = nc * ny;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int row = i / nc, c = i - row * nc;
        bufA[c * ny + row] = a[(size_t)row * nx + c0 + c];
    }
    C* res = lds_fft_g<DIR>(bufA, bufB, ny, nc, P, tw);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int row = i / nc, c = i - row * nc;
        a[(size_t)row * nx + c0 + c] = res[c * ny + row];
    }
}

// fftwcomplex_2d_sym (2d/functions1.c:492-521) in place on [ny][nx]: Im = 0
// at the four self-conjugate corners, row 0 and row ny/2 mirrored from their
// left halves, column 0 from its top half, and every column j > nx/2 of rows
// i >= 1 overwritten by conj(ny - i, nx - j).  Column nx/2 is kept.  The
// elements written and the elements read are disjoint sets, so one thread
// per element suffices.
template <typename C>
__global__ void k_proj_full(C* __restrict__ a, int nx, int ny) {
    const size_t n = (size_t)nx * ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q / nx), j = (int)(q - (size_t)i * nx);
        const int hx = nx / 2, hy = ny / 2;
        if ((i == 0 || i == hy) && (j == 0 || j == hx)) {
            a[q].y = 0;
        } else if ((i == 0 || i == hy) && j > hx) {
            const C s = a[(size_t)i * nx + (nx - j)];
            a[q] = cmk<C>(s.x, -s.y);
        } else if (j == 0 && i > hy) {
            const C s = a[(size_t)(ny - i) * nx];
            a[q] = cmk<C>(s.x, -s.y);
        } else if (i >= 1 && j > hx) {
            const C s = a[(size_t)(ny - i) * nx + (nx - j)];
            a[q] = cmk<C>(s.x, -s.y);
        }
    }
}

// f(q) over q = 0..n-1 (grid-stride); F is a device lambda
template <typename F>
__global__ void __launch_bounds__(256) k_map(size_t n, F f) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) f(q);
}

// ------------------------------------------------------------ host side ---
// A 2-D transform of [ny][nx] complex arrays of type C on one stream.
template <typename C> struct Fft2G {
    int nx = 0, ny = 0, B = 1, CW = 1;
    size_t lds_rows = 0, lds_cols = 0;
    PlanG px, py;
    C *twx = nullptr, *twy = nullptr;   // device tables (owned)

    ~Fft2G() {
        if (twx) (void)hipFree(twx);
        if (twy) (void)hipFree(twy);
    }
    // maximum dynamic LDS a block may ask for on gfx950
    static constexpr size_t LDS_MAX = 160 * 1024;

    // returns a message, or nullptr on success
    const char* init(int nx_, int ny_, hipStream_t s) {
        nx = nx_; ny = ny_;
        if (!plan_g(nx, px) || !plan_g(ny, py)) return "grid sides must factor into 2, 3 and 5";
        // rows per block: about 4096 elements, LDS for two buffers
        B = std::max(1, 4096 / nx);
        B = std::min(B, ny);
        while (B > 1 && 2 * (size_t)B * nx * sizeof(C) > 64 * 1024) --B;
        lds_rows = 2 * (size_t)B * nx * sizeof(C);
        CW = std::max(1, std::min(16, 4096 / ny));
        CW = std::min(CW, nx);
        while (CW > 1 && 2 * (size_t)CW * ny * sizeof(C) > 64 * 1024) --CW;
        lds_cols = 2 * (size_t)CW * ny * sizeof(C);
        if (lds_rows > LDS_MAX || lds_cols > LDS_MAX) return "grid side too long for the LDS transform (at most 4096 FP64, 8192 FP32)";
        std::vector<C> tx = twiddles_g<C>(nx), ty = twiddles_g<C>(ny);
        if (hipMalloc((void**)&twx, sizeof(C) * nx) != hipSuccess) return "hipMalloc failed";
        if (hipMalloc((void**)&twy, sizeof(C) * ny) != hipSuccess) return "hipMalloc failed";
        if (hipMemcpyAsync(twx, tx.data(), sizeof(C) * nx, hipMemcpyHostToDevice, s) != hipSuccess) return "copy failed";
        if (hipMemcpyAsync(twy, ty.data(), sizeof(C) * ny, hipMemcpyHostToDevice, s) != hipSuccess) return "copy failed";
        if (hipStreamSynchronize(s) != hipSuccess) return "stream synchronisation failed";
        // blocks above 64 KB of LDS (sides beyond 2048 FP64 / 4096 FP32) need
        // the kernels' limit raised, once, to the largest plan's need
        static size_t raised = 64 * 1024;
        const size_t need = std::max(lds_rows, lds_cols);
        if (need > raised) {
            const void* ks[4] = {(const void*)k_fftg_rows<C, -1>, (const void*)k_fftg_rows<C, 1>,
                                 (const void*)k_fftg_cols<C, -1>, (const void*)k_fftg_cols<C, 1>};
            for (const void* k : ks)
                if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need) != hipSuccess)
                    return "hipFuncSetAttribute (dynamic LDS) failed";
            raised = need;
        }
        return nullptr;
    }

    // in-place 2-D transform; dir = -1 forward, +1 backward
    void run(C* a, int dir, hipStream_t s) const {
        const dim3 gr((ny + B - 1) / B), gc((nx + CW - 1) / CW), blk(256);
        if (dir < 0) {
            k_fftg_rows<C, -1><<<gr, blk, lds_rows, s>>>(a, nx, ny, B, px, twx);
            k_fftg_cols<C, -1><<<gc, blk, lds_cols, s>>>(a, nx, ny, CW, py, twy);
        } else {
            k_fftg_rows<C, 1><<<gr, blk, lds_rows, s>>>(a, nx, ny, B, px, twx);
            k_fftg_cols<C, 1><<<gc, blk, lds_cols, s>>>(a, nx, ny, CW, py, twy);
        }
    }
};

}  // namespace hw
