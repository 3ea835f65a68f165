// hw_fft.h -- LDS-resident Stockham FFT building blocks for gfx950 (CDNA4).
//
// Layouts used by every 2-D kernel of the library:
//   physical field   P[y][x]              (x fastest, the reference's i*Nx+j)
//   half spectrum    S[kx][ky]            kx = 0..NX/2, ky fastest ("column major")
//   x-intermediate   X[kx][y]             after the x-direction real transform
// A 2-D forward transform is an x-pass (real FFT of each row, written
// transposed as X[kx][y]) followed by a y-pass (complex FFT of each contiguous
// column).  The inverse runs the same two passes backwards.  Forward transforms
// are unnormalised (exp(-i..)); inverse transforms are exp(+i..) and the
// library divides by NX*NY at the end of the x-pass so physical fields come out
// in physical units.
//
// The FFT inside a pass is a mixed radix-8/4/2 Stockham autosort working in
// LDS: each radix pass reads its butterfly inputs into registers, synchronises,
// applies twiddles (from an FP64-accurate table) and the radix-R DFT, and
// writes the outputs in natural order for the next pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hw {

template <typename T> struct ctype;
template <> struct ctype<float> { using type = float2; };
template <> struct ctype<double> { using type = double2; };

template <typename T> using cplx = typename ctype<T>::type;
template <typename T, int N> using carr = cplx<T>[N];

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) { C r; r.x = a; r.y = b; return r; }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) { return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
// multiply by i (DIR=+1) or -i (DIR=-1)
template <int DIR, typename C> __device__ __forceinline__ C cmul_i(C a) {
    return DIR > 0 ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x);
}

// ------------------------------------------------------------- radix DFTs --
template <int DIR, typename C> __device__ __forceinline__ void dft2(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR, typename C> __device__ __forceinline__ void dft4(C* v) {
    C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    C a2 = cadd(v[1], v[3]), a3 = cmul_i<DIR>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <int DIR, typename C> __device__ __forceinline__ void dft8(C* v) {
    using R = decltype(C::x);
    const R h = (R)0.70710678118654752440084436210485;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4<DIR>(e);
    dft4<DIR>(o);
    // W8^k = exp(DIR * 2 pi i k / 8) with DIR=-1 forward
    C w1 = cmk<C>(h * (o[1].x - (R)DIR * o[1].y) , h * (o[1].y + (R)DIR * o[1].x));
    C w2 = cmul_i<DIR>(o[2]);
    C w3 = cmk<C>(h * (-o[3].x - (R)DIR * o[3].y), h * (-o[3].y + (R)DIR * o[3].x));
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);   v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);   v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);   v[7] = csub(e[3], w3);
}

// 16-point DFT as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2; inner twiddles
// W16^(n2 k1), W16 = exp(DIR 2 pi i / 16)
template <int DIR, typename C> __device__ __forceinline__ void dft16(C* v) {
    using R = decltype(C::x);
    constexpr double cs[10] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0,
                               -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613, -1.0,
                               -0.92387953251128675613};
    constexpr double sn[10] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613, 1.0,
                               0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0,
                               -0.38268343236508977173};
    C a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        C t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft4<DIR>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            const int m = n2 * k1;
            if (m == 0) a[n2][k1] = t[k1];
            else a[n2][k1] = cmul(t[k1], cmk<C>((R)cs[m], (R)((double)DIR * sn[m])));
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft4<DIR>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dftR(C* v) {
    if constexpr (R == 16) dft16<DIR>(v);
    else if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v);
    else dft2<DIR>(v);
}

#ifndef HW_MAX_RADIX
#define HW_MAX_RADIX 8
#endif
constexpr int pass_radix(int rem) {
    return (HW_MAX_RADIX >= 16 && rem % 16 == 0) ? 16 : (rem % 8 == 0 ? 8 : (rem % 4 == 0 ? 4 : 2));
}

// ------------------------------------------------------- Stockham in LDS --
// B sequences of length L live at lds[b*LS + n].  tw[m*TS] = exp(-2 pi i m/L).
// All NT threads of the block must call it.  Ends with a barrier.
template <typename C, int L, int B, int NT, int DIR, int NS>
__device__ __forceinline__ void fft_lds_pass(C* lds, int LS, const C* __restrict__ tw, int TS, int tid) {
    if constexpr (NS < L) {
        constexpr int R = pass_radix(L / NS);
        constexpr int M = L / R;             // butterflies per sequence
        constexpr int NB = B * M;
        constexpr int PER = (NB + NT - 1) / NT;
        C v[PER][R];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            int bf = tid + p * NT;
            if (NB % NT == 0 || bf < NB) {
                int b = bf / M, j = bf % M;
#pragma unroll
                for (int r = 0; r < R; ++r) v[p][r] = lds[b * LS + j + r * M];
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            int bf = tid + p * NT;
            if (NB % NT == 0 || bf < NB) {
                int b = bf / M, j = bf % M;
                int k = j % NS;
                if constexpr (NS > 1) {
                    const int ex = k * (L / (NS * R));
                    if constexpr (R <= 4) {
#pragma unroll
                        for (int r = 1; r < R; ++r) {
                            C w = tw[(r * ex) * TS];
                            if (DIR > 0) w = cconj(w);
                            v[p][r] = cmul(v[p][r], w);
                        }
                    } else {
                        // w^r from w, w^4 (and w^8) read from the table:
                        // at most two roundings more than a table entry
                        C wp[R];
                        C w1 = tw[ex * TS], w4 = tw[4 * ex * TS];
                        if (DIR > 0) { w1 = cconj(w1); w4 = cconj(w4); }
                        wp[1] = w1;
                        wp[2] = cmul(w1, w1);
                        wp[3] = cmul(wp[2], w1);
                        wp[4] = w4;
                        wp[5] = cmul(w4, w1);
                        wp[6] = cmul(w4, wp[2]);
                        wp[7] = cmul(w4, wp[3]);
                        if constexpr (R == 16) {
                            C w8 = tw[8 * ex * TS];
                            if (DIR > 0) w8 = cconj(w8);
                            wp[8] = w8;
#pragma unroll
                            for (int r = 1; r < 8; ++r) wp[8 + r] = cmul(w8, wp[r]);
                        }
#pragma unroll
                        for (int r = 1; r < R; ++r) v[p][r] = cmul(v[p][r], wp[r]);
                    }
                }
                dftR<R, DIR>(v[p]);
                int idxD = (j / NS) * NS * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[b * LS + idxD + r * NS] = v[p][r];
            }
        }
        __syncthreads();
        fft_lds_pass<C, L, B, NT, DIR, NS * R>(lds, LS, tw, TS, tid);
    }
}

template <typename C, int L, int B, int NT, int DIR>
__device__ __forceinline__ void fft_lds(C* lds, int LS, const C* __restrict__ tw, int TS, int tid) {
    fft_lds_pass<C, L, B, NT, DIR, 1>(lds, LS, tw, TS, tid);
}

// ------------------------------------------------------- pass geometry ----
// Elements each thread holds between transforms: fewer elements per thread
// means fewer live registers and more waves in flight per CU.
#ifndef HW_X_RUN_BYTES
#define HW_X_RUN_BYTES 16
#endif
#ifndef HW_X_PER_THREAD
#define HW_X_PER_THREAD 8
#endif
#ifndef HW_Y_PER_THREAD
#define HW_Y_PER_THREAD 4
#endif
// x-pass: RR rows per block so that an RR-element run of one spectral column
// (the transposed write/read) is 32 bytes long; blocks are remapped so that
// the row groups sharing a 128-byte line run on the same XCD (same L2).
template <typename T, int NX> struct XGeo {
    using C = typename ctype<T>::type;
    static constexpr int RR = HW_X_RUN_BYTES / (int)sizeof(C);   // rows per block: 32-byte runs
    static_assert(RR >= 1, "run shorter than one element");
    static constexpr int L = NX / 2;                    // half-length complex FFT
    static constexpr int LS = L + 2;                    // LDS row stride (complex)
    static constexpr int NPAIR = RR * L;                // physical (x,x+1) pairs per block
    static constexpr int NT0 = NPAIR / HW_X_PER_THREAD;
    static constexpr int NT = NT0 > 1024 ? 1024 : (NT0 < 64 ? 64 : NT0);
    static constexpr int PP = (NPAIR + NT - 1) / NT;    // pairs per thread
    static constexpr int NSPEC = RR * (L + 1);          // spectral values per block
    static constexpr int PS = (NSPEC + NT - 1) / NT;
    static constexpr int LDS_ELEMS = RR * LS;
};

// y-pass: CC whole columns per block.
template <typename T, int NY> struct YGeo {
    using C = typename ctype<T>::type;
    static constexpr int CC = NY >= 1024 ? 1 : 1024 / NY;
    static constexpr int LS = NY;
    static constexpr int NE = CC * NY;
    static constexpr int NT0 = NE / HW_Y_PER_THREAD;
    static constexpr int NT = NT0 > 1024 ? 1024 : (NT0 < 64 ? 64 : NT0);
    static constexpr int PE = (NE + NT - 1) / NT;       // elements per thread
    static constexpr int LDS_ELEMS = CC * LS;
};

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.  This
// maps the hardware block index so that XCD x gets a contiguous range of
// logical blocks.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    if (nb % 8) return b;
    int per = nb / 8;
    return (b % 8) * per + b / 8;
}

// Twiddle tables handed to kernels (device pointers):
//   twx[m] = exp(-2 pi i m / NX), m < NX    (x-pass: the L=NX/2 FFT uses stride 2)
//   twy[m] = exp(-2 pi i m / NY), m < NY
template <typename C> struct Tw { const C* x; const C* y; };

// ----------------------------------------------------------- x-pass ops ---
// Inverse x-pass: spectral X[kx][y0..y0+RR) (kx = 0..NX/2) -> physical rows,
// delivered as (x[2n], x[2n+1]) pairs in registers, scaled by `scale`.
// Pair q of this thread is pos = tid + q*NT: row = pos / L, n = pos % L.
// Also returns, per row of the block, Im of the Nyquist bin times `scale`
// in nyq_im[row] (LDS scalar array of RR entries), which the adjust operator
// needs for the reference's complex inverse.
// The inverse x-pass in two halves so a kernel can issue the (transposed,
// 32-byte-run) global loads of the next field before running the FFT of the
// current one: x_spec_load only issues loads into registers; x_inverse_b
// consumes them.
template <typename T, int NX> using XSpec = carr<T, XGeo<T, NX>::PS>;

template <typename T, int NX, int NY>
__device__ __forceinline__ void x_spec_load(const typename ctype<T>::type* __restrict__ X, int y0, XSpec<T, NX>& b,
                                            int tid) {
    using G = XGeo<T, NX>;
#pragma unroll
    for (int q = 0; q < G::PS; ++q) {
        int idx = tid + q * G::NT;
        if (G::NSPEC % G::NT == 0 || idx < G::NSPEC) {
            int kx = idx / G::RR, r = idx % G::RR;
            b[q] = X[(size_t)kx * NY + y0 + r];
        }
    }
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void x_inverse_b(const XSpec<T, NX>& b, int y0,
                                            typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twx,
                                            carr<T, XGeo<T, NX>::PP>& out, T scale, T* nyq_im, int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    constexpr int L = G::L, RR = G::RR, LS = G::LS, NT = G::NT;
    (void)y0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < G::PS; ++q) {
        int idx = tid + q * NT;
        if (G::NSPEC % NT == 0 || idx < G::NSPEC) {
            int kx = idx / RR, r = idx % RR;
            lds[r * LS + kx] = b[q];
        }
    }
    __syncthreads();
    if (nyq_im && tid < RR) nyq_im[tid] = lds[tid * LS + L].y * scale;
    // pre-process pairs (k, L-k): Z[k] = (X[k] + conj X[L-k]) + i (X[k] - conj X[L-k]) W^-k
    constexpr int NPR = RR * (L / 2 + 1);
    for (int idx = tid; idx < NPR; idx += NT) {
        int r = idx / (L / 2 + 1), k = idx % (L / 2 + 1);
        C* row = lds + r * LS;
        if (k == 0) {
            T a = row[0].x, b = row[L].x;      // imaginary parts of DC / Nyquist ignored
            T e = a + b, o = a - b;
            row[0] = cmk<C>(e, o);
        } else {
            C xa = row[k], xb = row[L - k];
            C wa = cconj(twx[k]);              // W^-k = exp(+2 pi i k / NX)
            C ea = cadd(xa, cconj(xb));
            C oa = cmul(csub(xa, cconj(xb)), wa);
            C za = cadd(ea, cmul_i<1>(oa));
            if (k != L - k) {
                C wb = cconj(twx[L - k]);
                C eb = cadd(xb, cconj(xa));
                C ob = cmul(csub(xb, cconj(xa)), wb);
                row[L - k] = cadd(eb, cmul_i<1>(ob));
            }
            row[k] = za;
        }
    }
    __syncthreads();
    fft_lds<C, L, RR, NT, +1>(lds, LS, twx, 2, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        int pos = tid + q * NT;
        if (G::NPAIR % NT == 0 || pos < G::NPAIR) {
            int r = pos / L, n = pos % L;
            out[q] = cscale(lds[r * LS + n], scale);
        }
    }
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void x_inverse(const typename ctype<T>::type* __restrict__ X, int y0,
                                          typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twx,
                                          carr<T, XGeo<T, NX>::PP>& out, T scale, T* nyq_im, int tid) {
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(X, y0, b, tid);
    x_inverse_b<T, NX, NY>(b, y0, lds, twx, out, scale, nyq_im, tid);
}

// Forward x-pass of a real field given as (x[2n], x[2n+1]) pairs in registers.
// Leaves X[k], k = 0..L, of every row in LDS (row r at lds[r*LS + k]).
template <typename T, int NX>
__device__ __forceinline__ void x_forward_lds(const carr<T, XGeo<T, NX>::PP>& in,
                                              typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twx,
                                              int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    constexpr int L = G::L, RR = G::RR, LS = G::LS, NT = G::NT;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        int pos = tid + q * NT;
        if (G::NPAIR % NT == 0 || pos < G::NPAIR) {
            int r = pos / L, n = pos % L;
            lds[r * LS + n] = in[q];
        }
    }
    __syncthreads();
    fft_lds<C, L, RR, NT, -1>(lds, LS, twx, 2, tid);
    // post-process pairs: X[k] = (Z[k] + conj Z[L-k])/2 - i/2 W^k (Z[k] - conj Z[L-k])
    constexpr int NPR = RR * (L / 2 + 1);
    for (int idx = tid; idx < NPR; idx += NT) {
        int r = idx / (L / 2 + 1), k = idx % (L / 2 + 1);
        C* row = lds + r * LS;
        if (k == 0) {
            C z = row[0];
            row[0] = cmk<C>(z.x + z.y, (T)0);
            row[L] = cmk<C>(z.x - z.y, (T)0);
        } else {
            C za = row[k], zb = row[L - k];
            C ea = cscale(cadd(za, cconj(zb)), (T)0.5);
            C da = cscale(csub(za, cconj(zb)), (T)0.5);
            C xa = cadd(ea, cmul_i<-1>(cmul(twx[k], da)));
            if (k != L - k) {
                C eb = cscale(cadd(zb, cconj(za)), (T)0.5);
                C db = cscale(csub(zb, cconj(za)), (T)0.5);
                row[L - k] = cadd(eb, cmul_i<-1>(cmul(twx[L - k], db)));
            }
            row[k] = xa;
        }
    }
    __syncthreads();
}

// Transposed store of X[k] (k = 0..L) of the RR rows in LDS to X[kx][y0+r].
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_store_spec(typename ctype<T>::type* __restrict__ X, int y0,
                                             const typename ctype<T>::type* lds, int tid) {
    using G = XGeo<T, NX>;
    constexpr int RR = G::RR, LS = G::LS, NT = G::NT;
#pragma unroll
    for (int q = 0; q < G::PS; ++q) {
        int idx = tid + q * NT;
        if (G::NSPEC % NT == 0 || idx < G::NSPEC) {
            int kx = idx / RR, r = idx % RR;
            X[(size_t)kx * NY + y0 + r] = lds[r * LS + kx];
        }
    }
}

// Read the spectral values of the block (same ownership as x_store_spec) from LDS.
template <typename T, int NX>
__device__ __forceinline__ void x_spec_to_regs(carr<T, XGeo<T, NX>::PS>& v,
                                               const typename ctype<T>::type* lds, int tid) {
    using G = XGeo<T, NX>;
    constexpr int RR = G::RR, LS = G::LS, NT = G::NT;
#pragma unroll
    for (int q = 0; q < G::PS; ++q) {
        int idx = tid + q * NT;
        if (G::NSPEC % NT == 0 || idx < G::NSPEC) {
            int kx = idx / RR, r = idx % RR;
            v[q] = lds[r * LS + kx];
        }
    }
}

// ----------------------------------------------------------- y-pass ops ---
// Load CC columns (col0 ...) of a spectral array, multiply by a real
// per-entry multiplier (optional, same [kx][ky] layout), into LDS.
// Element e of this thread: idx = tid + e*NT -> c = idx / NY, ky = idx % NY.
template <typename T, int NY>
__device__ __forceinline__ void y_regs_to_lds(const carr<T, YGeo<T, NY>::PE>& v,
                                              typename ctype<T>::type* lds, int tid) {
    using G = YGeo<T, NY>;
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int idx = tid + e * G::NT;
        if (G::NE % G::NT == 0 || idx < G::NE) lds[idx] = v[e];
    }
}

template <typename T, int NY>
__device__ __forceinline__ void y_lds_to_regs(carr<T, YGeo<T, NY>::PE>& v,
                                              const typename ctype<T>::type* lds, int tid) {
    using G = YGeo<T, NY>;
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int idx = tid + e * G::NT;
        if (G::NE % G::NT == 0 || idx < G::NE) v[e] = lds[idx];
    }
}

// Hermitian projection of the reference (2d/functions1.c:492-521) restricted
// to the stored half: column kx=0 gets its rows ky>NY/2 from conj(rows NY-ky)
// and Im zeroed at ky=0, NY/2; column kx=NX/2 gets Im zeroed at ky=0, NY/2.
template <typename T, int NY>
__device__ __forceinline__ void y_project(typename ctype<T>::type* lds, int col0, int ncolx, int tid) {
    using G = YGeo<T, NY>;
    using C = typename ctype<T>::type;
    for (int c = 0; c < G::CC; ++c) {
        int kx = col0 + c;
        if (kx == 0) {
            C* col = lds + c * G::LS;
            for (int ky = NY / 2 + 1 + tid; ky < NY; ky += G::NT) col[ky] = cconj(col[NY - ky]);
            if (tid == 0) { col[0].y = 0; col[NY / 2].y = 0; }
        } else if (kx == ncolx - 1) {
            C* col = lds + c * G::LS;
            if (tid == 0) { col[0].y = 0; col[NY / 2].y = 0; }
        }
    }
}

template <typename T, int NY, int DIR>
__device__ __forceinline__ void y_fft(typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twy, int tid) {
    using G = YGeo<T, NY>;
    using C = typename ctype<T>::type;
    fft_lds<C, NY, G::CC, G::NT, DIR>(lds, G::LS, twy, 1, tid);
}

}  // namespace hw
