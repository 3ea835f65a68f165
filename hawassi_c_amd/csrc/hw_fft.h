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
// The FFT inside a pass is a mixed radix-8/4/2 (and 5/3) Stockham autosort working in
// LDS: each radix pass reads its butterfly inputs into registers, synchronises,
// applies twiddles (from an FP64-accurate table) and the radix-R DFT, and
// writes the outputs in natural order for the next pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cmath>
#include <vector>

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

// ------------------------------------------------------ small DFTs -------
// 3-point: y_q = sum_r v_r e^(DIR 2 pi i r q / 3)
template <int DIR, typename C> __device__ __forceinline__ void dft3g(C* v) {
    using R = decltype(C::x);
    const R s3 = (R)0.86602540378443864676;    // sin(2 pi / 3)
    C a = cadd(v[1], v[2]), b = csub(v[1], v[2]);
    C m = cmk<C>(v[0].x - (R)0.5 * a.x, v[0].y - (R)0.5 * a.y);
    C ib = cmul_i<DIR>(cscale(b, s3));           // i DIR sin(2 pi/3) (v1 - v2)
    v[0] = cadd(v[0], a);
    v[1] = cadd(m, ib);
    v[2] = csub(m, ib);
}

// 5-point with the symmetric pairs a1 = v1 + v4, b1 = v1 - v4, a2 = v2 + v3,
// b2 = v2 - v3
template <int DIR, typename C> __device__ __forceinline__ void dft5g(C* v) {
    using R = decltype(C::x);
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;   // cos(2 pi/5), cos(4 pi/5)
    const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;    // sin(2 pi/5), sin(4 pi/5)
    C a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    C a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    C m1 = cmk<C>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    C m2 = cmk<C>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    C n1 = cmul_i<DIR>(cmk<C>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    C n2 = cmul_i<DIR>(cmk<C>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dftR(C* v) {
    if constexpr (R == 16) dft16<DIR>(v);
    else if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v);
    else if constexpr (R == 5) dft5g<DIR>(v);
    else if constexpr (R == 3) dft3g<DIR>(v);
    else dft2<DIR>(v);
}

#ifndef HW_MAX_RADIX
#define HW_MAX_RADIX 8
#endif
// Radix-16 LDS passes are switched off: built with them, the 1-D right-hand
// side at n = 2048 (1024 threads, two elements each) took an illegal memory
// access on gfx950, and no out-of-bounds index has been found that explains
// it (the pass's LDS and twiddle indices stay in range in a host model).
#if HW_MAX_RADIX > 8
#error "radix-16 LDS passes are switched off (an unexplained device fault at n = 2048)"
#endif
// the radix of the LDS Stockham pass that leaves rem points per sequence:
// 16/8/4/2 while rem has a factor 2, then 5 and 3 (lengths 2^a 3^b 5^c)
constexpr int pass_radix(int rem) {
    return (HW_MAX_RADIX >= 16 && rem % 16 == 0) ? 16
         : rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 5 == 0 ? 5 : 3;
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
        // The twiddles a butterfly needs depend only on its index: their
        // loads are issued first, so they are in flight while the inputs come
        // from LDS and the block meets at the barrier.  Radix 2/4: every
        // w^r from the table; radix 8/16: w, w^4 (and w^8), the other powers
        // formed below (at most two roundings more than a table entry).
        constexpr int NTW = NS == 1 ? 1 : (R == 16 ? 3 : R == 8 ? 2 : R - 1);
        C wt[PER][NTW];
        if constexpr (NS > 1) {
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                int bf = tid + p * NT;
                if (NB % NT == 0 || bf < NB) {
                    const int k = (bf % M) % NS;
                    const int ex = k * (L / (NS * R));
                    if constexpr (R != 8 && R != 16) {
#pragma unroll
                        for (int r = 1; r < R; ++r) wt[p][r - 1] = tw[(r * ex) * TS];
                    } else {
                        wt[p][0] = tw[ex * TS];
                        wt[p][1] = tw[4 * ex * TS];
                        if constexpr (R == 16) wt[p][2] = tw[8 * ex * TS];
                    }
                }
            }
        }
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
                    if constexpr (R != 8 && R != 16) {
#pragma unroll
                        for (int r = 1; r < R; ++r) {
                            C w = wt[p][r - 1];
                            if (DIR > 0) w = cconj(w);
                            v[p][r] = cmul(v[p][r], w);
                        }
                    } else {
                        C wp[R];
                        C w1 = wt[p][0], w4 = wt[p][1];
                        if (DIR > 0) { w1 = cconj(w1); w4 = cconj(w4); }
                        wp[1] = w1;
                        wp[2] = cmul(w1, w1);
                        wp[3] = cmul(wp[2], w1);
                        wp[4] = w4;
                        wp[5] = cmul(w4, w1);
                        wp[6] = cmul(w4, wp[2]);
                        wp[7] = cmul(w4, wp[3]);
                        if constexpr (R == 16) {
                            C w8 = wt[p][2];
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

// ------------------------------------------- register-resident Stockham ---
// The 2-D passes keep the E elements a thread owns in registers for the whole
// transform: each radix-R pass applies twiddles and the radix-R DFT in
// registers, and only the hand-over between passes goes through LDS (one
// write, one barrier, one read).  Two exchange buffers alternate, so a pass
// needs one barrier; when the block's LDS budget cannot hold two, one buffer
// is used with a barrier before every write.
//
// Schedule: radix E (4 or 8) everywhere, with the remainder factor as the
// SECOND pass so that the first and last radix are equal.  Then the register
// layout of the input ("slot p*R0 + r of thread t holds element j + r*(L/R0)
// of sequence b, (b, j) = divmod(t + p*NT, L/R0)") is also the layout of the
// output, so an inverse transform leaves data exactly where the next forward
// transform expects it.
#ifndef HW_LDS_PAD
#define HW_LDS_PAD 5
#endif
// exchange-buffer index: one pad element every 2^HW_LDS_PAD (bank spread).
// pidx(a + b) == pidx(a) + pidx(b) whenever a is a multiple of 2^HW_LDS_PAD.
__host__ __device__ constexpr int pidx(int i) { return i + (i >> HW_LDS_PAD); }
constexpr int PADQ = 1 << HW_LDS_PAD;

constexpr int rs_log(int L, int E) { int a = 0; while (L > 1 && L % E == 0) { L /= E; ++a; } return a; }
constexpr int rs_rem(int L, int E) { int a = rs_log(L, E); for (int i = 0; i < a; ++i) L /= E; return L; }
constexpr int rs_np(int L, int E) {
    const int a = rs_log(L, E), s = rs_rem(L, E);
    if (s == 1) return a;
    if (a >= 2) return a + 1;
    if (a == 1) return (E == 8 && L == 16) ? 2 : 3;   // 16 = 4*4; 32 = 2*8*2; 8 = 2*2*2 (E = 4)
    return 1;
}
constexpr int rs_radix(int L, int E, int i) {
    const int a = rs_log(L, E), s = rs_rem(L, E);
    if (s == 1) return E;
    if (a >= 2) return i == 1 ? s : E;
    if (a == 1) {
        if (E == 8) return L == 16 ? 4 : (i == 1 ? 8 : 2);
        return 2;
    }
    return L;
}
constexpr int rs_ns(int L, int E, int i) { int ns = 1; for (int k = 0; k < i; ++k) ns *= rs_radix(L, E, k); return ns; }
// Per-pass twiddle tables: pass I >= 1 (radix R, NS = product of the earlier
// radices) needs w^(r k L/(NS R)), w = exp(-2 pi i / L), for r = 1..R-1 and
// k < NS; it is stored at [(r-1) NS + k] from rs_twoff(L, E, I), so a wave
// reads consecutive entries (no LDS bank conflicts) and no twiddle is formed
// by multiplication.
constexpr int rs_twoff(int L, int E, int I) {
    int o = 0;
    for (int i = 1; i < I; ++i) o += (rs_radix(L, E, i) - 1) * rs_ns(L, E, i);
    return o;
}
constexpr int rs_twsize(int L, int E) { return rs_twoff(L, E, rs_np(L, E)); }

// v[r] *= t[(r-1) NS] (conjugated for DIR = +1), t = the pass's table at k
template <typename C, int R, int NS, int DIR>
__device__ __forceinline__ void tw_apply(C* v, const C* __restrict__ t) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
        const C w = t[(r - 1) * NS];
        v[r] = DIR > 0 ? cmk<C>(v[r].x * w.x + v[r].y * w.y, v[r].y * w.x - v[r].x * w.y) : cmul(v[r], w);
    }
}

// Pass I of a length-L transform of B sequences by NT threads holding E
// elements each.  Exchange buffers: B rows of stride LSb at bufA / bufB
// (bufB == bufA when TWO is false); tw = the per-pass twiddle tables.
template <typename C, int L, int B, int NT, int E, int DIR, bool TWO, int I>
__device__ __forceinline__ void fr_pass(C (&v)[E], C* bufA, C* bufB, int LSb, const C* __restrict__ tw, int tid) {
    constexpr int NP = rs_np(L, E);
    static_assert(B * L == NT * E, "register layout must cover the block exactly");
    if constexpr (I < NP) {
        constexpr int R = rs_radix(L, E, I), NS = rs_ns(L, E, I), M = L / R;
#pragma unroll
        for (int p = 0; p < E / R; ++p) {
            const int j = (tid + p * NT) % M;
            if constexpr (NS > 1) tw_apply<C, R, NS, DIR>(&v[p * R], tw + rs_twoff(L, E, I) + j % NS);
            dftR<R, DIR>(&v[p * R]);
        }
        if constexpr (I + 1 < NP) {
            C* buf = (TWO && (I % 2)) ? bufB : bufA;
            if constexpr (!TWO) __syncthreads();
#pragma unroll
            for (int p = 0; p < E / R; ++p) {
                const int bf = tid + p * NT, b = bf / M, j = bf % M, k = j % NS;
                // element r goes to idxD + r*NS; that sum never carries across
                // a pad boundary that idxD itself does not (powers of two)
                C* dst = buf + b * LSb + pidx((j / NS) * NS * R + k);
#pragma unroll
                for (int r = 0; r < R; ++r) dst[pidx(r * NS)] = v[p * R + r];
            }
            __syncthreads();
            constexpr int R2 = rs_radix(L, E, I + 1), M2 = L / R2;
#pragma unroll
            for (int p = 0; p < E / R2; ++p) {
                const int bf = tid + p * NT, b = bf / M2, j = bf % M2;
                const C* src = buf + b * LSb;
                if constexpr (M2 % PADQ == 0) {
                    src += pidx(j);
#pragma unroll
                    for (int r = 0; r < R2; ++r) v[p * R2 + r] = src[pidx(r * M2)];
                } else {
#pragma unroll
                    for (int r = 0; r < R2; ++r) v[p * R2 + r] = src[pidx(j + r * M2)];
                }
            }
            fr_pass<C, L, B, NT, E, DIR, TWO, I + 1>(v, bufA, bufB, LSb, tw, tid);
        }
    }
}

// Unnormalised transform of B sequences held in registers (layout above).
// The caller guarantees that no thread still reads bufA/bufB from earlier work.
template <typename C, int L, int B, int NT, int E, int DIR, bool TWO>
__device__ __forceinline__ void fft_reg(C (&v)[E], C* bufA, C* bufB, int LSb, const C* __restrict__ tw, int tid) {
    fr_pass<C, L, B, NT, E, DIR, TWO, 0>(v, bufA, bufB, LSb, tw, tid);
}
// whether the last exchange of fft_reg used bufA
template <int L, int E> constexpr bool fr_last_a() { return rs_np(L, E) >= 2 && (rs_np(L, E) - 2) % 2 == 0; }

// slot s of thread tid -> (sequence b, index n)
template <int L, int NT, int E> __device__ __forceinline__ void fr_slot(int tid, int s, int& b, int& n) {
    constexpr int R0 = rs_radix(L, E, 0), M0 = L / R0;
    const int bf = tid + (s / R0) * NT;
    b = bf / M0;
    n = bf % M0 + (s % R0) * M0;
}

// ------------------------------------------------------- pass geometry ----
#ifndef HW_X_RUN_BYTES
#define HW_X_RUN_BYTES 64
#endif
#ifndef HW_X_PER_THREAD
#define HW_X_PER_THREAD 8
#endif
#ifndef HW_Y_PER_THREAD
#define HW_Y_PER_THREAD 8
#endif
// LDS budgets per block (gfx950 has 160 KB per CU).  An x-pass block is the
// only block on its CU (1024 threads), a y-pass block shares the CU.
#ifndef HW_X_LDS_BUDGET
#define HW_X_LDS_BUDGET 163840
#endif
#ifndef HW_Y_LDS_BUDGET
#define HW_Y_LDS_BUDGET 65536
#endif
// waves per SIMD the FP32 y kernels are compiled for
#ifndef HW_Y_WPE32
#define HW_Y_WPE32 3
#endif

// waves per SIMD the x kernels are compiled for (0: one block's waves per
// SIMD, NT / 256); more than that lets several x blocks share a CU when
// their LDS fits (HW_X_LDS_BUDGET), so that one block's loads overlap
// another's LDS transform phases
#ifndef HW_X_WPE
#define HW_X_WPE 0
#endif

// x-pass: RR rows per block, so that the transposed spectral runs
// X[kx][y0 .. y0+RR) are HW_X_RUN_BYTES long (at least one wave per block).
// Short runs cost far more than their bytes: the lines they touch are shared
// by several blocks, so each line crosses the L2 -> L1 path once per block and
// partial-line stores are merged in L2; 64-byte runs with streaming stores
// keep the transposed traffic close to its byte count.
// A row is a real FFT of length NX done as a complex FFT of L = NX/2 points
// z[n] = x[2n] + i x[2n+1]; each thread holds E such pairs.
template <typename T, int NX> struct XGeo {
    using C = typename ctype<T>::type;
    static constexpr int L = NX / 2;
    static constexpr int E = HW_X_PER_THREAD;
    static constexpr int RR0 = HW_X_RUN_BYTES / (int)sizeof(C) > 0 ? HW_X_RUN_BYTES / (int)sizeof(C) : 1;
    static constexpr int RRW = (64 * E + L - 1) / L;
    static constexpr int RR = RR0 > RRW ? RR0 : RRW;
    static constexpr int NT = RR * L / E;
    static constexpr int PP = E;                        // pairs per thread
    static constexpr int NPAIR = RR * L;
    static constexpr int LSb = pidx(L) + 2;             // buffer row: X[0..L] while staging
    static constexpr int NSPEC = RR * (L + 1);          // spectral values per block
    static constexpr int PS = (NSPEC + NT - 1) / NT;
    static constexpr int SK = NT / RR;                  // kx step between a thread's spectral slots
    // two exchange buffers when they fit beside the twiddle table, else one;
    // the twiddle table is staged in LDS when it fits
    // twiddles: the FFT's per-pass tables, then exp(-2 pi i k / NX), k = 0..L,
    // for the real/complex pre- and post-processing
    static constexpr int TWPP = rs_twsize(L, E);
    static constexpr int TWN = TWPP + L + 1;
    static constexpr bool TWO = (2 * RR * LSb + TWN) * (int)sizeof(C) + 64 <= HW_X_LDS_BUDGET;
    static constexpr int LDS_ELEMS = (TWO ? 2 : 1) * RR * LSb;
    static constexpr bool TWL = (LDS_ELEMS + TWN) * (int)sizeof(C) + 64 <= HW_X_LDS_BUDGET;
    // physical pair q of thread tid -> row * L + n
    __device__ static __forceinline__ int ppos(int tid, int q) {
        int b, n;
        fr_slot<L, NT, E>(tid, q, b, n);
        return b * L + n;
    }
    static constexpr int WPE = HW_X_WPE > 0 ? HW_X_WPE : (NT / 256 > 0 ? NT / 256 : 1);
    static_assert(NT % RR == 0 && NT <= 1024, "x-pass block shape");
};
// launch bounds of every x-pass kernel (T, NX in scope)
#define HW_XK_BOUNDS __launch_bounds__((XGeo<T, NX>::NT)) __attribute__((amdgpu_waves_per_eu(XGeo<T, NX>::WPE)))

// y-pass: CC whole columns per block (about 256 threads), E elements each.
template <typename T, int NY> struct YGeo {
    using C = typename ctype<T>::type;
    static constexpr int E = HW_Y_PER_THREAD;
    static constexpr int CC = NY >= 256 * E ? 1 : (256 * E) / NY;
    static constexpr int LS = pidx(NY - 1) + 1 + ((pidx(NY - 1) + 1) & 1);
    static constexpr int NE = CC * NY;
    static constexpr int NT = NE / E;
    static constexpr int PE = E;                        // elements per thread
    static constexpr int TWN = rs_twsize(NY, E);         // per-pass twiddle tables
    static constexpr bool TWO = (2 * CC * LS + TWN) * (int)sizeof(C) <= HW_Y_LDS_BUDGET;
    static constexpr int LDS_ELEMS = (TWO ? 2 : 1) * CC * LS;
    static constexpr bool TWL = (LDS_ELEMS + TWN) * (int)sizeof(C) <= HW_Y_LDS_BUDGET;
    // waves per SIMD the y kernels are compiled for: FP32 blocks fit three
    // to a CU by LDS, so registers should not be the tighter limit
    static constexpr int WPE = sizeof(T) == 4 ? HW_Y_WPE32 : 1;
    // element e of thread tid -> column * NY + ky within the block
    __device__ static __forceinline__ int eidx(int tid, int e) {
        int b, n;
        fr_slot<NY, NT, E>(tid, e, b, n);
        return b * NY + n;
    }
    static_assert(NT <= 1024 && NT >= 64, "y-pass block shape");
};

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.  This
// maps the hardware block index so that XCD x gets a contiguous range of
// logical blocks.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    if (nb % 8) return b;
    int per = nb / 8;
    return (b % 8) * per + b / 8;
}

// Twiddle tables handed to kernels (device pointers, built on the host in
// FP64 and rounded once):
//   twx: per-pass tables of the length-NX/2 FFT (rs_twoff), then
//        exp(-2 pi i k / NX) for k = 0..NX/2 from XGeo::TWPP
//   twy: per-pass tables of the length-NY FFT

// ----------------------------------------------------------- x-pass ops ---
// Spectral values of a block: slot q of thread tid is idx = tid + q*NT,
// kx = idx / RR, row idx % RR (runs of RR consecutive y per kx in global).
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
__device__ __forceinline__ void x_store_slot(typename ctype<T>::type* __restrict__ X, int y0, int q,
                                             typename ctype<T>::type v, int tid);

template <typename T, int NX, int NY>
__device__ __forceinline__ void x_store_spec(typename ctype<T>::type* __restrict__ X, int y0, const XSpec<T, NX>& b,
                                             int tid) {
    using G = XGeo<T, NX>;
#pragma unroll
    for (int q = 0; q < G::PS; ++q) {
        int idx = tid + q * G::NT;
        if (G::NSPEC % G::NT == 0 || idx < G::NSPEC) x_store_slot<T, NX, NY>(X, y0, q, b[q], tid);   // streaming
    }
}

// Inverse x-pass: spectral values (x_spec_load layout) -> physical pairs
// (x[2n], x[2n+1]) of this thread's slots (XGeo::ppos), scaled by `scale`.
// Also returns, per row of the block, Im of the Nyquist bin times `scale` in
// nyq_im[row] (LDS array of RR entries), which the adjust operator needs for
// the reference's complex inverse.  The imaginary parts of the DC and Nyquist
// bins are ignored (real-field inverse).
// In two halves: x_inv_stage hands the spectral registers to LDS (after which
// the caller may reuse them, e.g. to issue the next field's loads so that they
// are in flight during the FFT), x_inv_finish runs the transform.
template <typename T, int NX>
__device__ __forceinline__ void x_inv_stage(const XSpec<T, NX>& b, typename ctype<T>::type* lds, int tid) {
    using G = XGeo<T, NX>;
    constexpr int RR = G::RR, LSb = G::LSb, NT = G::NT;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < G::PS; ++q) {
        int idx = tid + q * NT;
        if (G::NSPEC % NT == 0 || idx < G::NSPEC) {
            int kx = idx / RR, r = idx % RR;
            lds[r * LSb + pidx(kx)] = b[q];
        }
    }
}

template <typename T, int NX>
__device__ __forceinline__ void x_inv_finish(typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twx,
                                             carr<T, XGeo<T, NX>::PP>& out, T scale, T* nyq_im, int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    constexpr int L = G::L, RR = G::RR, LSb = G::LSb, NT = G::NT, E = G::E;
    C* S = lds;                                  // staging, then the second exchange buffer
    C* A = G::TWO ? lds + RR * LSb : lds;
    __syncthreads();
    if (nyq_im && tid < RR) nyq_im[tid] = S[tid * LSb + pidx(L)].y * scale;
    // Z[k] = (X[k] + conj X[L-k]) + i (X[k] - conj X[L-k]) W^-k, W = exp(-2 pi i / NX)
    constexpr int R0 = rs_radix(L, E, 0), M0 = L / R0;
#pragma unroll
    for (int p = 0; p < E / R0; ++p) {
        const int bf = tid + p * NT, rb = bf / M0, j = bf % M0;
        const C* row = S + rb * LSb;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int k = j + r * M0;
            C xa, xb;
            if constexpr (M0 % PADQ == 0) {
                xa = row[pidx(j) + pidx(r * M0)];
                xb = row[pidx(M0 - j) + pidx(L - (r + 1) * M0)];
            } else {
                xa = row[pidx(k)];
                xb = row[pidx(L - k)];
            }
            if (r == 0 && j == 0) { xa.y = (T)0; xb.y = (T)0; }
            const C wa = cconj(twx[G::TWPP + k]);
            const C ea = cadd(xa, cconj(xb));
            const C oa = cmul(csub(xa, cconj(xb)), wa);
            out[p * R0 + r] = cadd(ea, cmul_i<1>(oa));
        }
    }
    fft_reg<C, L, RR, NT, E, +1, G::TWO>(out, A, S, LSb, twx, tid);
#pragma unroll
    for (int s = 0; s < E; ++s) out[s] = cscale(out[s], scale);
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void x_inverse_b(const XSpec<T, NX>& b, int y0,
                                            typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twx,
                                            carr<T, XGeo<T, NX>::PP>& out, T scale, T* nyq_im, int tid) {
    (void)y0;
    x_inv_stage<T, NX>(b, lds, tid);
    x_inv_finish<T, NX>(lds, twx, out, scale, nyq_im, tid);
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void x_inverse(const typename ctype<T>::type* __restrict__ X, int y0,
                                          typename ctype<T>::type* lds, const typename ctype<T>::type* __restrict__ twx,
                                          carr<T, XGeo<T, NX>::PP>& out, T scale, T* nyq_im, int tid) {
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(X, y0, b, tid);
    x_inverse_b<T, NX, NY>(b, y0, lds, twx, out, scale, nyq_im, tid);
}

// Forward x-pass of a real field given as pairs in registers (XGeo::ppos
// layout).  x_fwd_to_lds runs the transform and leaves Z (the half-length
// complex spectrum of every row) in LDS; x_post forms the real-field spectrum
// X[k] of one spectral slot from it, so callers can store slot by slot
// without holding the whole spectral array in registers.
template <typename T, int NX>
__device__ __forceinline__ const typename ctype<T>::type* x_fwd_to_lds(const carr<T, XGeo<T, NX>::PP>& f,
                                                                      typename ctype<T>::type* lds,
                                                                      const typename ctype<T>::type* __restrict__ twx,
                                                                      int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    constexpr int L = G::L, RR = G::RR, LSb = G::LSb, NT = G::NT, E = G::E;
    C* S = lds;
    C* A = G::TWO ? lds + RR * LSb : lds;
    C v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = f[s];
    __syncthreads();
    fft_reg<C, L, RR, NT, E, -1, G::TWO>(v, A, S, LSb, twx, tid);
    // Z to LDS in natural order (the buffer the last exchange did not use)
    C* O = G::TWO ? (fr_last_a<L, E>() ? S : A) : S;
    if constexpr (!G::TWO) __syncthreads();
    constexpr int R0 = rs_radix(L, E, 0), M0 = L / R0;
#pragma unroll
    for (int p = 0; p < E / R0; ++p) {
        const int bf = tid + p * NT, rb = bf / M0, j = bf % M0;
        C* row = O + rb * LSb;
        if constexpr (M0 % PADQ == 0) {
            row += pidx(j);
#pragma unroll
            for (int r = 0; r < R0; ++r) row[pidx(r * M0)] = v[p * R0 + r];
        } else {
#pragma unroll
            for (int r = 0; r < R0; ++r) row[pidx(j + r * M0)] = v[p * R0 + r];
        }
    }
    __syncthreads();
    return O;
}

// X[k] = (Z[k] + conj Z[L-k])/2 - i/2 W^k (Z[k] - conj Z[L-k]) for spectral
// slot q of this thread (kx = tid / RR + q * NT / RR, row tid % RR); the
// caller checks kx <= L.
template <typename T, int NX>
__device__ __forceinline__ typename ctype<T>::type x_post(const typename ctype<T>::type* O,
                                                          const typename ctype<T>::type* __restrict__ twx, int q, int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    constexpr int L = G::L, RR = G::RR, LSb = G::LSb, SK = G::SK;
    const int kx0 = tid / RR, kx = kx0 + q * SK;
    const C* row = O + (tid % RR) * LSb;
    if (kx == 0 || kx == L) {
        const C z = row[0];
        return kx == 0 ? cmk<C>(z.x + z.y, (T)0) : cmk<C>(z.x - z.y, (T)0);
    }
    C za, zb;
    if constexpr (SK % PADQ == 0) {
        za = row[pidx(kx0) + pidx(q * SK)];
        zb = row[pidx(SK - kx0) + pidx(L - (q + 1) * SK)];
    } else {
        za = row[pidx(kx)];
        zb = row[pidx(L - kx)];
    }
    const C ea = cscale(cadd(za, cconj(zb)), (T)0.5);
    const C da = cscale(csub(za, cconj(zb)), (T)0.5);
    return cadd(ea, cmul_i<-1>(cmul(twx[G::TWPP + kx], da)));
}

// spectral slot q of this thread holds a value (kx <= L)
template <typename T, int NX> __device__ __forceinline__ bool x_slot_ok(int q, int tid) {
    using G = XGeo<T, NX>;
    return G::NSPEC % G::NT == 0 || tid + q * G::NT < G::NSPEC;
}

template <typename T, int NX>
__device__ __forceinline__ void x_forward_regs(const carr<T, XGeo<T, NX>::PP>& f, XSpec<T, NX>& bs,
                                               typename ctype<T>::type* lds,
                                               const typename ctype<T>::type* __restrict__ twx, int tid) {
    using G = XGeo<T, NX>;
    const auto* O = x_fwd_to_lds<T, NX>(f, lds, twx, tid);
#pragma unroll
    for (int q = 0; q < G::PS; ++q)
        if (x_slot_ok<T, NX>(q, tid)) bs[q] = x_post<T, NX>(O, twx, q, tid);
}

// streaming store of one spectral value of slot q to X[kx][y0 + row]
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_store_slot(typename ctype<T>::type* __restrict__ X, int y0, int q,
                                             typename ctype<T>::type v, int tid) {
    using G = XGeo<T, NX>;
    const int idx = tid + q * G::NT;
    typename ctype<T>::type* p = X + (size_t)(idx / G::RR) * NY + y0 + idx % G::RR;
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// ----------------------------------------------------------- y-pass ops ---
// Hermitian projection of the reference (2d/functions1.c:492-521) restricted
// to the stored half, applied to a block's columns in registers: column kx=0
// gets its rows ky>NY/2 from conj(rows NY-ky) and Im zeroed at ky=0, NY/2;
// column kx=NX/2 gets Im zeroed at ky=0, NY/2.  `buf` is an LDS row of NY
// (padded) no thread reads until the next barrier.
template <typename T, int NX, int NY>
__device__ __forceinline__ void y_project_regs(carr<T, YGeo<T, NY>::PE>& v, typename ctype<T>::type* buf, int col0,
                                               int tid) {
    using G = YGeo<T, NY>;
    using C = typename ctype<T>::type;
    if (col0 == 0) {   // block-uniform
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int b, n;
            fr_slot<NY, G::NT, G::E>(tid, e, b, n);
            if (b == 0) buf[pidx(n)] = v[e];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int b, n;
            fr_slot<NY, G::NT, G::E>(tid, e, b, n);
            if (b == 0 && n > NY / 2) v[e] = cconj(buf[pidx(NY - n)]);
        }
    }
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int b, n;
        fr_slot<NY, G::NT, G::E>(tid, e, b, n);
        const int kx = col0 + b;
        if ((kx == 0 || kx == NX / 2) && (n == 0 || n == NY / 2)) v[e].y = (T)0;
    }
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void y_fwd_regs(carr<T, YGeo<T, NY>::PE>& v, typename ctype<T>::type* lds,
                                           const typename ctype<T>::type* twy, int tid) {
    using G = YGeo<T, NY>;
    using C = typename ctype<T>::type;
    __syncthreads();
    fft_reg<C, NY, G::CC, G::NT, G::E, -1, G::TWO>(v, lds, G::TWO ? lds + G::CC * G::LS : lds, G::LS, twy, tid);
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void y_inv_regs(carr<T, YGeo<T, NY>::PE>& v, typename ctype<T>::type* lds,
                                           const typename ctype<T>::type* twy, int col0, int tid) {
    using G = YGeo<T, NY>;
    using C = typename ctype<T>::type;
    __syncthreads();
    // the projection stages column 0 in the buffer the first exchange does not write
    C* pb = G::TWO ? lds + G::CC * G::LS : lds;
    y_project_regs<T, NX, NY>(v, pb, col0, tid);
    fft_reg<C, NY, G::CC, G::NT, G::E, +1, G::TWO>(v, lds, G::TWO ? lds + G::CC * G::LS : lds, G::LS, twy, tid);
}

// Host: twiddle tables of the 2-D passes (FP64 values rounded once to T).
template <typename T> std::vector<typename ctype<T>::type> fr_pass_tables(int L, int E) {
    std::vector<typename ctype<T>::type> t((size_t)rs_twsize(L, E));
    for (int I = 1; I < rs_np(L, E); ++I) {
        const int R = rs_radix(L, E, I), NS = rs_ns(L, E, I), o = rs_twoff(L, E, I);
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < NS; ++k) {
                const long m = (long)r * k * (L / (NS * R)) % L;
                const double a = -2.0 * M_PI * (double)m / (double)L;
                t[(size_t)o + (size_t)(r - 1) * NS + k].x = (T)std::cos(a);
                t[(size_t)o + (size_t)(r - 1) * NS + k].y = (T)std::sin(a);
            }
    }
    return t;
}
template <typename T, int NX> std::vector<typename ctype<T>::type> x_twiddles() {
    using G = XGeo<T, NX>;
    auto t = fr_pass_tables<T>(G::L, G::E);
    t.resize((size_t)G::TWN);
    for (int k = 0; k <= G::L; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)NX;
        t[(size_t)G::TWPP + k].x = (T)std::cos(a);
        t[(size_t)G::TWPP + k].y = (T)std::sin(a);
    }
    return t;
}
template <typename T, int NY> std::vector<typename ctype<T>::type> y_twiddles() {
    return fr_pass_tables<T>(NY, YGeo<T, NY>::E);
}

}  // namespace hw
