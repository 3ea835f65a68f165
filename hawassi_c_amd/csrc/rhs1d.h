// rhs1d.h -- the 1-D HS2 right-hand side (rhs_HsF2, 1d/rhs.c:96-241) and
// its RKF45 step control for a batch of independent realizations.
//
// One workgroup owns one realization (n <= 16384 complex points in FP32,
// <= 8192 in FP64) and evaluates its whole right-hand side: every transform
// is an LDS-resident Stockham FFT of length n, fields between transforms stay
// in registers or in a per-realization scratch area.  The reference takes the
// real part of full complex inverse transforms (no Hermitian projection in
// 1-D, 1d/functions.c:1642-1661); two such real-part inverses are done with
// one complex FFT of Herm(X) + i Herm(Y), and two forward transforms of real
// fields with one complex FFT of a + i b.  Thread t owns the spectral (and
// physical) indices k = t + e*NT (k < n/2) and their partners n - k (0 is
// paired with n/2), so the Hermitian splits are thread-local.
//
// Batches mix sizes: the realizations are grouped by n into size classes,
// each with its own kernel instance and thread count (nt_for_n).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hw_fft.h"

namespace hw {

// ----------------------------------------------------------- descriptors --
struct Real1D {                 // read-only model description of one realization
    int n, flat, src, prop;     // prop: 0 Uni+, 1 Uni-, 2 Bi, 3 none
    int brk, nt, cap, pad0;
    long long off;              // element offset of this realization (sum of earlier n)
    long long tw_off;           // offset of the length-n twiddle table in Batch1DArgs::tw
    long long sig_off;          // offset into the signal arrays
    long long cb_off;           // offset into the crest arrays (sum of earlier caps)
    double dx, KBC, TC, Tstar, MaxEtaInit, delb, eps_abs, eps_rel;
    int xbstart, xbend, xbendL, xbstartR;
    int lenIF, ILL, IFF, mid;   // mid: a third reference depth (Up3IP)
    int order, wall;            // order: 1 rhs_HsF1, 2 rhs_HsF2, 3 rhs_HsF3; wall: wall source on
    int idx_wall, pad3;
    double refl;                // reflection coefficient of the wall
};

struct Rk1D {                   // driver state of one realization (FP64)
    double t, h, h0, t1;        // time, suggested step, step of the current attempt, target
    int need_in, final_step, active, err;
    long long accepted, rejected, rhs;
};

struct Cb1 {                    // CB (1d/header.h:44-52); Bposition is not used by the logic
    int Bindex, BindexNow;
    double Ucrest;
    int DirProp, Bcount;
    double Btime;
};

struct Brk1D {
    int flag;                   // the reference's global `flag` (0 none, 1 first crests, 2 tracking)
    int overflow;
    int pad[2];
};

enum : int { R1_ETA = 0, R1_U, R1_M0U, R1_H, R1_P2, R1_FLUX, R1_AU, R1_AE, R1_AC, R1_NR };
enum : int { C1_EH = 0, C1_M0H, C1_EUH, C1_U2H, C1_HP, C1_UH, C1_NC };

template <typename T> struct Batch1DArgs {
    using C = typename ctype<T>::type;
    const Real1D* real;
    Rk1D* rk;
    const int* list;            // realization ids of this launch (blockIdx.x -> id)
    // per-element model arrays, offset `off`
    const T *k, *aal, *chi, *damp, *opa, *opb, *gm, *gp, *depth;
    const T *opc, *gc;          // third reference depth: Csq_mid, gam_mid (zero when unused)
    const C *gamx, *gamxs;
    const C* gamws;             // wall skew kernel gamW_skew (zero without a wall)
    double* wst;                // wall source state per realization: t_prev, S, previous refl*eta, calls
    // influx signals (natural spline: values and half second derivatives)
    const double *tsig, *sig, *sigc, *skw, *skwc;
    // state (FP64) and stage derivatives, offset 2*off: [eta^ (n)][u^ (n)]
    double2* y;
    C* K[7];                    // K[0..5]: rkf45 stages, K[6]: dydt_out / stand-alone output
    const double2* ext;         // stand-alone input (stage 0)
    double t_ext;
    // scratch: real fields at off*R1_NR + f*n, complex at off*C1_NC + f*n
    T* sr;
    C* sc;
    // breaking
    T* B;
    int* peaks;
    Cb1 *cb, *cbn, *cbp;
    int* ide;
    Brk1D* bs;
    const C* tw;                // twiddle tables of every size present: exp(-2 pi i m / n) at tw_off + m
    T g;
};

// stage input coefficients: input = y + h0 * sum_j a[s][j] K_j
__constant__ double c_rk1d[8][6] = {
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {1.0 / 4.0, 0, 0, 0, 0, 0},
    {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0, 0},
    {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0, 0},
    {8341.0 / 4104.0, -32832.0 / 4104.0, 29440.0 / 4104.0, -845.0 / 4104.0, 0, 0},
    {-6080.0 / 20520.0, 41040.0 / 20520.0, -28352.0 / 20520.0, 9295.0 / 20520.0, -5643.0 / 20520.0, 0},
    {902880.0 / 7618050.0, 0.0, 3953664.0 / 7618050.0, 3855735.0 / 7618050.0, -1371249.0 / 7618050.0,
     277020.0 / 7618050.0},
};
__constant__ double c_rk1d_ah[5] = {1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
__constant__ double c_rk1d_ec[6] = {1.0 / 360.0, 0.0, -128.0 / 4275.0, -2197.0 / 75240.0, 1.0 / 50.0, 2.0 / 55.0};

// ------------------------------------------------------- block helpers --
// red: LDS scratch of at least 2*NT/64 + 8 words.
// Blocks whose size is not a multiple of 64 (n = 2000: 1000 threads) end
// with a partial wave; its shuffles then skip partners beyond the block
// (lane ^ o >= active lanes), so that lane 0 still reduces every thread of
// its wave: the partners it reads, and their partners in turn, have lower
// lane numbers than any thread whose value they carry.
template <int NT> __device__ __forceinline__ bool shfl_ok(int o) {
    if constexpr (NT % 64 == 0) {
        (void)o;
        return true;
    } else {
        const int lane = threadIdx.x & 63, w0 = threadIdx.x & ~63;
        const int act = NT - w0 < 64 ? NT - w0 : 64;
        return (lane ^ o) < act;
    }
}
template <int NT> __device__ __forceinline__ int block_or(int v, int* red) {
    v = __any(v) ? 1 : 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int w = 0; w < (NT + 63) / 64; ++w) r |= red[w];
    __syncthreads();
    return r;
}

template <int NT> __device__ __forceinline__ int block_min(int v, int* red) {
    for (int o = 32; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, 64);
        if (shfl_ok<NT>(o)) v = u < v ? u : v;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = red[0];
#pragma unroll
    for (int w = 1; w < (NT + 63) / 64; ++w) r = red[w] < r ? red[w] : r;
    __syncthreads();
    return r;
}

template <int NT> __device__ __forceinline__ double block_max_d(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) {
        double u = __shfl_xor(v, o, 64);
        if (shfl_ok<NT>(o)) v = u > v ? u : v;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < (NT + 63) / 64; ++w) r = red[w] > r ? red[w] : r;
    __syncthreads();
    return r;
}

// first index of the maximum (ties -> smallest index); idx = INT_MAX if empty
template <int NT> __device__ __forceinline__ void block_argmax(double& v, int& idx, double* redv, int* redi) {
    for (int o = 32; o > 0; o >>= 1) {
        double u = __shfl_xor(v, o, 64);
        int j = __shfl_xor(idx, o, 64);
        if (shfl_ok<NT>(o) && (u > v || (u == v && j < idx))) { v = u; idx = j; }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { redv[threadIdx.x >> 6] = v; redi[threadIdx.x >> 6] = idx; }
    __syncthreads();
    double bv = redv[0];
    int bi = redi[0];
#pragma unroll
    for (int w = 1; w < (NT + 63) / 64; ++w)
        if (redv[w] > bv || (redv[w] == bv && redi[w] < bi)) { bv = redv[w]; bi = redi[w]; }
    __syncthreads();
    v = bv;
    idx = bi;
}

// ordered compaction: threads with f set get consecutive slots in thread
// order; returns this thread's slot (or -1) and the block total in *total
template <int NT> __device__ __forceinline__ int block_compact(int f, int* red, int* total) {
    unsigned long long m = __ballot(f);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int pre = __popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) red[w] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < (NT + 63) / 64; ++q) {
        if (q < w) base += red[q];
        tot += red[q];
    }
    __syncthreads();
    *total = tot;
    return f ? base + pre : -1;
}

// --------------------------------------------------------- pair mapping --
template <typename T, int N, int NT> struct P1 {
    using C = typename ctype<T>::type;
    static constexpr int E = N / NT;
    static constexpr int EH = E / 2;
    static_assert(E >= 2 && E <= 16, "n / NT must be in [2, 16]");
    __device__ static __forceinline__ int kl(int tid, int e) { return tid + e * NT; }
    __device__ static __forceinline__ int kh(int tid, int e) {
        int k = tid + e * NT;
        return k == 0 ? N / 2 : N - k;
    }

    __device__ static __forceinline__ void ldc(const C* __restrict__ p, C (&l)[EH], C (&h)[EH], int tid) {
#pragma unroll
        for (int e = 0; e < EH; ++e) { l[e] = p[kl(tid, e)]; h[e] = p[kh(tid, e)]; }
    }
    __device__ static __forceinline__ void stc(C* __restrict__ p, const C (&l)[EH], const C (&h)[EH], int tid) {
#pragma unroll
        for (int e = 0; e < EH; ++e) { p[kl(tid, e)] = l[e]; p[kh(tid, e)] = h[e]; }
    }
    __device__ static __forceinline__ void ldr(const T* __restrict__ p, T (&l)[EH], T (&h)[EH], int tid) {
#pragma unroll
        for (int e = 0; e < EH; ++e) { l[e] = p[kl(tid, e)]; h[e] = p[kh(tid, e)]; }
    }
    __device__ static __forceinline__ void str(T* __restrict__ p, const T (&l)[EH], const T (&h)[EH], int tid) {
#pragma unroll
        for (int e = 0; e < EH; ++e) { p[kl(tid, e)] = l[e]; p[kh(tid, e)] = h[e]; }
    }

    // Per-slot pieces of the paired transforms.  Slot e of a thread is the
    // index pair (kl, kh); outside the FFT itself a thread only touches its
    // own two LDS entries, so reading a transform's output and writing the
    // next transform's input at a slot needs no barrier, and no field has to
    // be held in registers across a transform.
    //   inverse: (a, b) = (Re F^-1 X, Re F^-1 Y) from one complex FFT of
    //            Herm(X) + i Herm(Y), Herm(X)[k] = (X[k] + conj X[-k]) / 2
    //   forward: (A, B) = (F a, F b) for real a, b from one FFT of a + i b
    __device__ static __forceinline__ void inv_in(C* lds, int tid, int e, C xl, C xh, C yl, C yh) {
        const T hf = (T)0.5;
        const bool z = kl(tid, e) == 0;
        C xml = z ? xl : xh, xmh = z ? xh : xl;
        C yml = z ? yl : yh, ymh = z ? yh : yl;
        T hxr = hf * (xl.x + xml.x), hxi = hf * (xl.y - xml.y);
        T hyr = hf * (yl.x + yml.x), hyi = hf * (yl.y - yml.y);
        lds[kl(tid, e)] = cmk<C>(hxr - hyi, hxi + hyr);
        hxr = hf * (xh.x + xmh.x); hxi = hf * (xh.y - xmh.y);
        hyr = hf * (yh.x + ymh.x); hyi = hf * (yh.y - ymh.y);
        lds[kh(tid, e)] = cmk<C>(hxr - hyi, hxi + hyr);
    }
    __device__ static __forceinline__ void inv_out(const C* lds, int tid, int e, T& al, T& ah, T& bl, T& bh) {
        const T s = (T)1 / (T)N;
        const C zl = lds[kl(tid, e)], zh = lds[kh(tid, e)];
        al = zl.x * s; bl = zl.y * s;
        ah = zh.x * s; bh = zh.y * s;
    }
    __device__ static __forceinline__ void fwd_in(C* lds, int tid, int e, T al, T ah, T bl, T bh) {
        lds[kl(tid, e)] = cmk<C>(al, bl);
        lds[kh(tid, e)] = cmk<C>(ah, bh);
    }
    __device__ static __forceinline__ void fwd_out(const C* lds, int tid, int e, C& Al, C& Ah, C& Bl, C& Bh) {
        const T hf = (T)0.5;
        const bool z = kl(tid, e) == 0;
        const C zl = lds[kl(tid, e)], zh = lds[kh(tid, e)];
        const C ml = z ? zl : zh, mh = z ? zh : zl;
        Al = cmk<C>(hf * (zl.x + ml.x), hf * (zl.y - ml.y));
        Bl = cmk<C>(hf * (zl.y + ml.y), -hf * (zl.x - ml.x));
        Ah = cmk<C>(hf * (zh.x + mh.x), hf * (zh.y - mh.y));
        Bh = cmk<C>(hf * (zh.y + mh.y), -hf * (zh.x - mh.x));
    }
    // the transform itself: inputs of every slot written -> outputs readable
    template <int DIR> __device__ static __forceinline__ void run(C* lds, const C* __restrict__ tw, int tid) {
        __syncthreads();
        fft_lds<C, N, 1, NT, DIR>(lds, N, tw, 1, tid);
    }
};

// ------------------------------------------------------------ breaking ----
// breaking_process + breaking_criterion (1d/breaking.c:63-450) for one
// realization, by the whole workgroup.  eta, u, H (Hilbert transform of eta)
// are physical fields in global memory; B is written.  The crest lists and
// `flag` are per-realization device state.  Sequential parts run on thread 0
// between barriers; scans over nodes are block-parallel reductions that keep
// the reference's first-index semantics.
// What breaking_process reads of the batch, handed over by value: a
// reference to the argument struct would make its address escape into this
// (non-inlined) call, keep the struct in scratch memory and turn every
// pointer the RHS reads from it into a generic (flat) one.
template <typename T> struct Brk1DView {
    const Real1D* real;
    const T* depth;
    int* peaks;
    Cb1 *cb, *cbn, *cbp;
    int* ide;
    Brk1D* bs;
};

template <typename T, int N, int NT>
__device__ __attribute__((noinline)) void brk1d_process(Brk1DView<T> a, const Real1D& R, double t, const T* __restrict__ eta,
                              const T* __restrict__ u, const T* __restrict__ H, T* __restrict__ B, int* sm,
                              double* smd) {
    const int tid = threadIdx.x;
    const int cap = R.cap;
    Cb1* CB = a.cb + R.cb_off;
    Cb1* New = a.cbn + R.cb_off;
    Cb1* Prev = a.cbp + R.cb_off;
    int* IDE = a.ide + R.cb_off;
    int* peaks = a.peaks + R.off;
    Brk1D* bs = a.bs + (&R - a.real);
    int* red = sm;              // >= 2*NT/64 words
    int* shv = sm + 40;         // shared scalars
    double* redv = smd;
    double* shd = smd + 20;

    for (int i = tid; i < N; i += NT) B[i] = (T)0;
    // ---- breaking_criterion: threshold test and peaks (ordered)
    int any = 0;
    for (int i = tid; i < R.lenIF; i += NT) any |= (double)eta[R.xbstart + i] > R.MaxEtaInit;
    any = block_or<NT>(any, red);
    int npk = 0;
    if (any) {
        const bool single = R.xbendL == 0 && R.xbstartR == 0;
        for (int w = 0; w < (single ? 1 : 2); ++w) {
            const int w0 = single ? R.xbstart : (w == 0 ? R.xbstart : R.xbstartR);
            const int len = single ? R.lenIF : (w == 0 ? R.ILL : R.IFF);
            for (int base = 1; base < len - 1; base += NT) {
                int j = base + tid, f = 0;
                if (j < len - 1) {
                    int q = w0 + j;
                    double v = (double)eta[q];
                    f = !(v < R.MaxEtaInit) && v > (double)eta[q + 1] && v > (double)eta[q - 1];
                }
                int tot;
                int slot = block_compact<NT>(f, red, &tot);
                if (slot >= 0) peaks[npk + slot] = w0 + j;
                npk += tot;
            }
        }
    }
    __syncthreads();
    // ---- crest test at each peak; passing peaks fill CB in peak order
    if (npk > 0) {
        const double dx = R.dx;
        int itcb = 0;
        for (int base = 0; base < npk; base += NT) {
            int p = base + tid, f = 0;
            double Uc = 0.0;
            int J = 0;
            if (p < npk) {
                J = peaks[p];
                auto grad = [&](const T* f_, int i) -> double {
                    if (i == 0) return ((double)f_[1] - (double)f_[0]) / dx;
                    if (i == N - 1) return ((double)f_[N - 1] - (double)f_[N - 2]) / dx;
                    return ((double)f_[i + 1] - (double)f_[i - 1]) / (2 * dx);
                };
                double e = eta[J], h = H[J];
                double Kx = (e * grad(H, J) - h * grad(eta, J)) / (e * e + h * h);
                double d = (double)a.depth[R.off + J] + e;
                double Cc = Kx == 0 ? sqrt(9.80665 * d) : (Kx < 0 ? -1.0 : 1.0) * sqrt(9.80665 * Kx * tanh(d * Kx)) / Kx;
                Uc = u[J];
                f = fabs(Uc) > R.KBC * fabs(Cc);
            }
            int tot;
            int slot = block_compact<NT>(f, red, &tot);
            if (slot >= 0 && itcb + slot < cap) {
                Cb1& c = CB[itcb + slot];
                c.Bindex = J; c.BindexNow = J; c.Ucrest = fabs(Uc);
                c.DirProp = (int)(Uc / fabs(Uc)); c.Btime = t;
            }
            itcb += tot;
        }
        if (itcb > cap) { if (tid == 0) bs->overflow = 1; itcb = cap; }
        __syncthreads();
        for (int j = tid; j < itcb; j += NT) CB[j].Bcount = itcb;
    }
    __syncthreads();
    // ---- crest list for this call (1d/breaking.c:212-300), thread 0
    if (tid == 0) {
        int flag = bs->flag;
        int NCB = CB[0].Bcount;
        if (NCB > 0 && flag == 0) {
            flag = 1;
            for (int j = 0; j < NCB; ++j) {
                New[j].Bindex = CB[j].Bindex; New[j].BindexNow = CB[j].BindexNow; New[j].Ucrest = CB[j].Ucrest;
                New[j].DirProp = CB[j].DirProp; New[j].Btime = CB[j].Btime; New[j].Bcount = CB[j].Bcount;
            }
        } else if (NCB == 0 && flag == 2) {
            int np = Prev[0].Bcount;
            for (int j = 0; j < np; ++j) New[j] = Prev[j];
        } else if (NCB > 0 && flag == 2) {
            int np = Prev[0].Bcount;
            for (int j = 0; j < np; ++j) New[j] = Prev[j];
            int add = 0;
            for (int kk = 0; kk < NCB; ++kk) {
                int isnew = 1;
                for (int l = 0; l < np; ++l) {
                    int dleft = abs(IDE[l] - Prev[l].BindexNow);
                    if (Prev[l].DirProp == 1) {
                        if (CB[kk].Bindex >= Prev[l].BindexNow - 3 * dleft && CB[kk].Bindex <= IDE[l]) { isnew = 0; break; }
                    } else {
                        if (CB[kk].Bindex >= Prev[l].BindexNow && CB[kk].Bindex <= IDE[l] + 3 * dleft) { isnew = 0; break; }
                    }
                }
                if (isnew) {
                    if (np + add >= cap) { bs->overflow = 1; break; }
                    Cb1& d = New[np + add];
                    d.Bindex = CB[kk].Bindex; d.BindexNow = CB[kk].Bindex; IDE[np + add] = CB[kk].Bindex;
                    d.Ucrest = CB[kk].Ucrest; d.Btime = CB[kk].Btime; d.DirProp = CB[kk].DirProp;
                    d.Bcount = CB[kk].Bcount;
                    ++add;
                }
            }
            int j = 0;
            while (j < cap && New[j].Bindex > 0) ++j;
            for (int i = 0; i < j; ++i) New[i].Bcount = j;
        }
        bs->flag = flag;
        shv[0] = flag;
        shv[1] = (flag == 1 || flag == 2) ? New[0].Bcount : 0;
    }
    __syncthreads();
    const int flag0 = shv[0];
    const int NCBnow = shv[1];
    if (flag0 == 1 || flag0 == 2) {
        if (tid == 0) shv[2] = 0;   // iterNprev
        for (int j = tid; j < cap; j += NT) {
            Prev[j].Bindex = 0; Prev[j].Ucrest = 0; Prev[j].DirProp = 0; Prev[j].Btime = 0; Prev[j].Bcount = 0;
        }
        __syncthreads();
        int anyterm = 0;   // IDTerminate == 1 seen
        for (int lx = 0; lx < NCBnow; ++lx) {
            const int Dir = New[lx].DirProp;
            int now;
            if (flag0 == 1) {
                now = New[lx].Bindex;
            } else {
                // idx_max over eta[start + j*Dir], j < lenloc (first maximum)
                const int start = New[lx].BindexNow;
                const int lenloc = abs(IDE[lx] - start);
                double bv = -HUGE_VAL;
                int bi = 0x7fffffff;
                for (int j = tid; j < lenloc; j += NT) {
                    double v = eta[start + j * Dir];
                    if (v > bv || (v == bv && j < bi)) { bv = v; bi = j; }
                }
                block_argmax<NT>(bv, bi, redv, red);
                now = start + (lenloc > 0 ? bi : 0) * Dir;
            }
            __syncthreads();
            if (tid == 0) New[lx].BindexNow = now;
            if (lx != 0 && now == New[lx - 1].BindexNow) { __syncthreads(); continue; }
            const double tb = New[lx].Btime, UI = New[lx].Ucrest, UF = R.TC * UI;
            double Us;
            if (t - tb >= R.Tstar) Us = UF;
            else if (t - tb >= 0 && t - tb < R.Tstar) Us = UI + ((t - tb) / R.Tstar) * (UF - UI);
            else Us = UI;
            const int indEND = Dir == 1 ? R.xbend : R.xbstart;
            // nodes ly = now + p*Dir while ly < indEND (and inside the grid)
            int plen;
            if (Dir == 1) plen = indEND - now;
            else if (Dir == -1) plen = now < indEND ? now + 1 : 0;
            else plen = now < indEND ? 1 : 0;
            if (plen < 0) plen = 0;
            int pstop = 0x7fffffff;
            for (int p = tid; p < plen; p += NT) {
                int ly = now + p * Dir;
                if (fabs((double)u[ly]) < UF) { pstop = p; break; }
            }
            pstop = block_min<NT>(pstop, red);
            const int pend = pstop < plen ? pstop : plen;
            int fl = 0;
            for (int p = tid; p < pend; p += NT) {
                int ly = now + p * Dir;
                double au = fabs((double)u[ly]);
                T bval;
                if (au > 2 * Us) bval = (T)1;
                else if (au <= Us) bval = (T)0;
                else bval = (T)(au / Us - 1);
                B[ly] = bval;
                fl |= au >= Us;
            }
            fl = block_or<NT>(fl, red);
            if (tid == 0 && fl) {
                int it = shv[2];
                IDE[it] = pstop < plen ? now + pstop * Dir : now;
                Prev[it].Bindex = New[lx].Bindex; Prev[it].BindexNow = New[lx].BindexNow;
                Prev[it].Ucrest = New[lx].Ucrest; Prev[it].Btime = New[lx].Btime;
                Prev[it].DirProp = New[lx].DirProp; Prev[it].Bcount = New[lx].Bcount;
                shv[2] = it + 1;
            }
            anyterm |= fl;
            __syncthreads();
        }
        if (tid == 0) {
            int j = 0;
            while (j < cap && Prev[j].Bindex > 0) ++j;
            for (int i = 0; i < j; ++i) Prev[i].Bcount = j;
            if (NCBnow > 0) bs->flag = anyterm ? 2 : 0;
        }
        __syncthreads();
    }
    for (int j = tid; j < cap; j += NT) {
        New[j].Bindex = 0; New[j].Ucrest = 0; New[j].DirProp = 0; New[j].Btime = 0; New[j].Bcount = 0;
        CB[j].Bindex = 0; CB[j].Ucrest = 0; CB[j].DirProp = 0; CB[j].Btime = 0; CB[j].Bcount = 0;
    }
    __syncthreads();
}

// ------------------------------------------------------------ the RHS ----
// MID: the realization has a third reference depth (its code is a separate
// instance so that the two-depth path keeps its register allocation)
template <typename T, int N, int NT, bool MID>
__device__ __attribute__((always_inline)) void rhs1d_body(const Batch1DArgs<T>& a, int r, int stage,
                                                           typename ctype<T>::type* lds, int* sm,
                           double* smd) {
    using C = typename ctype<T>::type;
    using P = P1<T, N, NT>;
    constexpr int EH = P::EH;
    const int tid = threadIdx.x;
    const Real1D& R = a.real[r];
    Rk1D* S = a.rk + r;
    const long long off = R.off;
    const C* tw = a.tw + R.tw_off;
    const T g = a.g;
    const T* kk = a.k + off;
    T* sr = a.sr + off * R1_NR;
    C* sc = a.sc + off * C1_NC;

    // ---- time and stage input (the RK combination is fused into the load)
    double t, h0 = S->h0;
    if (stage == 0) t = a.t_ext;
    else if (stage == 1) t = S->t;
    else if (stage == 7) t = S->t + h0;
    else t = S->t + c_rk1d_ah[stage - 2] * h0;
    if (tid == 0) {
        double pv = 0.0, pvs = 0.0;
        if (R.src && R.nt > 1) {
            const double* ts = a.tsig + R.sig_off;
            int lo = 0, hi = R.nt - 1;
            while (hi > lo + 1) { int m = (hi + lo) >> 1; if (ts[m] > t) hi = m; else lo = m; }
            double hh = ts[lo + 1] - ts[lo], dxs = t - ts[lo];
            const double *ys[2] = {a.sig + R.sig_off, a.skw + R.sig_off}, *cs[2] = {a.sigc + R.sig_off, a.skwc + R.sig_off};
            double out[2];
            for (int q = 0; q < 2; ++q) {
                double ci = cs[q][lo], cip = cs[q][lo + 1], dy = ys[q][lo + 1] - ys[q][lo];
                double b = dy / hh - hh * (cip + 2.0 * ci) / 3.0, d = (cip - ci) / (3.0 * hh);
                out[q] = ys[q][lo] + dxs * (b + dxs * (ci + dxs * d));
            }
            pv = out[0];
            pvs = out[1];
        }
        smd[30] = pv;
        smd[31] = pvs;
        S->rhs += 1;
    }
    // Every transform is a chain link in LDS: the per-slot epilogue of one
    // transform writes the next one's input in place, and fields needed
    // again later are kept in this realization's scratch (sr, sc) rather
    // than in registers.
    const T* opa = a.opa + off;
    const T* opb = a.opb + off;
    const T* gmv = a.gm + off;
    const T* gpv = a.gp + off;
    // third-depth arrays are addressed where used (a.opc[off + i], ...), so
    // the two-depth path keeps no extra pointers live
#define HW1_OPC(i) a.opc[off + (i)]
#define HW1_GC(i) a.gc[off + (i)]
#define HW1_AC(i) sr[R1_AC * N + (i)]
    T* s_eta = sr + R1_ETA * N;
    T* s_u = sr + R1_U * N;
    T* s_m0 = sr + R1_M0U * N;
    T* s_p = sr + R1_P2 * N;
    T* s_au = sr + R1_AU * N;
    T* s_ae = sr + R1_AE * N;
    C* c_eh = sc + C1_EH * N;
    C* c_uh = sc + C1_UH * N;
    C* c_m0h = sc + C1_M0H * N;
    C* c_euh = sc + C1_EUH * N;
    C* c_u2h = sc + C1_U2H * N;
    C* c_hp = sc + C1_HP * N;
    // ---- stage input (the RK combination is fused into the load); eta, u
    {
        const double2* base = (stage == 0 ? a.ext : a.y) + 2 * off;
        const int nk = stage >= 2 ? 6 : 0;
        double co[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) co[j] = c_rk1d[stage][j];
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            int idx[4] = {P::kl(tid, e), P::kh(tid, e), N + P::kl(tid, e), N + P::kh(tid, e)};
            C v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double2 b = base[idx[q]];
                double sx = 0.0, sy = 0.0;
                // fully unrolled, so every a.K[j] is a constant index (a
                // runtime index into the argument struct would copy it to
                // scratch and turn every pointer read from it generic)
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    if (j < nk && co[j] != 0.0) {
                        C kv = a.K[j][2 * off + idx[q]];
                        sx += co[j] * (double)kv.x;
                        sy += co[j] * (double)kv.y;
                    }
                }
                v[q] = cmk<C>((T)(b.x + h0 * sx), (T)(b.y + h0 * sy));
            }
            c_eh[idx[0]] = v[0]; c_eh[idx[1]] = v[1];
            c_uh[idx[0]] = v[2]; c_uh[idx[1]] = v[3];
            P::inv_in(lds, tid, e, v[0], v[1], v[2], v[3]);
        }
    }
    P::template run<1>(lds, tw, tid);
    // ---- M0u^ = deluH1 = -i k HSS(u^)/g (two depths) or -i k C^2 u^/g (flat);
    // its inverse shares a transform with the Hilbert transform of eta
    auto hilbert_in = [&](int e, C m0l, C m0h) {
        const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
        C xl = cmk<C>((T)0, (T)0), xh = xl;
        if (R.brk) {
            T s0 = kk[i0] == (T)0 ? (T)0 : (kk[i0] > (T)0 ? (T)1 : (T)-1);
            T s1 = kk[i1] == (T)0 ? (T)0 : (kk[i1] > (T)0 ? (T)1 : (T)-1);
            xl = cmul_i<-1>(cscale(c_eh[i0], s0));
            xh = cmul_i<-1>(cscale(c_eh[i1], s1));
        }
        P::inv_in(lds, tid, e, m0l, m0h, xl, xh);
    };
#pragma unroll 1
    for (int e = 0; e < EH; ++e) {
        const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
        T etl, eth, ul, uh;
        P::inv_out(lds, tid, e, etl, eth, ul, uh);
        s_eta[i0] = etl; s_eta[i1] = eth;
        s_u[i0] = ul; s_u[i1] = uh;
        const C u0 = c_uh[i0], u1 = c_uh[i1];
        if (!R.flat && MID) {
            // the middle depth first: F^-1(C2_mid u^) is kept in s_ac
            P::inv_in(lds, tid, e, cscale(u0, HW1_OPC(i0)), cscale(u1, HW1_OPC(i1)), cmk<C>((T)0, (T)0), cmk<C>((T)0, (T)0));
        } else if (!R.flat) {
            P::inv_in(lds, tid, e, cscale(u0, opa[i0]), cscale(u1, opa[i1]), cscale(u0, opb[i0]), cscale(u1, opb[i1]));
        } else {
            const C m0l = cmul_i<-1>(cscale(u0, kk[i0] * opa[i0] / g));
            const C m0h = cmul_i<-1>(cscale(u1, kk[i1] * opa[i1] / g));
            c_m0h[i0] = m0l; c_m0h[i1] = m0h;
            hilbert_in(e, m0l, m0h);
        }
    }
    if (!R.flat) {
        if (MID) {
            P::template run<1>(lds, tw, tid);
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                T acl, ach, z0, z1;
                P::inv_out(lds, tid, e, acl, ach, z0, z1);
                HW1_AC(i0) = acl; HW1_AC(i1) = ach;
                const C u0 = c_uh[i0], u1 = c_uh[i1];
                P::inv_in(lds, tid, e, cscale(u0, opa[i0]), cscale(u1, opa[i1]), cscale(u0, opb[i0]), cscale(u1, opb[i1]));
            }
        }
        P::template run<1>(lds, tw, tid);
        // (a-, a+) = F^-1(C2-+ u^); F(g- a- + g+ a+ [+ g_mid a_mid]), F(g- u)
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            T aml, amh, apl, aph;
            P::inv_out(lds, tid, e, aml, amh, apl, aph);
            const T g0 = gmv[i0], g1 = gmv[i1], q0 = gpv[i0], q1 = gpv[i1];
            T f0 = g0 * aml + q0 * apl, f1 = g1 * amh + q1 * aph;
            if (MID) { f0 += HW1_GC(i0) * HW1_AC(i0); f1 += HW1_GC(i1) * HW1_AC(i1); }
            P::fwd_in(lds, tid, e, f0, f1, g0 * s_u[i0], g1 * s_u[i1]);
        }
        P::template run<-1>(lds, tw, tid);
        // HP = (F(g- a- + g+ a+) + C- F(g- u)) / 2; then F(g+ u), F(eta u)
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            C xl, xh, yl, yh;
            P::fwd_out(lds, tid, e, xl, xh, yl, yh);
            c_hp[i0] = cscale(cadd(xl, cscale(yl, opa[i0])), (T)0.5);
            c_hp[i1] = cscale(cadd(xh, cscale(yh, opa[i1])), (T)0.5);
            const T u0 = s_u[i0], u1 = s_u[i1];
            P::fwd_in(lds, tid, e, gpv[i0] * u0, gpv[i1] * u1, s_eta[i0] * u0, s_eta[i1] * u1);
        }
        P::template run<-1>(lds, tw, tid);
        if (MID) {
            // + C2_mid F(g_mid u) / 2 needs one more forward transform
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                C xl, xh, yl, yh;
                P::fwd_out(lds, tid, e, xl, xh, yl, yh);
                c_euh[i0] = yl; c_euh[i1] = yh;   // F(eta u)
                c_hp[i0] = cadd(c_hp[i0], cscale(xl, (T)0.5 * opb[i0]));
                c_hp[i1] = cadd(c_hp[i1], cscale(xh, (T)0.5 * opb[i1]));
                P::fwd_in(lds, tid, e, HW1_GC(i0) * s_u[i0], HW1_GC(i1) * s_u[i1], (T)0, (T)0);
            }
            P::template run<-1>(lds, tw, tid);
        }
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            C xl, xh, yl, yh;
            P::fwd_out(lds, tid, e, xl, xh, yl, yh);
            C H0, H1;
            if (MID) {
                H0 = cadd(c_hp[i0], cscale(xl, (T)0.5 * HW1_OPC(i0)));
                H1 = cadd(c_hp[i1], cscale(xh, (T)0.5 * HW1_OPC(i1)));
            } else {
                c_euh[i0] = yl; c_euh[i1] = yh;   // F(eta u)
                H0 = cadd(c_hp[i0], cscale(xl, (T)0.5 * opb[i0]));
                H1 = cadd(c_hp[i1], cscale(xh, (T)0.5 * opb[i1]));
            }
            const C m0l = cmul_i<-1>(cscale(H0, kk[i0] / g));
            const C m0h = cmul_i<-1>(cscale(H1, kk[i1] / g));
            c_m0h[i0] = m0l; c_m0h[i1] = m0h;
            hilbert_in(e, m0l, m0h);
        }
    }
    P::template run<1>(lds, tw, tid);
    // ---- M0u and H(eta); products F(eta M0u), F(u^2 - M0u^2), F(eta u) (flat)
#pragma unroll 1
    for (int e = 0; e < EH; ++e) {
        const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
        T m0l, m0h, hl, hh;
        P::inv_out(lds, tid, e, m0l, m0h, hl, hh);
        s_m0[i0] = m0l; s_m0[i1] = m0h;
        if (R.brk) { sr[R1_H * N + i0] = hl; sr[R1_H * N + i1] = hh; }
        const T pl = s_eta[i0] * m0l, ph = s_eta[i1] * m0h;
        s_p[i0] = pl; s_p[i1] = ph;
        if (!R.flat) {
            const T u0 = s_u[i0], u1 = s_u[i1];
            P::fwd_in(lds, tid, e, pl, ph, u0 * u0 - m0l * m0l, u1 * u1 - m0h * m0h);
        } else {
            P::fwd_in(lds, tid, e, s_eta[i0] * s_u[i0], s_eta[i1] * s_u[i1], pl, ph);
        }
    }
    P::template run<-1>(lds, tw, tid);
    // ---- L(eta M0u): k^2 HSS(F(eta M0u))/g (two depths) or Lop F(eta M0u) (flat),
    // left in c_hp; u2 = F(u^2 - M0u^2) in c_u2h
    if (!R.flat) {
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            C eml, emh, u2l, u2h;
            P::fwd_out(lds, tid, e, eml, emh, u2l, u2h);
            c_u2h[i0] = u2l; c_u2h[i1] = u2h;
            if (MID) {
                // the middle depth first; F(eta M0u) waits in c_hp
                c_hp[i0] = eml; c_hp[i1] = emh;
                P::inv_in(lds, tid, e, cscale(eml, HW1_OPC(i0)), cscale(emh, HW1_OPC(i1)), cmk<C>((T)0, (T)0), cmk<C>((T)0, (T)0));
            } else {
                P::inv_in(lds, tid, e, cscale(eml, opa[i0]), cscale(emh, opa[i1]), cscale(eml, opb[i0]), cscale(emh, opb[i1]));
            }
        }
        if (MID) {
            P::template run<1>(lds, tw, tid);
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                T bcl, bch, z0, z1;
                P::inv_out(lds, tid, e, bcl, bch, z0, z1);
                HW1_AC(i0) = bcl; HW1_AC(i1) = bch;
                const C eml = c_hp[i0], emh = c_hp[i1];
                P::inv_in(lds, tid, e, cscale(eml, opa[i0]), cscale(emh, opa[i1]), cscale(eml, opb[i0]), cscale(emh, opb[i1]));
            }
        }
        P::template run<1>(lds, tw, tid);
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            T bml, bmh, bpl, bph;
            P::inv_out(lds, tid, e, bml, bmh, bpl, bph);
            const T g0 = gmv[i0], g1 = gmv[i1], q0 = gpv[i0], q1 = gpv[i1];
            T f0 = g0 * bml + q0 * bpl, f1 = g1 * bmh + q1 * bph;
            if (MID) { f0 += HW1_GC(i0) * HW1_AC(i0); f1 += HW1_GC(i1) * HW1_AC(i1); }
            P::fwd_in(lds, tid, e, f0, f1, g0 * s_p[i0], g1 * s_p[i1]);
        }
        P::template run<-1>(lds, tw, tid);
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            C xl, xh, yl, yh;
            P::fwd_out(lds, tid, e, xl, xh, yl, yh);
            c_hp[i0] = cscale(cadd(xl, cscale(yl, opa[i0])), (T)0.5);
            c_hp[i1] = cscale(cadd(xh, cscale(yh, opa[i1])), (T)0.5);
            // F(g+ p) and (middle depth) F(g_mid p)
            if (MID) P::fwd_in(lds, tid, e, gpv[i0] * s_p[i0], gpv[i1] * s_p[i1], HW1_GC(i0) * s_p[i0], HW1_GC(i1) * s_p[i1]);
            else P::fwd_in(lds, tid, e, gpv[i0] * s_p[i0], gpv[i1] * s_p[i1], (T)0, (T)0);
        }
        P::template run<-1>(lds, tw, tid);
    } else {
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            C eul, euh, eml, emh;
            P::fwd_out(lds, tid, e, eul, euh, eml, emh);
            c_euh[i0] = eul; c_euh[i1] = euh;
            c_hp[i0] = cscale(eml, opb[i0]);
            c_hp[i1] = cscale(emh, opb[i1]);
            const T u0 = s_u[i0], u1 = s_u[i1], m0 = s_m0[i0], m1 = s_m0[i1];
            P::fwd_in(lds, tid, e, u0 * u0 - m0 * m0, u1 * u1 - m1 * m1, (T)0, (T)0);
        }
        P::template run<-1>(lds, tw, tid);
    }
    // ---- rhs_HsF3's third-order terms (1d/rhs.c:243-399; ifft_fft_Hsf3,
    // 1d/functions.c:1568-1640, and ifft_fft_Hsf3_v, :3440-3555):
    //   deluH3   = L'(F(eta LM0u + eta^2 dxu / 2)) - k^2 F(eta^2 M0u) / 2
    //   deletaH3 = -i k F(M0u (eta dxu + LM0u))
    // with L' the operator of LetaM0u (flat: L; two depths: k^2 HSS / g).
    // They are added to the H2 terms before adjust_delH; on return the
    // adjust inputs are in the exchange buffer.
    auto hs3_terms = [&]() {
        T* s_x = sr + R1_P2 * N;     // eta LM0u + eta^2 dxu / 2 (physical)
        T* s_d = sr + R1_FLUX * N;   // M0u (eta dxu + LM0u)
        C* c_x = c_euh;              // its transform (two depths)
        P::template run<1>(lds, tw, tid);
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            T l0, l1, d0, d1;
            P::inv_out(lds, tid, e, l0, l1, d0, d1);
            const T e0 = s_eta[i0], e1 = s_eta[i1], m0 = s_m0[i0], m1 = s_m0[i1];
            const T x0 = e0 * l0 + (T)0.5 * e0 * e0 * d0, x1 = e1 * l1 + (T)0.5 * e1 * e1 * d1;
            s_d[i0] = m0 * (e0 * d0 + l0); s_d[i1] = m1 * (e1 * d1 + l1);
            s_x[i0] = x0; s_x[i1] = x1;
            P::fwd_in(lds, tid, e, x0, x1, e0 * e0 * m0, e1 * e1 * m1);
        }
        P::template run<-1>(lds, tw, tid);
        if (R.flat) {
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                const T k0 = kk[i0], k1 = kk[i1];
                C xl, xh, al, ah;
                P::fwd_out(lds, tid, e, xl, xh, al, ah);
                c_hp[i0] = cadd(c_hp[i0], csub(cscale(xl, opb[i0]), cscale(al, (T)0.5 * k0 * k0)));
                c_hp[i1] = cadd(c_hp[i1], csub(cscale(xh, opb[i1]), cscale(ah, (T)0.5 * k1 * k1)));
                P::fwd_in(lds, tid, e, s_d[i0], s_d[i1], (T)0, (T)0);
            }
            P::template run<-1>(lds, tw, tid);
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                C dl, dh, z0, z1;
                P::fwd_out(lds, tid, e, dl, dh, z0, z1);
                const C del_ = csub(c_u2h[i0], cmul_i<1>(cscale(dl, kk[i0])));
                const C deh = csub(c_u2h[i1], cmul_i<1>(cscale(dh, kk[i1])));
                P::inv_in(lds, tid, e, c_hp[i0], c_hp[i1], del_, deh);
            }
            return;
        }
        // two depths: HSS of X^ as for LetaM0u (middle depth first when present)
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            const T k0 = kk[i0], k1 = kk[i1];
            C xl, xh, al, ah;
            P::fwd_out(lds, tid, e, xl, xh, al, ah);
            c_hp[i0] = csub(c_hp[i0], cscale(al, (T)0.5 * k0 * k0));
            c_hp[i1] = csub(c_hp[i1], cscale(ah, (T)0.5 * k1 * k1));
            c_x[i0] = xl; c_x[i1] = xh;
            if (MID)
                P::inv_in(lds, tid, e, cscale(xl, HW1_OPC(i0)), cscale(xh, HW1_OPC(i1)), cmk<C>((T)0, (T)0), cmk<C>((T)0, (T)0));
            else
                P::inv_in(lds, tid, e, cscale(xl, opa[i0]), cscale(xh, opa[i1]), cscale(xl, opb[i0]), cscale(xh, opb[i1]));
        }
        if (MID) {
            P::template run<1>(lds, tw, tid);
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                T cl, ch, z0, z1;
                P::inv_out(lds, tid, e, cl, ch, z0, z1);
                HW1_AC(i0) = cl; HW1_AC(i1) = ch;
                const C xl = c_x[i0], xh = c_x[i1];
                P::inv_in(lds, tid, e, cscale(xl, opa[i0]), cscale(xh, opa[i1]), cscale(xl, opb[i0]), cscale(xh, opb[i1]));
            }
        }
        P::template run<1>(lds, tw, tid);
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            T bml, bmh, bpl, bph;
            P::inv_out(lds, tid, e, bml, bmh, bpl, bph);
            const T g0 = gmv[i0], g1 = gmv[i1], q0 = gpv[i0], q1 = gpv[i1];
            T f0 = g0 * bml + q0 * bpl, f1 = g1 * bmh + q1 * bph;
            if (MID) { f0 += HW1_GC(i0) * HW1_AC(i0); f1 += HW1_GC(i1) * HW1_AC(i1); }
            P::fwd_in(lds, tid, e, f0, f1, g0 * s_x[i0], g1 * s_x[i1]);
        }
        P::template run<-1>(lds, tw, tid);
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            C xl, xh, yl, yh;
            P::fwd_out(lds, tid, e, xl, xh, yl, yh);
            c_x[i0] = cscale(cadd(xl, cscale(yl, opa[i0])), (T)0.5);   // HSS partial
            c_x[i1] = cscale(cadd(xh, cscale(yh, opa[i1])), (T)0.5);
            P::fwd_in(lds, tid, e, gpv[i0] * s_x[i0], gpv[i1] * s_x[i1], s_d[i0], s_d[i1]);
        }
        P::template run<-1>(lds, tw, tid);
        if (MID) {
#pragma unroll 1
            for (int e = 0; e < EH; ++e) {
                const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
                C xl, xh, dl, dh;
                P::fwd_out(lds, tid, e, xl, xh, dl, dh);
                c_x[i0] = cadd(c_x[i0], cscale(xl, (T)0.5 * opb[i0]));
                c_x[i1] = cadd(c_x[i1], cscale(xh, (T)0.5 * opb[i1]));
                c_u2h[i0] = csub(c_u2h[i0], cmul_i<1>(cscale(dl, kk[i0])));
                c_u2h[i1] = csub(c_u2h[i1], cmul_i<1>(cscale(dh, kk[i1])));
                P::fwd_in(lds, tid, e, HW1_GC(i0) * s_x[i0], HW1_GC(i1) * s_x[i1], (T)0, (T)0);
            }
            P::template run<-1>(lds, tw, tid);
        }
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            const T k0 = kk[i0], k1 = kk[i1];
            C xl, xh, dl, dh;
            P::fwd_out(lds, tid, e, xl, xh, dl, dh);
            C h0, h1, del_, deh;
            if (MID) {
                h0 = cadd(c_x[i0], cscale(xl, (T)0.5 * HW1_OPC(i0)));
                h1 = cadd(c_x[i1], cscale(xh, (T)0.5 * HW1_OPC(i1)));
                del_ = c_u2h[i0]; deh = c_u2h[i1];
            } else {
                h0 = cadd(c_x[i0], cscale(xl, (T)0.5 * opb[i0]));
                h1 = cadd(c_x[i1], cscale(xh, (T)0.5 * opb[i1]));
                del_ = csub(c_u2h[i0], cmul_i<1>(cscale(dl, k0)));
                deh = csub(c_u2h[i1], cmul_i<1>(cscale(dh, k1)));
            }
            const C dul = cadd(c_hp[i0], cscale(h0, k0 * k0 / g));
            const C duh = cadd(c_hp[i1], cscale(h1, k1 * k1 / g));
            P::inv_in(lds, tid, e, dul, duh, del_, deh);
        }
    };
    // ---- delH2 terms (1d/rhs.c:141-147) and their adjustment (adjust_delH)
#pragma unroll 1
    for (int e = 0; e < EH; ++e) {
        const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
        const T k0 = kk[i0], k1 = kk[i1];
        C lel, leh, u2l, u2h;
        if (!R.flat) {
            C xl, xh, yl, yh;
            P::fwd_out(lds, tid, e, xl, xh, yl, yh);
            C hl = cadd(c_hp[i0], cscale(xl, (T)0.5 * opb[i0]));
            C hh = cadd(c_hp[i1], cscale(xh, (T)0.5 * opb[i1]));
            if (MID) { hl = cadd(hl, cscale(yl, (T)0.5 * HW1_OPC(i0))); hh = cadd(hh, cscale(yh, (T)0.5 * HW1_OPC(i1))); }
            lel = cscale(hl, k0 * k0 / g);
            leh = cscale(hh, k1 * k1 / g);
            u2l = c_u2h[i0]; u2h = c_u2h[i1];
        } else {
            C yl, yh;
            P::fwd_out(lds, tid, e, u2l, u2h, yl, yh);
            lel = c_hp[i0]; leh = c_hp[i1];
        }
        const C eul = c_euh[i0], euh = c_euh[i1];
        C dul = csub(cmk<C>(-lel.x, -lel.y), cmul_i<1>(cscale(eul, k0)));
        C duh = csub(cmk<C>(-leh.x, -leh.y), cmul_i<1>(cscale(euh, k1)));
        C del_ = cscale(cmul_i<1>(cscale(u2l, k0)), (T)-0.5);
        C deh = cscale(cmul_i<1>(cscale(u2h, k1)), (T)-0.5);
        if (R.order == 1) {   // rhs_HsF1: no nonlinear terms
            dul = duh = del_ = deh = cmk<C>((T)0, (T)0);
        }
        if (R.order == 3) {
            // rhs_HsF3: the H2 terms wait in c_hp / c_u2h; LM0u^ (= lel) and
            // dxu^ = i k u^ are inverted together
            c_hp[i0] = dul; c_hp[i1] = duh; c_u2h[i0] = del_; c_u2h[i1] = deh;
            P::inv_in(lds, tid, e, lel, leh, cmul_i<1>(cscale(c_uh[i0], k0)), cmul_i<1>(cscale(c_uh[i1], k1)));
        } else {
            P::inv_in(lds, tid, e, dul, duh, del_, deh);
        }
    }
    if (R.order == 3) hs3_terms();
    P::template run<1>(lds, tw, tid);
#pragma unroll 1
    for (int e = 0; e < EH; ++e) {
        const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
        T aul, auh, ael, aeh;
        P::inv_out(lds, tid, e, aul, auh, ael, aeh);
        const T c0 = a.chi[off + i0], c1 = a.chi[off + i1];
        s_au[i0] = aul * c0; s_au[i1] = auh * c1;
        s_ae[i0] = ael * c0; s_ae[i1] = aeh * c1;
    }
    // ---- breaking source Rb (rhs_breaking, 1d/rhs.c:617-657)
    T* Bp = a.B + off;
    T* flux = sr + R1_FLUX * N;
    if (R.brk) {
        brk1d_process<T, N, NT>(Brk1DView<T>{a.real, a.depth, a.peaks, a.cb, a.cbn, a.cbp, a.ide, a.bs}, R, t, s_eta, s_u,
                                  sr + R1_H * N, Bp, sm, smd);
        const T db2 = (T)(R.delb * R.delb);
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
            const T d0 = s_m0[i0] + s_au[i0], d1 = s_m0[i1] + s_au[i1];   // Re F^-1(deluH12)
            flux[i0] = -Bp[i0] * db2 * (a.depth[off + i0] + s_eta[i0]) * d0 * d0;
            flux[i1] = -Bp[i1] * db2 * (a.depth[off + i1] + s_eta[i1]) * d1 * d1;
        }
        __syncthreads();
    }
    // ---- wall source (fun_wallinfluxsource, 1d/rhs.c:1216-1239): the
    // trapezoid integral of refl * eta(x_wall) over the times of successive
    // calls, advanced only by a call at a later time than the last one
    if (R.wall && tid == 0) {
        double* W = a.wst + 4 * r;
        if (t > W[0]) {
            const double se = R.refl * (double)s_eta[R.idx_wall];
            W[1] = W[3] < 1.0 ? 0.0 : W[1] + (se + W[2]) * (t - W[0]) / 2;
            W[3] += 1.0;
            W[0] = t;
            W[2] = se;
        }
        smd[32] = 2.0 * W[1];
    }
    const T idx1 = (T)(1.0 / R.dx), idx2 = (T)(0.5 / R.dx);
    auto gradf = [&](int i) -> T {
        if (i == 0) return (flux[1] - flux[0]) * idx1;
        if (i == N - 1) return (flux[N - 1] - flux[N - 2]) * idx1;
        return (flux[i + 1] - flux[i - 1]) * idx2;
    };
    // ---- output transforms with damping folded in, forcing, mask
#pragma unroll 1
    for (int e = 0; e < EH; ++e) {
        const int i0 = P::kl(tid, e), i1 = P::kh(tid, e);
        const T d0 = a.damp[off + i0], d1 = a.damp[off + i1];
        T rbl = (T)0, rbh = (T)0;
        if (R.brk) {
            rbl = gradf(i0) / (a.depth[off + i0] + s_eta[i0]);
            rbh = gradf(i1) / (a.depth[off + i1] + s_eta[i1]);
        }
        P::fwd_in(lds, tid, e, s_au[i0] - d0 * s_eta[i0], s_au[i1] - d1 * s_eta[i1],
                  s_ae[i0] - d0 * s_u[i0] + rbl, s_ae[i1] - d1 * s_u[i1] + rbh);
    }
    P::template run<-1>(lds, tw, tid);
    {
        const T pv = (T)smd[30], pvs = (T)smd[31];
        const T sw = R.wall ? (T)smd[32] : (T)0;
        C* Ko;
        switch (stage) {   // constant indices into the argument struct (see the stage input)
            case 1: Ko = a.K[0]; break;
            case 2: Ko = a.K[1]; break;
            case 3: Ko = a.K[2]; break;
            case 4: Ko = a.K[3]; break;
            case 5: Ko = a.K[4]; break;
            case 6: Ko = a.K[5]; break;
            default: Ko = a.K[6]; break;
        }
        Ko += 2 * off;
#pragma unroll 1
        for (int e = 0; e < EH; ++e) {
            C oel, oeh, oul, ouh;
            P::fwd_out(lds, tid, e, oel, oeh, oul, ouh);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int i = q == 0 ? P::kl(tid, e) : P::kh(tid, e);
                C fo = cmk<C>((T)0, (T)0);
                if (R.src) {
                    C gx = a.gamx[off + i], gs = a.gamxs[off + i];
                    if (R.prop == 0) fo = cmk<C>(pv * gx.x - pvs * gs.x, pv * gx.y - pvs * gs.y);
                    else if (R.prop == 1) fo = cmk<C>(pv * gx.x + pvs * gs.x, pv * gx.y + pvs * gs.y);
                    else if (R.prop == 2) fo = cmk<C>((T)2 * pv * gx.x, (T)2 * pv * gx.y);
                }
                const T al = a.aal[off + i];
                const C oe = q == 0 ? oel : oeh;
                const C ou = q == 0 ? oul : ouh;
                if (R.wall) fo = cadd(fo, cscale(a.gamws[off + i], sw));
                C ke = cscale(cadd(cadd(c_m0h[i], fo), oe), al);
                C ku = cscale(cadd(cmul_i<-1>(cscale(c_eh[i], g * kk[i])), ou), al);
                Ko[i] = ke;
                Ko[N + i] = ku;
            }
        }
    }
}
#undef HW1_OPC
#undef HW1_GC
#undef HW1_AC

// One size class of the batch: the realizations of n = N (MID: with a
// third reference depth) listed at a.list[list0 ..], one block each.  Each
// class has a kernel of its own, so the body is compiled once per size with
// its own register allocation (one kernel switching over every size spilled
// several times more), and the classes of a batch run side by side on their
// own streams.
template <typename T, int NT, int N, bool MID>
__global__ void __launch_bounds__(NT) k1d_rhs_n(Batch1DArgs<T> a, int stage, int list0) {
    using C = typename ctype<T>::type;
    extern __shared__ unsigned char hw_lds1d[];
    __shared__ int sm[64];
    __shared__ double smd[40];
    const int r = a.list[list0 + (int)blockIdx.x];
    Rk1D* S = a.rk + r;
    if (stage >= 1) {
        if (!S->active) return;
        if (stage == 1 && !S->need_in) return;
    }
    rhs1d_body<T, N, NT, MID>(a, r, stage, reinterpret_cast<C*>(hw_lds1d), sm, smd);
}

// the block size of size class n: two elements per thread up to n = 2048,
// then 1024 threads with up to 16 elements each
constexpr int nt_for_n(int n) { return n / 2 < 1024 ? n / 2 : 1024; }

// ---------------------------------------------------- step control ----
// Per realization: y_new = y + h0 (c1 K1 + c3 K3 + c4 K4 + c5 K5 + c6 K6),
// err = h0 (ec1 K1 + ...), rmax = max |err| / (eps_abs + eps_rel |y_new|),
// then control_y's hadjust and gsl_odeiv_evolve_apply's accept / retry rule;
// on acceptance y <- y_new, t advances and the next evolve_apply call (which
// evaluates dydt_in) is armed; the next attempt's step is clipped to t1.
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k1d_control(Batch1DArgs<T> a, int list0) {
    using C = typename ctype<T>::type;
    __shared__ double red[20];
    __shared__ int dec[2];
    const int r = a.list[list0 + (int)blockIdx.x];
    const Real1D& R = a.real[r];
    Rk1D* S = a.rk + r;
    if (!S->active) return;
    const int n2 = 2 * R.n;
    const long long o = 2 * R.off;
    const double h0 = S->h0;
    const C *K1 = a.K[0] + o, *K3 = a.K[2] + o, *K4 = a.K[3] + o, *K5 = a.K[4] + o, *K6 = a.K[5] + o;
    double2* y = a.y + o;
    double rmax = 2.2250738585072014e-308;
    for (int i = threadIdx.x; i < n2; i += NT) {
        C k1 = K1[i], k3 = K3[i], k4 = K4[i], k5 = K5[i], k6 = K6[i];
        double2 yv = y[i];
        double dxv = c_rk1d[7][0] * (double)k1.x + c_rk1d[7][2] * (double)k3.x + c_rk1d[7][3] * (double)k4.x +
                     c_rk1d[7][4] * (double)k5.x + c_rk1d[7][5] * (double)k6.x;
        double dyv = c_rk1d[7][0] * (double)k1.y + c_rk1d[7][2] * (double)k3.y + c_rk1d[7][3] * (double)k4.y +
                     c_rk1d[7][4] * (double)k5.y + c_rk1d[7][5] * (double)k6.y;
        double nx = yv.x + h0 * dxv, ny = yv.y + h0 * dyv;
        double ex = h0 * (c_rk1d_ec[0] * (double)k1.x + c_rk1d_ec[2] * (double)k3.x + c_rk1d_ec[3] * (double)k4.x +
                          c_rk1d_ec[4] * (double)k5.x + c_rk1d_ec[5] * (double)k6.x);
        double ey = h0 * (c_rk1d_ec[0] * (double)k1.y + c_rk1d_ec[2] * (double)k3.y + c_rk1d_ec[3] * (double)k4.y +
                          c_rk1d_ec[4] * (double)k5.y + c_rk1d_ec[5] * (double)k6.y);
        double rx = fabs(ex) / fabs(R.eps_rel * fabs(nx) + R.eps_abs);
        double ry = fabs(ey) / fabs(R.eps_rel * fabs(ny) + R.eps_abs);
        if (rx > rmax) rmax = rx;
        if (ry > rmax) rmax = ry;
    }
    rmax = block_max_d<NT>(rmax, red);
    if (threadIdx.x == 0) {
        const double t0 = S->t;
        double hn = h0;
        int st = 0;
        if (rmax > 1.1) {
            double q = 0.9 / pow(rmax, 1.0 / 5.0);
            if (q < 0.2) q = 0.2;
            hn = q * h0;
            st = -1;
        } else if (rmax < 0.5) {
            double q = 0.9 / pow(rmax, 1.0 / 6.0);
            if (q > 5.0) q = 5.0;
            if (q < 1.0) q = 1.0;
            hn = q * h0;
            st = 1;
        }
        const double tn = S->final_step ? S->t1 : t0 + h0;
        int accept = 1;
        if (st < 0) {
            if (fabs(hn) < fabs(h0) && tn + hn != tn) accept = 0;
            else hn = h0;
        }
        if (accept) {
            S->t = tn;
            S->h = hn;
            S->need_in = 1;
            S->accepted += 1;
            S->active = tn < S->t1;
            hn = S->h;
        } else {
            S->rejected += 1;
            S->need_in = 0;
        }
        // arm the next attempt: h0 clipped to the remaining interval
        double dt = S->t1 - S->t;
        if ((dt >= 0.0 && hn > dt) || (dt < 0.0 && hn < dt)) { S->h0 = dt; S->final_step = 1; }
        else { S->h0 = hn; S->final_step = 0; }
        dec[0] = accept;
    }
    __syncthreads();
    if (dec[0]) {
        for (int i = threadIdx.x; i < n2; i += NT) {
            C k1 = K1[i], k3 = K3[i], k4 = K4[i], k5 = K5[i], k6 = K6[i];
            double2 yv = y[i];
            double dxv = c_rk1d[7][0] * (double)k1.x + c_rk1d[7][2] * (double)k3.x + c_rk1d[7][3] * (double)k4.x +
                         c_rk1d[7][4] * (double)k5.x + c_rk1d[7][5] * (double)k6.x;
            double dyv = c_rk1d[7][0] * (double)k1.y + c_rk1d[7][2] * (double)k3.y + c_rk1d[7][3] * (double)k4.y +
                         c_rk1d[7][4] * (double)k5.y + c_rk1d[7][5] * (double)k6.y;
            y[i] = make_double2(yv.x + h0 * dxv, yv.y + h0 * dyv);
        }
    }
}

// Start of a driver interval: every realization with t < t1 becomes active,
// and its first attempt uses the suggested step clipped to t1.
static __global__ void k1d_begin(Rk1D* rk, int nreal, double t1) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreal) return;
    Rk1D* S = rk + r;
    S->t1 = t1;
    S->active = S->t < t1;
    S->need_in = 1;
    double dt = t1 - S->t, h = S->h;
    if ((dt >= 0.0 && h > dt) || (dt < 0.0 && h < dt)) { S->h0 = dt; S->final_step = 1; }
    else { S->h0 = h; S->final_step = 0; }
}

// active count and crest-list overflows (for the host's loop condition):
// out[0] = realizations still short of t1, out[1] = realizations whose
// breaking crest list overflowed
static __global__ void k1d_count(const Rk1D* rk, const Brk1D* bs, int nreal, int* out) {
    __shared__ int c, o;
    if (threadIdx.x == 0) { c = 0; o = 0; }
    __syncthreads();
    for (int r = threadIdx.x; r < nreal; r += blockDim.x) {
        if (rk[r].active) atomicAdd(&c, 1);
        if (bs[r].overflow) atomicAdd(&o, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) { out[0] = c; out[1] = o; }
}

}  // namespace hw
