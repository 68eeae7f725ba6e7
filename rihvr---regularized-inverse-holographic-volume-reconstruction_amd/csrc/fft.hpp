// fft.hpp -- hand-written LDS Stockham FFT for gfx950 (device templates).
//
// One transform of length L is worked on by TPT threads of a workgroup; the
// workgroup holds several transforms (rows of a plane, or a 16-column strip).
// Passes are radix-R Stockham autosort steps (decimation in time): pass p with
// Ns = product of the previous radices takes butterfly b (0 <= b < L/R),
// loads elements b + r*L/R, multiplies them by W_{Ns*R}^{r*(b mod Ns)}, runs a
// radix-R DFT and stores to (b/Ns)*Ns*R + (b mod Ns) + r*Ns.  The first pass
// loads through a caller functor (prologue: global loads fused with the
// transfer function or phase factor); the last pass stores through a caller
// functor (epilogue: global stores, phase factor, z-sum accumulation).
// Intermediate passes exchange data through one LDS buffer per transform.
#pragma once
#include <utility>

#include "device_common.hpp"

namespace rihvr {

// ---------------------------------------------------------------------------
// Radix plans.  TPT = L / R0 threads per transform (one first-pass butterfly
// per thread).  Radices 2, 4, 8, 16 have hand-written butterflies; any other
// radix (3, 5, 7, 9, 13) uses a direct DFT with roots read from the twiddle
// table (W_R^m = W_L^{m L / R}).
// ---------------------------------------------------------------------------
template <int... R> struct Radices {};

template <int L> struct FftPlan;
#define RIHVR_PLAN(LEN, ...)                                              \
    template <> struct FftPlan<LEN> {                                     \
        using radices = Radices<__VA_ARGS__>;                             \
        static constexpr int first = FirstOf<__VA_ARGS__>::value;         \
        static constexpr int TPT = LEN / first;                           \
    };

template <int R0, int... Rest> struct FirstOf { static constexpr int value = R0; };

RIHVR_PLAN(16, 16)
RIHVR_PLAN(32, 16, 2)
RIHVR_PLAN(48, 16, 3)
RIHVR_PLAN(64, 8, 8)
RIHVR_PLAN(80, 16, 5)
RIHVR_PLAN(96, 16, 2, 3)
RIHVR_PLAN(128, 16, 8)
RIHVR_PLAN(192, 16, 4, 3)
RIHVR_PLAN(256, 16, 16)
RIHVR_PLAN(320, 16, 4, 5)
RIHVR_PLAN(384, 16, 8, 3)
RIHVR_PLAN(448, 7, 8, 8)
RIHVR_PLAN(512, 8, 8, 8)
RIHVR_PLAN(576, 9, 8, 8)
RIHVR_PLAN(640, 16, 8, 5)
RIHVR_PLAN(768, 16, 16, 3)
RIHVR_PLAN(832, 13, 8, 8)
RIHVR_PLAN(896, 16, 8, 7)
RIHVR_PLAN(1024, 32, 32)
#undef RIHVR_PLAN

// Plans for row transforms (one transform per 1 or more waves, several transforms per
// workgroup).  Where the row plan has a smaller first radix than the column plan, a
// thread holds fewer elements, so the row kernels -- which keep per-element data in
// registers across planes -- fit more waves per SIMD.  Column kernels use FftPlan (their
// workgroup of 16 transforms must stay within 1024 threads), 32 x 32 at 1024 (one LDS
// exchange instead of two).
template <int L> struct RowPlan : FftPlan<L> {};
template <> struct RowPlan<1024> {
    using radices = Radices<8, 8, 8, 2>;
    static constexpr int first = 8;
    static constexpr int TPT = 1024 / 8;
};

// Multiply by W4 = e^{SIGN * i pi/2}: -i for the forward transform, +i inverse.
template <int SIGN> __device__ __forceinline__ float2 mul_w4(float2 a) {
    return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int SIGN> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    float2 t2 = cadd(a1, a3), t3 = mul_w4<SIGN>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// multiply by W_16^m (m in 0..15) with compile-time m
template <int SIGN, int M> __device__ __forceinline__ float2 mul_w16(float2 a) {
    constexpr float c16[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                               0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                               -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                               0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
    constexpr float s16[16] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                               1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                               0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                               -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f};
    constexpr int m = M & 15;
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) return mul_w4<SIGN>(a);
    else if constexpr (m == 8) return make_float2(-a.x, -a.y);
    else if constexpr (m == 12) return mul_w4<-SIGN>(a);
    else {
        constexpr float c = c16[m];
        constexpr float s = SIGN < 0 ? -s16[m] : s16[m];
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <int R, int SIGN, int L> struct Dft;

template <int SIGN, int L> struct Dft<2, SIGN, L> {
    __device__ static __forceinline__ void run(float2 (&v)[2], const float2*) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <int SIGN, int L> struct Dft<4, SIGN, L> {
    __device__ static __forceinline__ void run(float2 (&v)[4], const float2*) { dft4<SIGN>(v[0], v[1], v[2], v[3]); }
};
template <int SIGN, int L> struct Dft<8, SIGN, L> {
    __device__ static __forceinline__ void run(float2 (&v)[8], const float2*) {
        // X[k1 + 2 k2] ... as two radix-4 DFTs (even/odd inputs) + radix-2
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<SIGN>(e0, e1, e2, e3);
        dft4<SIGN>(o0, o1, o2, o3);
        o1 = mul_w16<SIGN, 2>(o1);
        o2 = mul_w16<SIGN, 4>(o2);
        o3 = mul_w16<SIGN, 6>(o3);
        v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
    }
};
template <int SIGN, int L> struct Dft<16, SIGN, L> {
    // 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
    __device__ static __forceinline__ void run(float2 (&v)[16], const float2*) {
        dft4<SIGN>(v[0], v[4], v[8], v[12]);
        dft4<SIGN>(v[1], v[5], v[9], v[13]);
        dft4<SIGN>(v[2], v[6], v[10], v[14]);
        dft4<SIGN>(v[3], v[7], v[11], v[15]);
        // v[n2 + 4 k1] now holds A_{n2}[k1]; twiddle W16^{n2 k1}
        v[5] = mul_w16<SIGN, 1>(v[5]);
        v[9] = mul_w16<SIGN, 2>(v[9]);
        v[13] = mul_w16<SIGN, 3>(v[13]);
        v[6] = mul_w16<SIGN, 2>(v[6]);
        v[10] = mul_w16<SIGN, 4>(v[10]);
        v[14] = mul_w16<SIGN, 6>(v[14]);
        v[7] = mul_w16<SIGN, 3>(v[7]);
        v[11] = mul_w16<SIGN, 6>(v[11]);
        v[15] = mul_w16<SIGN, 9>(v[15]);
        dft4<SIGN>(v[0], v[1], v[2], v[3]);
        dft4<SIGN>(v[4], v[5], v[6], v[7]);
        dft4<SIGN>(v[8], v[9], v[10], v[11]);
        dft4<SIGN>(v[12], v[13], v[14], v[15]);
        // v[4 k1 + k2] holds X[k1 + 4 k2]; transpose to natural order
        float2 t[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = t[i];
    }
};
// multiply by W_32^m (m in 0..15) with compile-time m: even m are W_16^(m/2)
template <int SIGN, int M> __device__ __forceinline__ float2 mul_w32(float2 a) {
    if constexpr ((M & 1) == 0) {
        return mul_w16<SIGN, M / 2>(a);
    } else {
        // cos and sin of 2 pi m / 32 for odd m = 1, 3, ..., 15
        constexpr float c32[8] = {0.98078528040323043f, 0.83146961230254524f, 0.55557023301960218f,
                                  0.19509032201612825f, -0.19509032201612825f, -0.55557023301960218f,
                                  -0.83146961230254524f, -0.98078528040323043f};
        constexpr float s32[8] = {0.19509032201612825f, 0.55557023301960218f, 0.83146961230254524f,
                                  0.98078528040323043f, 0.98078528040323043f, 0.83146961230254524f,
                                  0.55557023301960218f, 0.19509032201612825f};
        constexpr float c = c32[M / 2];
        constexpr float s = SIGN < 0 ? -s32[M / 2] : s32[M / 2];
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}
template <int SIGN, int L> struct Dft<32, SIGN, L> {
    // radix 2 over two radix-16 DFTs: X[k] = E[k] + W32^k O[k], X[k + 16] = E[k] - W32^k O[k]
    template <int... K>
    __device__ static __forceinline__ void combine(float2 (&v)[32], const float2 (&e)[16], float2 (&o)[16],
                                                   std::integer_sequence<int, K...>) {
        ((o[K] = mul_w32<SIGN, K>(o[K])), ...);
        ((v[K] = cadd(e[K], o[K]), v[K + 16] = csub(e[K], o[K])), ...);
    }
    __device__ static __forceinline__ void run(float2 (&v)[32], const float2* tw) {
        float2 e[16], o[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        Dft<16, SIGN, L>::run(e, tw);
        Dft<16, SIGN, L>::run(o, tw);
        combine(v, e, o, std::make_integer_sequence<int, 16>{});
    }
};
// generic direct DFT (odd radices); roots from the L-point twiddle table
template <int R, int SIGN, int L> struct Dft {
    __device__ static __forceinline__ void run(float2 (&v)[R], const float2* tw) {
        float2 o[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            float2 acc = v[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[((r * k) % R) * (L / R)];
                acc = cadd(acc, cmul(v[r], w));
            }
            o[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = o[k];
    }
};

// ---------------------------------------------------------------------------
// Pass recursion.  src(i, r, j) returns element j of the transform for the
// first pass; dst(i, r, j, v) receives output element j of the last pass;
// lds(j) is this transform's LDS slot j.  tw: W_L^m table in LDS, m < L.
// TOFF < 0: every twiddle is read from tw.  TOFF >= 0: the pass's twiddles are
// tc[TOFF ...] (registers filled once by TwCache::load, for kernels that run the
// same transform positions over many planes: no LDS reads or index arithmetic
// per plane, the same values).
// ---------------------------------------------------------------------------
template <int TOFF, int L, int TPT, int SIGN, int NS, int R, int... REST> struct FftPasses {
    static constexpr int NBF = L / R;
    static constexpr int NB = (NBF + TPT - 1) / TPT;
    static constexpr bool EVEN = (NBF % TPT) == 0;
    // twiddles this pass takes from tc, and the next pass's offset
    static constexpr int NTW = NS > 1 ? NB * (R - 1) : 0;
    static constexpr int NEXT = TOFF < 0 ? -1 : TOFF + NTW;

    template <class Src, class Dst, class Lds>
    __device__ static __forceinline__ void run(const float2* __restrict__ tw, const float2* tc, int t, Src&& src,
                                               Dst&& dst, Lds&& lds) {
        float2 v[NB][R];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int b = t + i * TPT;
            if (EVEN || b < NBF) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = src(i, r, b + r * NBF);
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int b = t + i * TPT;
            if (EVEN || b < NBF) {
                const int k = b % NS;
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        if constexpr (TOFF >= 0)
                            v[i][r] = cmul(v[i][r], tc[TOFF + i * (R - 1) + (r - 1)]);
                        else
                            v[i][r] = cmul(v[i][r], tw[(r * k) * (L / (NS * R))]);
                    }
                }
                Dft<R, SIGN, L>::run(v[i], tw);
                const int base = (b / NS) * NS * R + k;
                if constexpr (sizeof...(REST) == 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r) dst(i, r, base + r * NS, v[i][r]);
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) lds(base + r * NS) = v[i][r];
                }
            }
        }
        if constexpr (sizeof...(REST) > 0) {
            __syncthreads();
            FftPasses<NEXT, L, TPT, SIGN, NS * R, REST...>::run(
                tw, tc, t, [&](int, int, int j) -> float2 { return lds(j); }, dst, lds);
        }
    }

    // the twiddles run() takes from tc at this pass and the later ones, in its order
    __device__ static __forceinline__ void load(const float2* __restrict__ tw, int t, float2* tc) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int b = t + i * TPT;
                const int k = (EVEN || b < NBF) ? b % NS : 0;
#pragma unroll
                for (int r = 1; r < R; ++r) tc[TOFF + i * (R - 1) + (r - 1)] = tw[(r * k) * (L / (NS * R))];
            }
        }
        if constexpr (sizeof...(REST) > 0) FftPasses<NEXT, L, TPT, SIGN, NS * R, REST...>::load(tw, t, tc);
    }
};

template <int L, int TPT, int NS, int... R> struct TwCount;
template <int L, int TPT, int NS> struct TwCount<L, TPT, NS> { static constexpr int value = 0; };
template <int L, int TPT, int NS, int R, int... REST> struct TwCount<L, TPT, NS, R, REST...> {
    static constexpr int NB = (L / R + TPT - 1) / TPT;
    static constexpr int value = (NS > 1 ? NB * (R - 1) : 0) + TwCount<L, TPT, NS * R, REST...>::value;
};

template <int L, int SIGN, class Rad, int TPT = FftPlan<L>::TPT> struct FftRun;
template <int L, int SIGN, int... R, int TPT> struct FftRun<L, SIGN, Radices<R...>, TPT> {
    template <class Src, class Dst, class Lds>
    __device__ static __forceinline__ void run(const float2* tw, int t, Src&& src, Dst&& dst, Lds&& lds) {
        FftPasses<-1, L, TPT, SIGN, 1, R...>::run(tw, nullptr, t, src, dst, lds);
    }
};

// A thread's twiddles of one transform position held in registers (N complex values), for
// kernels that loop over planes with the same positions: load() once after the table is in
// LDS, then run() per plane.
template <int L, int SIGN, class Rad, int TPT> struct TwCache;
template <int L, int SIGN, int... R, int TPT> struct TwCache<L, SIGN, Radices<R...>, TPT> {
    static constexpr int N = TwCount<L, TPT, 1, R...>::value;
    float2 w[N > 0 ? N : 1];
    __device__ __forceinline__ void load(const float2* tw, int t) {
        FftPasses<0, L, TPT, SIGN, 1, R...>::load(tw, t, w);
    }
    template <class Src, class Dst, class Lds>
    __device__ __forceinline__ void run(const float2* tw, int t, Src&& src, Dst&& dst, Lds&& lds) const {
        FftPasses<0, L, TPT, SIGN, 1, R...>::run(tw, w, t, src, dst, lds);
    }
};

// Number of output elements a thread holds in the last pass, and the position
// helper for register accumulators (used by the z-sum epilogue).
template <class Rad> struct LastRadix;
template <int R0> struct LastRadix<Radices<R0>> { static constexpr int value = R0; };
template <int R0, int R1, int... Rest> struct LastRadix<Radices<R0, R1, Rest...>> {
    static constexpr int value = LastRadix<Radices<R1, Rest...>>::value;
};
template <int L, class Plan = FftPlan<L>> struct LastPass {
    static constexpr int R = LastRadix<typename Plan::radices>::value;
    static constexpr int TPT = Plan::TPT;
    static constexpr int NBF = L / R;
    static constexpr int NB = (NBF + TPT - 1) / TPT;
};

}  // namespace rihvr
