// MRG32k3a (L'Ecuyer 1999; streams and substreams of L'Ecuyer, Simard, Chen and
// Kelton, Oper. Res. 50(6) 2002), the generator behind the reference's RngStream
// (RngStream.cpp:266-289). State: 6 integers, (s0 s1 s2) mod m1 and (s3 s4 s5) mod m2.
//
// Host side: exact jump-ahead matrices A^n mod m (3x3, 128-bit products).
// Device side: one draw, and applying a jump matrix to a state, with 32x32->64-bit
// products reduced by m = 2^32 - c folding. The draw returns the same double as
// RngStream::U01: ((p1 > p2) ? p1 - p2 : p1 - p2 + m1) * (1 / (m1 + 1)).
#pragma once
#include <stdint.h>

namespace mrg {

constexpr uint32_t M1 = 4294967087u;
constexpr uint32_t M2 = 4294944443u;
constexpr uint32_t C1 = 209u;     // 2^32 - M1
constexpr uint32_t C2 = 22853u;   // 2^32 - M2
constexpr uint32_t A12 = 1403580u;
constexpr uint32_t A13N = 810728u;
constexpr uint32_t A21 = 527612u;
constexpr uint32_t A23N = 1370589u;

// a 3x3 jump matrix per component, entries reduced mod m
struct Jump {
    uint32_t m[2][9];
};

// ---------------------------------------------------------------- host
inline uint32_t h_mod(unsigned __int128 v, uint32_t m) { return (uint32_t)(v % m); }

inline void h_matmul(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t m) {
    uint32_t T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            unsigned __int128 s = 0;
            for (int k = 0; k < 3; ++k) s += (unsigned __int128)A[3 * i + k] * B[3 * k + j];
            T[3 * i + j] = h_mod(s, m);
        }
    for (int i = 0; i < 9; ++i) C[i] = T[i];
}

inline void h_one_step(int comp, uint32_t *A) {
    for (int i = 0; i < 9; ++i) A[i] = 0;
    A[1] = 1;
    A[5] = 1;
    if (comp == 0) {
        A[6] = M1 - A13N;
        A[7] = A12;
    } else {
        A[6] = M2 - A23N;
        A[8] = A21;
    }
}

inline Jump h_identity() {
    Jump J{};
    for (int c = 0; c < 2; ++c) J.m[c][0] = J.m[c][4] = J.m[c][8] = 1;
    return J;
}

inline Jump h_mul(const Jump &A, const Jump &B) {
    Jump C;
    h_matmul(A.m[0], B.m[0], C.m[0], M1);
    h_matmul(A.m[1], B.m[1], C.m[1], M2);
    return C;
}

// A^n for both components
inline Jump h_pow(unsigned long long n) {
    Jump R = h_identity(), W;
    h_one_step(0, W.m[0]);
    h_one_step(1, W.m[1]);
    while (n) {
        if (n & 1ull) R = h_mul(W, R);
        W = h_mul(W, W);
        n >>= 1;
    }
    return R;
}

// W^n for a jump W
inline Jump h_pow_of(Jump W, unsigned long long n) {
    Jump R = h_identity();
    while (n) {
        if (n & 1ull) R = h_mul(W, R);
        W = h_mul(W, W);
        n >>= 1;
    }
    return R;
}

// A^(n * 2^127): the jump from stream k to stream k+n of one package seed (each new
// RngStream starts 2^127 draws after the previous one, RngStream.cpp:327-345)
inline Jump h_stream_jump(unsigned long long n) {
    Jump W = h_pow(1);
    for (int i = 0; i < 127; ++i) W = h_mul(W, W);
    return h_pow_of(W, n);
}

inline void h_apply(const Jump &J, const uint32_t *s, uint32_t *out) {
    uint32_t t[6];
    for (int c = 0; c < 2; ++c) {
        const uint32_t m = c == 0 ? M1 : M2;
        for (int i = 0; i < 3; ++i) {
            unsigned __int128 acc = 0;
            for (int k = 0; k < 3; ++k) acc += (unsigned __int128)J.m[c][3 * i + k] * s[3 * c + k];
            t[3 * c + i] = h_mod(acc, m);
        }
    }
    for (int i = 0; i < 6; ++i) out[i] = t[i];
}

inline bool h_valid_state(const uint32_t *s) {
    for (int i = 0; i < 3; ++i)
        if (s[i] >= M1) return false;
    for (int i = 3; i < 6; ++i)
        if (s[i] >= M2) return false;
    if (s[0] == 0 && s[1] == 0 && s[2] == 0) return false;
    if (s[3] == 0 && s[4] == 0 && s[5] == 0) return false;
    return true;
}

#ifdef __HIPCC__
// ---------------------------------------------------------------- device
// x mod m for x < 2^64, m = 2^32 - c, c < 2^15: two folds then one subtraction
__device__ __forceinline__ uint32_t d_mod(uint64_t x, uint32_t c, uint32_t m) {
    x = (x >> 32) * c + (x & 0xffffffffull);
    x = (x >> 32) * c + (x & 0xffffffffull);
    if (x >= m) x -= m;
    return (uint32_t)x;
}

__device__ __forceinline__ void d_apply_comp(const uint32_t *M, uint32_t *s, uint32_t c, uint32_t m) {
    uint32_t t[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        uint64_t acc = (uint64_t)d_mod((uint64_t)M[3 * i] * s[0], c, m) +
                       (uint64_t)d_mod((uint64_t)M[3 * i + 1] * s[1], c, m) +
                       (uint64_t)d_mod((uint64_t)M[3 * i + 2] * s[2], c, m);
        t[i] = d_mod(acc, c, m);
    }
    s[0] = t[0];
    s[1] = t[1];
    s[2] = t[2];
}

// state <- J * state
__device__ __forceinline__ void d_apply(const Jump &J, uint32_t *s) {
    d_apply_comp(J.m[0], s, C1, M1);
    d_apply_comp(J.m[1], s + 3, C2, M2);
}

// one RngStream::U01 draw
__device__ __forceinline__ double d_u01(uint32_t *s) {
    const double norm = 1.0 / (4294967087.0 + 1.0);
    uint64_t t1 = (uint64_t)d_mod((uint64_t)A12 * s[1], C1, M1) + d_mod((uint64_t)(M1 - A13N) * s[0], C1, M1);
    uint32_t p1 = (uint32_t)(t1 >= M1 ? t1 - M1 : t1);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = p1;
    uint64_t t2 = (uint64_t)d_mod((uint64_t)A21 * s[5], C2, M2) + d_mod((uint64_t)(M2 - A23N) * s[3], C2, M2);
    uint32_t p2 = (uint32_t)(t2 >= M2 ? t2 - M2 : t2);
    s[3] = s[4];
    s[4] = s[5];
    s[5] = p2;
    const long long d = (long long)p1 - (long long)p2;
    return (p1 > p2) ? (double)d * norm : (double)(d + (long long)M1) * norm;
}
#endif

}  // namespace mrg
