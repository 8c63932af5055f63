/*
 * C restatement of the reference's Hadamard transform (test infrastructure).
 *
 * Follows Hadamard._forwardC (Hadamard.pyx:164-230) and its typed butterfly
 * loop _hadamardCore (Hadamard.pyx:36-61): for every column, copy the input
 * into the output, then run `order` in-place radix-2 stages with butterfly
 * distance d = 1, 2, 4, ...; each butterfly stores (a + b, a - b).
 *
 * Integer types wrap modulo 2^bits exactly as the reference's C loop does
 * (sums are formed in unsigned arithmetic, which is the two's-complement
 * wrap that gcc produces for the reference's int/long arithmetic).  Float
 * types use plain IEEE adds in the same order, so float32/float64 outputs are
 * the reference's bit for bit.  Complex types are transformed as (re, im)
 * pairs with the same butterflies.
 *
 * Layout: column-major, column j of x at x + j*ldx (elements), length n = 2^order.
 */
#include <stdint.h>
#include <string.h>

#define DEF_INT(NAME, T, UT)                                                   \
    void NAME(const T *x, T *y, int order, int64_t k, int64_t ldx,             \
              int64_t ldy) {                                                   \
        int64_t n = (int64_t)1 << order;                                       \
        for (int64_t col = 0; col < k; ++col) {                                \
            T *v = y + col * ldy;                                              \
            memmove(v, x + col * ldx, (size_t)n * sizeof(T));                  \
            for (int64_t d = 1; d < n; d <<= 1)                                \
                for (int64_t g = 0; g < n; g += 2 * d)                         \
                    for (int64_t e = g; e < g + d; ++e) {                      \
                        UT a = (UT)v[e], b = (UT)v[e + d];                     \
                        v[e] = (T)(UT)(a + b);                                 \
                        v[e + d] = (T)(UT)(a - b);                             \
                    }                                                          \
        }                                                                      \
    }

DEF_INT(fwht_oracle_i8, int8_t, uint8_t)
DEF_INT(fwht_oracle_i16, int16_t, uint16_t)
DEF_INT(fwht_oracle_i32, int32_t, uint32_t)
DEF_INT(fwht_oracle_i64, int64_t, uint64_t)

/* `pair` = 1 for real, 2 for complex (re, im interleaved). */
#define DEF_FLT(NAME, T)                                                       \
    void NAME(const T *x, T *y, int order, int64_t k, int64_t ldx,             \
              int64_t ldy, int pair) {                                         \
        int64_t n = (int64_t)1 << order;                                       \
        for (int64_t col = 0; col < k; ++col) {                                \
            T *v = y + col * ldy * pair;                                       \
            memmove(v, x + col * ldx * pair, (size_t)(n * pair) * sizeof(T));  \
            for (int64_t d = 1; d < n; d <<= 1)                                \
                for (int64_t g = 0; g < n; g += 2 * d)                         \
                    for (int64_t e = g; e < g + d; ++e)                        \
                        for (int p = 0; p < pair; ++p) {                       \
                            T a = v[e * pair + p], b = v[(e + d) * pair + p];  \
                            v[e * pair + p] = a + b;                           \
                            v[(e + d) * pair + p] = a - b;                     \
                        }                                                      \
        }                                                                      \
    }

DEF_FLT(fwht_oracle_f32, float)
DEF_FLT(fwht_oracle_f64, double)
