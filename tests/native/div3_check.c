/* x / 3 through the reciprocal: q0 = x * fl(1/3), q = fma(fma(-3, q0, x), fl(1/3), q0),
 * compared with the IEEE division x / 3.0f for every float x in [+0, 2] (subnormals
 * included).  The shadow filter's div3_unit computes exactly this sequence.
 * Prints the number of mismatches and exits non-zero if there is one. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(void)
{
    const float y = 1.0f / 3.0f;
    long bad = 0;
    for (uint32_t u = 0; u <= 0x40000000u; u++) {
        float x, q0, q, e;
        uint32_t a, b;
        memcpy(&x, &u, 4);
        q0 = x * y;
        q = fmaf(fmaf(-3.0f, q0, x), y, q0);
        e = x / 3.0f;
        memcpy(&a, &q, 4);
        memcpy(&b, &e, 4);
        if (a != b) {
            if (bad < 5) printf("x=%a fast=%a ieee=%a\n", x, q, e);
            bad++;
        }
    }
    printf("%ld mismatches\n", bad);
    return bad != 0;
}
