/* Exhaustive check of the device's division by a constant (div_const in
 * pcr_amd/csrc/pcr_device.h): for every float x with 2^-100 <= |x| <= 2^100,
 *     q0 = x * rc,  r = fma(-c, q0, x),  q1 = fma(r, rc, q0)
 * with rc = RN(1/c) must equal the correctly rounded x / c bit for bit.
 * Usage: constdiv_check C  (C as a decimal or hex float).  Prints the count of
 * mismatches and the first one; exit status 0 when there are none. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float C, RC;
enum { kThreads = 16 };
typedef struct {
    uint64_t lo, hi, bad;
    uint32_t first;
} Job;

static float fbits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static uint32_t bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static void *run(void *p) {
    Job *j = p;
    for (uint64_t i = j->lo; i < j->hi; i++) {
        const float x = fbits((uint32_t)i), ax = fabsf(x);
        if (!(ax >= 0x1p-100f && ax <= 0x1p100f)) continue; /* the division itself */
        volatile float want = x / C;
        const float q0 = x * RC;
        const float r = fmaf(-C, q0, x);
        const float q1 = fmaf(r, RC, q0);
        if (bits(want) != bits(q1)) {
            if (!j->bad) j->first = (uint32_t)i;
            j->bad++;
        }
    }
    return NULL;
}
int main(int argc, char **argv) {
    if (argc != 2) return 2;
    C = strtof(argv[1], NULL);
    volatile float one = 1.0f;
    RC = one / C;
    pthread_t t[kThreads];
    Job J[kThreads];
    for (int k = 0; k < kThreads; k++) {
        J[k] = (Job){(1ull << 32) * k / kThreads, (1ull << 32) * (k + 1) / kThreads, 0, 0};
        pthread_create(&t[k], NULL, run, &J[k]);
    }
    uint64_t bad = 0;
    uint32_t first = 0;
    for (int k = 0; k < kThreads; k++) {
        pthread_join(t[k], NULL);
        if (J[k].bad && !bad) first = J[k].first;
        bad += J[k].bad;
    }
    printf("c=%a rc=%a mismatches=%llu", C, RC, (unsigned long long)bad);
    if (bad) printf(" first x=%a", fbits(first));
    printf("\n");
    return bad ? 1 : 0;
}
