/*
 * ref_mt_driver.c -- calls the reference's own MT19937 (src/mt19937ar-cok.c,
 * compiled unmodified from /root/reference by oracle/Makefile into
 * oracle/_ref/libref_mt.so).  Test infrastructure only: used to pin
 * or_mt_* in sn_oracle.c against the reference source itself.
 */
#include <stdint.h>

void init_genrand(unsigned long s);        /* src/mt19937ar-cok.c:63 */
unsigned long genrand_int32(void);         /* src/mt19937ar-cok.c:130 */
double genrand_real1(void);                /* src/mt19937ar-cok.c:164 */
double genrand_real2(void);                /* src/mt19937ar-cok.c:182 */

/* kind 0: genrand_int32, 1: genrand_real1, 2: genrand_real2, 3: cycle 0,1,2 */
void ref_mt_stream(uint32_t seed, int64_t n, int kind, double *out)
{
    init_genrand((unsigned long)seed);
    for (int64_t i = 0; i < n; i++) {
        int k = kind == 3 ? (int)(i % 3) : kind;
        if (k == 0) out[i] = (double)genrand_int32();
        else if (k == 1) out[i] = genrand_real1();
        else out[i] = genrand_real2();
    }
}
