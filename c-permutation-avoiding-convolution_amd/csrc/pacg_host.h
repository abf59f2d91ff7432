/* Host-side helpers shared by the plan layer and the drop-in entry points. */
#ifndef PACG_HOST_H
#define PACG_HOST_H

#include "pacg.h"

/* mixed-radix digit reversal: sum_q d_q * n / k_q (pure radix: reverse_index_r*) */
long long pacg_rev_index_mixed(long long p, long long n, const int* radix, int t);

/* error plumbing (thread-local message, see pacg_last_error) */
int pacg_fail(int code, const char* fmt, ...);
void pacg_clear_error(void);

#endif
