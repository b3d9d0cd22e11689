/* FP64 instantiation of the oracle (test infrastructure only). */
#define R double
#define FN(n) n##_f64
#include "oracle_impl.h"

/* Thread count of the OpenMP loops (the CPU baseline is timed with 1 thread and with
 * every host thread). */
#include <omp.h>
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_get_threads(void) { return omp_get_max_threads(); }
