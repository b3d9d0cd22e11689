/* FP64 instantiation of the oracle (test infrastructure only). */
#define R double
#define FN(n) n##_f64
#include "oracle_impl.h"
