/* FP32 instantiation of the oracle (test infrastructure only): same algorithm in
 * float, used to separate algorithm differences from precision drift. */
#define R float
#define FN(n) n##_f32
#include "oracle_impl.h"
