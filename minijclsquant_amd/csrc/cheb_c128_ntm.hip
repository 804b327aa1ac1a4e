// complex128 step kernels of the tuned shapes with non-temporal matrix loads
// (launch variants 24..47); the dispatch source is shared with cheb_c128.hip.
#define JKPM_STEP_NTM_UNIT 1
#include "cheb_c128.hip"
