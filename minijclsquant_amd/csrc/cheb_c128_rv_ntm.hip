// complex128 real-value twins of the tuned step kernels with non-temporal
// matrix loads (variants 24..47); the dispatch source is shared with cheb_c128.hip.
#define JKPM_STEP_NTM_UNIT 1
#define JKPM_STEP_RV_UNIT 1
#include "cheb_c128.hip"
