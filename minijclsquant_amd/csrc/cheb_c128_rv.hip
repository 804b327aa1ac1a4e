// complex128 real-value twins of the tuned step kernels (the handle's values are
// all real); the dispatch source is shared with cheb_c128.hip.
#define JKPM_STEP_RV_UNIT 1
#include "cheb_c128.hip"
