/* CPU oracle for the learned disparity-proposal network -- TEST INFRASTRUCTURE ONLY. */
#include "cnn_oracle.h"
int orc_cnn_version(void) { return 0; }
