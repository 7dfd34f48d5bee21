// Stage kernels with no optional source terms (FEAT = 0).
#define RKG_FEAT 0
#include "stage_feat.inc"
